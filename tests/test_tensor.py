"""Tensor White-Bear (TensorWhiteBear, src/haskell/WhiteBear.hs:44-51, src/HardSpheres.cpp:30-67)
on the HIP path against the FP64 oracle: eleven weighted fields, the same tolerances as
the Mark I parity tests."""

import math

import numpy as np
import pytest

from oracle import fmt_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [((16, 16, 16), np.diag([3.2, 3.2, 3.2])),
          ((32, 48, 20), np.diag([3.1, 4.9, 2.3])),
          ((20, 16, 32), np.array([[2.0, 0.0, 0.1], [0.2, 1.9, 0.0], [0.0, 0.3, 3.4]])),
          ((21, 16, 21), np.diag([3.1, 2.4, 3.0])),
          ((64, 64, 64), np.diag([6.4, 6.4, 6.4]))]


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _fluid(dm, shape, cell, eta=0.3, vext=None, prec=1, seed=None, amp=0.3):
    g = O.Grid(shape, cell)
    of = O.make_fluid(g, eta, vext=vext, functional="wbt")
    if seed is not None:
        of.x += amp * np.random.default_rng(seed).standard_normal(g.shape)
    f = dm.WhiteBearFluidVeff(shape, cell, R=of.R, kT=of.kT, mu=of.mu, fft_precision=prec,
                              functional=dm.WHITE_BEAR_TENSOR)
    f.Vext = of.vext
    f.Veff = of.x
    return g, of, f


def _wall_vext(g):
    vext = O.planar_wall_vext(g, 1.0)
    return np.where(np.isfinite(vext), 0.5 * np.sin(g.positions()[..., 1]), vext).astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 3e-6)])
@pytest.mark.parametrize("shape,cell", SHAPES[:4])
def test_tensor_weighted_densities(gpu, shape, cell, prec, tol):
    print("tensor weighted", shape, prec, flush=True)
    g, of, f = _fluid(gpu, shape, cell, prec=prec, seed=3)
    got = f.weighted_densities()
    ref = of.weighted_densities(of.density())
    assert got.shape[0] == 11
    for a in range(11):
        assert _rel(got[a], ref[a]) < tol, a


@pytest.mark.parametrize("prec,etol,gtol", [(1, 1e-12, 1e-10), (0, 2e-7, 2e-5)])
@pytest.mark.parametrize("shape,cell", SHAPES)
def test_tensor_energy_and_gradient(gpu, shape, cell, prec, etol, gtol):
    print("tensor energy/grad", shape, prec, flush=True)
    g = O.Grid(shape, cell)
    g, of, f = _fluid(gpu, shape, cell, vext=_wall_vext(g), prec=prec, seed=11)
    E = f.energy()
    Eo = of.energy()
    assert abs(E - Eo) <= etol * abs(Eo)
    for k, v in of.last_terms.items():
        assert abs(f.last_terms[k] - v) <= etol * abs(Eo) + 1e-300, k
    E2, gr, pg = f.energy_grad_and_precond()
    Eo2, gro, pgo = of.energy_grad_and_precond()
    assert E2 == E
    assert abs(E2 - Eo2) <= etol * abs(Eo2)
    assert _rel(gr, gro) < gtol
    assert _rel(pg, pgo) < gtol


def test_tensor_uniform_fluid_is_bulk(gpu):
    eta = 0.4
    nb = O.bulk_density(eta, 1.0)
    mu = gpu.find_chemical_potential(nb, functional=gpu.WHITE_BEAR_TENSOR)
    p = gpu.pressure(nb, functional=gpu.WHITE_BEAR_TENSOR)
    assert abs(p - O.pressure(nb, 1.0)) < 1e-12 * p
    f = gpu.WhiteBearFluidVeff((32, 32, 32), np.diag([3.2] * 3), mu=mu, fft_precision=1,
                               functional=gpu.WHITE_BEAR_TENSOR)
    f.Veff = -math.log(nb)
    E, gr, pg = f.energy_grad_and_precond()
    V = 3.2 ** 3
    assert abs(E + p * V) < 1e-12 * p * V
    assert np.max(np.abs(pg)) < 1e-12


def test_tensor_cg_trajectory_fp64(gpu):
    shape, cell = (32, 32, 32), np.diag([3.2, 3.2, 3.2])
    g = O.Grid(shape, cell)
    g, of, f = _fluid(gpu, shape, cell, eta=0.3, vext=O.planar_wall_vext(g, 1.0))
    mo = O.Minimize(of)
    mo.set_relative_precision(1e-12)
    mo.precondition(True)
    m = gpu.Minimize(f)
    m.set_relative_precision(1e-12)
    m.precondition(True)
    for it in range(12):
        a = mo.improve_energy()
        b = m.improve_energy()
        assert a == b
        assert abs(m.energy() - mo.energy()) <= 1e-11 * abs(mo.energy()), it
        if not a:
            break
    assert m.stats()["iter"] == mo.iter
    assert _rel(f.Veff, of.x) < 1e-9


def test_tensor_ladder_reaches_fp64_answer_and_switching(gpu):
    shape, cell = (32, 32, 32), np.diag([3.2, 3.2, 3.2])
    g = O.Grid(shape, cell)
    _, of, f = _fluid(gpu, shape, cell, eta=0.35, vext=O.planar_wall_vext(g, 1.0), prec=1)
    m = gpu.Minimize(f)
    m.precondition(True)
    m.set_relative_precision(1e-13)
    m.run()
    E64 = m.energy()
    m.close()
    f.Veff = of.x
    f.fft_precision = 0
    m = gpu.Minimize(f)
    m.precondition(True)
    m.set_relative_precision(1e-11)
    m.set_mixed_precision(1e-5, 1)
    m.run()
    assert m.stats()["stop_reason"] == "converged"
    assert abs(m.energy() / E64 - 1) < 1e-9
    m.close()
    # the plan goes back to the five-field functional and agrees with it
    f.functional = gpu.WHITE_BEAR
    f.fft_precision = 1
    o1 = O.make_fluid(g, 0.35, vext=of.vext)
    o1.x[...] = f.Veff
    assert abs(f.energy() / o1.energy() - 1) < 1e-12
    assert f.weighted_densities().shape[0] == 5


def test_tensor_plans_are_not_batched(gpu):
    shape, cell = (32, 32, 32), np.diag([3.2] * 3)
    _, _, f1 = _fluid(gpu, shape, cell)
    _, _, f2 = _fluid(gpu, shape, cell)
    f2.functional = gpu.WHITE_BEAR
    with pytest.raises(gpu.DeftError):
        gpu.Batch([f1, f2])
    b = gpu.Batch([f2])
    with pytest.raises(gpu.DeftError):
        f2.functional = gpu.WHITE_BEAR_TENSOR
    b.close()
