"""White-Bear Mark II on the HIP path against the FP64 oracle, in both of the reference's
forms: WhiteBearMarkII (src/haskell/WhiteBear.hs:140-159) and HardSpheresWBm2slow
(src/HardSpheres.cpp:161-185, Phi3 over 36 pi).  Same tolerances as the Mark I tests."""

import math

import numpy as np
import pytest

from oracle import fmt_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [((16, 16, 16), np.diag([3.2, 3.2, 3.2])),
          ((32, 48, 20), np.diag([3.1, 4.9, 2.3])),
          ((20, 16, 32), np.array([[2.0, 0.0, 0.1], [0.2, 1.9, 0.0], [0.0, 0.3, 3.4]])),
          ((64, 64, 64), np.diag([6.4, 6.4, 6.4]))]


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


FORMS = [("wb2", 1), ("wb2slow", 2)]


def _fluid(dm, shape, cell, eta=0.3, vext=None, prec=1, seed=None, amp=0.3, form=FORMS[0]):
    g = O.Grid(shape, cell)
    of = O.make_fluid(g, eta, vext=vext, functional=form[0])
    if seed is not None:
        of.x += amp * np.random.default_rng(seed).standard_normal(g.shape)
    f = dm.WhiteBearFluidVeff(shape, cell, R=of.R, kT=of.kT, mu=of.mu, fft_precision=prec,
                              functional=form[1])
    f.Vext = of.vext
    f.Veff = of.x
    return g, of, f


def _wall_vext(g):
    vext = O.planar_wall_vext(g, 1.0)
    return np.where(np.isfinite(vext), 0.5 * np.sin(g.positions()[..., 1]), vext).astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("prec,etol,gtol", [(1, 1e-12, 1e-10), (0, 2e-7, 2e-5)])
@pytest.mark.parametrize("shape,cell", SHAPES)
def test_mark_ii_energy_and_gradient(gpu, shape, cell, prec, etol, gtol, form):
    print("mark II energy/grad", form, shape, prec, flush=True)
    g = O.Grid(shape, cell)
    g, of, f = _fluid(gpu, shape, cell, vext=_wall_vext(g), prec=prec, seed=11, form=form)
    assert f.functional == form[1]
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


@pytest.mark.parametrize("form", FORMS)
def test_mark_ii_low_density_series_branch(gpu, form):
    """A dilute fluid puts every n3 below the series cut-off (|n3| < 0.05)."""
    shape, cell = (32, 32, 32), np.diag([6.4, 6.4, 6.4])
    g = O.Grid(shape, cell)
    g, of, f = _fluid(gpu, shape, cell, eta=0.01, vext=_wall_vext(g), prec=1, seed=5, amp=0.5, form=form)
    n3 = of.weighted_densities(of.density())[0]
    assert np.max(np.abs(n3)) < 0.05
    E, gr, _ = f.energy_grad_and_precond()
    Eo, gro, _ = of.energy_grad_and_precond()
    assert abs(E - Eo) <= 1e-12 * abs(Eo)
    assert _rel(gr, gro) < 1e-10


def test_mark_ii_uniform_fluid_is_bulk(gpu):
    eta = 0.4
    nb = O.bulk_density(eta, 1.0)
    mu = gpu.find_chemical_potential(nb)
    p = gpu.pressure(nb)
    assert abs(p - O.pressure(nb, 1.0, functional="wb2")) < 1e-12 * p
    f = gpu.WhiteBearFluidVeff((32, 32, 32), np.diag([3.2] * 3), mu=mu, fft_precision=1,
                               functional=gpu.WHITE_BEAR_MARK_II)
    f.Veff = -math.log(nb)
    E, gr, pg = f.energy_grad_and_precond()
    V = 3.2 ** 3
    assert abs(E + p * V) < 1e-12 * p * V
    assert np.max(np.abs(pg)) < 1e-12


def test_switching_functional_on_a_plan(gpu):
    shape, cell = (32, 32, 32), np.diag([3.2] * 3)
    g = O.Grid(shape, cell)
    g, of, f = _fluid(gpu, shape, cell, vext=O.planar_wall_vext(g, 1.0), seed=2)
    e2 = f.energy()
    f.functional = gpu.WHITE_BEAR
    e1 = f.energy()
    o1 = O.make_fluid(g, 0.3, vext=of.vext)
    o1.x[...] = of.x
    assert abs(e1 - o1.energy()) <= 1e-12 * abs(e1)
    f.functional = gpu.WHITE_BEAR_MARK_II
    assert f.energy() == e2
    with pytest.raises(gpu.DeftError):
        f.functional = 7
    with pytest.raises(gpu.DeftError):
        gpu.Batch([f])  # batches take White-Bear plans only


def test_mark_ii_cg_trajectory_fp64(gpu):
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
