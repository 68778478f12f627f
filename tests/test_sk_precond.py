"""The oracle's structure-factor preconditioner (oracle/fmt_oracle.py, no reference
counterpart) pinned by what the mathematics fixes, on the CPU:

* with S(k) = 1 it is the new API's grad/n (src/haskell/NewCode.hs:229-242);
* on the uniform fluid it inverts the Hessian: (kT/dV) M [g(V_b + e d) - g(V_b - e d)]/(2e)
  returns d (the Newton step is exact for the quadratic part), so a dropped term, a wrong
  sign or factor in H_ex(k) fails;
* it is symmetric and positive definite on a non-uniform fluid with walls and
  constrained points, and zero on both;
* minimizations preconditioned with it reach the same minimum as with grad/n.
"""

import math

import numpy as np
import pytest

from oracle import fmt_oracle as O


def _uniform(shape, L, eta, functional="wb"):
    g = O.Grid(shape, np.diag(L))
    f = O.make_fluid(g, eta, functional=functional)
    return g, f, O.bulk_density(eta, 1.0)


def test_unit_structure_factor_is_the_new_api_preconditioner():
    g = O.Grid((16, 18, 20), np.diag([2.0, 2.2, 2.4]))
    vext = O.planar_wall_vext(g, 0.5)
    f = O.make_fluid(g, 0.35, vext=vext)
    f.x = f.x + 0.3 * np.random.default_rng(2).standard_normal(g.shape)
    E, gr, pg = f.energy_grad_and_precond()
    nb = O.bulk_density(0.35, 1.0)
    ones = np.ones_like(O.structure_factor(g, nb, 1.0))
    got = O.sk_precondition(f, gr, ones, nb)
    assert np.max(np.abs(got - pg)) <= 1e-12 * np.max(np.abs(pg))


@pytest.mark.parametrize("eta,functional", [(0.2, "wb"), (0.4, "wb"), (0.45, "wb2")])
def test_uniform_fluid_newton_step_is_exact(eta, functional):
    g, f, nb = _uniform((24, 20, 18), [2.4, 2.0, 1.8], eta, functional)
    xb = f.x.copy()
    rng = np.random.default_rng(5)
    d = rng.standard_normal(g.shape)
    e = 1e-4
    gp = f.grad(xb + e * d)
    gm = f.grad(xb - e * d)
    hd = (gp - gm) / (2 * e)
    S = O.structure_factor(g, nb, 1.0, 1.0, functional)
    step = O.sk_precondition(f, hd, S, nb, xb) * (f.kT / g.dV)
    assert np.max(np.abs(step - d)) < 1e-6 * np.max(np.abs(d))
    # without S (the new API's grad/n) the step is far off at these packings
    plain = O.sk_precondition(f, hd, np.ones_like(S), nb, xb) * (f.kT / g.dV)
    assert np.max(np.abs(plain - d)) > 0.01 * np.max(np.abs(d))


def test_symmetric_positive_definite_with_walls_and_constraints():
    g = O.Grid((20, 16, 24), np.diag([2.0, 1.6, 2.4]))
    vext = O.planar_wall_vext(g, 0.6)
    vext = np.where(np.isfinite(vext), 0.5 * np.sin(g.positions()[..., 1]), vext)
    vext[3] = np.nan
    f = O.make_fluid(g, 0.4, vext=vext)
    f.x = f.x + 0.5 * np.random.default_rng(8).standard_normal(g.shape)
    nb = O.bulk_density(0.4, 1.0)
    S = O.structure_factor(g, nb, 1.0)
    rng = np.random.default_rng(9)
    u, v = rng.standard_normal(g.shape), rng.standard_normal(g.shape)
    Mu, Mv = O.sk_precondition(f, u, S, nb), O.sk_precondition(f, v, S, nb)
    assert abs(np.vdot(u, Mv) - np.vdot(Mu, v)) <= 1e-12 * abs(np.vdot(u, Mv))
    assert np.vdot(u, Mu) > 0 and np.vdot(v, Mv) > 0
    assert np.all(Mu[~np.isfinite(vext)] == 0)
    # S(k) of the hard-sphere fluid: 1 at large k, the compressibility value at k = 0
    assert abs(S.flat[-1] - 1) < 1e-3
    p_dn = O.chemical_potential(nb * (1 + 1e-6), 1.0) - O.chemical_potential(nb * (1 - 1e-6), 1.0)
    s0 = (1.0 / nb) / (p_dn / (2e-6 * nb))  # kT / (n dmu/dn)
    assert abs(S.flat[0] / s0 - 1) < 1e-5


def test_preconditioned_minimization_reaches_the_same_minimum():
    g = O.Grid.cubic(32, 3.2)
    vext = O.spherical_solute_vext(g, 1.0)
    results = []
    for sk in (False, True):
        f = O.make_fluid(g, 0.4, vext=vext)
        ff = O.StructureFactorFluid(f, O.bulk_density(0.4, 1.0)) if sk else f
        m = O.Minimize(ff)
        m.precondition(True)
        m.set_relative_precision(1e-13)
        while m.improve_energy():
            pass
        results.append((m.iter, m.energy()))
    (it0, E0), (it1, E1) = results
    assert abs(E1 / E0 - 1) < 1e-11
    assert it1 < it0


# --- the product's DEFT_PGRAD_STRUCTURE_FACTOR on the GPU against this oracle ---

SK_SHAPES = [((32, 48, 20), np.diag([3.1, 4.9, 2.3])),
             ((20, 16, 32), np.array([[2.0, 0.0, 0.1], [0.2, 1.9, 0.0], [0.0, 0.3, 3.4]])),
             ((18, 25, 30), np.diag([2.1, 2.6, 3.3]))]  # the last: generic edges


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [(1, 1e-9), (0, 3e-5)])
@pytest.mark.parametrize("shape,cell", SK_SHAPES)
def test_gpu_structure_factor_pgrad_matches_oracle(gpu, shape, cell, prec, tol):
    print("sk pgrad", shape, prec, flush=True)
    g = O.Grid(shape, cell)
    vext = O.planar_wall_vext(g, 0.6)
    vext = np.where(np.isfinite(vext), 0.4 * np.sin(g.positions()[..., 1]), vext)
    vext[shape[0] // 2] = np.nan
    if prec == 0:
        vext = vext.astype(np.float32).astype(np.float64)
    eta = 0.4
    nb = O.bulk_density(eta, 1.0)
    of = O.make_fluid(g, eta, vext=vext)
    of.x = of.x + 0.3 * np.random.default_rng(4).standard_normal(g.shape)
    f = gpu.WhiteBearFluidVeff(shape, cell, R=of.R, kT=of.kT, mu=of.mu, fft_precision=prec)
    f.Vext = vext
    f.Veff = of.x
    assert abs(f.bulk_density() / nb - 1) < 1e-12
    E, gr, pg = f.sk_precond()
    Eo, gro, _ = of.energy_grad_and_precond()
    pgo = O.sk_precondition(of, gro, O.structure_factor(g, nb, 1.0), nb)
    assert np.max(np.abs(pg - pgo)) < tol * np.max(np.abs(pgo))
    assert np.all(pg[~np.isfinite(vext)] == 0)
    f.close()


def _c2_fluid(dm):
    import json
    import os
    import deft_amd.workloads as W
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_c2.json")))
    nb = 0.40 / (4 * math.pi / 3)
    f = dm.WhiteBearFluidVeff((128, 128, 128), np.diag([12.8] * 3), mu=dm.find_chemical_potential(nb),
                              fft_precision=dm.FFT_FP32)
    f.Vext = W.c2_solute_vext(128, 12.8, 3.0)
    return f, gold, nb


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["new", "sk"])
def test_gpu_headline_ladder_with_each_preconditioner(gpu, kind):
    """The 128^3 solute at eta 0.40 through the FP32 -> FP64 ladder and the |p V| 1e-10
    rule with either preconditioner: both meet the FP64 oracle's golden minimum to
    |dF/F| < 1e-9; the structure-factor one needs at most half the iterations."""
    f, gold, nb = _c2_fluid(gpu)
    f.fill_Veff(-math.log(nb))
    m = gpu.Minimize(f)
    m.precondition(True)
    if kind == "sk":
        m.set_preconditioner(gpu.PGRAD_STRUCTURE_FACTOR)
    m.set_precision(abs(gpu.pressure(nb) * 12.8 ** 3) * 1e-10)
    m.set_mixed_precision(1e-5, gpu.FFT_FP64)
    m.set_maxiter(1000)
    m.run()
    s = m.stats()
    rel = abs(m.energy() - gold["energy"]) / abs(gold["energy"])
    n = f.density()
    dn = float(np.max(np.abs(n[:, 0, 0] - np.array(gold["density_x_axis"]))) / gold["density_max"])
    print("c2", kind, s["iter"], s["stop_reason"], "dF/F", rel, "dn", dn, flush=True)
    assert s["stop_reason"] == "converged"
    assert rel < 1e-9 and dn < 3e-5
    if kind == "sk":
        assert s["iter"] <= 20
    m.close()
    f.close()
