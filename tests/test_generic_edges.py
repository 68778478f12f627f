"""Grid edges without fused per-edge kernels: the generic line passes
(deft_amd/csrc/generic.hpp) against the FP64 oracle.

The reference takes any N (FFTW, src/GridDescription.cpp:26-30): its old-API drivers
build 65^3 grids (6.4/0.1), odd 201^3 ones and 1 x 1 x 3601 planar-wall lines
(papers/contact/figs/walls.cpp:91-93, 36/0.01 + 1 = 3601 = 13 * 277).  Tolerances as
the fused passes' (tests/test_gpu_parity.py): FP64 FFTs to round-off (1e-12 on
energies, 1e-10 of the field scale), FP32 to FP32 round-off of the transforms.
"""

import math

import numpy as np
import pytest

from oracle import fmt_oracle as O

pytestmark = pytest.mark.gpu

GEN_SHAPES = [((65, 65, 65), np.diag([6.4, 6.4, 6.4])),
              ((18, 25, 30), np.diag([2.1, 2.6, 3.3])),
              ((7, 22, 13), np.array([[1.9, 0.0, 0.2], [0.1, 2.5, 0.0], [0.0, 0.3, 1.6]])),
              ((12, 1, 45), np.diag([2.4, 0.01, 4.5])),
              ((1, 1, 3601), np.diag([0.001, 0.001, 36.0]))]
# 3601 = 13 * 277 (direct-DFT stages), 65 = 5 * 13, 22 = 2 * 11, 7 and 13 prime


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _fluid(dm, shape, cell, eta=0.3, vext=None, prec=1, seed=None, amp=0.3):
    g = O.Grid(shape, cell)
    of = O.make_fluid(g, eta, vext=vext)
    if seed is not None:
        of.x = of.x + amp * np.random.default_rng(seed).standard_normal(g.shape)
    f = dm.WhiteBearFluidVeff(shape, cell, R=of.R, kT=of.kT, mu=of.mu, fft_precision=prec)
    f.Vext = of.vext
    f.Veff = of.x
    return g, of, f


def _wall_vext(g, width):
    """Hard wall |z| < width on the periodic cell (nearest image), sin bumps elsewhere."""
    z = g.positions()[..., 2]
    Lz = float(np.linalg.norm(g.cell[2]))
    z = z - Lz * np.round(z / Lz)
    v = 0.3 * np.sin(2 * math.pi * g.positions()[..., 2] / Lz * 3)
    return np.where(np.abs(z) < width, np.inf, v)


def test_edges_are_supported_and_generic(gpu):
    for n in (1, 7, 13, 22, 25, 65, 191, 277, 3300, 3601, 4096):
        assert gpu.supported_length(n), n
        assert not gpu.fused_length(n), n
    for n in (16, 20, 21, 32, 48, 64, 96, 128, 201, 256):
        assert gpu.supported_length(n) and gpu.fused_length(n), n
    for n in (0, 4097, 2 * 521, 1031):  # beyond 4096, prime factors above 512
        assert not gpu.supported_length(n), n


@pytest.mark.parametrize("prec,tol", [(1, 1e-13), (0, 3e-6)])
@pytest.mark.parametrize("shape,cell", GEN_SHAPES)
def test_fft_seam_matches_numpy(gpu, shape, cell, prec, tol):
    print("generic fft", shape, prec, flush=True)
    g = O.Grid(shape, cell)
    f = gpu.WhiteBearFluidVeff(shape, cell, fft_precision=prec)
    a = np.random.default_rng(7).standard_normal(shape)
    F = f.fft(a)
    assert _rel(F, O.fft(g, a)) < tol
    assert _rel(f.ifft(F), a) < 2 * tol
    f.close()


@pytest.mark.parametrize("prec,etol,gtol", [(1, 1e-12, 1e-10), (0, 2e-7, 2e-5)])
@pytest.mark.parametrize("shape,cell", GEN_SHAPES)
def test_energy_and_gradient(gpu, shape, cell, prec, etol, gtol):
    print("generic energy/grad", shape, prec, flush=True)
    g = O.Grid(shape, cell)
    vext = _wall_vext(g, 0.6) if shape[2] >= 30 else None
    if vext is not None and prec == 0:
        vext = vext.astype(np.float32).astype(np.float64)
    g, of, f = _fluid(gpu, shape, cell, vext=vext, prec=prec, seed=11)
    got = f.weighted_densities()
    ref = of.weighted_densities(of.density())
    # FP32: the direct-DFT stages (radix 7, 11, 13, 277) sum up to p terms in FP32;
    # the vector densities, which cancel to ~1e-4 of n3, see 5e-6 of their maximum
    for a in range(5):
        assert _rel(got[a], ref[a]) < (1e-11 if prec == 1 else 1e-5), a
    E = f.energy()
    Eo = of.energy()
    # relative to the terms' magnitudes: at eta 0.3 on the wall line they cancel to 1/40
    scale = sum(abs(v) for v in of.last_terms.values())
    assert abs(E - Eo) <= etol * scale
    for k, v in of.last_terms.items():
        assert abs(f.last_terms[k] - v) <= etol * scale, k
    E2, gr, pg = f.energy_grad_and_precond()
    Eo2, gro, pgo = of.energy_grad_and_precond()
    assert abs(E2 - Eo2) <= etol * scale
    assert _rel(gr, gro) < gtol
    assert _rel(pg, pgo) < gtol
    if vext is not None:
        wall = ~np.isfinite(vext)
        assert np.all(gr[wall] == 0) and np.all(pg[wall] == 0)
    f.close()


@pytest.mark.parametrize("functional", ["mark2", "tensor"])
def test_other_functionals(gpu, functional):
    shape, cell = GEN_SHAPES[1]
    print("generic", functional, flush=True)
    g = O.Grid(shape, cell)
    of = O.make_fluid(g, 0.35, functional="wb2" if functional == "mark2" else "wbt")
    of.x = of.x + 0.2 * np.random.default_rng(5).standard_normal(g.shape)
    f = gpu.WhiteBearFluidVeff(shape, cell, R=of.R, kT=of.kT, mu=of.mu, fft_precision=1,
                               functional=gpu.WHITE_BEAR_MARK_II if functional == "mark2" else gpu.WHITE_BEAR_TENSOR)
    f.Veff = of.x
    E, gr, pg = f.energy_grad_and_precond()
    Eo, gro, pgo = of.energy_grad_and_precond()
    assert abs(E - Eo) <= 1e-12 * abs(Eo)
    assert _rel(gr, gro) < 1e-10
    f.close()


def test_every_radix_schedule_gives_the_same_answer(gpu):
    """65 = 5 * 13, 30 = 2 * 3 * 5, 3601 = 13 * 277 in every candidate order: FP64
    results agree to round-off; an invalid order is refused; autotune keeps one of them."""
    for shape, cell in (GEN_SHAPES[0], GEN_SHAPES[1], GEN_SHAPES[4]):
        g, of, f = _fluid(gpu, shape, cell, seed=3)
        E0, g0, _ = f.energy_grad_and_precond()
        for axis in range(3):
            L = shape[axis]
            if L == 1:
                continue
            cands = [f.radix_schedule(1, axis)]
            fac = O_factors(L)
            cands.append(sorted(fac, reverse=True))
            cands.append(sorted(fac))
            for c in cands:
                f.set_radix_schedule(1, axis, c)
                assert f.radix_schedule(1, axis) == c
                E, gr, _ = f.energy_grad_and_precond()
                assert abs(E - E0) <= 1e-13 * abs(E0), (shape, axis, c)
                assert _rel(gr, g0) < 1e-11, (shape, axis, c)
            with pytest.raises(gpu.DeftError):
                f.set_radix_schedule(1, axis, [L + 1] if L > 1 else [2])
        f.autotune()
        for axis in range(3):
            s = f.radix_schedule(1, axis)
            assert int(np.prod(s)) == shape[axis] if s else shape[axis] == 1
        E, gr, _ = f.energy_grad_and_precond()
        assert abs(E - E0) <= 1e-13 * abs(E0)
        f.close()


def O_factors(n):
    out, p = [], 2
    while p * p <= n:
        while n % p == 0:
            out.append(p)
            n //= p
        p += 1
    if n > 1:
        out.append(n)
    return out


def test_fused_plans_refuse_a_radix_change(gpu):
    f = gpu.WhiteBearFluidVeff((32, 32, 32), np.diag([3.2] * 3), fft_precision=1)
    assert f.radix_schedule(1, 2) == [8, 4]
    with pytest.raises(gpu.DeftError):
        f.set_radix_schedule(1, 2, [4, 8])
    f.close()


@pytest.mark.parametrize("eta,functional", [(0.3, "wb"), (0.4, "wb2"), (0.2, "wbt")])
def test_wall_line_minimum_matches_oracle(gpu, eta, functional):
    """The 1 x 1 x 3601 planar-wall line of papers/contact/figs/walls.cpp (constrained
    slab |z| <= 3, Veff from -ln(eta constraint/(4 pi/3) + 1e-4 n_b)) minimized to
    relative 1e-13 with the new-API Minimize at FP64 against the oracle's minimum of the
    same problem: energy to 1e-11, density to 1e-6 of its maximum, and the contact
    density at the slab's edge against the bulk pressure."""
    print("wall line", eta, functional, flush=True)
    g = O.Grid((1, 1, 3601), np.diag([0.001, 0.001, 36.0]))
    z = g.positions()[..., 2]
    z = z - 36.0 * np.floor(z / 36.0 + 0.5)
    constraint = (np.abs(z) > 3).astype(float)
    nb = eta / (4 * math.pi / 3)
    mu = O.chemical_potential(nb, 1.0, 1.0, functional)
    vext = np.where(constraint == 0, np.nan, 0.0)
    x0 = -np.log((eta * constraint + 1e-4 * eta) / (4 * math.pi / 3))
    of = O.WhiteBearFluidVeff(g, 1.0, 1.0, mu, vext, x0.copy(), functional)
    om = O.Minimize(of)
    om.precondition(True)
    om.set_relative_precision(1e-13)
    while om.improve_energy():
        pass
    kind = {"wb": gpu.WHITE_BEAR, "wb2": gpu.WHITE_BEAR_MARK_II, "wbt": gpu.WHITE_BEAR_TENSOR}[functional]
    f = gpu.WhiteBearFluidVeff((1, 1, 3601), np.diag([0.001, 0.001, 36.0]), mu=mu, fft_precision=1, functional=kind)
    f.Vext = vext
    f.Veff = x0
    m = gpu.Minimize(f)
    m.precondition(True)
    m.set_relative_precision(1e-13)
    m.set_maxiter(5000)
    m.run()
    s = m.stats()
    print("wall line", functional, s["iter"], s["stop_reason"], m.energy(), om.energy(), flush=True)
    assert s["stop_reason"] == "converged"
    assert abs(m.energy() / om.energy() - 1) < 1e-11
    # the density is first order in the remaining error, the energy second order
    n, no = f.density(), of.density()
    assert np.max(np.abs(n - no)) < 1e-6 * no.max()
    # hard-wall contact theorem n(contact) = p/kT, to the grid's first order (dz = 0.01)
    k = int(np.argmax(z[0, 0] > 3.0))
    p = O.pressure(nb, 1.0, functional=functional)
    print("contact", z[0, 0, k], n[0, 0, k], "p/kT", p, flush=True)
    assert abs(n[0, 0, k] / p - 1) < 0.03
    m.close()
    f.close()


@pytest.mark.parametrize("prec,etol,gtol", [(1, 1e-12, 1e-10), (0, 2e-7, 2e-5)])
def test_generic_engine_on_fused_edges_meets_oracle(gpu, prec, etol, gtol):
    """The engine choice of the autotuner: on edges that have fused kernels the generic
    line passes (forced by deft_plan_set_engine) give the oracle's energy and gradient
    as the fused passes do."""
    shape, cell = (32, 48, 20), np.diag([3.1, 4.9, 2.3])
    g, of, f = _fluid(gpu, shape, cell, vext=_wall_vext(O.Grid(shape, cell), 0.5), prec=prec, seed=3)
    E_f, g_f, _ = f.energy_grad_and_precond()
    f.set_engine(True)
    assert f.engine()[0] == "generic"
    E_g, g_g, _ = f.energy_grad_and_precond()
    Eo, go, _ = of.energy_grad_and_precond()
    for E, gr in ((E_f, g_f), (E_g, g_g)):
        assert abs(E - Eo) <= etol * abs(Eo)
        assert np.max(np.abs(gr - go)) <= gtol * np.max(np.abs(go))
    f.set_engine(False)
    assert f.engine()[0] == "fused"
    f.close()


def test_autotune_keeps_the_faster_engine_and_the_minimizer_state(gpu):
    """deft_plan_autotune times one FP32 and one FP64 energy+gradient evaluation with each
    engine and keeps the faster; run in the middle of a minimization it leaves Veff and
    the gradient buffers as they were, so the iterates that follow are the same as
    without it."""
    shape, cell = (64, 64, 64), np.diag([6.4, 6.4, 6.4])
    g = O.Grid(shape, cell)
    runs = []
    for tune in (False, True):
        _, of, f = _fluid(gpu, shape, cell, vext=O.planar_wall_vext(g, 1.0), prec=1, seed=2, amp=0.2)
        m = gpu.Minimize(f)
        m.precondition(True)
        for _ in range(3):
            m.improve_energy()
        if tune:
            eng, tf, tg = f.engine()
            assert tf == -1.0 and tg == -1.0
            f.autotune()
            eng, tf, tg = f.engine()
            print("engines: fused", tf, "generic", tg, "->", eng, flush=True)
            assert tf > 0 and tg > 0
            assert eng == ("generic" if tg < tf else "fused")
        Es = []
        for _ in range(4):
            m.improve_energy()
            Es.append(m.energy())
        runs.append(Es)
        m.close()
        f.close()
    assert np.allclose(runs[0], runs[1], rtol=1e-11, atol=0)
