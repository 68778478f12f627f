"""Parity of the HIP path (through the C ABI) against the FP64 oracle.

Tolerances: FP64 FFT mode must agree with the oracle to round-off
(relative 1e-12 on energies, 1e-11 of the field scale on fields); FP32 FFT mode
to FP32 round-off of the transforms (relative 2e-6 on fields, 1e-7 on energies).
"""

import json
import math
import os

import numpy as np
import pytest

from oracle import fmt_oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _fluid(dm, shape, cell, eta=0.3, vext=None, prec=1, seed=None, amp=0.3):
    g = O.Grid(shape, cell)
    of = O.make_fluid(g, eta, vext=vext)
    if seed is not None:
        rng = np.random.default_rng(seed)
        of.x += amp * rng.standard_normal(g.shape)
    f = dm.WhiteBearFluidVeff(shape, cell, R=of.R, kT=of.kT, mu=of.mu, fft_precision=prec)
    f.Vext = of.vext
    f.Veff = of.x
    return g, of, f


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


SHAPES = [((16, 16, 16), np.diag([3.2, 3.2, 3.2])),
          ((32, 48, 20), np.diag([3.1, 4.9, 2.3])),
          ((20, 16, 32), np.array([[2.0, 0.0, 0.1], [0.2, 1.9, 0.0], [0.0, 0.3, 3.4]])),
          ((64, 64, 64), np.diag([6.4, 6.4, 6.4])),
          ((96, 32, 128), np.diag([9.6, 3.2, 12.8]))]
# odd edges (old-API grids): no Nyquist plane along z, direct-DFT radix-7 stage
ODD_SHAPES = [((21, 16, 20), np.diag([3.1, 2.4, 3.0])),
              ((16, 21, 21), np.array([[2.2, 0.0, 0.0], [0.3, 3.0, 0.0], [0.0, 0.2, 3.1]]))]


@pytest.mark.parametrize("prec,tol", [(1, 1e-13), (0, 3e-6)])
@pytest.mark.parametrize("shape,cell", SHAPES[:4] + ODD_SHAPES)
def test_fft_seam_matches_numpy(gpu, shape, cell, prec, tol):
    print("fft", shape, prec, flush=True)
    g = O.Grid(shape, cell)
    f = gpu.WhiteBearFluidVeff(shape, cell, fft_precision=prec)
    rng = np.random.default_rng(7)
    a = rng.standard_normal(shape)
    F = f.fft(a)
    ref = O.fft(g, a)
    assert _rel(F, ref) < tol
    back = f.ifft(ref)
    assert _rel(back, a) < tol
    # round trip of the reference seam: ifft(fft(f)) = f (tests/new-fftinverse.cpp:21-43)
    assert _rel(f.ifft(F), a) < 2 * tol


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 3e-6)])
@pytest.mark.parametrize("shape,cell", SHAPES + ODD_SHAPES)
def test_weighted_densities(gpu, shape, cell, prec, tol):
    print("weighted", shape, prec, flush=True)
    g, of, f = _fluid(gpu, shape, cell, prec=prec, seed=3)
    got = f.weighted_densities()
    ref = of.weighted_densities(of.density())
    for a in range(5):
        assert _rel(got[a], ref[a]) < tol, a


@pytest.mark.parametrize("prec,etol,gtol", [(1, 1e-12, 1e-10), (0, 2e-7, 2e-5)])
@pytest.mark.parametrize("shape,cell", SHAPES + ODD_SHAPES)
def test_energy_and_gradient(gpu, shape, cell, prec, etol, gtol):
    print("energy/grad", shape, prec, flush=True)
    vext = None
    g = O.Grid(shape, cell)
    if shape[0] >= 32:
        vext = O.planar_wall_vext(g, 1.0)
        vext = np.where(np.isfinite(vext), 0.5 * np.sin(g.positions()[..., 1]), vext).astype(np.float32).astype(np.float64)
    g, of, f = _fluid(gpu, shape, cell, vext=vext, prec=prec, seed=11)
    E = f.energy()
    Eo = of.energy()
    assert abs(E - Eo) <= etol * abs(Eo)
    for k, v in of.last_terms.items():
        assert abs(f.last_terms[k] - v) <= etol * abs(Eo) + 1e-300, k
    E2, gr, pg = f.energy_grad_and_precond()
    Eo2, gro, pgo = of.energy_grad_and_precond()
    assert abs(E2 - Eo2) <= etol * abs(Eo2)
    assert _rel(gr, gro) < gtol
    assert _rel(pg, pgo) < gtol
    if vext is not None:
        wall = ~np.isfinite(vext)
        assert np.all(gr[wall] == 0) and np.all(pg[wall] == 0)


@pytest.mark.parametrize("shape,cell", [SHAPES[1], SHAPES[3]])
def test_fp64_vext_is_not_rounded(gpu, shape, cell):
    """A smooth external potential that FP32 cannot represent (eight Gaussians, as the
    batched workload draws them): the FP64 passes use it unrounded, so energy and
    gradient meet the FP64 oracle on the exact potential at round-off; rounding it to
    FP32 first would move the energy by far more than that."""
    print("fp64 vext", shape, flush=True)
    g = O.Grid(shape, cell)
    rng = np.random.default_rng(21)
    pos = g.positions()
    L = np.diag(np.asarray(cell, dtype=np.float64)) if np.allclose(cell, np.diag(np.diag(cell))) else None
    vext = np.zeros(g.shape)
    for _ in range(8):
        c = rng.uniform(0, 1, 3) * np.diag(cell)
        d = pos - c
        if L is not None:
            d -= L * np.round(d / L)
        vext += rng.uniform(-2, 2) * np.exp(-np.sum(d * d, axis=-1) / (2 * rng.uniform(0.5, 2) ** 2))
    assert np.max(np.abs(vext - vext.astype(np.float32))) > 1e-9
    g, of, f = _fluid(gpu, shape, cell, vext=vext, prec=1, seed=4, amp=0.1)
    E, gr, pg = f.energy_grad_and_precond()
    Eo, gro, pgo = of.energy_grad_and_precond()
    assert abs(E - Eo) <= 1e-12 * abs(Eo)
    assert _rel(gr, gro) < 1e-10 and _rel(pg, pgo) < 1e-10
    of32 = O.make_fluid(g, 0.3, vext=vext.astype(np.float32).astype(np.float64))
    of32.x = of.x
    assert abs(of32.energy() - Eo) > 10 * abs(E - Eo)


def test_white_bear_known_answer(gpu):
    """tests/new-hard-spheres.cpp:52-63: uniform n = 0.1, R = 1, cell 2^3, Nx = 20 -> 2.72225468848892."""
    for prec, tol in ((1, 2e-11), (0, 1e-6)):
        f = gpu.WhiteBearFluidVeff((20, 20, 20), np.diag([2.0, 2.0, 2.0]), R=1.0, kT=1.0, mu=0.0, fft_precision=prec)
        f.Veff = -math.log(0.1)
        f.energy()
        t = f.last_terms
        wb = t["kTphi1"] + t["kTphi2"] + t["kTphi3"]
        assert abs(wb / 2.72225468848892 - 1) < tol


def test_uniform_fluid_is_bulk(gpu):
    """Omega = -p V and zero gradient for the bulk density at the bulk mu."""
    eta = 0.4
    nb = O.bulk_density(eta, 1.0)
    mu = gpu.find_chemical_potential(nb)
    assert abs(mu - O.chemical_potential(nb, 1.0)) < 1e-12 * abs(mu)
    p = gpu.pressure(nb)
    assert abs(p - O.pressure(nb, 1.0)) < 1e-12 * p
    f = gpu.WhiteBearFluidVeff((32, 32, 32), np.diag([3.2] * 3), mu=mu, fft_precision=1)
    f.Veff = -math.log(nb)
    E, gr, pg = f.energy_grad_and_precond()
    V = 3.2 ** 3
    assert abs(E + p * V) < 1e-12 * p * V
    assert np.max(np.abs(pg)) < 1e-12


def _trajectory(gpu, prec):
    shape, cell = (32, 32, 32), np.diag([3.2, 3.2, 3.2])
    g = O.Grid(shape, cell)
    vext = O.planar_wall_vext(g, 1.0)
    g, of, f = _fluid(gpu, shape, cell, eta=0.3, vext=vext, prec=prec)
    mo = O.Minimize(of)
    mo.set_relative_precision(1e-12)
    mo.precondition(True)
    m = gpu.Minimize(f)
    m.set_relative_precision(1e-12)
    m.precondition(True)
    return of, mo, f, m


def test_cg_trajectory_fp64(gpu):
    """FP64 mode follows the oracle's Minimize iterate by iterate."""
    of, mo, f, m = _trajectory(gpu, 1)
    for it in range(12):
        a = mo.improve_energy()
        b = m.improve_energy()
        assert a == b
        assert abs(m.energy() - mo.energy()) <= 1e-11 * abs(mo.energy()), it
        if not a:
            break
    s = m.stats()
    assert s["iter"] == mo.iter
    assert abs(s["step"] - mo.step) <= 1e-6 * abs(mo.step)
    assert _rel(f.Veff, of.x) < 1e-9


@pytest.mark.parametrize("prec", [0, 1])
def test_energy_only_and_gradient_evaluations_agree_bitwise(gpu, prec):
    """The line search compares energies from both evaluation kinds, so they must round
    identically (a difference of FP32 round-off size stalls the line search)."""
    shape, cell = (32, 32, 32), np.diag([3.2, 3.2, 3.2])
    g = O.Grid(shape, cell)
    for seed in range(3):
        _, of, f = _fluid(gpu, shape, cell, eta=0.35, vext=O.planar_wall_vext(g, 1.0), prec=prec, seed=seed)
        e = f.energy()
        eg, _, _ = f.energy_grad_and_precond()
        assert e == eg, (seed, e, eg, e - eg)


@pytest.mark.parametrize("polish_grad", [0, 1])
@pytest.mark.parametrize("amp", [1e-3, 1e-2])
def test_mixed_precision_from_near_converged_starts(gpu, polish_grad, amp):
    """Warm starts close to the answer: the FP32 phase hits its round-off floor within a
    few iterations; the FP64 phases must still converge to the FP64 minimum."""
    shape, cell = (32, 32, 32), np.diag([3.2, 3.2, 3.2])
    g = O.Grid(shape, cell)
    _, of, f = _fluid(gpu, shape, cell, eta=0.35, vext=O.planar_wall_vext(g, 1.0), prec=1)
    m = gpu.Minimize(f)
    m.precondition(True)
    m.set_relative_precision(1e-13)
    m.run()
    E_star, v_star = m.energy(), f.Veff.copy()
    m.close()
    rng = np.random.default_rng(int(amp * 1e4) + polish_grad)
    for trial in range(3):
        f.fft_precision = 0
        f.Veff = v_star + amp * rng.standard_normal(shape)
        m = gpu.Minimize(f)
        m.precondition(True)
        m.set_relative_precision(1e-11)
        m.set_mixed_precision(1e-6, polish_grad)
        m.run()
        s = m.stats()
        assert s["stop_reason"] == "converged", (trial, s, m.gradient_check(), m.energy() / E_star - 1)
        assert abs(m.energy() / E_star - 1) < 1e-9, (trial, m.energy(), E_star)
        m.close()


def test_cg_trajectory_fp64_fused_line_search(gpu):
    """Energy and gradient at the quadratic step's trial point in one evaluation: the
    iterates stay the oracle's; only the evaluation counts drop."""
    of, mo, f, m = _trajectory(gpu, 1)
    m.set_fused_line_search(1)
    for it in range(12):
        a = mo.improve_energy()
        b = m.improve_energy()
        assert a == b
        assert abs(m.energy() - mo.energy()) <= 1e-11 * abs(mo.energy()), it
        if not a:
            break
    s = m.stats()
    assert s["iter"] == mo.iter
    assert _rel(f.Veff, of.x) < 1e-9
    # against the unfused device run from the same start: same energies, fewer evaluations
    _, _, f2, m2 = _trajectory(gpu, 1)
    m2.set_fused_line_search(0)
    for it in range(s["iter"]):
        m2.improve_energy()
    s2 = m2.stats()
    assert s2["iter"] == s["iter"]
    assert abs(m2.energy() - m.energy()) <= 1e-13 * abs(m.energy())
    # Every gradient evaluation also counts as an energy evaluation, so the work in
    # energy-only units is num_energy + num_grad.  In this wall case the quadratic
    # trial point is rejected at every iteration, so fusing stops after the first
    # wasted gradient.
    assert s["num_energy_calcs"] + s["num_grad_calcs"] <= s2["num_energy_calcs"] + s2["num_grad_calcs"] + 1


@pytest.mark.parametrize("polish_grad", [0, 1])
def test_cg_mixed_precision_reaches_fp64_answer(gpu, polish_grad):
    """FP32 phase, then FP64 line-search energies (FP32 or FP64 gradients): |dF/F| < 1e-9."""
    of, mo, f, m = _trajectory(gpu, 0)
    while mo.improve_energy():
        pass
    m.set_relative_precision(1e-11)
    m.set_mixed_precision(1e-6, polish_grad)
    m.run()
    assert m.fp32_iterations() > 0
    assert m.phase_iterations()[2] == 2  # finished in the all-FP64 phase
    assert f.fft_precision == 0  # the plan's own setting is untouched
    assert abs(m.energy() / mo.energy() - 1) < 1e-9
    # the reported energy is an FP64 energy of the final Veff
    f.fft_precision = 1
    assert abs(f.energy() / m.energy() - 1) < 1e-14


@pytest.mark.slow
def test_c1_wall_matches_oracle_golden(gpu):
    gold = json.load(open(os.path.join(GOLDEN, "oracle_c1.json")))
    g = O.Grid.cubic(64, 6.4)
    vext = O.planar_wall_vext(g, 1.0)
    g, of, f = _fluid(gpu, (64, 64, 64), np.diag([6.4] * 3), eta=0.3, vext=vext, prec=1)
    m = gpu.Minimize(f)
    m.set_relative_precision(1e-12)
    m.precondition(True)
    m.run()
    E = m.energy()
    print("c1", E, gold["energy"], m.stats(), flush=True)
    assert abs(E / gold["energy"] - 1) < 1e-9
    n = f.density()
    assert np.max(np.abs(n[:, 0, 0] - np.array(gold["density_x_axis"]))) < 1e-6 * gold["density_max"]


def test_bad_sizes_raise(gpu):
    """Edges above 4096 or with a prime factor above 512 have no FFT pass."""
    with pytest.raises(gpu.DeftError):
        gpu.WhiteBearFluidVeff((1031, 1, 1), np.diag([1.0] * 3))
    with pytest.raises(gpu.DeftError):
        gpu.WhiteBearFluidVeff((1, 1, 4097), np.diag([1.0] * 3))
    with pytest.raises(gpu.DeftError):
        gpu.WhiteBearFluidVeff((0, 16, 16), np.diag([1.0] * 3))


def test_nan_veff_stops_minimizer(gpu):
    f = gpu.WhiteBearFluidVeff((16, 16, 16), np.diag([3.2] * 3), mu=0.0, fft_precision=1)
    v = np.full((16, 16, 16), 2.0)
    v[3, 4, 5] = np.nan
    f.Veff = v
    m = gpu.Minimize(f)
    assert m.improve_energy() is False
    assert m.stats()["stop_reason"] == "nan-energy"


@pytest.mark.parametrize("prec,etol,gtol", [(1, 1e-12, 1e-10), (0, 2e-7, 2e-5)])
def test_every_tuning_variant_matches_oracle(gpu, prec, etol, gtol):
    """All z-block / y-tile / x-tile variants the autotuner may pick give the oracle's answer."""
    shape, cell = (48, 32, 64), np.diag([4.8, 3.2, 6.4])
    g = O.Grid(shape, cell)
    vext = O.planar_wall_vext(g, 1.0)
    g, of, f = _fluid(gpu, shape, cell, vext=vext, prec=prec, seed=5)
    Eo, gro, pgo = of.energy_grad_and_precond()
    for zb in (512, 1024):
        for yt in (8192, 16384):
            for xt in (8192, 16384):
                f.set_tuning(prec, zb, yt, xt)
                E, gr, pg = f.energy_grad_and_precond()
                assert abs(E - Eo) <= etol * abs(Eo), (zb, yt, xt)
                assert _rel(gr, gro) < gtol, (zb, yt, xt)
    t = f.autotune()
    assert t[prec]["zblock"] in (512, 1024) and t[prec]["xtile"] in (8192, 16384)
    assert t[prec]["zphi"] in (0, 1) and t[prec]["xfwd"] in (0, 1)
    assert t[prec]["zphi"] == 0 or t[prec]["zblock"] == 512  # the wave Phi pass runs at 512
    E, gr, pg = f.energy_grad_and_precond()
    assert abs(E - Eo) <= etol * abs(Eo)
    assert np.array_equal(f.Veff, of.x)  # tuning never touches the state


@pytest.mark.parametrize("prec,etol,gtol", [(1, 1e-12, 1e-10), (0, 2e-7, 2e-5)])
@pytest.mark.parametrize("shape", [(48, 32, 64), (21, 20, 128), (16, 16, 256), (20, 21, 16), (32, 21, 32)])
def test_wave_phi_pass_equals_block_pass(gpu, shape, prec, etol, gtol):
    """The one-field-per-wave Phi pass runs the block pass's operations on the same
    values: energy, gradient and preconditioned gradient agree bit for bit, and with
    the oracle.  The line counts 420 and 672 leave the last block partly empty
    (8, 64 and 32 lines per block at z edges 128, 16 and 32)."""
    print("wave phi", shape, prec, flush=True)
    cell = np.diag([0.1 * n for n in shape])
    g, of, f = _fluid(gpu, shape, cell, prec=prec, seed=11)
    Eo, gro, pgo = of.energy_grad_and_precond()
    t = f.tuning(prec)
    f.set_tuning(prec, 512, t["ytile"], t["xtile"], zphi=0)
    Eb, grb, pgb = f.energy_grad_and_precond()
    eb = f.energy()
    f.set_tuning(prec, 512, t["ytile"], t["xtile"], zphi=1)
    assert f.tuning(prec)["zphi"] == 1
    Ew, grw, pgw = f.energy_grad_and_precond()
    ew = f.energy()
    assert Ew == Eb and ew == eb
    assert np.array_equal(grw, grb) and np.array_equal(pgw, pgb)
    assert abs(Ew - Eo) <= etol * abs(Eo)
    assert _rel(grw, gro) < gtol


@pytest.mark.parametrize("prec,etol,gtol", [(1, 1e-12, 1e-10), (0, 2e-7, 2e-5)])
@pytest.mark.parametrize("shape", [(128, 20, 32), (64, 21, 64), (16, 16, 20), (256, 16, 16), (32, 48, 128)])
def test_wave_x_pass_equals_block_pass(gpu, shape, prec, etol, gtol):
    """The one-tile-per-wave forward x pass runs the block pass's operations on the same
    values, alone and with the one-field-per-wave Phi pass.  The compiler fuses a few
    products into FMAs differently in the two kernels, so they agree to rounding (a few
    points differ in the last bits) rather than bit for bit: energies within 4e-16
    (FP64) / 2e-7 (FP32) relative, gradients within 2e-15 / 1e-6 of their maximum, and
    both with the oracle.  The column counts leave idle waves in the last block and tiles
    that span two y rows (4, 8, 32, 2 and 16 columns per tile at x edges 128, 64, 16,
    256, 32)."""
    print("wave x", shape, prec, flush=True)
    cell = np.diag([0.1 * n for n in shape])
    g, of, f = _fluid(gpu, shape, cell, prec=prec, seed=13)
    Eo, gro, pgo = of.energy_grad_and_precond()
    t = f.tuning(prec)
    f.set_tuning(prec, 512, t["ytile"], t["xtile"], zphi=0, xfwd=0)
    Eb, grb, pgb = f.energy_grad_and_precond()
    eb = f.energy()
    zphis = (0, 1) if shape[2] in (16, 32, 64, 128, 256) else (0,)
    for zphi in zphis:
        f.set_tuning(prec, 512, t["ytile"], t["xtile"], zphi=zphi, xfwd=1)
        assert f.tuning(prec)["xfwd"] == 1
        Ew, grw, pgw = f.energy_grad_and_precond()
        ew = f.energy()
        rt, gt = (4e-16, 2e-15) if prec == 1 else (2e-7, 1e-6)
        assert abs(Ew - Eb) <= rt * abs(Eb) and abs(ew - eb) <= rt * abs(eb), zphi
        assert _rel(grw, grb) <= gt and _rel(pgw, pgb) <= gt, zphi
    assert abs(Ew - Eo) <= etol * abs(Eo)
    assert _rel(grw, gro) < gtol


def test_wave_phi_pass_is_refused_where_it_does_not_apply(gpu):
    """Odd z (x) edges keep the block Phi (x) pass."""
    g, of, f = _fluid(gpu, (16, 16, 21), np.diag([1.6, 1.6, 2.1]), prec=1)
    with pytest.raises(gpu.DeftError):
        f.set_tuning(1, 512, 16384, 8192, zphi=1)
    with pytest.raises(gpu.DeftError):
        f.set_tuning(1, 512, 16384, 8192, zphi=2)
    assert f.tuning(1)["zphi"] == 0
    g, of, f = _fluid(gpu, (21, 16, 16), np.diag([2.1, 1.6, 1.6]), prec=1)
    with pytest.raises(gpu.DeftError):
        f.set_tuning(1, 512, 16384, 8192, xfwd=1)
    assert f.tuning(1)["xfwd"] == 0


def test_copy_bandwidth_is_a_sane_yardstick(gpu):
    """The streaming-copy figure the bench reports next to the spec peak."""
    g = gpu.copy_bandwidth(1 << 28, 5)
    assert 500.0 < g < 8000.0, g
    with pytest.raises(gpu.DeftError):
        gpu.copy_bandwidth(1000, 5)


@pytest.mark.parametrize("prec,etol,gtol", [(1, 1e-12, 1e-10), (0, 2e-7, 2e-5)])
def test_constrained_points_match_oracle(gpu, prec, etol, gtol):
    """NaN in Vext: the old API's constrain(mask, f) (density counts, gradient zero)."""
    shape, cell = (32, 32, 32), np.diag([4.0, 4.0, 4.0])
    g = O.Grid(shape, cell)
    wall = O.spherical_solute_vext(g, 1.3, center=(2.0, 2.0, 2.0))
    vext = np.where(np.isinf(wall), np.nan, 0.3 * np.cos(g.positions()[..., 0]))
    vext = vext.astype(np.float32).astype(np.float64)
    g, of, f = _fluid(gpu, shape, cell, vext=vext, prec=prec, seed=8)
    E, gr, pg = f.energy_grad_and_precond()
    Eo, gro, pgo = of.energy_grad_and_precond()
    assert abs(E - Eo) <= etol * abs(Eo)
    assert _rel(gr, gro) < gtol and _rel(pg, pgo) < gtol
    inside = np.isnan(vext)
    assert np.all(gr[inside] == 0) and np.all(pg[inside] == 0)
    n = f.density()
    assert np.all(n[inside] > 0)
