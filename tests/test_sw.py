"""The square-well liquid (src/haskell/SW_liquid.hs): oracle pins, the
product's point code on the host, GPU parity, and the reference's published
radial profiles.

The oracle (oracle/sw_oracle.py) is pinned by
* the five kernels' closed forms against the radial Fourier integrals of
  -eps ((r - sigma)/sigma)^i over the well, and their k = 0 integrals;
* its gSigmaA against the contact-density oracle's (whose columns match the
  reference's published contact profiles);
* its gradient against Richardson-extrapolated central differences of its
  energy;
* a uniform liquid: energy = volume x the homogeneous free energy density and
  zero gradient at the homogeneous chemical potential.
The product is checked against the oracle and, on the GPU, against the
reference's published profiles papers/square-well-fluid/data/radial-sw-*.dat
(tests/golden/sw_radial.json) through the reference driver's protocol.
"""

import json
import math
import os
import subprocess

import numpy as np
import pytest
from scipy.integrate import quad

from oracle import fmt_oracle as O
from oracle import sw_oracle as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


# ---------------------------------------------------------------------------
# oracle pins (CPU)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("lam", [1.3, 1.6])
def test_kernels_are_radial_transforms_of_the_well(lam):
    sigma, eps = 2.0, 0.7
    for k in (0.3, 1.1, 2.7, 6.0):
        xi = [float(x[0]) for x in W.sw_xi(np.array([k]), sigma, lam, eps)]
        for i in range(5):
            q = quad(lambda r: -eps * 4 * math.pi * r * r * np.sinc(k * r / math.pi) * ((r - sigma) / sigma) ** i,
                     sigma, lam * sigma, epsabs=1e-14, epsrel=1e-12)[0]
            assert abs(xi[i] - q) <= 1e-9 * max(1.0, abs(q)), (k, i)
    zero = W.sw_xi_at_zero(sigma, lam, eps)
    for i in range(5):
        q = quad(lambda r: -eps * 4 * math.pi * r * r * ((r - sigma) / sigma) ** i, sigma, lam * sigma)[0]
        assert abs(zero[i] - q) <= 1e-12 * abs(q), i


def _case(shape=(16, 14, 18), kT=3.0, ff=0.2, seed=1):
    R, eps, lam = 1.0, 1.0, 1.3
    grid = O.Grid(shape, np.array([[6.0, 0, 0], [0.5, 5.6, 0], [0, 0.4, 7.2]]))
    nb = ff / (math.pi * 8 / 6)
    mu = W.homogeneous_sw_dn(nb, R, eps, lam, kT)
    pos = grid.positions()
    vext = 0.8 * kT * np.cos(2 * math.pi * pos[..., 2] / 7.2) + 0.5 * np.sin(2 * math.pi * pos[..., 0] / 6.0)
    rng = np.random.default_rng(seed)
    x = -kT * math.log(nb) + 0.3 * kT * rng.normal(size=grid.shape)
    return grid, R, eps, lam, kT, nb, mu, vext, x


def test_gsigma_is_the_contact_estimator():
    grid, R, eps, lam, kT, nb, mu, vext, x = _case()
    f = W.SWLiquidVeff(grid, R, eps, lam, kT, mu, vext, x)
    g = f.sw_parts(x)["g"]
    gA = O.contact_values(f, functional="wb")[4]
    assert np.max(np.abs(g - gA)) <= 1e-13 * np.max(np.abs(gA))


def test_gradient_matches_richardson_differences():
    grid, R, eps, lam, kT, nb, mu, vext, x = _case()
    f = W.SWLiquidVeff(grid, R, eps, lam, kT, mu, vext, x)
    sd = f.sw_dn(x)
    n = f.density(x)
    rng = np.random.default_rng(4)
    for _ in range(4):
        i = tuple(int(rng.integers(0, s)) for s in grid.shape)

        def fd(h):
            xp, xm = x.copy(), x.copy()
            xp[i] += h
            xm[i] -= h
            return (f.sw_energy(xp) - f.sw_energy(xm)) / (2 * h)

        rich = (4 * fd(2e-3) - fd(4e-3)) / 3
        an = sd[i] * grid.dV * (-n[i] / kT)
        assert abs(rich - an) <= 2e-8 * abs(an) + 1e-13, (i, rich, an)


def test_uniform_liquid_is_homogeneous():
    grid, R, eps, lam, kT, nb, mu, _, _ = _case()
    x = np.full(grid.shape, -kT * math.log(nb))
    f = W.SWLiquidVeff(grid, R, eps, lam, kT, mu, np.zeros(grid.shape), x)
    E = f.energy(x)
    assert abs(E / (grid.volume * W.homogeneous_sw(nb, R, eps, lam, kT, mu)) - 1) < 1e-13
    assert np.max(np.abs(f.grad(x))) < 1e-14


# ---------------------------------------------------------------------------
# the product's point code on the host (deft_amd/csrc/sw.hpp)
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("sw") / "sw_harness")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "--cuda-host-only",
                    os.path.join(ROOT, "tests", "host", "sw_harness.cpp"), "-o", exe], check=True,
                   stderr=subprocess.DEVNULL)
    return exe


def _run(exe, text):
    return subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout.split("\n")


def test_kernels_match_oracle(harness):
    ks = np.concatenate([[0.0], np.geomspace(0.2, 60.0, 120)])
    for R, eps, lam in ((1.0, 1.0, 1.3), (0.8, 0.5, 1.7)):
        out = _run(harness, "".join(f"x {R!r} {eps!r} {lam!r} {float(k)!r}\n" for k in ks))
        got = np.array([[float(t) for t in l.split()] for l in out[:len(ks)]])
        ref = np.stack(W.sw_xi(ks, 2 * R, lam, eps), axis=1)
        # the closed forms cancel k^-7-sized terms at small k (in the reference too)
        tol = 1e-13 * np.max(np.abs(ref), axis=0) + 1e-15 * (1 + (2 * R * np.maximum(ks, 0.2)) ** -7)[:, None] * 1e3
        assert np.all(np.abs(got - ref) <= tol), (R, lam, np.max(np.abs(got - ref)))


def test_afields_and_jacobian_match_oracle(harness):
    rng = np.random.default_rng(2)
    pts = [(rng.uniform(0.01, 0.6), rng.uniform(0.2, 4.0), *rng.normal(scale=0.3, size=3)) for _ in range(30)]
    pts += [(0.02, 0.3, 0.01, -0.02, 0.0), (1e-4, 1e-3, 0.0, 0.0, 0.0)]  # the series branch of f3
    R = 1.0
    text = "".join(f"a {R!r} " + " ".join(repr(float(v)) for v in p) + "\n" for p in pts)
    text += "".join(f"j {R!r} " + " ".join(repr(float(v)) for v in p) + "\n" for p in pts)
    out = _run(harness, text)
    for i, p in enumerate(pts):
        a = [float(t) for t in out[i].split()]
        ref = [float(np.real(v)) for v in W.afields(*[np.float64(v) for v in p], R)]
        for x, y in zip(a, ref):
            assert abs(x - y) <= 1e-13 * max(1.0, abs(y)), (i, x, y)
        jac = np.array([float(t) for t in out[len(pts) + i].split()]).reshape(5, 8)
        h = 1e-30
        for m in range(5):
            w = [complex(v) for v in p]
            w[m] += 1j * h
            col = [np.imag(v) / h for v in W.afields(*[np.complex128(v) for v in w], R)]
            for j in range(8):
                assert abs(jac[m, j] - col[j]) <= 1e-12 * max(1.0, abs(col[j])), (i, m, j)


def test_homogeneous_matches_oracle(harness):
    cases = [(1.0, 1.0, 1.3, 3.0, 0.05), (1.0, 1.0, 1.3, 1.0, 0.2), (0.7, 2.0, 1.5, 0.8, 0.4)]
    text = "".join(f"h {R!r} {e!r} {l!r} {kT!r} {n!r} {math.log(O.nQ(kT))!r}\n" for R, e, l, kT, n in cases)
    out = _run(harness, text)
    for (R, e, l, kT, n), line in zip(cases, out):
        f, d = map(float, line.split())
        assert f == pytest.approx(W.homogeneous_sw(n, R, e, l, kT), rel=1e-13)
        assert d == pytest.approx(W.homogeneous_sw_dn(n, R, e, l, kT), rel=1e-12)


def test_homogeneous_c_abi_matches_oracle():
    import deft_amd
    for R, e, l, kT, n, mu in ((1.0, 1.0, 1.3, 3.0, 0.05, 0.3), (0.7, 2.0, 1.5, 0.8, 0.4, -1.0)):
        hf = deft_amd.HomogeneousSW_liquid(R=R, epsilon=e, lambda_=l, kT=kT, n=n, mu=mu)
        assert hf.energy() == pytest.approx(W.homogeneous_sw(n, R, e, l, kT, mu), rel=1e-13)
        assert hf.d_by_dn() == pytest.approx(W.homogeneous_sw_dn(n, R, e, l, kT, mu), rel=1e-12, abs=1e-13)


# ---------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol_e,tol_g", [(1, 1e-12, 1e-10), (0, 2e-6, 2e-4)])
def test_energy_and_gradient_match_oracle(gpu, prec, tol_e, tol_g):
    grid, R, eps, lam, kT, nb, mu, vext, x = _case()
    fo = W.SWLiquidVeff(grid, R, eps, lam, kT, mu, vext, x)
    f = gpu.SWLiquidVeff(grid.shape, grid.cell, R=R, epsilon=eps, lambda_=lam, kT=kT, mu=mu, fft_precision=prec)
    f.Vext = vext
    f.Veff = x
    E, Eo = f.energy(), fo.energy()
    assert abs(E - Eo) <= tol_e * abs(Eo), (E, Eo)
    for k, v in fo.last_terms.items():
        assert abs(f.last_terms[k] - v) <= tol_e * max(abs(Eo), 1.0), (k, f.last_terms[k], v)
    g, go = f.grad(), fo.grad(x)
    assert np.max(np.abs(g - go)) <= tol_g * np.max(np.abs(go))
    f.close()


@pytest.mark.gpu
def test_uniform_liquid_on_gpu(gpu):
    grid, R, eps, lam, kT, nb, mu, _, _ = _case()
    f = gpu.SWLiquidVeff(grid.shape, grid.cell, R=R, epsilon=eps, lambda_=lam, kT=kT, mu=mu, fft_precision=1)
    f.fill_Veff(-kT * math.log(nb))
    hf = gpu.HomogeneousSW_liquid(R=R, epsilon=eps, lambda_=lam, kT=kT, n=nb, mu=mu)
    assert abs(f.energy() / (grid.volume * hf.energy()) - 1) < 1e-12
    assert np.max(np.abs(f.grad())) < 1e-12
    f.close()


GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "sw_radial.json")))
# the published runs whose far field is back at the bulk packing fraction
RADIAL = [c for c in GOLD["cases"] if abs(c["eta"][-1] - c["ff"]) < 1e-3 * c["ff"] + 1e-4]


def driver_setup(kT, lam, ff, edge=18.0, dx=0.1, R=1.0, eps=1.0):
    """papers/square-well-fluid/figs/new-radial-sw.cpp:97-151: a test particle at the
    origin (Vext = 400 kT for r <= sigma, -eps for sigma < r < lambda sigma), mu from
    the homogeneous liquid, Veff = -kT ln n_b (+ 400 kT inside)."""
    sigma = 2 * R
    N = int(math.ceil(edge / dx / 2)) * 2
    idx = np.arange(N)
    c = np.where(idx > N // 2, idx - N, idx) * edge / N
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    nb = ff / (math.pi * sigma ** 3 / 6)
    vmax = 400 * kT
    vext = np.where(r <= sigma, vmax, np.where((r > sigma) & (r < sigma * lam), -eps, 0.0))
    veff = -kT * math.log(nb) + np.where(r <= sigma, vmax, 0.0)
    return N, vext, veff, nb


@pytest.mark.gpu
@pytest.mark.parametrize("case", RADIAL, ids=lambda c: f"kT{c['kT']:g}-ff{c['ff']:g}")
def test_published_radial_profile(gpu, case):
    kT, lam, ff = case["kT"], case["lambda"], case["ff"]
    R, eps = 1.0, 1.0
    N, vext, veff, nb = driver_setup(kT, lam, ff)
    hf = gpu.HomogeneousSW_liquid(R=R, epsilon=eps, lambda_=lam, kT=kT, n=nb, mu=0.0)
    mu = hf.d_by_dn()
    f = gpu.SWLiquidVeff((N, N, N), np.diag([18.0] * 3), R=R, epsilon=eps, lambda_=lam, kT=kT, mu=mu,
                         fft_precision=1)
    f.Vext = vext
    f.Veff = veff
    m = gpu.Minimize(f)
    m.set_relative_precision(0)
    m.set_maxiter(500)
    m.precondition(True)
    m.run()
    eta = f.density()[0, 0, :N // 2] * math.pi * (2 * R) ** 3 / 6
    pub = np.array(case["eta"])
    outside = np.array(case["r_over_sigma"]) > 1.0
    err = np.max(np.abs(eta[outside] - pub[outside]) / pub[outside])
    print(f"kT {kT} ff {ff}: {m.stats()['iter']} iterations, max relative deviation {err:.2e}")
    assert err < 1e-4, err
    m.close()
    f.close()


@pytest.mark.gpu
def test_ladder_reaches_the_fp64_minimum(gpu):
    """The FP32 -> FP64 precision ladder on the square-well liquid (a 64^3 test
    particle set-up at dx = 0.2) ends at the pure-FP64 minimum: the same energy to
    1e-9 relative and densities within 1e-6 of the largest."""
    kT, lam, ff, R, eps = 3.0, 1.3, 0.3, 1.0, 1.0
    sigma, edge, M = 2 * R, 12.8, 64
    idx = np.arange(M)
    c = np.where(idx > M // 2, idx - M, idx) * edge / M
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    nb = ff / (math.pi * sigma ** 3 / 6)
    vext = np.where(r <= sigma, 400 * kT, np.where(r < sigma * lam, -eps, 0.0))
    mu = gpu.HomogeneousSW_liquid(R=R, epsilon=eps, lambda_=lam, kT=kT, n=nb).d_by_dn()
    out = []
    for ladder in (False, True):
        f = gpu.SWLiquidVeff((M,) * 3, np.diag([edge] * 3), R=R, epsilon=eps, lambda_=lam, kT=kT, mu=mu,
                             fft_precision=0 if ladder else 1)
        f.Vext = vext
        f.fill_Veff(-kT * math.log(nb))
        m = gpu.Minimize(f)
        m.precondition(True)
        # densities are fixed to about sqrt(energy precision) at a minimum, so
        # 1e-6 of the largest needs the energy to 1e-12
        m.set_relative_precision(1e-12)
        m.set_maxiter(600)
        if ladder:
            m.set_mixed_precision(1e-5, gpu.FFT_FP64)
        m.run()
        assert m.stats()["stop_reason"] == "converged", (ladder, m.stats())
        f.fft_precision = 1
        out.append((f.energy(), f.density()))
        m.close()
        f.close()
    (E64, n64), (El, nl) = out
    assert abs(El - E64) <= 1e-9 * abs(E64), (El, E64)
    assert np.max(np.abs(nl - n64)) <= 1e-6 * np.max(n64)
