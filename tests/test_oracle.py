"""Pins the FP64 oracle against the reference's known answers and against the mathematics.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign or
normalisation, a transposed index) fails it.
"""

import math

import numpy as np
import pytest

from oracle import fmt_oracle as O


def test_white_bear_known_answer():
    """tests/new-hard-spheres.cpp:52-63: uniform n = 0.1, R = 1, kT = 1, cell 2^3, Nx = 20."""
    g = O.Grid.cubic(20, 2.0)
    f = O.WhiteBearFluidVeff(g, 1.0, 1.0, 0.0, np.zeros(g.shape), np.full(g.shape, -math.log(0.1)))
    f.energy()
    t = f.last_terms
    wb = t["kTphi1"] + t["kTphi2"] + t["kTphi3"]
    assert abs(wb / 2.72225468848892 - 1) < 2e-11
    # homogeneous value = energy / a^3 (tests/new-hard-spheres.cpp:73-77)
    assert abs(O.homogeneous_whitebear(0.1, 1.0) / (2.72225468848892 / 8) - 1) < 2e-11


@pytest.mark.parametrize("eta", [0.1, 0.3, 0.4, 0.45])
def test_bulk_reduces_to_carnahan_starling(eta):
    """White-Bear's bulk limit is Carnahan-Starling: f_ex, mu_ex and p in closed form."""
    n = O.bulk_density(eta, 1.0)
    f_cs = n * (4 * eta - 3 * eta ** 2) / (1 - eta) ** 2
    mu_cs = (8 * eta - 9 * eta ** 2 + 3 * eta ** 3) / (1 - eta) ** 3
    p_cs = n * (1 + eta + eta ** 2 - eta ** 3) / (1 - eta) ** 3
    assert abs(O.homogeneous_whitebear(n, 1.0) / f_cs - 1) < 1e-13
    assert abs((O.chemical_potential(n, 1.0) - math.log(n / O.nQ(1.0))) / mu_cs - 1) < 1e-13
    assert abs(O.pressure(n, 1.0) / p_cs - 1) < 1e-12


def test_golden_numbers_of_the_survey():
    n3, n4 = O.bulk_density(0.3, 1.0), O.bulk_density(0.4, 1.0)
    assert abs(O.homogeneous_whitebear(n3, 1.0) - 0.135931313640731) < 1e-14
    assert abs(O.homogeneous_whitebear(n4, 1.0) - 0.297089227104872) < 1e-14
    assert abs(O.chemical_potential(n3, 1.0) - math.log(n3 / O.nQ(1.0)) - 4.87172011661808) < 1e-12
    assert abs(O.chemical_potential(n4, 1.0) - math.log(n4 / O.nQ(1.0)) - 9.03703703703704) < 1e-12
    assert abs(O.pressure(n3, 1.0) - 0.284599662814618) < 1e-13
    assert abs(O.pressure(n4, 1.0) - 0.661377207959654) < 1e-13
    # 1/nQ at kT = 1 (src/new/SW_liquidFast.cpp:47)
    assert abs(1 / O.nQ(1.0) / 2.646476976618268e-6 - 1) < 1e-14


@pytest.mark.parametrize("shape", [(4, 6, 8), (5, 3, 7)])
def test_fft_convention_against_brute_force_dft(shape):
    """fft = dV * sum_r f(r) exp(-i k.r) on the half spectrum; ifft = (1/V) sum_k (Hermitian).
    Odd edges (old-API grids) keep every kz < Nz/2 + 1 and have no Nyquist plane."""
    g = O.Grid(shape, np.diag([1.3, 2.1, 0.7]))
    rng = np.random.default_rng(0)
    f = rng.standard_normal(g.shape)
    F = O.fft(g, f)
    Nx, Ny, Nz = g.shape
    assert F.shape == (Nx, Ny, Nz // 2 + 1)
    x = np.arange(Nx)[:, None, None]
    y = np.arange(Ny)[None, :, None]
    z = np.arange(Nz)[None, None, :]
    for (a, b, c) in [(0, 0, 0), (1, 2, 3), (3, 2, 2), (2, 1, 0), (Nx - 1, Ny - 1, Nz // 2)]:
        ph = np.exp(-2j * np.pi * (a * x / Nx + b * y / Ny + c * z / Nz))
        assert abs(F[a, b, c] - g.dV * np.sum(f * ph)) < 1e-13
    back = O.ifft(g, F)
    assert np.max(np.abs(back - f)) < 1e-14


def test_fft_round_trip_reference_size():
    """tests/new-fftinverse.cpp:21-43: 100 x 98 x 96, ifft(fft(f)) = f to 1e-15-ish."""
    g = O.Grid((100, 98, 96), np.diag([5.0, 4.9, 4.8]))
    rng = np.random.default_rng(1)
    f = rng.random(g.shape)
    assert np.max(np.abs(O.ifft(g, O.fft(g, f)) - f)) < 1e-13


def test_weights_at_k_zero_and_series_branch():
    g = O.Grid.cubic(16, 3.2)
    for R in (0.7, 1.0, 2.0):
        w3, w2, wv = O.weights(g, R)
        assert abs(w3[0, 0, 0] - 4 * math.pi / 3 * R ** 3) < 1e-14 * R ** 3
        assert abs(w2[0, 0, 0] - 4 * math.pi * R * R) < 1e-14 * R * R
        # the vector weight is odd: i wv k_j vanishes at k = 0
        assert abs(wv[0, 0, 0] * g.kx[0, 0, 0]) == 0.0
    # series vs closed form on both sides of kR = 1e-3 (src/ReciprocalOperators.h:98-116)

    class _G:
        pass

    for kR in (0.999e-3, 1.001e-3):
        gg = _G()
        gg.k = np.array([kR])
        gg.dV = 1e-30  # no smearing
        w3, w2, wv = O.weights(gg, 1.0)
        k = kR
        exact3 = 4 * math.pi * (math.sin(k) - k * math.cos(k)) / k ** 3
        exact2 = 4 * math.pi * math.sin(k) / k
        exactv = 4 * math.pi * (math.cos(k) - math.sin(k) / k) / k ** 2
        assert abs(w3[0] / exact3 - 1) < 1e-9
        assert abs(w2[0] / exact2 - 1) < 1e-12
        assert abs(wv[0] / exactv - 1) < 1e-6


def test_weight_integrals():
    """tests/convolve.cpp:79-266: int step(f) = 4 pi R^3/3 int f, int shell(f) = 4 pi R^2 int f,
    vector shell integrates to 0 and |n2v| <= n2."""
    g = O.Grid.cubic(48, 6.0)
    r = O.minimum_image(g, (3.0, 3.0, 3.0))
    f = np.exp(-4.0 * r * r)
    fk = np.fft.rfftn(f)
    w3, w2, wv = O.weights(g, 1.0)
    n3 = np.fft.irfftn(fk * w3, s=g.shape, axes=(0, 1, 2))
    n2 = np.fft.irfftn(fk * w2, s=g.shape, axes=(0, 1, 2))
    vx = np.fft.irfftn(fk * 1j * wv * g.kx, s=g.shape, axes=(0, 1, 2))
    vy = np.fft.irfftn(fk * 1j * wv * g.ky, s=g.shape, axes=(0, 1, 2))
    vz = np.fft.irfftn(fk * 1j * wv * g.kz, s=g.shape, axes=(0, 1, 2))
    I = f.sum()
    assert abs(n3.sum() / I - 4 * math.pi / 3) < 1e-10
    assert abs(n2.sum() / I - 4 * math.pi) < 1e-10
    assert abs(vx.sum()) < 1e-10 * I and abs(vy.sum()) < 1e-10 * I
    assert np.all(np.sqrt(vx ** 2 + vy ** 2 + vz ** 2) <= n2 + 1e-9 * n2.max())
    # the vector weight points from the blob towards the evaluation point: n2v_x < 0 for x > centre
    # (w2v = i wv k, wv < 0 for small k: the gradient of the shell-smeared blob)
    x = g.positions()[..., 0]
    sel = (np.abs(x - 4.0) < 0.05) & (r < 1.2)
    assert np.all(vx[sel] > 0) or np.all(vx[sel] < 0)


def test_convolution_adjoint_parity():
    """<a, W b> = <W~ a, b>, W~(k) = W(-k): even weights are self-adjoint, odd ones flip sign."""
    g = O.Grid((16, 20, 32), np.diag([3.0, 3.5, 4.0]))
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal(g.shape), rng.standard_normal(g.shape)
    w3, w2, wv = O.weights(g, 1.0)
    conv = lambda W, f: np.fft.irfftn(np.fft.rfftn(f) * W, s=g.shape, axes=(0, 1, 2))
    for W, par in ((w3, 1), (w2, 1), (1j * wv * g.kx, -1), (1j * wv * g.kz, -1)):
        lhs = np.sum(a * conv(W, b))
        rhs = np.sum(conv(par * W, a) * b)
        assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + 1)
        if par < 0:  # and the sign does matter
            assert abs(np.sum(conv(W, a) * b) + lhs) < 1e-12 * (abs(lhs) + 1)


def test_f3_series_matches_closed_form_and_derivative():
    ns = np.array([-0.04, -1e-3, 1e-8, 1e-3, 0.02, 0.0499, 0.0501, 0.2, 0.45])
    f3, df3 = O.f3_and_derivative(ns)
    # closed form where it is well conditioned
    for n, v in zip(ns, f3):
        if abs(n) >= 0.01:
            L = math.log1p(-n)
            exact = (n + (1 - n) ** 2 * L) / (n * n * (1 - n) ** 2)
            assert abs(v / exact - 1) < 1e-12
    assert abs(O.f3_and_derivative(np.array([0.0]))[0][0] - 1.5) < 1e-15
    assert abs(O.f3_and_derivative(np.array([0.0]))[1][0] - 8.0 / 3.0) < 1e-14
    h = 1e-6
    for n, d in zip(ns, df3):
        fp = O.f3_and_derivative(np.array([n + h]))[0][0]
        fm = O.f3_and_derivative(np.array([n - h]))[0][0]
        assert abs((fp - fm) / (2 * h) - d) < 1e-7 * max(1, abs(d))


def test_phi_partials_by_finite_differences():
    rng = np.random.default_rng(4)
    R = 1.0
    for _ in range(20):
        n3 = rng.uniform(0.01, 0.5)
        n2 = rng.uniform(0.1, 3.0)
        v = rng.uniform(-0.5, 0.5, 3)
        args = [n3, n2, *v]
        _, _, _, P = O.whitebear_phi(*[np.array([a]) for a in args], R)
        for i in range(5):
            h = 1e-6
            ap, am = list(args), list(args)
            ap[i] += h
            am[i] -= h
            fp = sum(O.whitebear_phi(*[np.array([a]) for a in ap], R, want_partials=False)[:3])
            fm = sum(O.whitebear_phi(*[np.array([a]) for a in am], R, want_partials=False)[:3])
            assert abs((fp - fm)[0] / (2 * h) - P[i][0]) < 1e-6 * max(1, abs(P[i][0]))


@pytest.mark.parametrize("shape", [(16, 16, 20), (15, 21, 17)])
def test_gradient_by_finite_differences(shape):
    """Central differences along a random direction (src/Functional.cpp:159-212)."""
    g = O.Grid(shape, np.diag([3.2, 3.0, 3.4]))
    rng = np.random.default_rng(1)
    vext = O.planar_wall_vext(g, 1.0)
    vext = np.where(np.isfinite(vext), 0.3 * np.cos(g.positions()[..., 2]), vext)
    f = O.make_fluid(g, 0.3, vext=vext)
    f.x += 0.3 * rng.standard_normal(g.shape)
    E, gr, pg = f.energy_grad_and_precond()
    d = rng.standard_normal(g.shape)
    eps = 1e-4
    fd = (f.energy(f.x + eps * d) - f.energy(f.x - eps * d)) / (2 * eps)
    assert abs(fd - np.vdot(gr, d)) < 1e-7 * abs(fd)
    n = f.density()
    inside = np.isfinite(vext)
    assert np.allclose(pg[inside] * n[inside], gr[inside], rtol=1e-13, atol=0)
    assert np.all(gr[~inside] == 0) and np.all(pg[~inside] == 0)


def test_uniform_fluid_is_bulk():
    eta = 0.4
    g = O.Grid.cubic(16, 3.2)
    f = O.make_fluid(g, eta)
    E, gr, pg = f.energy_grad_and_precond()
    p = O.pressure(O.bulk_density(eta, 1.0), 1.0)
    assert abs(E + p * g.volume) < 1e-12 * p * g.volume
    assert np.max(np.abs(pg)) < 1e-12


class SqrSum:
    """tests/newcode.cpp:9-44: E = sum_i (i+1) x_i^2, preconditioner grad/(i+1)."""

    def __init__(self, x):
        self.x = np.array(x, dtype=np.float64)
        self.k = np.arange(1, len(x) + 1, dtype=np.float64)

    def energy(self):
        return float(np.sum(self.k * self.x * self.x))

    def grad(self):
        return 2 * self.k * self.x

    def energy_grad_and_precond(self):
        g = self.grad()
        return self.energy(), g, g / self.k

    def have_preconditioner(self):
        return True


def test_minimize_exact_minimum_of_quadratic():
    """tests/newcode.cpp:46-70: the default minimizer reaches E <= 1e-300."""
    f = SqrSum(np.arange(1000) * 0.1)
    m = O.Minimize(f)
    while m.improve_energy():
        pass
    assert m.energy() <= 1e-300


def test_minimize_precision_and_iteration_bound():
    """tests/newcode.cpp:72-98: precision 1e-9 within N = 1000 iterations."""
    f = SqrSum(np.arange(1000) * 0.1)
    m = O.Minimize(f)
    m.set_precision(1e-9)
    m.set_miniter(0)
    while m.improve_energy():
        pass
    assert m.energy() <= 1e-9
    assert m.iter <= 1000


def test_minimize_preconditioned_three_iterations():
    """tests/newcode.cpp:100-132: exact preconditioning converges within 3 iterations."""
    f = SqrSum(np.arange(1000) * 0.1)
    m = O.Minimize(f)
    m.set_precision(1e-9)
    m.set_miniter(0)
    m.precondition(True)
    while m.improve_energy():
        pass
    assert m.energy() <= 1e-9
    assert m.iter <= 3


def test_minimize_stops_on_nan():
    """src/new/Minimize.cpp:22-30: a NaN energy stops the minimizer."""
    f = SqrSum([1.0, float("nan"), 2.0])
    m = O.Minimize(f)
    assert m.improve_energy() is False
    assert m.stop_reason == "nan-energy"


def test_vector_weight_is_minus_step_weight():
    """wv = s 4 pi (R cos kR - sin kR/k)/k^2 equals -w3 = -s 4 pi (sin kR - kR cos kR)/k^3
    (identity used by the device kernels to skip a third radial function)."""
    g = O.Grid((32, 48, 20), np.diag([3.1, 4.9, 2.3]))
    w3, w2, wv = O.weights(g, 1.0)
    assert np.max(np.abs(wv + w3)) <= 1e-13 * np.max(np.abs(w3))


# ---------------------------------------------------------------------------
# White-Bear Mark II (src/haskell/WhiteBear.hs:140-159)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("eta", [0.1, 0.3, 0.4, 0.45])
def test_mark_ii_bulk_is_carnahan_starling(eta):
    """Mark II keeps the Carnahan-Starling bulk of Mark I (Hansen-Goos and Roth 2006)."""
    n = O.bulk_density(eta, 1.0)
    f_cs = n * (4 * eta - 3 * eta ** 2) / (1 - eta) ** 2
    mu_cs = (8 * eta - 9 * eta ** 2 + 3 * eta ** 3) / (1 - eta) ** 3
    p_cs = n * (1 + eta + eta ** 2 - eta ** 3) / (1 - eta) ** 3
    assert abs(O.homogeneous_whitebear(n, 1.0, functional="wb2") / f_cs - 1) < 1e-13
    assert abs((O.chemical_potential(n, 1.0, functional="wb2") - math.log(n / O.nQ(1.0))) / mu_cs - 1) < 1e-13
    assert abs(O.pressure(n, 1.0, functional="wb2") / p_cs - 1) < 1e-12


@pytest.mark.parametrize("eta", [0.02, 0.1, 0.3, 0.45])
def test_mark_ii_scaled_particle_relation(eta):
    """Mark II is built so that dPhi/dn3 at bulk weighted densities is the Carnahan-Starling
    pressure (the scaled-particle relation for an infinitely large particle); Mark I misses it.
    A slip in either phi2 or phi3 breaks this."""
    for R in (1.0, 0.7):
        n = O.bulk_density(eta, R)
        p_cs = n * (1 + eta + eta ** 2 - eta ** 3) / (1 - eta) ** 3
        n3, n2, z = np.array([eta]), np.array([4 * math.pi * R * R * n]), np.zeros(1)
        d2 = O.whitebear_phi(n3, n2, z, z, z, R, functional="wb2")[3][0][0]
        d1 = O.whitebear_phi(n3, n2, z, z, z, R, functional="wb")[3][0][0]
        assert abs(d2 / p_cs - 1) < 1e-12
        assert abs(d1 / p_cs - 1) > 1e-4


def test_mark_ii_low_density_limit_and_series():
    """phi2, phi3 -> 1 as n3 -> 0 (Rosenfeld's low-density limit: g2 = 1, g3 = 3/2, g2' = 1 + 2/9 * 0,
    dS2/dn = 2n, dS3/dn = 1); the series branch below |n3| = 0.05 meets the closed form."""
    g2, g2p, g3, g3p = O.markii_factors(np.array([0.0]))
    assert g2[0] == 1.0 and abs(g3[0] - 1.5) < 1e-15
    # g2' = d/dn[(1 + S2/9)/(1-n)] = 1 at 0;  g3' = 1.5 (2 - 4/9) = 7/3 at 0
    assert abs(g2p[0] - 1.0) < 1e-15 and abs(g3p[0] - 7.0 / 3.0) < 1e-14
    for c in (0.05, -0.05):
        lo, hi = O.markii_factors(np.array([c * (1 - 1e-12)])), O.markii_factors(np.array([c * (1 + 1e-12)]))
        for a, b in zip(lo, hi):
            assert abs(a[0] - b[0]) < 1e-12 * max(1.0, abs(a[0]))
    # closed forms of phi2 and phi3 at moderate n, written from the Haskell expressions
    for n in (0.1, 0.3, 0.49):
        L = math.log(1 - n)
        phi2 = (6 * n - 3 * n * n + 6 * (1 - n) * L) / n ** 3
        phi3 = (6 * n - 9 * n * n + 6 * n ** 3 + 6 * (1 - n) ** 2 * L) / (4 * n ** 3)
        g2, _, g3, _ = O.markii_factors(np.array([n]))
        assert abs(g2[0] - (1 + n * n * phi2 / 9) / (1 - n)) < 1e-13
        assert abs(g3[0] - 1.5 * (1 - 4 * n * phi3 / 9) / (1 - n) ** 2) < 1e-13
    h = 1e-6
    for n in (-0.04, 0.01, 0.0499, 0.0501, 0.2, 0.45):
        f = O.markii_factors(np.array([n]))
        fp, fm = O.markii_factors(np.array([n + h])), O.markii_factors(np.array([n - h]))
        for val, deriv in ((0, 1), (2, 3)):
            fd = (fp[val][0] - fm[val][0]) / (2 * h)
            assert abs(fd - f[deriv][0]) < 1e-7 * max(1.0, abs(fd))


def test_mark_ii_phi_partials_by_finite_differences():
    rng = np.random.default_rng(5)
    R = 1.0
    for _ in range(20):
        n3 = rng.uniform(-0.03, 0.5)
        n2 = rng.uniform(0.1, 3.0)
        v = rng.uniform(-0.5, 0.5, 3)
        args = [n3, n2, *v]
        _, _, _, P = O.whitebear_phi(*[np.array([a]) for a in args], R, functional="wb2")
        for i in range(5):
            h = 1e-6
            ap, am = list(args), list(args)
            ap[i] += h
            am[i] -= h
            fp = sum(O.whitebear_phi(*[np.array([a]) for a in ap], R, want_partials=False, functional="wb2")[:3])
            fm = sum(O.whitebear_phi(*[np.array([a]) for a in am], R, want_partials=False, functional="wb2")[:3])
            assert abs((fp - fm)[0] / (2 * h) - P[i][0]) < 1e-6 * max(1, abs(P[i][0]))


def test_mark_ii_gradient_and_uniform_fluid():
    g = O.Grid((16, 16, 20), np.diag([3.2, 3.0, 3.4]))
    rng = np.random.default_rng(2)
    vext = O.planar_wall_vext(g, 1.0)
    f = O.make_fluid(g, 0.35, vext=vext, functional="wb2")
    f.x += 0.3 * rng.standard_normal(g.shape)
    E, gr, pg = f.energy_grad_and_precond()
    d = rng.standard_normal(g.shape)
    eps = 1e-4
    fd = (f.energy(f.x + eps * d) - f.energy(f.x - eps * d)) / (2 * eps)
    assert abs(fd - np.vdot(gr, d)) < 1e-7 * abs(fd)
    # the wall makes the two functionals differ (vector terms and phi2, phi3)
    f1 = O.make_fluid(g, 0.35, vext=vext)
    f1.x[...] = f.x
    assert abs(f1.energy() / E - 1) > 1e-5
    u = O.make_fluid(O.Grid.cubic(16, 3.2), 0.4, functional="wb2")
    Eu, _, pgu = u.energy_grad_and_precond()
    p = O.pressure(O.bulk_density(0.4, 1.0), 1.0, functional="wb2")
    assert abs(Eu + p * u.grid.volume) < 1e-12 * p * u.grid.volume
    assert np.max(np.abs(pgu)) < 1e-12


def test_mark_ii_slow_matches_the_cpp_expressions():
    """HardSpheresWBm2slow (src/HardSpheres.cpp:161-185) written out as the C++ does,
    phi2_n3 = (1 + (6 - 3n + 6(1-n) ln(1-n)/n)/9)/(1-n),
    phi3_n3 = (1 - (6 - 9n + 6n^2 + 6(1-n)^2 ln(1-n)/n)/(9n))/(1-n)^2, Phi3 = phi3_n3 VTT/(36 pi),
    against the oracle's S2/S3 parametrisation; its bulk is not Carnahan-Starling."""
    for n in (0.07, 0.2, 0.35, 0.5):
        L = math.log(1 - n)
        phi2_n3 = (1 + (6 - 3 * n + 6 * (1 - n) * L / n) / 9) / (1 - n)
        phi3_n3 = (1 - (6 - 9 * n + 6 * n * n + 6 * (1 - n) ** 2 * L / n) / 9 / n) / (1 - n) ** 2
        g2, _, g3, _ = O.markii_factors(np.array([n]), g3_scale=1.0)
        assert abs(g2[0] / phi2_n3 - 1) < 1e-13
        assert abs(g3[0] / phi3_n3 - 1) < 1e-12
    for eta in (0.1, 0.4):
        nb = O.bulk_density(eta, 1.0)
        f_cs = nb * (4 * eta - 3 * eta ** 2) / (1 - eta) ** 2
        assert abs(O.homogeneous_whitebear(nb, 1.0, functional="wb2slow") / f_cs - 1) > 1e-3
    rng = np.random.default_rng(6)
    for _ in range(10):
        args = [rng.uniform(0.01, 0.5), rng.uniform(0.1, 3.0), *rng.uniform(-0.5, 0.5, 3)]
        _, _, _, P = O.whitebear_phi(*[np.array([a]) for a in args], 1.0, functional="wb2slow")
        for i in range(5):
            h = 1e-6
            ap, am = list(args), list(args)
            ap[i] += h
            am[i] -= h
            fp = sum(O.whitebear_phi(*[np.array([a]) for a in ap], 1.0, want_partials=False, functional="wb2slow")[:3])
            fm = sum(O.whitebear_phi(*[np.array([a]) for a in am], 1.0, want_partials=False, functional="wb2slow")[:3])
            assert abs((fp - fm)[0] / (2 * h) - P[i][0]) < 1e-6 * max(1, abs(P[i][0]))


# ---------------------------------------------------------------------------
# Tensor White-Bear (src/haskell/WhiteBear.hs:44-51, src/HardSpheres.cpp:30-67)
# ---------------------------------------------------------------------------

def _cpp_tensor_op(kvec, R, dr, i, j):
    """xxshell_op / xyshell_op ... of src/ReciprocalOperators.h:292-404, written out as the C++ does."""
    k = math.sqrt(sum(c * c for c in kvec))
    kR = k * R
    s = math.exp(-6.0 * k * k * dr * dr)
    if i == j:
        cos2 = kvec[i] * kvec[i] / (k * k)
        return s * (4 * math.pi * R * R) * (math.sin(kR) / kR * (-1.0 / 3 + cos2 - 3 * cos2 / (kR * kR) + 1 / (kR * kR))
                                             + math.cos(kR) / (kR * kR) * (-1 + 3 * cos2))
    return -s * (4 * math.pi) * kvec[i] * kvec[j] / k ** 4 * ((3 - kR * kR) * math.sin(kR) / kR - 3 * math.cos(kR))


def test_tensor_weights_match_the_reference_operators():
    g = O.Grid((12, 10, 14), np.array([[3.0, 0.0, 0.0], [0.4, 2.6, 0.0], [0.0, 0.3, 3.3]]))
    R = 1.1
    T = O.tensor_weights(g, R)
    dr = g.dV ** (1.0 / 3.0)
    rng = np.random.default_rng(3)
    picks = [(int(a), int(b), int(c)) for a, b, c in zip(rng.integers(0, 12, 40), rng.integers(0, 10, 40),
                                                           rng.integers(0, 8, 40))]
    for (a, b, c) in picks:
        kv = (g.kx[a, b, c], g.ky[a, b, c], g.kz[a, b, c])
        if g.k[a, b, c] == 0:
            assert all(t[a, b, c] == 0 for t in T)
            continue
        # the closed form cancels at small kR: compare at the accuracy it keeps
        tol = 1e-12 if g.k[a, b, c] * R > 1 else 1e-12 / (g.k[a, b, c] * R) ** 4
        for t, (i, j) in zip(T, O.TENSOR_COMPONENTS):
            ref = _cpp_tensor_op(kv, R, dr, i, j)
            assert abs(t[a, b, c] - ref) < tol * 4 * math.pi * R * R, (a, b, c, i, j)
    # traceless, and the amplitude identity s 4 pi R^2 A = w2 - 3 w3 / R
    assert np.max(np.abs(T[0] + T[1] + T[2])) < 1e-14
    w3, w2, _ = O.weights(g, R)
    kh2 = np.where(g.k > 0, (g.kx / np.where(g.k > 0, g.k, 1)) ** 2 - 1.0 / 3.0, 0.0)
    big = g.k * R > 1
    assert np.allclose(T[0][big], ((w2 - 3 * w3 / R) * kh2)[big], rtol=1e-12, atol=1e-14)


def test_tensor_weight_is_the_shell_tensor_in_real_space():
    """Along x the planar projection of delta(R - r)(x^2/r^2 - 1/3) is 2 pi R (x^2/R^2 - 1/3) on |x| < R;
    its cosine transform (quadrature, no formula shared with the oracle) is W_xx / smear."""
    from scipy.integrate import quad
    R = 0.9
    g = O.Grid((40, 8, 8), np.diag([12.0, 3.0, 3.0]))
    T = O.tensor_weights(g, R)
    dr = g.dV ** (1.0 / 3.0)
    for m in (1, 2, 3, 7, 13):
        k = g.kx[m, 0, 0]
        F = quad(lambda x: 2 * math.pi * R * (x * x / (R * R) - 1.0 / 3.0) * math.cos(k * x), -R, R,
                 epsabs=1e-12, epsrel=1e-12, limit=200)[0]
        s = math.exp(-6.0 * k * k * dr * dr)
        assert abs(T[0][m, 0, 0] / s - F) < 1e-10 * 4 * math.pi * R * R, m
        assert abs(T[1][m, 0, 0] + 0.5 * T[0][m, 0, 0]) < 1e-15 and T[3][m, 0, 0] == 0


def test_tensor_phi_partials_and_limits():
    rng = np.random.default_rng(8)
    R = 1.0
    for _ in range(15):
        args = [rng.uniform(-0.03, 0.5), rng.uniform(0.1, 3.0), *rng.uniform(-0.5, 0.5, 3), *rng.uniform(-0.4, 0.4, 6)]
        _, _, _, P = O.whitebear_tensor_phi(*[np.array([a]) for a in args], R)
        for i in range(11):
            h = 1e-6
            ap, am = list(args), list(args)
            ap[i] += h
            am[i] -= h
            fp = sum(O.whitebear_tensor_phi(*[np.array([a]) for a in ap], R, want_partials=False)[:3])
            fm = sum(O.whitebear_tensor_phi(*[np.array([a]) for a in am], R, want_partials=False)[:3])
            assert abs((fp - fm)[0] / (2 * h) - P[i][0]) < 1e-6 * max(1, abs(P[i][0])), i
        # with T = 0 the tensor functional is White-Bear
        z = [np.array([a]) for a in args[:5]] + [np.zeros(1)] * 6
        assert np.allclose(O.whitebear_tensor_phi(*z, R)[:3], O.whitebear_phi(*z[:5], R)[:3], rtol=1e-15, atol=0)
    # a uniaxial tensor aligned with n2v (planar symmetry): X = n2^3 - 3 n2 v^2 + 9 (v^2 t - 3 t^3/4) for
    # T = t diag(1, -1/2, -1/2), v = (v, 0, 0): tr T^3 = t^3 (1 - 1/8 - 1/8) = 3 t^3 / 4
    n2, v, t = 1.3, 0.4, 0.2
    X = n2 ** 3 - 3 * n2 * v * v + 9 * (v * v * t - 3 * t ** 3 / 8)
    n3 = 0.3
    f3 = O.f3_and_derivative(np.array([n3]))[0][0]
    a = [np.array([q]) for q in (n3, n2, v, 0.0, 0.0, t, -t / 2, -t / 2, 0.0, 0.0, 0.0)]
    assert abs(O.whitebear_tensor_phi(*a, 1.0)[2][0] - f3 * X / (36 * math.pi)) < 1e-15


@pytest.mark.parametrize("shape", [(16, 16, 20), (15, 21, 17)])
def test_tensor_gradient_and_uniform_fluid(shape):
    g = O.Grid(shape, np.diag([3.2, 3.0, 3.4]))
    rng = np.random.default_rng(12)
    f = O.make_fluid(g, 0.35, vext=O.planar_wall_vext(g, 1.0), functional="wbt")
    f.x += 0.3 * rng.standard_normal(g.shape)
    E, gr, pg = f.energy_grad_and_precond()
    d = rng.standard_normal(g.shape)
    eps = 1e-4
    fd = (f.energy(f.x + eps * d) - f.energy(f.x - eps * d)) / (2 * eps)
    assert abs(fd - np.vdot(gr, d)) < 1e-7 * abs(fd)
    u = O.make_fluid(O.Grid.cubic(16, 3.2), 0.4, functional="wbt")
    Eu, _, pgu = u.energy_grad_and_precond()
    p = O.pressure(O.bulk_density(0.4, 1.0), 1.0)
    assert abs(Eu + p * u.grid.volume) < 1e-12 * p * u.grid.volume
    assert np.max(np.abs(pgu)) < 1e-12


@pytest.mark.parametrize("fn", ["wb", "wb2"])
@pytest.mark.parametrize("eta", [0.1, 0.3, 0.45])
def test_contact_values_bulk_identities(fn, eta):
    """In a uniform fluid dA/dR is the R-derivative of the bulk free energy, so both
    gSigma estimators (and Yu-Wu, Gross) give the Carnahan-Starling contact value
    (1 - eta/2)/(1 - eta)^3 through the virial route; n0 = nA = n."""
    g = O.Grid((16, 12, 20), np.diag([3.2, 2.4, 4.0]))
    f = O.make_fluid(g, eta, functional=fn)
    n0, gS, yuwu, nA, gA, gross, gS_cpp = O.contact_values(f)
    cs = (1 - eta / 2) / (1 - eta) ** 3
    nb = O.bulk_density(eta, 1.0)
    for v in (gS, yuwu, gA, gross, gS_cpp):
        assert np.max(np.abs(v / cs - 1)) < 1e-12
    assert np.max(np.abs(n0 / nb - 1)) < 1e-12 and np.max(np.abs(nA / nb - 1)) < 1e-12


def test_constrained_points_follow_the_reference_constraint():
    """NaN in Vext is constrain(mask, f) with mask 0 (src/Functional.cpp:759-838): the energy
    is that of the unconstrained functional (no external potential there), the gradient
    is the unconstrained one times the mask."""
    g = O.Grid((16, 16, 20), np.diag([3.2, 3.0, 3.4]))
    rng = np.random.default_rng(21)
    mask = O.spherical_solute_vext(g, 1.1, center=(1.6, 1.5, 1.7))  # inf inside
    vc = np.where(np.isinf(mask), np.nan, 0.0)
    fc = O.make_fluid(g, 0.3, vext=vc)
    fu = O.make_fluid(g, 0.3, vext=np.zeros(g.shape))
    x = fc.x + 0.3 * rng.standard_normal(g.shape)
    fc.x[...] = x
    fu.x[...] = x
    Ec, gc, pc = fc.energy_grad_and_precond()
    Eu, gu, pu = fu.energy_grad_and_precond()
    assert Ec == Eu
    inside = np.isnan(vc)
    assert inside.any() and np.all(gc[inside] == 0) and np.all(pc[inside] == 0)
    assert np.array_equal(gc[~inside], gu[~inside]) and np.array_equal(pc[~inside], pu[~inside])
