"""FP64 CPU oracle for the White-Bear hard-sphere fluid and its conjugate-gradient minimizer.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``deft_amd``) never calls it and fails loudly
when its HIP library is missing.

It restates, in numpy/scipy FP64, the reference algorithm of deft
(github.com/droundy/deft) for this path:

* grid layout, FFT conventions and k-vectors
  - row-major real fields ``i = x*Ny*Nz + y*Nz + z``            (src/Grid.h:37-42)
  - forward transform r2c times dV                              (src/Grid.cpp:470-478, src/new/Vector.h:327-350)
  - inverse transform c2r times 1/V                             (src/ReciprocalGrid.cpp:52-65, src/new/Vector.h:352-373)
  - k-vector of a flat half-spectrum index, x and y wrapped
    to (-N/2, N/2]                                              (src/ReciprocalOperators.h:20-41)
* the fundamental-measure weights (step, shell, vector shell) with the
  Gaussian smearing exp(-6 k^2 dr^2), dr = dV^(1/3), and their small-kR
  series branches                                              (src/ReciprocalOperators.h:90-235, src/haskell/FMT.hs:21-65)
* the White-Bear excess free energy density without tensor
  (Phi1 + Phi2 + Phi3)                                          (src/HardSpheres.cpp:137-159, src/haskell/WhiteBear.hs:47-60)
  and its closed-form partial derivatives                       (src/haskell/WhiteBear.hs:114-121)
* White-Bear Mark II in both of the reference's forms           (src/haskell/WhiteBear.hs:140-159, src/HardSpheres.cpp:161-185)
* tensor White-Bear: the traceless shell tensor weight and the
  tensor third term                                             (src/haskell/FMT.hs:82-88, src/ReciprocalOperators.h:292-404,
                                                                 src/haskell/WhiteBear.hs:44-51, src/HardSpheres.cpp:30-67)
* the ideal gas with the water-mass thermal density nQ           (src/IdealGas.cpp:54-72, src/haskell/IdealGas.hs:7-20)
* the external potential and chemical potential terms           (src/ExternalPotential.cpp:19-21, src/ChemicalPotential.cpp:19-22)
* the effective-potential variable n = exp(-Veff/kT), its chain
  rule and the new-API preconditioner grad/n                    (src/EffectivePotentialToDensity.cpp:28-90,
                                                                 src/haskell/create_generators.py:21-22,
                                                                 src/haskell/NewCode.hs:229-242,306-319)
* the homogeneous path used for the chemical potential and
  pressure                                                      (src/equation-of-state.cpp:164-192, src/Functional.h:382-392)
* the new-API Polak-Ribiere conjugate gradient with quadratic
  line search, cached energies and the geometric error
  estimate                                                      (src/new/Minimize.cpp:9-313, src/new/Minimize.h:13-199)
* the structure-factor preconditioner of the product (no reference
  counterpart): grad -> n^-1/2 F^-1 S(k) F n^-1/2 grad with the uniform
  fluid's S(k); its pins are mathematical (tests/test_sk_precond.py)
* the old-API minimizer stack: ConjugateGradient and
  PreconditionedConjugateGradient (pgrad (dF/dn)/(Veff - kT))
  with QuadraticLineMinimizer, Precision and MaxIter            (src/ConjugateGradient.cpp:59-171,
                                                                 src/QuadraticLineMinimizer.cpp:27-215,
                                                                 src/Precision.cpp:23-156, src/Minimizer.cpp,
                                                                 src/EffectivePotentialToDensity.cpp:62-71)

One deliberate numerical difference from the reference: the Phi3 prefactor
f3(n3) = (n3 + (1-n3)^2 ln(1-n3)) / (n3^2 (1-n3)^2) is evaluated with its Taylor
series for |n3| < 0.05, where the closed form cancels catastrophically (it is
0/0 at n3 = 0, e.g. inside a hard wall).  The series is the same function;
tests check that both branches agree where they overlap.

Hard walls are V_ext = +inf; there the density is exactly zero, the point
contributes nothing to the energy and its gradient and preconditioned
gradient are zero (the limit of the reference's V_ext = V_eff = 500 kT wall,
papers/fuzzy-fmt/figs/new-bh-walls.cpp:136-145).
"""

import math

import numpy as np
from scipy import fft as _sfft

# Molecular mass of water in atomic units, used by deft for nQ (src/IdealGas.cpp:54-59).
WATER_MASS = 18.01528 * 1822.8885
SPREADING = 6.0  # src/ReciprocalOperators.h:90
F3_SERIES_CUTOFF = 0.05
F3_SERIES_TERMS = 18


def nQ(kT):
    """Thermal density (m kT / 2 pi)^(3/2) (src/IdealGas.cpp:54-59)."""
    return (WATER_MASS * kT / (2 * math.pi)) ** 1.5


def _rfftn(a):
    return _sfft.rfftn(a, workers=1)


def _irfftn(a, shape):
    return _sfft.irfftn(a, s=shape, workers=1)


# ---------------------------------------------------------------------------
# Grid and k-space weights
# ---------------------------------------------------------------------------

class Grid:
    """Grid metadata: shape, cell, dV and half-spectrum k-vectors.

    ``cell`` holds the lattice vectors a1, a2, a3 as rows; the real-space point
    (x, y, z) sits at x*a1/Nx + y*a2/Ny + z*a3/Nz (src/GridDescription.cpp:6-16,
    src/Grid.h:37-42).  dV = |det(cell)|/N.  k = 2 pi inv(cell)^T m with m the
    wrapped integer frequency (src/lattice.h:105-107, src/ReciprocalOperators.h:23-38).
    """

    def __init__(self, shape, cell):
        self.shape = tuple(int(s) for s in shape)
        Nx, Ny, Nz = self.shape
        # The new API asserts even edges (src/new/Vector.h:328-330); the old-API Grid
        # takes any edge (src/GridDescription.cpp:28, r2c of any size in src/Grid.cpp:470-478).
        if min(self.shape) < 1:
            raise ValueError("grid edges must be positive")
        self.cell = np.array(cell, dtype=np.float64).reshape(3, 3)
        self.volume = abs(float(np.linalg.det(self.cell)))
        self.N = Nx * Ny * Nz
        self.dV = self.volume / self.N
        self.kshape = (Nx, Ny, Nz // 2 + 1)
        recip = 2 * math.pi * np.linalg.inv(self.cell).T  # rows b1, b2, b3
        self.recip = recip

        def wrap(n):
            m = np.arange(n)
            return np.where(m > n // 2, m - n, m).astype(np.float64)

        mx = wrap(Nx)[:, None, None]
        my = wrap(Ny)[None, :, None]
        mz = np.arange(Nz // 2 + 1, dtype=np.float64)[None, None, :]
        self.kx = mx * recip[0, 0] + my * recip[1, 0] + mz * recip[2, 0]
        self.ky = mx * recip[0, 1] + my * recip[1, 1] + mz * recip[2, 1]
        self.kz = mx * recip[0, 2] + my * recip[1, 2] + mz * recip[2, 2]
        self.k = np.sqrt(self.kx ** 2 + self.ky ** 2 + self.kz ** 2)

    @classmethod
    def cubic(cls, n, length):
        return cls((n, n, n), np.diag([length] * 3))

    def positions(self):
        """Cartesian positions of every grid point, shape (Nx, Ny, Nz, 3), cell-origin based."""
        Nx, Ny, Nz = self.shape
        fx = (np.arange(Nx) / Nx)[:, None, None, None]
        fy = (np.arange(Ny) / Ny)[None, :, None, None]
        fz = (np.arange(Nz) / Nz)[None, None, :, None]
        c = self.cell
        return fx * c[0] + fy * c[1] + fz * c[2]


def weights(grid, R):
    """Fourier transforms of the fundamental-measure weights on the half spectrum.

    Returns (w3, w2, wv) with
      step          w3(k) = s 4 pi (sin kR - kR cos kR)/k^3        (src/ReciprocalOperators.h:94-118)
      shell         w2(k) = s 4 pi R sin(kR)/k                     (src/ReciprocalOperators.h:120-142)
      vector shell  w2v_j(k) = i wv(k) k_j,
                    wv(k) = s 4 pi (R cos kR - sin kR/k)/k^2       (src/ReciprocalOperators.h:162-235)
    s = exp(-6 k^2 dr^2), dr = dV^(1/3); for kR <= 1e-3 the reference's
    power series (without smearing) is used.  n0 = n2/(4 pi R^2),
    n1 = n2/(4 pi R), n1v = n2v/(4 pi R) (src/haskell/FMT.hs:34-39).
    """
    k = grid.k
    dr = grid.dV ** (1.0 / 3.0)
    kR = k * R
    small = kR <= 1e-3
    ks = np.where(small, 1.0, k)
    kRs = ks * R
    s = np.exp(-SPREADING * ks * ks * dr * dr)
    kR2 = kR * kR
    w3 = np.where(small,
                  (4 * math.pi / 3) * R ** 3 * (kR2 ** 3 * (-1.0 / 15120) + kR2 ** 2 * (1.0 / 280) + kR2 * (-1.0 / 10) + 1),
                  s * 4 * math.pi * (np.sin(kRs) - kRs * np.cos(kRs)) / ks ** 3)
    w2 = np.where(small,
                  4 * math.pi * R * R * (-kR2 ** 3 * (1.0 / 120 / 6 / 7) + kR2 ** 2 * (1.0 / 120) - kR2 * (1.0 / 6) + 1),
                  s * 4 * math.pi * R * np.sin(kRs) / ks)
    wv = np.where(small,
                  4 * math.pi * R ** 3 * (kR2 ** 2 * (1.0 / 5040 - 1.0 / 720) + kR2 * (1.0 / 24 - 1.0 / 120) - 1.0 / 3),
                  s * 4 * math.pi * (R * np.cos(kRs) - np.sin(kRs) / ks) / ks ** 2)
    return w3, w2, wv


TENSOR_SERIES_CUTOFF = 1.0
TENSOR_SERIES_TERMS = 12
TENSOR_COMPONENTS = ((0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0))  # xx, yy, zz, xy, yz, zx


def tensor_weights(grid, R):
    """The shell tensor weight of src/haskell/FMT.hs:82-88 (tshell), equal to the C++
    xx/yy/zz/xy/yz/zxshell_op of src/ReciprocalOperators.h:292-404:
      W_ij(k) = s 4 pi [((kR^2 - 3) sin kR/kR + 3 cos kR)/k^4 k_i k_j
                        + (sin kR/kR - kR sin kR/3 - cos kR)/k^2 delta_ij].
    The two brackets are A(kR) R^2 k^_i k^_j and -A(kR) R^2 delta_ij / 3 with
    A(x) = ((x^2 - 3) sin x/x + 3 cos x)/x^2, so W is traceless:
      W_ij = s 4 pi R^2 A(kR) (k^_i k^_j - delta_ij/3),
    and s 4 pi R^2 A(kR) = w2(k) - 3 w3(k)/R.  For kR < 1 A uses its series
    A(x) = sum_{m>=2} (-1)^(m+1) 4 m (m-1) x^(2m-2)/(2m+1)!  (closed form cancels);
    W = 0 at k = 0.  Returns the six components in TENSOR_COMPONENTS order."""
    k = grid.k
    dr = grid.dV ** (1.0 / 3.0)
    x = k * R
    small = x < TENSOR_SERIES_CUTOFF
    xb = np.where(small, 1.5, x)
    A_closed = ((xb * xb - 3) * np.sin(xb) / xb + 3 * np.cos(xb)) / (xb * xb)
    xs = np.where(small, x, 0.0)
    A_ser = np.zeros_like(xs)
    for m in range(2, TENSOR_SERIES_TERMS + 2):
        A_ser += (-1) ** (m + 1) * 4 * m * (m - 1) * xs ** (2 * m - 2) / math.factorial(2 * m + 1)
    A = np.where(small, A_ser, A_closed)
    amp = np.exp(-SPREADING * k * k * dr * dr) * 4 * math.pi * R * R * A
    ks = np.where(k > 0, k, 1.0)
    kh = (grid.kx / ks, grid.ky / ks, grid.kz / ks)
    out = []
    for i, j in TENSOR_COMPONENTS:
        out.append(np.where(k > 0, amp * (kh[i] * kh[j] - (1.0 / 3.0 if i == j else 0.0)), 0.0))
    return out


def fft(grid, f):
    """Forward transform with deft's normalisation: r2c times dV (src/Grid.cpp:470-478)."""
    return _rfftn(f) * grid.dV


def ifft(grid, F):
    """Inverse transform with deft's normalisation: c2r times 1/V (src/ReciprocalGrid.cpp:52-65)."""
    return _irfftn(F, grid.shape) / grid.dV


# ---------------------------------------------------------------------------
# White-Bear free-energy density and partials
# ---------------------------------------------------------------------------

def f3_and_derivative(n3):
    """f3(n) = (n + (1-n)^2 ln(1-n)) / (n^2 (1-n)^2) and df3/dn.

    Closed form as in src/haskell/WhiteBear.hs:50 (src/HardSpheres.cpp:153) for
    |n| >= 0.05; below that the Taylor series f3 = g/(1-n)^2 with
    g(n) = 3/2 - sum_{k>=3} 2 n^(k-2) / (k (k-1) (k-2)).
    """
    n = np.asarray(n3, dtype=np.float64)
    small = np.abs(n) < F3_SERIES_CUTOFF
    omn = 1.0 - n
    # closed form (guard the small branch against 0/0)
    nb = np.where(small, 0.5, n)
    omb = 1.0 - nb
    L = np.log1p(-nb)
    num = nb + omb * omb * L
    den = nb * nb * omb * omb
    f3_big = num / den
    dnum = nb - 2.0 * omb * L
    dden = 2.0 * nb * omb * (1.0 - 2.0 * nb)
    df3_big = dnum / den - num * dden / (den * den)
    # series
    ns = np.where(small, n, 0.0)
    g = np.full_like(ns, 1.5)
    gp = np.zeros_like(ns)
    p = np.ones_like(ns)  # n^(k-3)
    for kk in range(3, F3_SERIES_TERMS + 3):
        # term of g: 2 n^(kk-2) / (kk (kk-1) (kk-2)); of g': 2 (kk-2) n^(kk-3) / (...)
        c = 2.0 / (kk * (kk - 1) * (kk - 2))
        gp -= c * (kk - 2) * p
        p = p * ns
        g -= c * p
    oms = 1.0 - ns
    f3_small = g / (oms * oms)
    df3_small = gp / (oms * oms) + 2.0 * g / (oms * oms * oms)
    return np.where(small, f3_small, f3_big), np.where(small, df3_small, df3_big)


WB2_SERIES_CUTOFF = 0.05
WB2_SERIES_TERMS = 24


def markii_factors(n3, g3_scale=1.5):
    """White-Bear Mark II factors of src/haskell/WhiteBear.hs:140-159 (Hansen-Goos and
    Roth 2006, eqs. 16-17):
      Phi2 = (1 + n3^2 phi2(n3)/9) (n1 n2 - n1v.n2v)/(1-n3),
      Phi3 = (1 - 4 n3 phi3(n3)/9) (n2^3 - 3 n2 n2v.n2v)/(24 pi (1-n3)^2),
      phi2 = (6n - 3n^2 + 6(1-n) ln(1-n))/n^3, phi3 = (6n - 9n^2 + 6n^3 + 6(1-n)^2 ln(1-n))/(4 n^3).
    Returned as g2 = (1 + S2/9)/(1-n) and g3 = 1.5 (1 - 4 S3/9)/(1-n)^2 with their n-derivatives,
    so that Phi2 = g2 (n2^2 - n2v^2)/(4 pi R) and Phi3 = g3 (n2^3 - 3 n2 n2v^2)/(36 pi) as for
    White-Bear (g2 = 1/(1-n), g3 = f3).  S2 = n^2 phi2 and S3 = n phi3 use their power series
    below |n| = 0.05:  S2 = sum_{m>=3} 6 n^(m-1)/(m(m-1)),
    S3 = n - sum_{m>=4} 3 n^(m-2)/(m(m-1)(m-2)).
    g3_scale = 1 gives HardSpheresWBm2slow's Phi3 (src/HardSpheres.cpp:176-177), whose
    prefactor is 1/(36 pi) instead of 1/(24 pi)."""
    n = np.asarray(n3, dtype=np.float64)
    small = np.abs(n) < WB2_SERIES_CUTOFF
    nb = np.where(small, 0.5, n)
    L = np.log1p(-nb)
    N2 = 6 * nb - 3 * nb ** 2 + 6 * (1 - nb) * L
    S2b = N2 / nb
    S2pb = (-3 * nb ** 2 - 6 * nb - 6 * L) / nb ** 2
    N3 = 6 * nb - 9 * nb ** 2 + 6 * nb ** 3 + 6 * (1 - nb) ** 2 * L
    N3p = -12 * nb + 18 * nb ** 2 - 12 * (1 - nb) * L
    S3b = N3 / (4 * nb ** 2)
    S3pb = (N3p * nb - 2 * N3) / (4 * nb ** 3)
    ns = np.where(small, n, 0.0)
    S2s = np.zeros_like(ns)
    S2ps = np.zeros_like(ns)
    S3s = ns.copy()
    S3ps = np.ones_like(ns)
    for m in range(3, WB2_SERIES_TERMS + 3):
        S2s += 6 * ns ** (m - 1) / (m * (m - 1))
        S2ps += 6 * ns ** (m - 2) / m
    for m in range(4, WB2_SERIES_TERMS + 4):
        S3s -= 3 * ns ** (m - 2) / (m * (m - 1) * (m - 2))
        S3ps -= 3 * ns ** (m - 3) / (m * (m - 1))
    S2 = np.where(small, S2s, S2b)
    S2p = np.where(small, S2ps, S2pb)
    S3 = np.where(small, S3s, S3b)
    S3p = np.where(small, S3ps, S3pb)
    io = 1.0 / (1.0 - n)
    g2 = io * (1 + S2 / 9)
    g2p = io * io * (1 + S2 / 9) + io * S2p / 9
    g3 = g3_scale * (1 - 4 * S3 / 9) * io * io
    g3p = g3_scale * (-(4.0 / 9) * S3p * io * io + (1 - 4 * S3 / 9) * 2 * io ** 3)
    return g2, g2p, g3, g3p


def whitebear_phi(n3, n2, vx, vy, vz, R, want_partials=True, functional="wb"):
    """Phi = Phi1 + Phi2 + Phi3 of the no-tensor White-Bear functional and its partials.

    Phi1 = -n0 ln(1-n3); Phi2 = (n1 n2 - n1v.n2v)/(1-n3);
    Phi3 = f3(n3)/(36 pi) n2 (n2^2 - 3 n2v.n2v)   (src/haskell/WhiteBear.hs:47-60,
    src/HardSpheres.cpp:137-159).  Partials with respect to (n3, n2, n2v_x, n2v_y, n2v_z),
    n0, n1, n1v being folded in through n0 = n2/(4 pi R^2) etc. (src/haskell/WhiteBear.hs:114-121).
    Returns (phi1, phi2, phi3, partials or None).
    """
    fpR = 4 * math.pi * R
    fpR2 = 4 * math.pi * R * R
    omn = 1.0 - n3
    L = np.log1p(-n3)
    v2 = vx * vx + vy * vy + vz * vz
    if functional == "wb":
        f3, df3 = f3_and_derivative(n3)
        g2, g2p = 1.0 / omn, 1.0 / (omn * omn)
    elif functional == "wb2":
        g2, g2p, f3, df3 = markii_factors(n3)
    elif functional == "wb2slow":
        # HardSpheresWBm2slow (src/HardSpheres.cpp:161-185): Phi3 over 36 pi, not 24 pi
        g2, g2p, f3, df3 = markii_factors(n3, g3_scale=1.0)
    else:
        raise ValueError(functional)
    t3 = n2 * (n2 * n2 - 3.0 * v2)
    phi1 = -n2 * L / fpR2
    phi2 = (n2 * n2 - v2) * g2 / fpR
    phi3 = f3 * t3 / (36 * math.pi)
    if not want_partials:
        return phi1, phi2, phi3, None
    d3 = n2 / (fpR2 * omn) + (n2 * n2 - v2) * g2p / fpR + df3 * t3 / (36 * math.pi)
    d2 = -L / fpR2 + 2.0 * n2 * g2 / fpR + f3 * (3.0 * n2 * n2 - 3.0 * v2) / (36 * math.pi)
    cv = -2.0 * g2 / fpR - f3 * 6.0 * n2 / (36 * math.pi)
    return phi1, phi2, phi3, (d3, d2, cv * vx, cv * vy, cv * vz)


# ---------------------------------------------------------------------------
# Homogeneous path: free energy per volume, chemical potential, pressure
# ---------------------------------------------------------------------------

def whitebear_tensor_phi(n3, n2, vx, vy, vz, txx, tyy, tzz, txy, tyz, tzx, R, want_partials=True):
    """Tensor White-Bear (TensorWhiteBear, src/haskell/WhiteBear.hs:44-51,
    src/HardSpheres.cpp:30-67): Phi1 and Phi2 as White-Bear,
    Phi3t = f3(n3)/(36 pi) X,  X = n2^3 - 3 n2 n2v.n2v + 9 (n2v.T.n2v - tr(T^3)/2)
    with T the (traceless) shell tensor density.  Partials with respect to the 11 fields
    (n3, n2, n2v_x, n2v_y, n2v_z, T_xx, T_yy, T_zz, T_xy, T_yz, T_zx), each off-diagonal
    component standing for both of its symmetric entries."""
    fpR = 4 * math.pi * R
    fpR2 = 4 * math.pi * R * R
    omn = 1.0 - n3
    L = np.log1p(-n3)
    v2 = vx * vx + vy * vy + vz * vz
    f3, df3 = f3_and_derivative(n3)
    Tv = (txx * vx + txy * vy + tzx * vz, txy * vx + tyy * vy + tyz * vz, tzx * vx + tyz * vy + tzz * vz)
    vTv = vx * Tv[0] + vy * Tv[1] + vz * Tv[2]
    tr3 = (6 * txy * tyz * tzx + txx * (txx * txx + 3 * txy * txy + 3 * tzx * tzx)
           + tyy * (3 * txy * txy + tyy * tyy + 3 * tyz * tyz) + tzz * (3 * tzx * tzx + 3 * tyz * tyz + tzz * tzz))
    X = n2 * n2 * n2 - 3 * n2 * v2 + 9 * (vTv - tr3 / 2)
    phi1 = -n2 * L / fpR2
    phi2 = (n2 * n2 - v2) / (fpR * omn)
    phi3 = f3 * X / (36 * math.pi)
    if not want_partials:
        return phi1, phi2, phi3, None
    c = f3 / (36 * math.pi)
    d3 = n2 / (fpR2 * omn) + (n2 * n2 - v2) / (fpR * omn * omn) + df3 * X / (36 * math.pi)
    d2 = -L / fpR2 + 2.0 * n2 / (fpR * omn) + c * (3 * n2 * n2 - 3 * v2)
    dv = [-2.0 * v / (fpR * omn) + c * (-6 * n2 * v + 18 * t) for v, t in zip((vx, vy, vz), Tv)]
    sq = (txx * txx + txy * txy + tzx * tzx, txy * txy + tyy * tyy + tyz * tyz, tzx * tzx + tyz * tyz + tzz * tzz,
          txx * txy + txy * tyy + tzx * tyz, txy * tzx + tyy * tyz + tyz * tzz, txx * tzx + txy * tyz + tzx * tzz)
    dt = [c * (9 * vx * vx - 13.5 * sq[0]), c * (9 * vy * vy - 13.5 * sq[1]), c * (9 * vz * vz - 13.5 * sq[2]),
          c * (18 * vx * vy - 27 * sq[3]), c * (18 * vy * vz - 27 * sq[4]), c * (18 * vz * vx - 27 * sq[5])]
    return phi1, phi2, phi3, (d3, d2, *dv, *dt)


def homogeneous_whitebear(n, R, kT=1.0, functional="wb"):
    """Excess free energy per volume kT Phi(n) of a uniform fluid (k=0 weights, src/Functional.h:382-392)."""
    n3 = np.float64(4 * math.pi / 3 * R ** 3 * n)
    n2 = np.float64(4 * math.pi * R * R * n)
    z = np.float64(0.0)
    # tensor White-Bear reduces to White-Bear in bulk (T = 0, n2v = 0)
    fn = "wb" if functional == "wbt" else functional
    p1, p2, p3, _ = whitebear_phi(n3, n2, z, z, z, R, want_partials=False, functional=fn)
    return float(kT * (p1 + p2 + p3))


def homogeneous_whitebear_dn(n, R, kT=1.0, functional="wb"):
    """d(kT Phi)/dn of a uniform fluid, through the partials (chain rule over n3, n2)."""
    n3 = np.float64(4 * math.pi / 3 * R ** 3 * n)
    n2 = np.float64(4 * math.pi * R * R * n)
    z = np.float64(0.0)
    fn = "wb" if functional == "wbt" else functional
    _, _, _, d = whitebear_phi(n3, n2, z, z, z, R, functional=fn)
    return float(kT * (d[0] * (4 * math.pi / 3 * R ** 3) + d[1] * (4 * math.pi * R * R)))


def chemical_potential(n, R, kT=1.0, functional="wb"):
    """mu = d f/dn of ideal gas + White-Bear at bulk density n (src/equation-of-state.cpp:189-192)."""
    return float(kT * math.log(n / nQ(kT)) + homogeneous_whitebear_dn(n, R, kT, functional))


def pressure(n, R, kT=1.0, functional="wb"):
    """p = n df/dn - f (src/equation-of-state.cpp:164-170)."""
    f = kT * n * (math.log(n / nQ(kT)) - 1) + homogeneous_whitebear(n, R, kT, functional)
    return float(n * chemical_potential(n, R, kT, functional) - f)


def bulk_density(eta, R):
    return eta / (4 * math.pi / 3 * R ** 3)


# ---------------------------------------------------------------------------
# The fluid functional of Veff: energy, gradient, preconditioned gradient
# ---------------------------------------------------------------------------

class WhiteBearFluidVeff:
    """F[Veff] = kT int Phi + int kT n (ln(n/nQ) - 1) + int Vext n - mu int n, n = exp(-Veff/kT).

    Mirrors deft's generated ``WhiteBearFluidVeff`` (src/haskell/WhiteBear.hs:29-31,
    src/haskell/create_generators.py:21-22) and the old-API composition
    OfEffectivePotential(HardSpheresWBnotensor + IdealGas + ChemicalPotential)
    (papers/contact/figs/walls.cpp:86-89).  ``x`` is the Veff field (modified
    in place by the minimizer, like the reference's data vector).
    """

    def __init__(self, grid, R, kT, mu, vext, x, functional="wb"):
        self.grid = grid
        self.functional = functional
        self.R = float(R)
        self.kT = float(kT)
        self.mu = float(mu)
        self.vext = np.asarray(vext, dtype=np.float64).reshape(grid.shape)
        # +inf: hard wall (n = 0); NaN: constrained point of the old API's
        # constrain(mask, f) (src/Functional.cpp:759-838): density and energy as
        # usual without external potential, gradient and pgrad zero
        self.finite = np.isfinite(self.vext)
        self.present = ~np.isinf(self.vext)
        self.vext_f = np.where(self.finite, self.vext, 0.0)
        self.x = np.array(x, dtype=np.float64).reshape(grid.shape)
        w3, w2, wv = weights(grid, self.R)
        self.W = (w3, w2, 1j * wv * grid.kx, 1j * wv * grid.ky, 1j * wv * grid.kz)
        self.parity = (1.0, 1.0, -1.0, -1.0, -1.0)  # w(-k) = +/- w(k) (src/Functional.h:407-417)
        if functional == "wbt":
            self.W = self.W + tuple(tensor_weights(grid, self.R))
            self.parity = self.parity + (1.0,) * 6
        self.last_terms = {}
        self.n_energy = 0
        self.n_grad = 0

    def density(self, x=None):
        x = self.x if x is None else x
        return np.where(self.present, np.exp(-x / self.kT), 0.0)

    def _phi(self, nal, partials):
        if self.functional == "wbt":
            return whitebear_tensor_phi(*nal, self.R, want_partials=partials)
        return whitebear_phi(*nal, self.R, want_partials=partials, functional=self.functional)

    def weighted_densities(self, n):
        nk = _rfftn(n)
        return [_irfftn(nk * W, self.grid.shape) for W in self.W]

    def _local_terms(self, x, n):
        kT = self.kT
        ideal = np.where(self.present, -n * (x + kT * math.log(nQ(kT)) + kT), 0.0)
        ext = self.vext_f * n
        mun = -self.mu * n
        return ideal, ext, mun

    def energy(self, x=None):
        """Grand free energy (sum times dV); records each term like ``printme``."""
        x = self.x if x is None else x
        n = self.density(x)
        nal = self.weighted_densities(n)
        p1, p2, p3, _ = self._phi(nal, False)
        ideal, ext, mun = self._local_terms(x, n)
        dV = self.grid.dV
        kT = self.kT
        t = {
            "kTphi1": kT * float(np.sum(p1)) * dV,
            "kTphi2": kT * float(np.sum(p2)) * dV,
            "kTphi3": kT * float(np.sum(p3)) * dV,
            "ideal": float(np.sum(ideal)) * dV,
            "int_V_n": float(np.sum(ext)) * dV,
            "minus_mu_N": float(np.sum(mun)) * dV,
        }
        self.last_terms = t
        self.n_energy += 1
        return (t["kTphi1"] + t["kTphi2"] + t["kTphi3"] + t["ideal"] + t["int_V_n"] + t["minus_mu_N"])

    def dF_dn(self, x=None):
        """Gradient of F with respect to the density values (including dV)."""
        x = self.x if x is None else x
        n = self.density(x)
        nal = self.weighted_densities(n)
        _, _, _, P = self._phi(nal, True)
        acc = None
        for Pa, W, par in zip(P, self.W, self.parity):
            term = _rfftn(Pa) * (W * par)
            acc = term if acc is None else acc + term
        conv = _irfftn(acc, self.grid.shape)
        kT = self.kT
        local = -x - kT * math.log(nQ(kT)) + self.vext_f - self.mu
        return np.where(self.finite, (kT * conv + local) * self.grid.dV, 0.0), n

    def grad(self, x=None):
        """dF/dVeff = dF/dn * (-n/kT) (src/EffectivePotentialToDensity.cpp:62-79)."""
        g, n = self.dF_dn(x)
        self.n_grad += 1
        return -g * n / self.kT

    def energy_grad_and_precond(self, x=None):
        """(energy, grad, precond) with precond = grad / n = -dF/dn / kT
        (src/haskell/NewCode.hs:229-242,306-319; zero on hard-wall points)."""
        E = self.energy(x)
        g, n = self.dF_dn(x)
        self.n_grad += 1
        return E, -g * n / self.kT, -g / self.kT

    def have_preconditioner(self):
        return True

    def grad_and_pgrad(self, x=None):
        """Old-API integralgrad with pgrad (src/EffectivePotentialToDensity.cpp:62-71):
        grad = -(n/kT) dF/dn, pgrad = (dF/dn)/(Veff - kT); zero on hard-wall and
        constrained points (the constraint masks both, src/Functional.cpp:809-824)."""
        x = self.x if x is None else x
        g, n = self.dF_dn(x)
        self.n_grad += 1
        with np.errstate(divide="ignore", invalid="ignore"):
            pg = np.where(self.finite, g / (x - self.kT), 0.0)
        return -g * n / self.kT, pg


# ---------------------------------------------------------------------------
# Structure-factor preconditioner (no counterpart in the reference; the
# product's DEFT_PGRAD_STRUCTURE_FACTOR).  In a uniform fluid at n_b the
# Hessian of F over Veff is (n_b/kT) (kT/n_b + H_ex(k)) (n_b/kT) per
# Fourier mode, H_ex(k) = kT [P33 w3^2 + 2 P32 w3 w2 + P22 w2^2 + cv k^2 w3^2]
# with the second derivatives P_ab of Phi at the bulk weighted densities
# and cv the n2v.n2v coefficient (the vector weight is -i k w3).  Its inverse
# is D F^-1 S(k) F D with D = n^-1/2 and S(k) = (kT/n_b)/(kT/n_b + H_ex(k)),
# the fluid's structure factor; with S = 1 it is the new API's grad/n.  On a
# non-uniform fluid D takes the local density.  Points below 1e-3 n_b (the
# dilute side of walls and soft repulsions) keep grad/n, where S ~ 1 and
# 1/sqrt(n) would be large; hard-wall and constrained points get zero.  The
# operator is symmetric positive definite, so Polak-Ribiere CG stays valid.
# ---------------------------------------------------------------------------

SK_FLOOR = 1e-3   # fraction of n_b below which a point keeps grad/n
SK_CLAMP = 1e3    # S(k) is kept in [1/SK_CLAMP, SK_CLAMP]


def phi_bulk_hessian(nb, R, functional="wb"):
    """(P33, P32, P22, cv) at the bulk weighted densities n3 = 4/3 pi R^3 nb,
    n2 = 4 pi R^2 nb, n2v = 0, by central differences (relative step 1e-5) of
    the partials; cv = d(dPhi/dn2v_x)/dn2v_x."""
    n3 = 4 * math.pi / 3 * R ** 3 * nb
    n2 = 4 * math.pi * R * R * nb
    fn = "wb" if functional == "wbt" else functional

    def part(a3, a2, vx):
        _, _, _, d = whitebear_phi(np.float64(a3), np.float64(a2), np.float64(vx), np.float64(0.0),
                                   np.float64(0.0), R, functional=fn)
        return [float(v) for v in d]

    h3, h2, hv = 1e-5 * n3, 1e-5 * n2, 1e-5 * n2
    p, m = part(n3 + h3, n2, 0.0), part(n3 - h3, n2, 0.0)
    P33 = (p[0] - m[0]) / (2 * h3)
    P32a = (p[1] - m[1]) / (2 * h3)
    p, m = part(n3, n2 + h2, 0.0), part(n3, n2 - h2, 0.0)
    P22 = (p[1] - m[1]) / (2 * h2)
    P32b = (p[0] - m[0]) / (2 * h2)
    p, m = part(n3, n2, hv), part(n3, n2, -hv)
    cv = (p[2] - m[2]) / (2 * hv)
    return P33, 0.5 * (P32a + P32b), P22, cv


def structure_factor(grid, nb, R, kT=1.0, functional="wb"):
    """S(k) of the uniform fluid on the half spectrum, with the grid's smeared weights."""
    w3, w2, _ = weights(grid, R)
    P33, P32, P22, cv = phi_bulk_hessian(nb, R, functional)
    hex_ = kT * (P33 * w3 * w3 + 2 * P32 * w3 * w2 + P22 * w2 * w2 + cv * grid.k ** 2 * w3 * w3)
    S = (kT / nb) / (kT / nb + hex_)
    S = np.where(np.isfinite(S) & (S > 0), S, SK_CLAMP)
    return np.clip(S, 1.0 / SK_CLAMP, SK_CLAMP)


def sk_precondition(fluid, grad, S, nb, x=None):
    """The structure-factor preconditioned gradient of ``grad`` (dF/dVeff) at ``x``."""
    x = fluid.x if x is None else x
    n = fluid.density(x)
    dense = fluid.finite & (n >= SK_FLOOR * nb)
    rn = np.where(dense, 1.0 / np.sqrt(np.where(dense, n, 1.0)), 0.0)
    conv = _irfftn(S * _rfftn(grad * rn), fluid.grid.shape)
    with np.errstate(divide="ignore", invalid="ignore"):
        dilute = np.where(fluid.finite & ~dense & (n > 0), grad / np.where(n > 0, n, 1.0), 0.0)
    return np.where(dense, rn * conv, dilute)


class StructureFactorFluid:
    """A fluid whose energy_grad_and_precond returns the structure-factor
    preconditioned gradient (for the oracle's Minimize)."""

    def __init__(self, fluid, nb):
        self.fluid = fluid
        self.nb = float(nb)
        self.S = structure_factor(fluid.grid, nb, fluid.R, fluid.kT, fluid.functional)

    def __getattr__(self, name):
        return getattr(self.fluid, name)

    @property
    def x(self):
        return self.fluid.x

    @x.setter
    def x(self, v):
        self.fluid.x = v

    def energy_grad_and_precond(self, x=None):
        E, g, _ = self.fluid.energy_grad_and_precond(x)
        return E, g, sk_precondition(self.fluid, g, self.S, self.nb, x)


# ---------------------------------------------------------------------------
# New-API minimizer (src/new/Minimize.cpp:9-313, src/new/Minimize.h:13-199)
# ---------------------------------------------------------------------------

def _better(a, b):
    """a is better than b (src/new/Minimize.cpp:5-7): a < b or b is NaN."""
    return a < b or b != b


class Minimize:
    """Polak-Ribiere CG with deft's quadratic line search and cached energies.

    ``f`` provides ``x`` (modified in place), ``energy()``, ``grad()``,
    ``energy_grad_and_precond()`` and ``have_preconditioner()``.
    """

    def __init__(self, f):
        self.f = f
        self.iter = 0
        self.maxiter = 10000000
        self.miniter = 6
        self.precision = 0.0
        self.relative_precision = 1e-15
        self.use_preconditioning = False
        self.step = 0.1
        self.orig_stepsize = 0.1
        self.oldgradsqr = 1.0
        self.deltaE = 0.0
        self.dEdn = 0.0
        self.log_dEdn_ratio_average = 0.0
        self.error_estimate = 0.0
        self.direction = None
        self.oldgrad = None
        self._E = None
        self._g = None
        self._pg = None
        self.num_energy_calcs = 0
        self.num_grad_calcs = 0
        self.log = []
        self.stop_reason = None

    # configuration (src/new/Minimize.h:56-89)
    def set_precision(self, p):
        self.precision = p

    def set_relative_precision(self, p):
        self.relative_precision = p

    def set_maxiter(self, m):
        self.maxiter = m

    def set_miniter(self, m):
        self.miniter = m

    def precondition(self, u):
        self.use_preconditioning = bool(u)

    def invalidate_cache(self):
        self._E = None
        self._g = None
        self._pg = None

    def energy(self):
        if self._E is None:
            self._E = self.f.energy()
            self.num_energy_calcs += 1
        return self._E

    def grad(self):
        if self._g is None:
            self._g = self.f.grad()
            self.num_grad_calcs += 1
        return self._g

    def pgrad(self):
        if self._pg is None:
            if self.use_preconditioning and self.f.have_preconditioner():
                self.invalidate_cache()
                E, g, pg = self.f.energy_grad_and_precond()
                self.num_energy_calcs += 1
                self.num_grad_calcs += 1
                self._E, self._g, self._pg = E, g, pg
            else:
                self.grad()
                self._pg = self._g
        return self._pg

    def _move(self, s):
        """x += s*direction (invalidates nothing by itself, as in the reference)."""
        self.f.x += s * self.direction

    def improve_energy(self):
        self.iter += 1
        if self.iter >= self.maxiter:
            self.stop_reason = "maxiter"
            return False
        E0 = self.energy()
        old_deltaE = self.deltaE
        if E0 != E0:
            self.stop_reason = "nan-energy"
            return False
        pg = self.pgrad()
        g = self.grad()
        self.invalidate_cache()
        if self.oldgrad is None:
            beta = float(np.vdot(pg, g)) / self.oldgradsqr
        else:
            beta = float(np.vdot(pg, g - self.oldgrad)) / self.oldgradsqr
        self.oldgrad = g
        if beta < 0 or beta != beta or self.oldgradsqr == 0:
            beta = 0.0
        self.oldgradsqr = float(np.vdot(pg, g))
        if self.direction is None:
            self.direction = -pg
        else:
            self.direction = -pg + beta * self.direction
        gdotd = float(np.vdot(g, self.direction))
        if gdotd > 0:
            self.direction = -g
            gdotd = float(np.vdot(g, self.direction))
        del g, pg

        slope = -gdotd
        if E0 != E0:
            self.stop_reason = "nan-energy"
            return False
        if E0 >= np.finfo(np.float64).max or E0 <= -np.finfo(np.float64).max:
            self.stop_reason = "infinite-energy"
            return False
        if slope != slope:
            self.stop_reason = "nan-slope"
            return False
        if slope == 0:
            self.stop_reason = "zero-slope"
            return False
        if self.step == self.orig_stepsize and abs(E0) > 1e-200 and abs(slope) > 1e-100:
            self.step = 0.001 * abs(E0) / slope
        if slope * self.step < 0:
            self.step *= -1
        step1 = self.step
        self._move(step1)
        Etried = -137.0
        while True:
            step1 *= 0.5
            self._move(-step1)
            self.invalidate_cache()
            if self.energy() == Etried:
                break
            Etried = self.energy()
            if not _better(E0, self.energy()):
                break
        E1 = self.energy()
        curvature = 2.0 * (E1 - E0 + step1 * slope) / (step1 * step1)
        step2 = slope / curvature if curvature != 0 else math.copysign(math.inf, slope) * math.copysign(1.0, curvature)
        if curvature <= 0.0:
            if _better(E0, E1):
                self.invalidate_cache()
                self._move(-step1)
                self.stop_reason = "negative-curvature"
                return False
            Ebest = E0
            while self.energy() <= Ebest:
                Ebest = self.energy()
                self._move(step1)
                self.invalidate_cache()
                step1 *= 2
            step1 *= 0.5
            self._move(-step1)
            self.step = step1
            self.invalidate_cache()
        elif E1 == E0:
            while True:
                self.invalidate_cache()
                self._move(-step1)
                step1 *= 2
                self._move(step1)
                if self.energy() != E0:
                    break
            if self.energy() > E0:
                self._move(-step1)
                self.invalidate_cache()
                step1 = 0
            elif self.energy() != self.energy():
                self._move(-step1)
                self.invalidate_cache()
                step1 = 0
        else:
            self.step = step2
            self.invalidate_cache()
            self._move(step2 - step1)
            if _better(E1, self.energy()) and _better(E1, E0):
                self.invalidate_cache()
                self._move(-(step2 - step1))
                self.step = step1
            elif self.energy() > E0:
                self.invalidate_cache()
                self._move(-step2)
                self.stop_reason = "energy-rose"
                return False

        newE = self.energy()
        self.deltaE = newE - E0
        w = 0.01
        self.dEdn = (1 - w) * self.dEdn + w * self.deltaE
        self.dEdn = max(self.deltaE, old_deltaE)
        if self.deltaE and old_deltaE:
            new_log = math.log(abs(self.deltaE / old_deltaE))
        else:
            new_log = 0.0
        self.log_dEdn_ratio_average = (1 - w) * self.log_dEdn_ratio_average + w * new_log
        ratio_avg = math.exp(self.log_dEdn_ratio_average)
        if self.log_dEdn_ratio_average != 0:
            error_guess = abs(max(self.dEdn, self.deltaE) / self.log_dEdn_ratio_average)
        else:
            error_guess = math.inf
        if ratio_avg >= 1:
            error_guess = self.precision + abs(newE)
        self.error_estimate = 2 * error_guess
        self.log.append((self.iter, newE, self.deltaE, self.error_estimate, self.step))
        if self.deltaE == 0 and old_deltaE == 0:
            self.stop_reason = "no-change"
            return self.iter < self.miniter
        if not (self.error_estimate < self.precision) and not (self.error_estimate < self.relative_precision * abs(newE)):
            return True
        self.stop_reason = "converged"
        return self.iter < self.miniter


# ---------------------------------------------------------------------------
# Old-API minimizers (src/ConjugateGradient.cpp, src/QuadraticLineMinimizer.cpp,
# src/Precision.cpp, src/Minimizer.cpp).  ``f`` provides ``x`` (modified in
# place), ``energy()``, ``grad()`` and ``grad_and_pgrad()``.  As in the
# reference, energy() is evaluated afresh on every call.
# ---------------------------------------------------------------------------

class OldConjugateGradient:
    """ConjugateGradientType (src/ConjugateGradient.cpp:59-108) or, with
    ``preconditioned``, PreconditionedConjugateGradientType (:115-160), each
    with QuadraticLineMinimizer and the remembered step (default 10,
    src/LineMinimizer.h:18-21)."""

    def __init__(self, f, preconditioned=False, stepsize=10.0):
        self.f = f
        self.preconditioned = preconditioned
        self.step = float(stepsize)
        self.orig_step = float(stepsize)
        self.iter = 0
        self.direction = np.zeros_like(f.x)
        self.oldgrad = np.zeros_like(f.x)
        self.oldgradsqr = 0.0
        self.linmin_reason = None
        self.log = []

    def energy(self):
        return self.f.energy()

    def minimize(self):
        """minimize(f, gd, x): src/ConjugateGradient.cpp:34-44."""
        self.step = self.orig_step
        self.iter = 0
        self.direction = np.zeros_like(self.f.x)
        self.oldgrad = np.zeros_like(self.f.x)
        self.oldgradsqr = 0.0

    def improve_energy(self):
        self.iter += 1
        E0 = self.energy()
        if E0 != E0:
            return False
        if not self.preconditioned:
            g = -self.f.grad()
            beta = float(np.vdot(g, g - self.oldgrad)) / self.oldgradsqr if self.oldgradsqr else math.nan
            self.oldgrad = g
            if beta < 0 or beta != beta or self.oldgradsqr == 0:
                beta = 0.0
            self.oldgradsqr = float(np.vdot(self.oldgrad, self.oldgrad))
            self.direction = g + beta * self.direction
            gdotd = float(np.vdot(self.oldgrad, self.direction))
            if gdotd < 0:
                self.direction = self.oldgrad.copy()
                gdotd = float(np.vdot(self.oldgrad, self.direction))
        else:
            grad, pgrad = self.f.grad_and_pgrad()
            beta = (float(np.vdot(-pgrad, -grad - self.oldgrad)) / self.oldgradsqr if self.oldgradsqr
                    else math.nan)
            self.oldgrad = -grad
            if beta < 0 or beta != beta or self.oldgradsqr == 0:
                beta = 0.0
            self.oldgradsqr = float(np.vdot(-pgrad, self.oldgrad))
            self.direction = -pgrad + beta * self.direction
            gdotd = float(np.vdot(self.oldgrad, self.direction))
        self.linmin_reason = quadratic_line_minimize(self, -gdotd)
        E = self.energy()
        self.log.append((self.iter, E, self.step))
        if self.preconditioned:
            return E < E0 or beta != 0
        return E < E0


def quadratic_line_minimize(cg, slope):
    """QuadraticLineMinimizerType::improve_energy (src/QuadraticLineMinimizer.cpp:27-198)
    along cg.direction with slope = grad.direction, updating cg.step; returns why
    it returned early (None on the ordinary paths)."""
    f, d = cg.f, cg.direction
    E0 = f.energy()
    if E0 != E0:
        return "nan-energy"
    if E0 > 1e307 or E0 < -1e307:
        return "infinite-energy"
    if slope != slope:
        return "nan-slope"
    if slope == 0:
        return "zero-slope"
    if slope * cg.step > 0:
        cg.step *= -1
    step1 = cg.step
    f.x += step1 * d
    Etried = -137.0
    while True:
        step1 *= 0.5
        f.x -= step1 * d
        if f.energy() == Etried:
            break
        Etried = f.energy()
        if not _better(E0, f.energy()):
            break
    E1 = f.energy()
    curvature = 2.0 * (E1 - E0 - step1 * slope) / (step1 * step1)
    step2 = -slope / curvature if curvature != 0 else math.inf
    if curvature <= 0.0:
        if _better(E0, E1):
            f.x -= step1 * d
            return "negative-curvature"
        Ebest = E0
        while f.energy() <= Ebest:
            Ebest = f.energy()
            f.x += step1 * d
            step1 *= 2
        step1 *= 0.5
        f.x -= step1 * d
        cg.step = step1
    elif E1 == E0:
        while True:
            f.x -= step1 * d
            step1 *= 2
            f.x += step1 * d
            if f.energy() != E0:
                break
        if f.energy() > E0:
            f.x -= step1 * d
        elif f.energy() != f.energy():
            f.x -= step1 * d
    else:
        cg.step = step2
        f.x += (step2 - step1) * d
        E2 = f.energy()
        if _better(E1, E2) and _better(E1, E0):
            f.x -= (step2 - step1) * d
        elif E2 > E0:
            f.x -= step2 * d
            return "energy-rose"
    return None


class OldPrecision:
    """PrecisionMinimizer (src/Precision.cpp:23-156) around a minimizer, with
    MaxIterType (src/Minimizer.cpp) outside it when ``maxiter`` is given."""

    def __init__(self, precision, m, maxiter=None):
        self.m = m
        self.criterion = float(precision)
        self.error_estimate = max(1.0, 2 * precision)
        self.dEdn = 0.0
        self.log_avg = 0.0
        self.deltaE = 0.0
        self.w = 0.1
        self.maxiter = maxiter
        self.iter = 0
        self.stop_reason = None

    def energy(self):
        return self.m.energy()

    def improve_energy(self):
        if self.maxiter is not None and self.iter > self.maxiter:
            self.stop_reason = "maxiter"
            return False
        self.iter += 1
        old_deltaE = self.deltaE
        old_energy = self.m.energy()
        not_done = self.m.improve_energy()
        new_energy = self.m.energy()
        self.deltaE = old_energy - new_energy
        w = self.w
        self.dEdn = (1 - w) * self.dEdn + w * self.deltaE
        self.dEdn = max(self.deltaE, old_deltaE)
        new_log = math.log(abs(self.deltaE / old_deltaE)) if (self.deltaE and old_deltaE) else 0.0
        self.log_avg = (1 - w) * self.log_avg + w * new_log
        ratio_avg = math.exp(self.log_avg)
        error_guess = abs(max(self.dEdn, self.deltaE) / self.log_avg) if self.log_avg else math.inf
        if ratio_avg >= 1:
            error_guess = self.criterion + abs(new_energy)
        self.error_estimate = 2 * error_guess
        if self.deltaE == 0 and old_deltaE == 0:
            self.stop_reason = "no-change"
            return False
        if not_done and not (self.error_estimate < self.criterion):
            return True
        self.stop_reason = "converged" if not_done else "not-improved"
        return False


# ---------------------------------------------------------------------------
# Problem builders for the configurations this path is measured on
# ---------------------------------------------------------------------------

def minimum_image(grid, center):
    """Distance of each grid point from ``center`` under periodic minimum image (orthorhombic cells)."""
    pos = grid.positions() - np.asarray(center, dtype=np.float64)
    L = np.diag(grid.cell)
    pos = pos - L * np.round(pos / L)
    return np.sqrt(np.sum(pos * pos, axis=-1))


def planar_wall_vext(grid, R):
    """V = inf for x < R (first cell axis), else 0."""
    x = grid.positions()[..., 0]
    return np.where(x < R, np.inf, 0.0)


def spherical_solute_vext(grid, radius, center=(0.0, 0.0, 0.0)):
    """V = inf for sphere centres within ``radius`` of ``center`` (periodic), else 0
    (papers/contact/figs/inner-sphere.cpp:29-37)."""
    r = minimum_image(grid, center)
    return np.where(r < radius, np.inf, 0.0)


def make_fluid(grid, eta, R=1.0, kT=1.0, vext=None, functional="wb"):
    """Fluid at bulk packing fraction eta, mu from the bulk equation of state, Veff = -kT ln n_b."""
    nb = bulk_density(eta, R)
    mu = chemical_potential(nb, R, kT, functional)
    if vext is None:
        vext = np.zeros(grid.shape)
    x = np.full(grid.shape, -kT * math.log(nb))
    return WhiteBearFluidVeff(grid, R, kT, mu, vext, x, functional)


def minimize(f, relative_precision=1e-12, maxiter=100000, precondition=True, miniter=6, callback=None):
    m = Minimize(f)
    m.set_relative_precision(relative_precision)
    m.set_maxiter(maxiter)
    m.set_miniter(miniter)
    m.precondition(precondition)
    while m.improve_energy():
        if callback is not None:
            callback(m)
    return m


# ---------------------------------------------------------------------------
# Contact-density estimators of the contact paper (post-processing of a
# converged density; src/haskell/WhiteBear.hs:68-138, src/ContactDensity.cpp:255-400)
# ---------------------------------------------------------------------------

def _independent_partials(n3, n2, vx, vy, vz, R, functional):
    """dPhi/dn0, dn1, dn2, dn3 and the vector partials dPhi/dn1v, dPhi/dn2v with
    n0, n1, n2, n3, n1v, n2v independent (the Haskell `derive a 1 phit` of
    src/haskell/WhiteBear.hs:86-87), for Phi2 = g2 (n1 n2 - n1v.n2v) and
    Phi3 = g3 n2 (n2^2 - 3 n2v.n2v)/(36 pi): White-Bear g2 = 1/(1-n3), g3 = f3;
    Mark II its factors."""
    omn = 1.0 - n3
    L = np.log1p(-n3)
    if functional == "wb":
        g3, g3p = f3_and_derivative(n3)
        g2, g2p = 1.0 / omn, 1.0 / (omn * omn)
    elif functional in ("wb2", "wb2slow"):
        g2, g2p, g3, g3p = markii_factors(n3, g3_scale=1.0 if functional == "wb2slow" else 1.5)
    else:
        raise ValueError(functional)
    n0 = n2 / (4 * math.pi * R * R)
    n1 = n2 / (4 * math.pi * R)
    v = (vx, vy, vz)
    v1 = [c / (4 * math.pi * R) for c in v]
    v2 = vx * vx + vy * vy + vz * vz
    vtt = n2 * (n2 * n2 - 3 * v2)
    c36 = 1.0 / (36 * math.pi)
    dn0 = -L
    dn1 = g2 * n2
    dn2 = g2 * n1 + g3 * (3 * n2 * n2 - 3 * v2) * c36
    dn3 = n0 / omn + g2p * (n1 * n2 - (v1[0] * vx + v1[1] * vy + v1[2] * vz)) + g3p * vtt * c36
    dn1v = [-g2 * c for c in v]
    dn2v = [-g2 * a - g3 * 6 * n2 * c * c36 for a, c in zip(v1, v)]
    return dn0, dn1, dn2, dn3, dn1v, dn2v


def contact_values(fluid, functional=None):
    """The six fields the contact-paper drivers write next to the density
    (papers/contact/figs/inner-sphere.cpp:220-233): n0, gSigmaS, YuWu_S, nA,
    gSigmaA and the Gross estimate, in that order, then gSigmaS in the C++
    by-hand form (gSigmaScpp / gSigmaSm2cpp, src/ContactDensity.cpp:231-255,
    269-355): dA/dR = n2 (dn3 - dn1/(4 pi R^2) - dn0/(2 pi R^3))
    - n2' (dn2 + dn1/(4 pi R) + dn0/(4 pi R^2)) - sum_j n2v_j (c2 n2' + c1 n2'/(4 pi R)
    + c1 n2v_j/(4 pi R^2)) with n2' = -2 n2/R for every component and c2, c1 the
    vector partials over n2v.

    gSigmaS = dA/dR / (kT n0^2 4 pi (2R)^2) with the local
      dA/dR/kT = dn3 n2 + 2 n2 dn2/R + n0 dn1 + (2/R) dn2v.n2v + 3 dn1v.n2v/(4 pi R^2)
      (src/haskell/WhiteBear.hs:76-87 with n2' = -2 n2/R, n2v' = -2 n2v/R; the dn0 term cancels);
    gSigmaA = [shell(A) - shell'(B) + vshell'.(C) + vshell.(D)] / shell_2R(n)
      with A = dn3 - dn1/(4 pi R^2) - dn0/(2 pi R^3), B = dn2 + dn1/(4 pi R) + dn0/(4 pi R^2),
      C = dn2v + dn1v/(4 pi R), D = dn1v/(4 pi R^2)
      (src/haskell/WhiteBear.hs:119-138; shell' = s 4 pi (-sin kR/k - R cos kR),
      vshell'_j = i s 4 pi R sin(kR)/k k_j, src/haskell/FMT.hs:66-81);
    YuWu_S = gHS zeta^2 with gHS of Fu and Wu 2005 eq. 13 (src/ContactDensity.cpp:362-400);
    Gross = (1 - eta/2)/(1 - eta)^3, eta = step_2R(n)/8 (src/haskell/WhiteBear.hs:140-142);
    nA = shell_2R(n)/(4 pi (2R)^2).
    """
    functional = fluid.functional if functional is None else functional
    g = fluid.grid
    R = fluid.R
    n = fluid.density()
    n3, n2, vx, vy, vz = fluid.weighted_densities(n)[:5]
    nk = _rfftn(n)
    w3d, w2d, _ = weights(g, 2 * R)
    shell2 = _irfftn(nk * w2d, g.shape)
    step2 = _irfftn(nk * w3d, g.shape)
    dn0, dn1, dn2, dn3, dn1v, dn2v = _independent_partials(n3, n2, vx, vy, vz, R, functional)
    n0 = n2 / (4 * math.pi * R * R)
    fpR, fpR2 = 4 * math.pi * R, 4 * math.pi * R * R
    v = (vx, vy, vz)
    dAdR = (dn3 * n2 + 2 * n2 * dn2 / R + n0 * dn1
            + (2 / R) * sum(a * c for a, c in zip(dn2v, v)) + 3 * sum(a * c for a, c in zip(dn1v, v)) / (fpR * R))
    gS = dAdR / (n0 * n0 * 4 * math.pi * (2 * R) ** 2)
    # kernels of the A-estimator
    k = g.k
    dr = g.dV ** (1.0 / 3.0)
    s = np.exp(-SPREADING * k * k * dr * dr)
    kR = k * R
    ks = np.where(k > 0, k, 1.0)
    w3, w2, wv = weights(g, R)
    shellp = np.where(k > 0, s * 4 * math.pi * (-np.sin(kR) / ks - R * np.cos(kR)), -8 * math.pi * R)
    vsp = np.where(k > 0, s * 4 * math.pi * R * np.sin(kR) / ks, 0.0)
    kv = (g.kx, g.ky, g.kz)
    A = dn3 - dn1 / fpR2 - dn0 / (2 * math.pi * R ** 3)
    B = dn2 + dn1 / fpR + dn0 / fpR2
    acc = _rfftn(A) * w2 - _rfftn(B) * shellp
    for j in range(3):
        C = dn2v[j] + dn1v[j] / fpR
        D = dn1v[j] / fpR2
        acc = acc + _rfftn(C) * (1j * vsp * kv[j]) + _rfftn(D) * (1j * wv * kv[j])
    gA = _irfftn(acc, g.shape) / shell2
    invd = 1.0 / (1.0 - n3)
    zeta2 = (R / 3.0) * n2
    zeta = 1.0 - (vx * vx + vy * vy + vz * vz) / (n2 * n2)
    ghs = invd * (1.0 + 0.5 * (invd * zeta2) * zeta * (3.0 + invd * zeta2))
    yuwu = ghs * zeta * zeta
    eta = step2 / 8.0
    gross = (1 - 0.5 * eta) / (1 - eta) ** 3
    nA = shell2 / (4 * math.pi * (2 * R) ** 2)
    n2p = -2 * n2 / R
    # dn1v = c1 n2v, dn2v = c2 n2v with the scalar factors
    if functional == "wb":
        g2 = 1.0 / (1.0 - n3)
        g3 = f3_and_derivative(n3)[0]
    else:
        g2, _, g3, _ = markii_factors(n3, g3_scale=1.0 if functional == "wb2slow" else 1.5)
    c1 = -g2
    c2 = -g2 / fpR - g3 * 6 * n2 / (36 * math.pi)
    cpp = n2 * (dn3 - dn1 / fpR2 - dn0 / (2 * math.pi * R ** 3)) - n2p * (dn2 + dn1 / fpR + dn0 / fpR2)
    for c in v:
        cpp = cpp - c * (c2 * n2p + c1 * n2p / fpR + c1 * c / fpR2)
    gS_cpp = cpp / (4 * math.pi * (2 * R) ** 2) / (n0 * n0)
    return n0, gS, yuwu, nA, gA, gross, gS_cpp
