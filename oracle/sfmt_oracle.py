"""FP64 CPU oracle for soft fundamental-measure theory (SFMT).

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/`` may
import it.  The product path (``deft_amd``) never calls it.

It restates, in numpy/scipy FP64, deft's soft-FMT functional
(src/haskell/SFMT.hs, the source of the generated SFMTFluid / SFMTFluidVeff /
HomogeneousSFMTFluid classes of src/haskell/create_generators.py:13-27):

* the softness parameters: alpha = sigma (2/(1 + sqrt(ln 2 / (epsilon/kT))))^(1/6)
  and the user-set Xi                                          (src/haskell/SFMT.hs:43-56)
* the Gaussian-smeared weights w3, w2, w1, w0, w2v, w1v and the
  transverse tensor weight, with Re erf of a complex argument   (src/haskell/SFMT.hs:103-133)
  - Re erf is Faddeeva::erf(z).real() in the generated code     (src/haskell/Expression.hs:316;
    the Faddeeva package, src/Faddeeva.cpp).  scipy.special.erf of a complex
    argument is the same package; tests compare it with the reference's own
    src/Faddeeva.cpp compiled into oracle/_ref.
  - the k = 0 values are the limits the generator's setZero takes
    (series of Re erf, src/haskell/Expression.hs:357-363)
  - the tensor components are w2t (1 - 3 k_i^2/k^2) on the diagonal and
    +3 w2t k_i k_j/k^2 off it, exactly as src/haskell/SFMT.hs:64-69 writes them
* Phi1 = -n0 ln(1-n3), Phi2 = (n2 n1 - n1v.n2v)/(1-n3),
  Phi3 = (n2^3 - 3 n2 n2v.n2v + 9/2 (n2v.T.n2v - tr T^3))/(24 pi (1-n3)^2)
  (per kT)                                                       (src/haskell/SFMT.hs:13-27,71-91)
* the homogeneous fluid (k = 0 weights), its d/dn and pressure   (src/haskell/SFMT.hs:39-41,
                                                                  papers/fuzzy-fmt/figs/homogeneous.cpp:52-57)

Two properties of the reference's weights that the tests pin: w2v equals
i k w3 (the gradient of n3), and in the hard-sphere limit Xi -> 0 every weight
tends to its Rosenfeld form with R = alpha/2 except w0, which tends to half of
the Rosenfeld n0 weight (w0(0) = 1/2 for every Xi): SFMT.hs:109 lacks the
factor 2 that w1v and w2t carry.  The oracle keeps the reference's w0.
"""

import math

import numpy as np
from scipy import special

from . import fmt_oracle as F

SQRT2 = math.sqrt(2.0)
# field order of the sixteen weighted densities
FIELDS = ("n3", "n2", "n1", "n0", "n2vx", "n2vy", "n2vz", "n1vx", "n1vy", "n1vz",
          "Txx", "Tyy", "Tzz", "Txy", "Tyz", "Tzx")


def sfmt_alpha(sigma, epsilon, kT):
    """alpha of src/haskell/SFMT.hs:55-56."""
    betaeps = epsilon / kT
    return sigma * (2.0 / (1.0 + math.sqrt(math.log(2.0) / betaeps))) ** (1.0 / 6.0)


def erf_term(u, v):
    """exp(-v^2) Re erf(u + i v) for u >= 0, v > 0.

    As the generated code writes it (exp(-alpha^2/(2 Xi^2)) times
    Faddeeva::erf(z).real()) where erf(u + i v) is finite in FP64 (v < 25),
    otherwise through the Faddeeva function w = wofz:
    exp(-v^2) Re erf(u + iv) = exp(-v^2) - exp(-u^2) Re[exp(-2iuv) conj(w(v + iu))]."""
    u = np.asarray(u, dtype=np.float64)
    if v < 25.0:
        return np.exp(-v * v) * special.erf(u + 1j * v).real
    w = special.wofz(v + 1j * u)
    c, s = np.cos(2 * u * v), np.sin(2 * u * v)
    return np.exp(-v * v) - np.exp(-u * u) * (c * w.real - s * w.imag)


def sfmt_radial(k, alpha, xi):
    """The seven radial functions (w3, w2, w1, w0, W2v, W1v, W2t) of
    src/haskell/SFMT.hs:103-133 at |k| = k (array); the vector weights are
    i k_j W2v and i k_j W1v, the tensor weights W2t (delta_ij - 3 kh_i kh_j)
    on the diagonal and 3 W2t kh_i kh_j off it."""
    k = np.asarray(k, dtype=np.float64)
    z = k == 0
    ks = np.where(z, 1.0, k)
    E = np.exp(-(xi * ks) ** 2 / 8)
    c = np.cos(ks * alpha / 2)
    s = np.sin(ks * alpha / 2)
    G = erf_term(ks * xi / (2 * SQRT2), alpha / (SQRT2 * xi))
    w3 = 4 * math.pi / ks ** 2 * E * ((1 + xi * xi * ks * ks / 4) * s / ks - alpha / 2 * c)
    w2 = 2 * math.pi * E * (xi * xi / 2 * c + alpha * s / ks)
    w1 = E * s / ks
    w0 = math.sqrt(math.pi / 2) / (ks * xi) * G
    W2v = 4 * math.pi / ks ** 2 * E * ((xi * xi * ks / 4 + 1 / ks) * s - alpha / 2 * c)
    W1v = (math.sqrt(math.pi / 2) / (ks * xi) * 2 * G - E * c) / ks ** 2
    W2t = ((-math.pi * xi * xi / 3 - 4 * math.pi / ks ** 2) * E * c - 2 * math.pi * alpha / (3 * ks) * E * s
           + (2 * math.pi) ** 1.5 / (ks ** 3 * xi) * 2 * G)
    lim = sfmt_radial_at_zero(alpha, xi)
    out = []
    for f, l in zip((w3, w2, w1, w0, W2v, W1v, W2t), lim):
        out.append(np.where(z, l, f))
    return tuple(out)


def sfmt_radial_at_zero(alpha, xi):
    """k -> 0 limits: w3 = pi alpha Xi^2/2 + pi alpha^3/6, w2 = pi (Xi^2 + alpha^2),
    w1 = alpha/2, w0 = 1/2; the vector and tensor weights vanish there (their
    k = 0 values; W2v and W1v multiply k_j, W2t tends to 0)."""
    return (math.pi * alpha * xi * xi / 2 + math.pi * alpha ** 3 / 6, math.pi * (xi * xi + alpha * alpha),
            alpha / 2, 0.5, 0.0, 0.0, 0.0)


def sfmt_weights(grid, alpha, xi):
    """The sixteen k-space weights on grid's half spectrum, in FIELDS order, and
    their parities w(-k) = +/- w(k)."""
    w3, w2, w1, w0, W2v, W1v, W2t = sfmt_radial(grid.k, alpha, xi)
    kv = (grid.kx, grid.ky, grid.kz)
    ks = np.where(grid.k > 0, grid.k, 1.0)
    kh = [np.where(grid.k > 0, c / ks, 0.0) for c in kv]
    W = [w3, w2, w1, w0]
    W += [1j * W2v * c for c in kv]
    W += [1j * W1v * c for c in kv]
    W += [W2t * (1 - 3 * kh[0] ** 2), W2t * (1 - 3 * kh[1] ** 2), W2t * (1 - 3 * kh[2] ** 2),
          W2t * 3 * kh[0] * kh[1], W2t * 3 * kh[1] * kh[2], W2t * 3 * kh[2] * kh[0]]
    parity = (1.0,) * 4 + (-1.0,) * 6 + (1.0,) * 6
    return W, parity


def sfmt_phi(v, want_partials=True):
    """(Phi1, Phi2, Phi3) per kT of the sixteen weighted densities (FIELDS order)
    and, with want_partials, the sixteen partials of Phi1 + Phi2 + Phi3."""
    n3, n2, n1, n0 = v[0], v[1], v[2], v[3]
    ux, uy, uz = v[4], v[5], v[6]      # n2v
    qx, qy, qz = v[7], v[8], v[9]      # n1v
    txx, tyy, tzz, txy, tyz, tzx = v[10:16]
    omn = 1.0 - n3
    L = np.log1p(-n3)
    u2 = ux * ux + uy * uy + uz * uz
    uq = ux * qx + uy * qy + uz * qz
    Tu = (txx * ux + txy * uy + tzx * uz, txy * ux + tyy * uy + tyz * uz, tzx * ux + tyz * uy + tzz * uz)
    uTu = ux * Tu[0] + uy * Tu[1] + uz * Tu[2]
    tr3 = (txx ** 3 + tyy ** 3 + tzz ** 3 + 3 * txx * txy * txy + 3 * txx * tzx * tzx + 3 * txy * txy * tyy
           + 3 * tyy * tyz * tyz + 3 * tzx * tzx * tzz + 3 * tyz * tyz * tzz + 6 * txy * tyz * tzx)
    X = n2 ** 3 - 3 * n2 * u2 + 4.5 * (uTu - tr3)
    c3 = 1.0 / (24 * math.pi * omn * omn)
    phi1 = -n0 * L
    phi2 = (n2 * n1 - uq) / omn
    phi3 = X * c3
    if not want_partials:
        return phi1, phi2, phi3, None
    d3 = n0 / omn + (n2 * n1 - uq) / (omn * omn) + 2 * X * c3 / omn
    d2 = n1 / omn + (3 * n2 * n2 - 3 * u2) * c3
    d1 = n2 / omn
    d0 = -L
    du = [-q / omn + (-6 * n2 * u + 9 * t) * c3 for u, q, t in zip((ux, uy, uz), (qx, qy, qz), Tu)]
    dq = [-u / omn for u in (ux, uy, uz)]
    sq = (txx * txx + txy * txy + tzx * tzx, txy * txy + tyy * tyy + tyz * tyz, tzx * tzx + tyz * tyz + tzz * tzz,
          txx * txy + txy * tyy + tzx * tyz, txy * tzx + tyy * tyz + tyz * tzz, txx * tzx + txy * tyz + tzx * tzz)
    dt = [4.5 * c3 * (ux * ux - 3 * sq[0]), 4.5 * c3 * (uy * uy - 3 * sq[1]), 4.5 * c3 * (uz * uz - 3 * sq[2]),
          4.5 * c3 * (2 * ux * uy - 6 * sq[3]), 4.5 * c3 * (2 * uy * uz - 6 * sq[4]),
          4.5 * c3 * (2 * uz * ux - 6 * sq[5])]
    return phi1, phi2, phi3, (d3, d2, d1, d0, *du, *dq, *dt)


# ---------------------------------------------------------------------------
# Homogeneous fluid (HomogeneousSFMTFluid)
# ---------------------------------------------------------------------------

def homogeneous_sfmt(n, sigma, epsilon, xi, kT, mu=0.0):
    """(free energy per volume, its d/dn) of the uniform fluid: the k = 0 weights,
    kT Phi + kT n (ln(n/nQ) - 1) - mu n (src/haskell/SFMT.hs:39-41)."""
    alpha = sfmt_alpha(sigma, epsilon, kT)
    w3, w2, w1, w0 = sfmt_radial_at_zero(alpha, xi)[:4]
    z = np.float64(0.0)
    v = [np.float64(w3 * n), np.float64(w2 * n), np.float64(w1 * n), np.float64(w0 * n)] + [z] * 12
    p1, p2, p3, d = sfmt_phi(v)
    f = kT * float(p1 + p2 + p3) + kT * n * (math.log(n / F.nQ(kT)) - 1) - mu * n
    dfdn = kT * float(d[0] * w3 + d[1] * w2 + d[2] * w1 + d[3] * w0) + kT * math.log(n / F.nQ(kT)) - mu
    return f, dfdn


def sfmt_chemical_potential(n, sigma, epsilon, xi, kT):
    """mu = d f/dn at mu = 0 (the drivers' hf.mu() = hf.d_by_dn(),
    papers/fuzzy-fmt/figs/new-radial-wca.cpp:122-123)."""
    return homogeneous_sfmt(n, sigma, epsilon, xi, kT, 0.0)[1]


def sfmt_pressure(n, sigma, epsilon, xi, kT):
    """p = n d f/dn - f at mu = 0 (papers/fuzzy-fmt/figs/homogeneous.cpp:52-57)."""
    f, d = homogeneous_sfmt(n, sigma, epsilon, xi, kT, 0.0)
    return n * d - f


# ---------------------------------------------------------------------------
# The fluid functional of Veff (SFMTFluidVeff)
# ---------------------------------------------------------------------------

class SFMTFluidVeff(F.WhiteBearFluidVeff):
    """F[Veff] = kT int (Phi1 + Phi2 + Phi3) + ideal gas + int (Vext - mu) n,
    n = exp(-Veff/kT) (src/haskell/SFMT.hs:35-37): the White-Bear oracle's
    density, ideal-gas, external and chain-rule terms with SFMT's sixteen
    weights and Phi."""

    def __init__(self, grid, sigma, epsilon, xi, kT, mu, vext, x):
        super().__init__(grid, 1.0, kT, mu, vext, x)
        self.functional = "sfmt"
        self.sigma, self.epsilon, self.xi = float(sigma), float(epsilon), float(xi)
        self.alpha = sfmt_alpha(self.sigma, self.epsilon, self.kT)
        W, parity = sfmt_weights(grid, self.alpha, self.xi)
        self.W = tuple(W)
        self.parity = parity

    def _phi(self, nal, partials):
        return sfmt_phi(nal, want_partials=partials)
