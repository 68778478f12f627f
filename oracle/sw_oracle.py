"""FP64 CPU oracle for deft's square-well liquid functional.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/`` may
import it.  The product path (``deft_amd``) never calls it.

It restates src/haskell/SW_liquid.hs (the source of the generated SW_liquid,
SW_liquidVeff and HomogeneousSW_liquid classes, src/new/SW_liquidFast.cpp):

  F = F_sw + F_WB + F_ideal + int n (Vext - mu),   n = exp(-Veff/kT)
  F_sw = int (epsilon/2) n(r) [sum_i Xi_i * (n P_i(g))](r) dr      (SW_liquid.hs:59-64)

* g = gSigmaA of the White-Bear functional at the same density
  (src/haskell/WhiteBear.hs:119-138): [shell(A) - shell'(B) + vshell'.C +
  vshell.D] / shell_2R(n), the estimator the contact oracle
  (fmt_oracle.contact_values) computes and the reference's published contact
  columns pin;
* P_0(g) = g, P_i(g) = sum_j k_ji (g - 1)^j for i = 1..4 with the table of
  SW_liquid.hs:67-90;
* Xi_i(k) = -4 pi epsilon x (closed forms of SW_liquid.hs:92-118), the
  Fourier transforms of -epsilon ((r - sigma)/sigma)^i on sigma < r < lambda sigma
  (tests check the closed forms against that radial integral), sigma = 2R;
  their k = 0 values are the integrals
  -4 pi epsilon sigma^3 [T^(i+1)/(i+1) + 2 T^(i+2)/(i+2) + T^(i+3)/(i+3)], T = lambda - 1.

The gradient of F_sw with respect to n (its functional derivative) is
  (eps/2) [c + sum_i P_i(g) (Xi_i * n)]                     direct terms
  - shell_2R^T * (B g / D)                                  through the denominator D
  + sum_m W_m^T * [sum_j (d a_j/d w_m) (K_j^T * (B / D))]   through the numerator
with c = sum_i Xi_i * (n P_i(g)), B = (eps/2) n sum_i (Xi_i * n) P_i'(g), a_j the
eight numerator fields (A, B, C_xyz, D_xyz) as functions of the five
White-Bear weighted densities w_m, K_j their kernels and W_m the White-Bear
weights; the Jacobian d a_j/d w_m is taken by complex-step differentiation,
exact to round-off (the a_j are analytic in w).
"""

import math

import numpy as np

from . import fmt_oracle as F

# P_i(g) = sum_j SW_COEF[i-1][j-1] (g - 1)^j, i, j = 1..4 (SW_liquid.hs:67-90)
SW_COEF = ((-1.754, 0.027, 0.838, -0.178),
           (-2.243, 4.403, -2.48, 0.363),
           (0.207, 0.712, -1.952, 1.046),
           (-0.002, -0.164, 0.324, -0.162))


def sw_xi_at_zero(sigma, lam, eps):
    T = lam - 1.0
    return [-4 * math.pi * eps * sigma ** 3 * (T ** (i + 1) / (i + 1) + 2 * T ** (i + 2) / (i + 2) + T ** (i + 3) / (i + 3))
            for i in range(5)]


def sw_xi(k, sigma, lam, eps):
    """The five kernels Xi_0..Xi_4 at |k| = k (SW_liquid.hs:92-118)."""
    k = np.asarray(k, dtype=np.float64)
    z = k == 0
    k = np.where(z, 1.0, k)
    ks, lks = k * sigma, lam * k * sigma
    e4 = eps * 4 * math.pi
    s1, c1, s2, c2 = np.sin(ks), np.cos(ks), np.sin(lks), np.cos(lks)
    k3 = k ** 3
    x0 = -e4 * ((s2 - s1) / k3 + (sigma * c1 - lam * sigma * c2) / k ** 2)
    x1 = -e4 * ((2 - lks * ks * (lam - 1)) * c2 / (k3 * ks) - ks * (s1 + (1 - 2 * lam) * s2) / (k3 * ks)
                - 2 * c1 / (k3 * ks))
    x2 = -e4 * (6 * s1 / (k3 * ks ** 2) + (ks ** 2 * (1 - 4 * lam + 3 * lam ** 2) - 6) * s2 / (k3 * ks ** 2)
                - ks * (4 + lam * (ks ** 2 * (lam - 1) ** 2 - 6)) * c2 / (k3 * ks ** 2) - 2 * ks * c1 / (k3 * ks ** 2))
    x3 = -e4 * (ks * (6 * s1 + (18 - 24 * lam + ks ** 2 * (lam - 1) ** 2 * (4 * lam - 1)) * s2) / (k3 * ks ** 3)
                + (6 * ks ** 2 * (lam - 1) * (2 * lam - 1) - lks * ks ** 3 * (lam - 1) ** 3 - 24) * c2 / (k3 * ks ** 3)
                + 24 * c1 / (k3 * ks ** 3))
    x4 = -e4 * (ks ** 2 * (lam - 1) * (36 - 60 * lam + ks ** 2 * (lam - 1) ** 2 * (5 * lam - 1)) * s2 / (k3 * ks ** 4)
                + 120 * (s2 - s1) / (k3 * ks ** 4)
                + ks * (24 * c1 - (24 * (5 * lam - 4) - 4 * ks ** 2 * (lam - 1) ** 2 * (5 * lam - 2)
                                   + lks * ks ** 3 * (lam - 1) ** 4) * c2) / (k3 * ks ** 4))
    lim = sw_xi_at_zero(sigma, lam, eps)
    return [np.where(z, l, x) for x, l in zip((x0, x1, x2, x3, x4), lim)]


def sw_P(g):
    """(P_0..P_4, dP_0..dP_4) at g."""
    P, dP = [g], [np.ones_like(g) if isinstance(g, np.ndarray) else 1.0]
    h = g - 1.0
    for row in SW_COEF:
        P.append(row[0] * h + row[1] * h ** 2 + row[2] * h ** 3 + row[3] * h ** 4)
        dP.append(row[0] + 2 * row[1] * h + 3 * row[2] * h ** 2 + 4 * row[3] * h ** 3)
    return P, dP


def f3_complex(n):
    """f3 and f3' of the White-Bear oracle (fmt_oracle.f3_and_derivative) for
    complex arguments (the complex step): closed form for |Re n| >= 0.05, the
    same Taylor series below."""
    n = np.asarray(n, dtype=np.complex128)
    small = np.abs(n.real) < F.F3_SERIES_CUTOFF
    nb = np.where(small, 0.5, n)
    omb = 1.0 - nb
    L = np.log(omb)
    num = nb + omb * omb * L
    den = nb * nb * omb * omb
    f_big = num / den
    dnum = nb - 2.0 * omb * L
    dden = 2.0 * nb * omb * (1.0 - 2.0 * nb)
    df_big = dnum / den - num * dden / (den * den)
    ns = np.where(small, n, 0.0)
    g = np.full_like(ns, 1.5)
    gp = np.zeros_like(ns)
    p = np.ones_like(ns)
    for kk in range(3, F.F3_SERIES_TERMS + 3):
        c = 2.0 / (kk * (kk - 1) * (kk - 2))
        gp -= c * (kk - 2) * p
        p = p * ns
        g -= c * p
    oms = 1.0 - ns
    return np.where(small, g / (oms * oms), f_big), np.where(small, gp / (oms * oms) + 2.0 * g / oms ** 3, df_big)


def afields(n3, n2, vx, vy, vz, R):
    """The eight gSigmaA numerator fields (A, B, C_x, C_y, C_z, D_x, D_y, D_z) of
    fmt_oracle.contact_values, as analytic functions of the White-Bear weighted
    densities (complex arguments allowed)."""
    fpR, fpR2 = 4 * math.pi * R, 4 * math.pi * R * R
    omn = 1.0 - n3
    L = np.log(omn + 0j) if np.iscomplexobj(n3) else np.log1p(-n3)
    g3, g3p = f3_complex(n3)
    if not np.iscomplexobj(n3) and not np.iscomplexobj(n2):
        g3, g3p = g3.real, g3p.real
    g2, g2p = 1.0 / omn, 1.0 / (omn * omn)
    n0, n1 = n2 / fpR2, n2 / fpR
    v = (vx, vy, vz)
    v2 = vx * vx + vy * vy + vz * vz
    c36 = 1.0 / (36 * math.pi)
    dn0 = -L
    dn1 = g2 * n2
    dn2 = g2 * n1 + g3 * (3 * n2 * n2 - 3 * v2) * c36
    dn3 = n0 / omn + g2p * (n1 * n2 - v2 / fpR) + g3p * n2 * (n2 * n2 - 3 * v2) * c36
    dn1v = [-g2 * c for c in v]
    dn2v = [-g2 * c / fpR - g3 * 6 * n2 * c * c36 for c in v]
    A = dn3 - dn1 / fpR2 - dn0 / (2 * math.pi * R ** 3)
    B = dn2 + dn1 / fpR + dn0 / fpR2
    C = [a + b / fpR for a, b in zip(dn2v, dn1v)]
    D = [b / fpR2 for b in dn1v]
    return [A, B] + C + D


def homogeneous_sw(n, R, eps, lam, kT, mu=0.0):
    """Free energy per volume of the uniform liquid, HomogeneousSW_liquid
    (SW_liquid.hs:45-47): White-Bear + ideal gas - mu n + (eps/2) n^2 sum_i Xi_i(0) P_i(g),
    g = (4 pi R^2 A + 8 pi R B)/(16 pi R^2 n) (shell(0) = 4 pi R^2, shell'(0) = -8 pi R,
    shell_2R(0) = 16 pi R^2).  Complex n is allowed (the derivative is its complex step)."""
    n3 = 4 * math.pi / 3 * R ** 3 * n
    n2 = 4 * math.pi * R * R * n
    z = 0.0 * n
    a = afields(n3, n2, z, z, z, R)
    g = (4 * math.pi * R * R * a[0] + 8 * math.pi * R * a[1]) / (16 * math.pi * R * R * n)
    P, _ = sw_P(g)
    xi0 = sw_xi_at_zero(2 * R, lam, eps)
    fsw = 0.5 * eps * n * n * sum(x * p for x, p in zip(xi0, P))
    omn = 1.0 - n3
    g3, _ = f3_complex(n3)
    if not np.iscomplexobj(n):
        g3 = g3.real
    L = np.log(omn + 0j) if np.iscomplexobj(n) else math.log1p(-n3)
    phi = -n2 / (4 * math.pi * R * R) * L + n2 * n2 / (4 * math.pi * R * omn) + g3 * n2 ** 3 / (36 * math.pi)
    lg = np.log(n / F.nQ(kT) + 0j) if np.iscomplexobj(n) else math.log(n / F.nQ(kT))
    return kT * phi + kT * n * (lg - 1) - mu * n + fsw


def homogeneous_sw_dn(n, R, eps, lam, kT, mu=0.0):
    h = 1e-20 * n
    return float(np.imag(homogeneous_sw(n + 1j * h, R, eps, lam, kT, mu)) / h)


class SWLiquidVeff(F.WhiteBearFluidVeff):
    """SW_liquidVeff: the White-Bear fluid of Veff plus F_sw."""

    def __init__(self, grid, R, epsilon, lam, kT, mu, vext, x):
        super().__init__(grid, R, kT, mu, vext, x)  # the White-Bear part ("wb")
        self.epsilon, self.lam = float(epsilon), float(lam)
        self.Xi = sw_xi(grid.k, 2 * R, self.lam, self.epsilon)
        k = grid.k
        dr = grid.dV ** (1.0 / 3.0)
        s = np.exp(-F.SPREADING * k * k * dr * dr)
        kR = k * R
        ks = np.where(k > 0, k, 1.0)
        _, w2, wv = F.weights(grid, R)
        shellp = np.where(k > 0, s * 4 * math.pi * (-np.sin(kR) / ks - R * np.cos(kR)), -8 * math.pi * R)
        vsp = np.where(k > 0, s * 4 * math.pi * R * np.sin(kR) / ks, 0.0)
        kv = (grid.kx, grid.ky, grid.kz)
        # numerator kernels of the a-fields: shell(A) - shell'(B) + vshell'.C + vshell.D
        self.K = [w2, -shellp] + [1j * vsp * c for c in kv] + [1j * wv * c for c in kv]
        self.shell2 = F.weights(grid, 2 * R)[1]

    def sw_parts(self, x):
        n = self.density(x)
        nk = F._rfftn(n)
        w5 = [F._irfftn(nk * W, self.grid.shape) for W in self.W[:5]]
        a = afields(*w5, self.R)
        acc = 0.0
        for aj, Kj in zip(a, self.K):
            acc = acc + F._rfftn(aj) * Kj
        num = F._irfftn(acc, self.grid.shape)
        den = F._irfftn(nk * self.shell2, self.grid.shape)
        g = num / den
        P, dP = sw_P(g)
        cacc = 0.0
        for Pi, Xi in zip(P, self.Xi):
            cacc = cacc + F._rfftn(n * Pi) * Xi
        c = F._irfftn(cacc, self.grid.shape)
        return dict(n=n, nk=nk, w5=w5, a=a, num=num, den=den, g=g, P=P, dP=dP, c=c)

    def sw_energy(self, x=None):
        x = self.x if x is None else x
        p = self.sw_parts(x)
        return 0.5 * self.epsilon * float(np.sum(p["n"] * p["c"])) * self.grid.dV

    def energy(self, x=None):
        x = self.x if x is None else x
        E = super().energy(x)
        Esw = self.sw_energy(x)
        self.last_terms["sw"] = Esw
        return E + Esw

    def sw_dn(self, x=None):
        """Functional derivative of F_sw with respect to n."""
        x = self.x if x is None else x
        p = self.sw_parts(x)
        shape = self.grid.shape
        eps = self.epsilon
        n, g, den = p["n"], p["g"], p["den"]
        kn = [F._irfftn(p["nk"] * Xi, shape) for Xi in self.Xi]
        direct = 0.5 * eps * (p["c"] + sum(Pi * ki for Pi, ki in zip(p["P"], kn)))
        Bf = 0.5 * eps * n * sum(ki * dPi for ki, dPi in zip(kn, p["dP"]))
        u = Bf / den
        t = Bf * g / den
        uk = F._rfftn(u)
        q = [F._irfftn(uk * np.conj(Kj), shape) for Kj in self.K]
        # r_m = sum_j (d a_j / d w_m) q_j by complex steps
        h = 1e-30
        r = []
        for m in range(5):
            wc = [wi.astype(np.complex128) for wi in p["w5"]]
            wc[m] = wc[m] + 1j * h
            ac = afields(*wc, self.R)
            r.append(sum(np.imag(aj) / h * qj for aj, qj in zip(ac, q)))
        back = 0.0
        for rm, Wm in zip(r, self.W[:5]):
            back = back + F._rfftn(rm) * np.conj(Wm)
        back = back - F._rfftn(t) * self.shell2
        return direct + F._irfftn(back, shape)

    def dF_dn(self, x=None):
        x = self.x if x is None else x
        g, n = super().dF_dn(x)
        extra = self.sw_dn(x) * self.grid.dV
        return np.where(self.finite, g + extra, 0.0), n
