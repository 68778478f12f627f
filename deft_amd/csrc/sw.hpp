// The square-well liquid (deft src/haskell/SW_liquid.hs): White-Bear plus the
// attraction
//   F_sw = int (eps/2) n(r) [sum_i Xi_i * (n P_i(g))](r) dr,
// g = gSigmaA of the White-Bear functional at the same density
// (src/haskell/WhiteBear.hs:119-138), P_0(g) = g and P_i(g) = sum_j k_ji (g-1)^j
// (SW_liquid.hs:67-90), Xi_i the Fourier transforms of -eps ((r - sigma)/sigma)^i
// on sigma < r < lambda sigma, sigma = 2R (the closed forms of SW_liquid.hs:92-118).
//
// Host and device code (the host side computes the homogeneous liquid).  The
// gradient through g needs the Jacobian of gSigmaA's eight numerator fields
// a_j (A, B, C_xyz, D_xyz: combinations of the White-Bear partials with n0,
// n1, n1v independent) with respect to the five White-Bear weighted
// densities; it is taken in forward mode with dual numbers (value,
// derivative), one pass per weighted density, so no second derivative of
// Phi is written out by hand.
#pragma once

#include <cmath>

#include "wb_point.hpp"

namespace deft {

// P_i(g) = sum_j kSWCoef[i-1][j-1] (g - 1)^j, i, j = 1..4 (SW_liquid.hs:67-90)
constexpr double kSWCoef[4][4] = {{-1.754, 0.027, 0.838, -0.178},
                                  {-2.243, 4.403, -2.48, 0.363},
                                  {0.207, 0.712, -1.952, 1.046},
                                  {-0.002, -0.164, 0.324, -0.162}};

struct SWConst {
  double R, sigma, eps, lambda;
};

inline SWConst make_sw_const(double R, double eps, double lambda) { return SWConst{R, 2.0 * R, eps, lambda}; }

// Xi_0..Xi_4 at |k| = k; at k = 0 the integrals
// -4 pi eps sigma^3 [T^(i+1)/(i+1) + 2 T^(i+2)/(i+2) + T^(i+3)/(i+3)], T = lambda - 1
__host__ __device__ inline void sw_xi(double k, const SWConst& c, double (&x)[5]) {
  const double pi = 3.14159265358979323846;
  const double sig = c.sigma, lam = c.lambda, e4 = c.eps * 4.0 * pi;
  if (k == 0.0) {
    const double T = lam - 1.0;
    double Tp = T;  // T^(i+1)
    for (int i = 0; i < 5; ++i) {
      x[i] = -e4 * sig * sig * sig * (Tp / (i + 1) + 2.0 * Tp * T / (i + 2) + Tp * T * T / (i + 3));
      Tp *= T;
    }
    return;
  }
  const double ks = k * sig, lks = lam * k * sig;
  const double s1 = sin(ks), c1 = cos(ks), s2 = sin(lks), c2 = cos(lks);
  const double k2 = k * k, k3 = k2 * k;
  const double lm1 = lam - 1.0;
  const double ks2 = ks * ks, ks3 = ks2 * ks, ks4 = ks2 * ks2;
  x[0] = -e4 * ((s2 - s1) / k3 + (sig * c1 - lam * sig * c2) / k2);
  x[1] = -e4 * ((2.0 - lks * ks * lm1) * c2 / (k3 * ks) - ks * (s1 + (1.0 - 2.0 * lam) * s2) / (k3 * ks) -
                2.0 * c1 / (k3 * ks));
  x[2] = -e4 * (6.0 * s1 / (k3 * ks2) + (ks2 * (1.0 - 4.0 * lam + 3.0 * lam * lam) - 6.0) * s2 / (k3 * ks2) -
                ks * (4.0 + lam * (ks2 * lm1 * lm1 - 6.0)) * c2 / (k3 * ks2) - 2.0 * ks * c1 / (k3 * ks2));
  x[3] = -e4 * (ks * (6.0 * s1 + (18.0 - 24.0 * lam + ks2 * lm1 * lm1 * (4.0 * lam - 1.0)) * s2) / (k3 * ks3) +
                (6.0 * ks2 * lm1 * (2.0 * lam - 1.0) - lks * ks3 * lm1 * lm1 * lm1 - 24.0) * c2 / (k3 * ks3) +
                24.0 * c1 / (k3 * ks3));
  x[4] = -e4 * (ks2 * lm1 * (36.0 - 60.0 * lam + ks2 * lm1 * lm1 * (5.0 * lam - 1.0)) * s2 / (k3 * ks4) +
                120.0 * (s2 - s1) / (k3 * ks4) +
                ks * (24.0 * c1 - (24.0 * (5.0 * lam - 4.0) - 4.0 * ks2 * lm1 * lm1 * (5.0 * lam - 2.0) +
                                   lks * ks3 * lm1 * lm1 * lm1 * lm1) * c2) /
                    (k3 * ks4));
}

// P_0..P_4 and their derivatives at g
__host__ __device__ inline void sw_P(double g, double (&P)[5], double (&dP)[5]) {
  const double h = g - 1.0;
  P[0] = g;
  dP[0] = 1.0;
  for (int i = 0; i < 4; ++i) {
    const double* c = kSWCoef[i];
    P[i + 1] = h * (c[0] + h * (c[1] + h * (c[2] + h * c[3])));
    dP[i + 1] = c[0] + h * (2.0 * c[1] + h * (3.0 * c[2] + h * 4.0 * c[3]));
  }
}

// --- forward-mode dual numbers ---------------------------------------------
struct Dual {
  double v, d;
};
__host__ __device__ inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__host__ __device__ inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__host__ __device__ inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__host__ __device__ inline Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__host__ __device__ inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__host__ __device__ inline Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__host__ __device__ inline Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__host__ __device__ inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__host__ __device__ inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__host__ __device__ inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__host__ __device__ inline Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return {q, -q * b.d / b.v};
}
__host__ __device__ inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
__host__ __device__ inline double dual_value(double x) { return x; }
__host__ __device__ inline double dual_value(Dual x) { return x.v; }
// log(1 - x)
__host__ __device__ inline double sw_log1m(double x) { return log1m(x); }
__host__ __device__ inline Dual sw_log1m(Dual x) { return {log1m(x.v), -x.d / (1.0 - x.v)}; }

// f3(n) = (n + (1-n)^2 ln(1-n))/(n^2 (1-n)^2) and f3'(n): the closed form for
// |n| >= 0.05, below that f3 = g/(1-n)^2 with g(n) = 3/2 - sum_{k>=3} 2 n^(k-2)/(k (k-1) (k-2))
// (18 terms), the White-Bear oracle's split
template <typename S>
__host__ __device__ inline void sw_f3(S n, S& f3, S& df3) {
  if (fabs(dual_value(n)) < 0.05) {
    S g = n * 0.0 + 1.5, gp = n * 0.0, p = n * 0.0 + 1.0;
    for (int kk = 3; kk < 21; ++kk) {
      const double c = 2.0 / (kk * (kk - 1) * (kk - 2));
      gp = gp - c * (kk - 2) * p;
      p = p * n;
      g = g - c * p;
    }
    const S om = 1.0 - n;
    f3 = g / (om * om);
    df3 = gp / (om * om) + 2.0 * g / (om * om * om);
    return;
  }
  const S om = 1.0 - n;
  const S L = sw_log1m(n);
  const S num = n + om * om * L;
  const S den = n * n * om * om;
  f3 = num / den;
  const S dnum = n - 2.0 * om * L;
  const S dden = 2.0 * n * om * (1.0 - 2.0 * n);
  df3 = dnum / den - num * dden / (den * den);
}

// gSigmaA's eight numerator fields from the White-Bear weighted densities
// (n3, n2, n2v): A = dn3 - dn1/(4 pi R^2) - dn0/(2 pi R^3),
// B = dn2 + dn1/(4 pi R) + dn0/(4 pi R^2), C = dn2v + dn1v/(4 pi R),
// D = dn1v/(4 pi R^2), with the partials of n0, n1, n2, n3, n1v, n2v independent
template <typename S>
__host__ __device__ inline void sw_afields(S n3, S n2, S vx, S vy, S vz, double R, S (&a)[8]) {
  const double pi = 3.14159265358979323846;
  const double fpR = 4.0 * pi * R, fpR2 = 4.0 * pi * R * R, c36 = 1.0 / (36.0 * pi);
  const S om = 1.0 - n3;
  const S L = sw_log1m(n3);
  S g3, g3p;
  sw_f3(n3, g3, g3p);
  const S g2 = 1.0 / om, g2p = g2 * g2;
  const S n0 = n2 / fpR2, n1 = n2 / fpR;
  const S v2 = vx * vx + vy * vy + vz * vz;
  const S dn0 = -L;
  const S dn1 = g2 * n2;
  const S dn2 = g2 * n1 + g3 * (3.0 * n2 * n2 - 3.0 * v2) * c36;
  const S dn3 = n0 / om + g2p * (n1 * n2 - v2 / fpR) + g3p * n2 * (n2 * n2 - 3.0 * v2) * c36;
  a[0] = dn3 - dn1 / fpR2 - dn0 / (2.0 * pi * R * R * R);
  a[1] = dn2 + dn1 / fpR + dn0 / fpR2;
  const S v[3] = {vx, vy, vz};
  for (int c = 0; c < 3; ++c) {
    const S d1v = -g2 * v[c];
    const S d2v = -g2 * v[c] / fpR - g3 * 6.0 * n2 * v[c] * c36;
    a[2 + c] = d2v + d1v / fpR;
    a[5 + c] = d1v / fpR2;
  }
}

// The uniform liquid at density n (HomogeneousSW_liquid, SW_liquid.hs:45-47):
// White-Bear + ideal gas + (eps/2) n^2 sum_i Xi_i(0) P_i(g),
// g = (4 pi R^2 A + 8 pi R B)/(16 pi R^2 n), and its derivative (dual numbers)
inline void sw_bulk(const SWConst& c, double n, double kT, double lognq, double* f, double* dfdn) {
  const double pi = 3.14159265358979323846;
  const double R = c.R;
  const Dual nd{n, 1.0};
  const Dual n3 = (4.0 * pi / 3.0 * R * R * R) * nd, n2 = (4.0 * pi * R * R) * nd;
  const Dual z{0.0, 0.0};
  Dual a[8];
  sw_afields(n3, n2, z, z, z, R, a);
  const Dual g = (4.0 * pi * R * R * a[0] + 8.0 * pi * R * a[1]) / (16.0 * pi * R * R * nd);
  double xi0[5];
  sw_xi(0.0, c, xi0);
  const Dual h = g - 1.0;
  Dual s = xi0[0] * g;
  for (int i = 0; i < 4; ++i) {
    const double* k = kSWCoef[i];
    s = s + xi0[i + 1] * (h * (k[0] + h * (k[1] + h * (k[2] + h * k[3]))));
  }
  const Dual fsw = 0.5 * c.eps * nd * nd * s;
  Dual g3, g3p;
  sw_f3(n3, g3, g3p);
  const Dual om = 1.0 - n3;
  const Dual phi = -(n2 / (4.0 * pi * R * R)) * sw_log1m(n3) + n2 * n2 / (4.0 * pi * R * om) + g3 * n2 * n2 * n2 / (36.0 * pi);
  const double lg = std::log(n) - lognq;
  *f = kT * phi.v + kT * n * (lg - 1.0) + fsw.v;
  *dfdn = kT * phi.d + kT * lg + fsw.d;
}

}  // namespace deft
