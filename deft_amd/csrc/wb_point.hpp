// Pointwise White-Bear (no tensor) free-energy density and its partials, FP64.
//
// Phi1 = -n0 ln(1-n3), Phi2 = (n1 n2 - n1v.n2v)/(1-n3),
// Phi3 = f3(n3)/(36 pi) n2 (n2^2 - 3 n2v.n2v),
// f3(n) = (n + (1-n)^2 ln(1-n)) / (n^2 (1-n)^2)
// (deft src/haskell/WhiteBear.hs:47-60, src/HardSpheres.cpp:137-159), with
// n0 = n2/(4 pi R^2), n1 = n2/(4 pi R), n1v = n2v/(4 pi R) (src/haskell/FMT.hs:34-39).
// Partials are taken with respect to (n3, n2, n2v_x, n2v_y, n2v_z), the
// n0/n1/n1v dependence folded in (closed forms of src/haskell/WhiteBear.hs:114-121).
//
// For |n3| < 0.05 f3 and f3' come from the Taylor series
// f3 = g/(1-n)^2, g = 3/2 - sum_{k>=3} 2 n^(k-2) / (k (k-1) (k-2)),
// since the closed form is 0/0 at n3 = 0 (inside hard walls).
//
// White-Bear Mark II (src/haskell/WhiteBear.hs:140-159): Phi2 and Phi3 take the
// factors (1 + n3^2 phi2/9) and (1 - 4 n3 phi3/9) of Hansen-Goos and Roth,
// phi2 = (6n - 3n^2 + 6(1-n) ln(1-n))/n^3, phi3 = (6n - 9n^2 + 6n^3 +
// 6(1-n)^2 ln(1-n))/(4 n^3), with Phi3's prefactor 1/(24 pi (1-n)^2).  Both
// are written as Phi2 = g2 (n2^2 - n2v^2)/(4 pi R), Phi3 = g3 n2 (n2^2 - 3 n2v^2)
// /(36 pi): White-Bear has g2 = 1/(1-n), g3 = f3; Mark II g2 = (1 + S2/9)/(1-n),
// g3 = 3/2 (1 - 4 S3/9)/(1-n)^2 with S2 = n^2 phi2, S3 = n phi3, whose series
// S2 = sum_{m>=3} 6 n^(m-1)/(m(m-1)), S3 = n - sum_{m>=4} 3 n^(m-2)/(m(m-1)(m-2))
// replace the closed forms below |n3| = 0.05.
// The slower C++ version, HardSpheresWBm2slow (src/HardSpheres.cpp:161-185),
// has 1/(36 pi) for Phi3's prefactor, i.e. g3 = (1 - 4 S3/9)/(1-n)^2: its bulk
// is not Carnahan-Starling and it lacks Rosenfeld's low-density limit, but the
// reference's published Mark II profiles (papers/contact/figs/
// inner-sphereWBm2-*.dat) were computed with it; it is kept as a third choice.
//
// Tensor White-Bear (TensorWhiteBear, src/haskell/WhiteBear.hs:44-51,
// src/HardSpheres.cpp:30-67): Phi1, Phi2 as White-Bear and
// Phi3 = f3/(36 pi) X, X = n2^3 - 3 n2 n2v.n2v + 9 (n2v.T.n2v - tr(T^3)/2),
// T the traceless shell tensor density; eleven fields (n3, n2, n2v_xyz,
// T_xx, T_yy, T_zz, T_xy, T_yz, T_zx), off-diagonal partials counting both
// symmetric entries.
#pragma once

#include <hip/hip_runtime.h>

namespace deft {

struct WBConst {
  double fpR;       // 4 pi R
  double fpR2;      // 4 pi R^2
  double inv_fpR;   // 1/(4 pi R)
  double inv_fpR2;  // 1/(4 pi R^2)
  int mark2;        // 0: White-Bear (or tensor), 1: Mark II, 2: Mark II with the 1/(36 pi) Phi3
  double g3_scale;  // Mark II: g3 = g3_scale (1 - 4 S3/9)/(1-n)^2 (3/2, or 1 for mark2 = 2)
  double R;
};

// functional: 0 White-Bear, 1 Mark II, 2 Mark II (36 pi form), 3 tensor White-Bear
inline WBConst make_wb_const(double R, int functional = 0) {
  const double pi = 3.14159265358979323846;
  const int mark2 = (functional == 1 || functional == 2) ? functional : 0;
  return WBConst{4 * pi * R, 4 * pi * R * R, 1.0 / (4 * pi * R), 1.0 / (4 * pi * R * R), mark2,
                 mark2 == 2 ? 1.0 : 1.5, R};
}

// log(1 - x) in FP64 for the pointwise kernels (the library's log1p costs
// about 110 FP64 instructions per call).  1 - x = 2^k (1 + f) with 1 + f in
// [sqrt(1/2), sqrt(2)); for |x| < 1 - sqrt(1/2) this is k = 0 and f = -x
// exactly, so small arguments keep full relative accuracy (no 1 - x rounding).
// log(1 + f) = 2 atanh(s), s = f/(2 + f), is evaluated as f - s (f - R) with
// R = 2s^2/3 + 2s^4/5 + ... + 2s^20/21 (the atanh Taylor series; |s| <= 0.1716,
// so the first omitted term is below 2e-17 relative), keeping the exact f as
// the leading term.  k ln 2 is added with ln 2 split in two, the high part
// having 32 significant bits so that k ln2_hi is exact.  x = 1 gives -inf,
// x > 1 and NaN give NaN.  Error within 2 ulp of log1p(-x) over the range the
// kernels use (pinned against numpy in the host tests).
__host__ __device__ inline double log1m(double x) {
#pragma clang fp contract(off)
  constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000
  constexpr double kLn2Lo = 1.90821492927058770002e-10;  // ln 2 - kLn2Hi
  const double u = 1.0 - x;
  int k = 0;
  double m = frexp(u, &k);  // u = m 2^k, m in [1/2, 1)
  const bool low = m < 0.70710678118654752440;
  m = low ? 2.0 * m : m;
  k = low ? k - 1 : k;
  const bool near = fabs(x) < 0.29289321881345247560;  // 1 - sqrt(1/2)
  const double f = near ? -x : m - 1.0;
  const double kk = near ? 0.0 : (double)k;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double P = 2.0 / 21.0;
#pragma unroll
  for (int j = 8; j >= 0; --j) P = fma(P, z, 2.0 / (2 * j + 3));
  const double lg = f - s * (f - z * P);
  const double r = fma(kk, kLn2Hi, fma(kk, kLn2Lo, lg));
  return u > 0.0 ? r : (u == 0.0 ? -__builtin_huge_val() : __builtin_nan(""));
}

constexpr double kWB2SeriesCutoff = 0.05;
constexpr int kWB2SeriesTerms = 24;

// Mark II factors g2, g2', g3, g3' (see above) given n, io = 1/(1-n), L = ln(1-n)
// and g3's prefactor s3.
__host__ __device__ inline void markii_of(double n, double io, double L, double s3, double& g2, double& g2p,
                                          double& g3, double& g3p) {
#pragma clang fp contract(off)
  double S2, S2p, S3, S3p;
  if (fabs(n) < kWB2SeriesCutoff) {
    // S2 = sum_{m>=3} 6 n^(m-1)/(m(m-1)), S3 = n - sum_{m>=4} 3 n^(m-2)/(m(m-1)(m-2)),
    // Horner in n with explicit FMAs (terms m < kWB2SeriesTerms + 3)
    double a = 0.0, ap = 0.0, b = 0.0, bp = 0.0;
#pragma unroll
    for (int m = kWB2SeriesTerms + 2; m >= 3; --m) {
      a = fma(a, n, 6.0 / (m * (m - 1.0)));         // sum 6 n^(m-3)/(m(m-1))
      ap = fma(ap, n, 6.0 / m);                      // sum 6 n^(m-3)/m
    }
#pragma unroll
    for (int m = kWB2SeriesTerms + 3; m >= 4; --m) {
      b = fma(b, n, 3.0 / (m * (m - 1.0) * (m - 2.0)));  // sum 3 n^(m-4)/(m(m-1)(m-2))
      bp = fma(bp, n, 3.0 / (m * (m - 1.0)));            // sum 3 n^(m-4)/(m(m-1))
    }
    const double n2_ = n * n;
    S2 = a * n2_;
    S2p = ap * n;
    S3 = n - b * n2_;
    S3p = 1.0 - bp * n;
  } else {
    const double om = 1.0 - n;
    const double N2 = 6.0 * n - 3.0 * n * n + 6.0 * om * L;
    S2 = N2 / n;
    S2p = (-3.0 * n * n - 6.0 * n - 6.0 * L) / (n * n);
    const double N3 = 6.0 * n - 9.0 * n * n + 6.0 * n * n * n + 6.0 * om * om * L;
    const double N3p = -12.0 * n + 18.0 * n * n - 12.0 * om * L;
    S3 = N3 / (4.0 * n * n);
    S3p = (N3p * n - 2.0 * N3) / (4.0 * n * n * n);
  }
  const double a2 = 1.0 + S2 / 9.0;
  g2 = io * a2;
  g2p = io * io * a2 + io * S2p / 9.0;
  const double a3 = 1.0 - 4.0 * S3 / 9.0;
  g3 = s3 * a3 * io * io;
  g3p = s3 * (-(4.0 / 9.0) * S3p * io * io + 2.0 * a3 * io * io * io);
}

constexpr double kInv36Pi = 1.0 / (36.0 * 3.14159265358979323846);
constexpr double kF3SeriesCutoff = 0.05;
constexpr int kF3SeriesTerms = 14;

// f3 and f3' given om = 1-n, io = 1/(1-n) and L = ln(1-n).
__host__ __device__ inline void f3_of(double n, double om, double io, double L, double& f3, double& df3) {
  // no FMA contraction: the energy-only and the energy+gradient kernels must
  // round the energy density identically (the line search compares them)
#pragma clang fp contract(off)
  if (fabs(n) < kF3SeriesCutoff) {
    // g = 3/2 - sum_{k>=3} c_k n^(k-2), g' = -sum c_k (k-2) n^(k-3), c_k = 2/(k(k-1)(k-2)),
    // Horner in n with explicit FMAs
    double t = 0.0, tp = 0.0;
#pragma unroll
    for (int k = kF3SeriesTerms + 2; k >= 3; --k) {
      t = fma(t, n, 2.0 / (k * (k - 1.0) * (k - 2.0)));
      tp = fma(tp, n, 2.0 / (k * (k - 1.0)));
    }
    const double g = fma(-n, t, 1.5), gp = -tp;
    const double io2 = io * io;
    f3 = g * io2;
    df3 = (gp + 2.0 * g * io) * io2;
  } else {
    const double num = n + om * om * L;
    const double iden = io * io / (n * n);
    f3 = num * iden;
    const double dnum = n - 2.0 * om * L;
    const double dden = 2.0 * n * om * (1.0 - 2.0 * n);
    df3 = (dnum - num * dden * iden) * iden;
  }
}

__host__ __device__ inline void f3_and_derivative(double n, double& f3, double& df3) {
  const double om = 1.0 - n;
  f3_of(n, om, 1.0 / om, log1m(n), f3, df3);
}

// Energy density terms only; M2 selects the Mark II factors at compile time
// (their series would otherwise cost the White-Bear kernels registers).
template <bool M2>
__host__ __device__ inline void wb_phi_k(double n3, double n2, double vx, double vy, double vz, const WBConst& c,
                                         double& p1, double& p2, double& p3) {
  // no FMA contraction: every kernel that sums the energy density calls this
  // function, so energies of the same Veff agree to the last bit
#pragma clang fp contract(off)
  const double om = 1.0 - n3;
  const double io = 1.0 / om;
  const double L = log1m(n3);
  const double v2 = vx * vx + vy * vy + vz * vz;
  double f3, df3, g2 = io;
  if constexpr (M2) {
    double g2p;
    markii_of(n3, io, L, c.g3_scale, g2, g2p, f3, df3);
  } else {
    f3_of(n3, om, io, L, f3, df3);
  }
  const double n2s = n2 * n2;
  p1 = -n2 * L * c.inv_fpR2;
  p2 = (n2s - v2) * c.inv_fpR * g2;
  p3 = f3 * n2 * (n2s - 3.0 * v2) * kInv36Pi;
}

// Energy density terms and the five partials (d3, d2, dvx, dvy, dvz).
template <bool M2>
__host__ __device__ inline void wb_phi_partials_k(double n3, double n2, double vx, double vy, double vz,
                                                  const WBConst& c, double& p1, double& p2, double& p3,
                                                  double* d) {
  const double om = 1.0 - n3;
  const double io = 1.0 / om;
  const double L = log1m(n3);
  const double v2 = vx * vx + vy * vy + vz * vz;
  double f3, df3;
  const double n2s = n2 * n2;
  const double q = n2s - v2;
  const double t3 = n2 * (n2s - 3.0 * v2);
  double a, ap;  // Phi2 = q a, dPhi2/dn3 = q ap
  if constexpr (M2) {
    double g2, g2p;
    markii_of(n3, io, L, c.g3_scale, g2, g2p, f3, df3);
    a = c.inv_fpR * g2;
    ap = c.inv_fpR * g2p;
  } else {
    f3_of(n3, om, io, L, f3, df3);
    a = c.inv_fpR * io;  // 1/(4 pi R (1-n3))
    ap = a * io;
  }
  p1 = -n2 * L * c.inv_fpR2;
  p2 = q * a;
  p3 = f3 * t3 * kInv36Pi;
  d[0] = n2 * c.inv_fpR2 * io + q * ap + df3 * t3 * kInv36Pi;
  d[1] = -L * c.inv_fpR2 + 2.0 * n2 * a + f3 * 3.0 * q * kInv36Pi;
  const double cv = -2.0 * a - f3 * 6.0 * n2 * kInv36Pi;
  d[2] = cv * vx;
  d[3] = cv * vy;
  d[4] = cv * vz;
}

// Runtime dispatch on the plan's functional (host code and tests).
__host__ __device__ inline void wb_phi(double n3, double n2, double vx, double vy, double vz, const WBConst& c,
                                       double& p1, double& p2, double& p3) {
  if (c.mark2)
    wb_phi_k<true>(n3, n2, vx, vy, vz, c, p1, p2, p3);
  else
    wb_phi_k<false>(n3, n2, vx, vy, vz, c, p1, p2, p3);
}
__host__ __device__ inline void wb_phi_partials(double n3, double n2, double vx, double vy, double vz,
                                                const WBConst& c, double& p1, double& p2, double& p3, double* d) {
  if (c.mark2)
    wb_phi_partials_k<true>(n3, n2, vx, vy, vz, c, p1, p2, p3, d);
  else
    wb_phi_partials_k<false>(n3, n2, vx, vy, vz, c, p1, p2, p3, d);
}

// Tensor White-Bear energy density terms; v[11] = n3, n2, n2v_x/y/z, T_xx, T_yy, T_zz, T_xy, T_yz, T_zx.
__host__ __device__ inline void wbt_phi(const double* v, const WBConst& c, double& p1, double& p2, double& p3) {
#pragma clang fp contract(off)
  const double n3 = v[0], n2 = v[1], vx = v[2], vy = v[3], vz = v[4];
  const double txx = v[5], tyy = v[6], tzz = v[7], txy = v[8], tyz = v[9], tzx = v[10];
  const double om = 1.0 - n3;
  const double io = 1.0 / om;
  const double L = log1m(n3);
  const double v2 = vx * vx + vy * vy + vz * vz;
  double f3, df3;
  f3_of(n3, om, io, L, f3, df3);
  const double tvx = txx * vx + txy * vy + tzx * vz, tvy = txy * vx + tyy * vy + tyz * vz,
               tvz = tzx * vx + tyz * vy + tzz * vz;
  const double vtv = vx * tvx + vy * tvy + vz * tvz;
  const double tr3 = 6.0 * txy * tyz * tzx + txx * (txx * txx + 3.0 * txy * txy + 3.0 * tzx * tzx) +
                     tyy * (3.0 * txy * txy + tyy * tyy + 3.0 * tyz * tyz) +
                     tzz * (3.0 * tzx * tzx + 3.0 * tyz * tyz + tzz * tzz);
  const double X = n2 * n2 * n2 - 3.0 * n2 * v2 + 9.0 * (vtv - 0.5 * tr3);
  p1 = -n2 * L * c.inv_fpR2;
  p2 = (n2 * n2 - v2) * c.inv_fpR * io;
  p3 = f3 * X * kInv36Pi;
}

// Tensor White-Bear terms and the eleven partials.
__host__ __device__ inline void wbt_phi_partials(const double* v, const WBConst& c, double& p1, double& p2,
                                                 double& p3, double* d) {
  const double n3 = v[0], n2 = v[1], vx = v[2], vy = v[3], vz = v[4];
  const double txx = v[5], tyy = v[6], tzz = v[7], txy = v[8], tyz = v[9], tzx = v[10];
  const double om = 1.0 - n3;
  const double io = 1.0 / om;
  const double L = log1m(n3);
  const double v2 = vx * vx + vy * vy + vz * vz;
  double f3, df3;
  f3_of(n3, om, io, L, f3, df3);
  const double tvx = txx * vx + txy * vy + tzx * vz, tvy = txy * vx + tyy * vy + tyz * vz,
               tvz = tzx * vx + tyz * vy + tzz * vz;
  const double vtv = vx * tvx + vy * tvy + vz * tvz;
  const double tr3 = 6.0 * txy * tyz * tzx + txx * (txx * txx + 3.0 * txy * txy + 3.0 * tzx * tzx) +
                     tyy * (3.0 * txy * txy + tyy * tyy + 3.0 * tyz * tyz) +
                     tzz * (3.0 * tzx * tzx + 3.0 * tyz * tyz + tzz * tzz);
  const double X = n2 * n2 * n2 - 3.0 * n2 * v2 + 9.0 * (vtv - 0.5 * tr3);
  const double q = n2 * n2 - v2;
  const double a = c.inv_fpR * io;
  p1 = -n2 * L * c.inv_fpR2;
  p2 = q * a;
  p3 = f3 * X * kInv36Pi;
  const double cf = f3 * kInv36Pi;
  d[0] = n2 * c.inv_fpR2 * io + q * a * io + df3 * X * kInv36Pi;
  d[1] = -L * c.inv_fpR2 + 2.0 * n2 * a + cf * 3.0 * q;
  d[2] = -2.0 * a * vx + cf * (18.0 * tvx - 6.0 * n2 * vx);
  d[3] = -2.0 * a * vy + cf * (18.0 * tvy - 6.0 * n2 * vy);
  d[4] = -2.0 * a * vz + cf * (18.0 * tvz - 6.0 * n2 * vz);
  // (T^2) components
  const double sxx = txx * txx + txy * txy + tzx * tzx, syy = txy * txy + tyy * tyy + tyz * tyz,
               szz = tzx * tzx + tyz * tyz + tzz * tzz;
  const double sxy = txx * txy + txy * tyy + tzx * tyz, syz = txy * tzx + tyy * tyz + tyz * tzz,
               szx = txx * tzx + txy * tyz + tzx * tzz;
  d[5] = cf * (9.0 * vx * vx - 13.5 * sxx);
  d[6] = cf * (9.0 * vy * vy - 13.5 * syy);
  d[7] = cf * (9.0 * vz * vz - 13.5 * szz);
  d[8] = cf * (18.0 * vx * vy - 27.0 * sxy);
  d[9] = cf * (18.0 * vy * vz - 27.0 * syz);
  d[10] = cf * (18.0 * vz * vx - 27.0 * szx);
}

// Amplitude a of the traceless shell tensor weight W_ij = a (k^_i k^_j - delta_ij/3)
// (src/haskell/FMT.hs:82-88, src/ReciprocalOperators.h:292-404):
// a = s 4 pi R^2 A(kR), A(x) = ((x^2 - 3) sin x/x + 3 cos x)/x^2, which equals
// w2 - 3 w3/R from the scalar weights; below kR = 1 (where that difference
// cancels) the series A = sum_{m>=2} (-1)^(m+1) 4 m (m-1) x^(2m-2)/(2m+1)!.
// w3, w2 and the result carry the same 1/N factor `scale`.
constexpr double kTensorSeriesCutoff = 1.0;
template <typename T>
__device__ __forceinline__ T tensor_amplitude(T k, T R, T six_dr2, T w3, T w2, T scale) {
  const T x = k * R;
  if (x >= (T)kTensorSeriesCutoff) return w2 - (T)3 * w3 / R;
  const T x2 = x * x;
  // terms m = 2..13, Horner in x^2: coefficients c_m = (-1)^(m+1) 4 m (m-1)/(2m+1)!
  T A = (T)0;
#pragma unroll
  for (int m = 13; m >= 2; --m) {
    double f = 1.0;
    for (int i = 2; i <= 2 * m + 1; ++i) f *= i;
    const double cm = ((m & 1) ? 1.0 : -1.0) * 4.0 * m * (m - 1) / f;
    A = A * x2 + (T)cm;
  }
  A *= x2;  // the m = 2 term is -x^2/15
  const T s = exp(-six_dr2 * k * k);
  return s * (T)(4.0 * 3.14159265358979323846) * R * R * A * scale;
}

// Fourier transforms of the fundamental-measure weights at wave vector k
// (deft src/ReciprocalOperators.h:94-235, src/haskell/FMT.hs:21-65):
//   w3 = s 4 pi (sin kR - kR cos kR)/k^3, w2 = s 4 pi R sin(kR)/k,
//   w2v_j = i wv k_j with wv = s 4 pi (R cos kR - sin kR/k)/k^2,
// s = exp(-6 k^2 dr^2); for kR <= 1e-3 the reference's unsmeared power series.
__device__ __forceinline__ void fm_weights(double k, double R, double six_dr2, double& w3, double& w2,
                                           double& wv) {
  const double kR = k * R;
  const double fp = 4.0 * 3.14159265358979323846;
  if (kR > 1e-3) {
    double sn, cs;
    sincos(kR, &sn, &cs);
    const double s = exp(-six_dr2 * k * k);
    w3 = s * fp * (sn - kR * cs) / (k * k * k);
    w2 = s * fp * R * sn / k;
    wv = s * fp * (R * cs - sn / k) / (k * k);
  } else {
    const double kR2 = kR * kR;
    const double R3 = R * R * R;
    w3 = (fp / 3.0) * R3 * (kR2 * kR2 * kR2 * (-1.0 / 15120) + kR2 * kR2 * (1.0 / 280) + kR2 * (-1.0 / 10) + 1);
    w2 = fp * (R * R) * (-kR2 * kR2 * kR2 * (1.0 / 120 / 6 / 7) + kR2 * kR2 * (1.0 / 120) - kR2 * (1.0 / 6) + 1);
    wv = fp * R3 * (kR2 * kR2 * (1.0 / 5040 - 1.0 / 720) + kR2 * (1.0 / 24 - 1.0 / 120) - 1.0 / 3);
  }
}

}  // namespace deft
