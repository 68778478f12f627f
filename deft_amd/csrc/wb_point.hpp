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
#pragma once

#include <hip/hip_runtime.h>

namespace deft {

struct WBConst {
  double fpR;    // 4 pi R
  double fpR2;   // 4 pi R^2
};

constexpr double kInv36Pi = 1.0 / (36.0 * 3.14159265358979323846);
constexpr double kF3SeriesCutoff = 0.05;
constexpr int kF3SeriesTerms = 18;

__host__ __device__ inline void f3_and_derivative(double n, double& f3, double& df3) {
  if (fabs(n) < kF3SeriesCutoff) {
    double g = 1.5, gp = 0.0, p = 1.0;  // p = n^(k-3)
#pragma unroll
    for (int k = 3; k < kF3SeriesTerms + 3; ++k) {
      const double c = 2.0 / (double(k) * (k - 1) * (k - 2));
      gp -= c * (k - 2) * p;
      p = p * n;
      g -= c * p;
    }
    const double om = 1.0 - n;
    f3 = g / (om * om);
    df3 = gp / (om * om) + 2.0 * g / (om * om * om);
  } else {
    const double om = 1.0 - n;
    const double L = log1p(-n);
    const double num = n + om * om * L;
    const double den = n * n * om * om;
    f3 = num / den;
    const double dnum = n - 2.0 * om * L;
    const double dden = 2.0 * n * om * (1.0 - 2.0 * n);
    df3 = dnum / den - num * dden / (den * den);
  }
}

// Energy density terms only.
__host__ __device__ inline void wb_phi(double n3, double n2, double vx, double vy, double vz, const WBConst& c,
                                       double& p1, double& p2, double& p3) {
  const double omn = 1.0 - n3;
  const double L = log1p(-n3);
  const double v2 = vx * vx + vy * vy + vz * vz;
  double f3, df3;
  f3_and_derivative(n3, f3, df3);
  const double t3 = n2 * (n2 * n2 - 3.0 * v2);
  p1 = -n2 * L / c.fpR2;
  p2 = (n2 * n2 - v2) / (c.fpR * omn);
  p3 = f3 * t3 * kInv36Pi;
}

// Energy density terms and the five partials (d3, d2, dvx, dvy, dvz).
__host__ __device__ inline void wb_phi_partials(double n3, double n2, double vx, double vy, double vz,
                                                const WBConst& c, double& p1, double& p2, double& p3,
                                                double* d) {
  const double omn = 1.0 - n3;
  const double L = log1p(-n3);
  const double v2 = vx * vx + vy * vy + vz * vz;
  double f3, df3;
  f3_and_derivative(n3, f3, df3);
  const double t3 = n2 * (n2 * n2 - 3.0 * v2);
  p1 = -n2 * L / c.fpR2;
  p2 = (n2 * n2 - v2) / (c.fpR * omn);
  p3 = f3 * t3 * kInv36Pi;
  d[0] = n2 / (c.fpR2 * omn) + (n2 * n2 - v2) / (c.fpR * omn * omn) + df3 * t3 * kInv36Pi;
  d[1] = -L / c.fpR2 + 2.0 * n2 / (c.fpR * omn) + f3 * (3.0 * n2 * n2 - 3.0 * v2) * kInv36Pi;
  const double cv = -2.0 / (c.fpR * omn) - f3 * 6.0 * n2 * kInv36Pi;
  d[2] = cv * vx;
  d[3] = cv * vy;
  d[4] = cv * vz;
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
