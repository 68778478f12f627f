// Soft fundamental-measure theory (deft src/haskell/SFMT.hs): the softness
// constants, the Gaussian-smeared k-space weights and the sixteen-field Phi
// with its partials.  Host and device code: the host harness
// tests/host/sfmt_harness.cpp compiles this header with g++.
//
// Weighted densities, in this order (field index):
//   0 n3, 1 n2, 2 n1, 3 n0, 4-6 n2v (x, y, z), 7-9 n1v, 10-15 the tensor
//   (xx, yy, zz, xy, yz, zx).
// k-space weights (SFMT.hs:103-133), E = exp(-Xi^2 k^2/8), c = cos(k alpha/2),
// s = sin(k alpha/2), G = exp(-alpha^2/(2 Xi^2)) Re erf(k Xi/(2 sqrt 2) + i alpha/(sqrt 2 Xi)):
//   w3  = 4 pi/k^2 E ((1 + Xi^2 k^2/4) s/k - alpha/2 c)
//   w2  = 2 pi E (Xi^2/2 c + alpha s/k)
//   w1  = E s/k
//   w0  = sqrt(pi/2)/(k Xi) G
//   n2v: i k_j W2v, W2v = 4 pi/k^2 E ((Xi^2 k/4 + 1/k) s - alpha/2 c)   (= w3)
//   n1v: i k_j W1v, W1v = (sqrt(pi/2)/(k Xi) 2G - E c)/k^2
//   tensor: W2t (1 - 3 kh_i^2) on the diagonal, 3 W2t kh_i kh_j off it (SFMT.hs:64-69),
//           W2t = (-pi Xi^2/3 - 4 pi/k^2) E c - 2 pi alpha/(3k) E s + (2 pi)^1.5/(k^3 Xi) 2G
// with their k -> 0 limits at k = 0.
//
// G comes from the Faddeeva function w(z) = exp(-z^2) erfc(-iz):
//   exp(-v^2) Re erf(u + iv) = exp(-v^2) - exp(-u^2) Re[exp(-2iuv) conj(w(v + iu))],
// finite for every v where the generated code's exp(-v^2) erf(u + iv) overflows.
// w in the upper half plane is Weideman's rational approximation (J. A. C.
// Weideman, "Computation of the complex error function", SIAM J. Numer. Anal.
// 31 (1994) 1497-1518): w(z) ~ 2 p(Z)/(L - iz)^2 + pi^-1/2/(L - iz),
// Z = (L + iz)/(L - iz), L = 2^-1/4 sqrt(N), p of degree N-1 with coefficients
// from the discrete cosine sums of f(t) = exp(-t^2)(L^2 + t^2) at t = L tan(theta/2).
// With N = 40 terms it agrees with the Faddeeva package to ~2e-14 relative
// over the upper half plane.
#pragma once

#include <cmath>

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#define __device__
#endif
#endif

namespace deft {

constexpr int kWeidemanTerms = 40;

struct SFMTConst {
  double alpha, xi;
  double L;                      // Weideman's scale
  double a[kWeidemanTerms];      // coefficients a_1 .. a_N of p
};

// alpha = sigma (2/(1 + sqrt(ln 2 / (epsilon/kT))))^(1/6) (SFMT.hs:55-56)
inline double sfmt_alpha(double sigma, double epsilon, double kT) {
  const double betaeps = epsilon / kT;
  return sigma * std::pow(2.0 / (1.0 + std::sqrt(std::log(2.0) / betaeps)), 1.0 / 6.0);
}

inline SFMTConst make_sfmt_const(double sigma, double epsilon, double kT, double xi) {
  SFMTConst c;
  c.alpha = sfmt_alpha(sigma, epsilon, kT);
  c.xi = xi;
  const int N = kWeidemanTerms, M = 2 * N;
  c.L = std::sqrt(N / std::sqrt(2.0));
  // a_n = (1/2M) sum_{j=-M+1}^{M-1} f(L tan(j pi/2M)) cos(j n pi/M)
  for (int n = 1; n <= N; ++n) {
    double s = 0.0;
    for (int j = -M + 1; j <= M - 1; ++j) {
      const double t = c.L * std::tan(j * M_PI / (2.0 * M));
      s += std::exp(-t * t) * (c.L * c.L + t * t) * std::cos(j * n * M_PI / M);
    }
    c.a[n - 1] = s / (2.0 * M);
  }
  return c;
}

// w(x + iy), y >= 0 (Weideman): re, im
__host__ __device__ inline void faddeeva_w(double x, double y, const SFMTConst& c, double& re, double& im) {
  // L - iz = (L + y) - ix, L + iz = (L - y) + ix
  const double dr = c.L + y, di = -x;
  const double nr = c.L - y, ni = x;
  const double den = dr * dr + di * di;
  // Z = (L + iz)/(L - iz)
  const double Zr = (nr * dr + ni * di) / den, Zi = (ni * dr - nr * di) / den;
  double pr = c.a[kWeidemanTerms - 1], pi = 0.0;
  for (int n = kWeidemanTerms - 2; n >= 0; --n) {
    const double tr = pr * Zr - pi * Zi + c.a[n];
    pi = pr * Zi + pi * Zr;
    pr = tr;
  }
  // 1/(L - iz) and its square
  const double qr = dr / den, qi = -di / den;
  const double q2r = qr * qr - qi * qi, q2i = 2.0 * qr * qi;
  const double rsp = 0.56418958354775628695;  // 1/sqrt(pi)
  re = 2.0 * (pr * q2r - pi * q2i) + rsp * qr;
  im = 2.0 * (pr * q2i + pi * q2r) + rsp * qi;
}

// exp(-v^2) Re erf(u + iv), u >= 0
__host__ __device__ inline double sfmt_erf_term(double u, double v, const SFMTConst& c) {
  double wr, wi;
  faddeeva_w(v, u, c, wr, wi);
  const double cs = cos(2.0 * u * v), sn = sin(2.0 * u * v);
  return exp(-v * v) - exp(-u * u) * (cs * wr - sn * wi);
}

// the seven radial functions (w3, w2, w1, w0, W2v, W1v, W2t) at |k| = k
__host__ __device__ inline void sfmt_radial(double k, const SFMTConst& c, double (&r)[7]) {
  const double alpha = c.alpha, xi = c.xi;
  const double pi = 3.14159265358979323846;
  if (k == 0.0) {
    r[0] = pi * alpha * xi * xi / 2 + pi * alpha * alpha * alpha / 6;
    r[1] = pi * (xi * xi + alpha * alpha);
    r[2] = alpha / 2;
    r[3] = 0.5;
    r[4] = r[5] = r[6] = 0.0;
    return;
  }
  const double E = exp(-xi * xi * k * k / 8);
  const double cs = cos(k * alpha / 2), sn = sin(k * alpha / 2);
  const double sq2 = 1.41421356237309504880;
  const double G = sfmt_erf_term(k * xi / (2 * sq2), alpha / (sq2 * xi), c);
  const double k2 = k * k;
  const double rh = 1.25331413731550025121;  // sqrt(pi/2)
  r[0] = 4 * pi / k2 * E * ((1 + xi * xi * k2 / 4) * sn / k - alpha / 2 * cs);
  r[1] = 2 * pi * E * (xi * xi / 2 * cs + alpha * sn / k);
  r[2] = E * sn / k;
  r[3] = rh / (k * xi) * G;
  r[4] = 4 * pi / k2 * E * ((xi * xi * k / 4 + 1 / k) * sn - alpha / 2 * cs);
  r[5] = (rh / (k * xi) * 2 * G - E * cs) / k2;
  const double tp15 = 15.74960994572241892278;  // (2 pi)^1.5
  r[6] = (-pi * xi * xi / 3 - 4 * pi / k2) * E * cs - 2 * pi * alpha / (3 * k) * E * sn +
         tp15 / (k2 * k * xi) * 2 * G;
}

// log(1 - x): the same function the White-Bear pass uses (wb_point.hpp); here
// through log1p, which both compilers provide for host and device
__host__ __device__ inline double sfmt_log1m(double x) { return log1p(-x); }

// Phi1, Phi2, Phi3 per kT (SFMT.hs:13-27,71-91)
__host__ __device__ inline void sfmt_phi(const double (&v)[16], double& p1, double& p2, double& p3) {
  const double n3 = v[0], n2 = v[1], n1 = v[2], n0 = v[3];
  const double ux = v[4], uy = v[5], uz = v[6], qx = v[7], qy = v[8], qz = v[9];
  const double txx = v[10], tyy = v[11], tzz = v[12], txy = v[13], tyz = v[14], tzx = v[15];
  const double omn = 1.0 - n3;
  const double u2 = ux * ux + uy * uy + uz * uz, uq = ux * qx + uy * qy + uz * qz;
  const double tux = txx * ux + txy * uy + tzx * uz, tuy = txy * ux + tyy * uy + tyz * uz,
               tuz = tzx * ux + tyz * uy + tzz * uz;
  const double uTu = ux * tux + uy * tuy + uz * tuz;
  const double tr3 = txx * (txx * txx + 3 * txy * txy + 3 * tzx * tzx) + tyy * (tyy * tyy + 3 * txy * txy + 3 * tyz * tyz) +
                     tzz * (tzz * tzz + 3 * tzx * tzx + 3 * tyz * tyz) + 6 * txy * tyz * tzx;
  const double X = n2 * (n2 * n2 - 3 * u2) + 4.5 * (uTu - tr3);
  const double c3 = 1.0 / (24 * 3.14159265358979323846 * omn * omn);
  p1 = -n0 * sfmt_log1m(n3);
  p2 = (n2 * n1 - uq) / omn;
  p3 = X * c3;
}

// the sixteen partials of Phi1 + Phi2 + Phi3
__host__ __device__ inline void sfmt_phi_partials(const double (&v)[16], double (&d)[16]) {
  const double n3 = v[0], n2 = v[1], n1 = v[2], n0 = v[3];
  const double ux = v[4], uy = v[5], uz = v[6], qx = v[7], qy = v[8], qz = v[9];
  const double txx = v[10], tyy = v[11], tzz = v[12], txy = v[13], tyz = v[14], tzx = v[15];
  const double omn = 1.0 - n3, io = 1.0 / omn;
  const double u2 = ux * ux + uy * uy + uz * uz, uq = ux * qx + uy * qy + uz * qz;
  const double tux = txx * ux + txy * uy + tzx * uz, tuy = txy * ux + tyy * uy + tyz * uz,
               tuz = tzx * ux + tyz * uy + tzz * uz;
  const double uTu = ux * tux + uy * tuy + uz * tuz;
  const double tr3 = txx * (txx * txx + 3 * txy * txy + 3 * tzx * tzx) + tyy * (tyy * tyy + 3 * txy * txy + 3 * tyz * tyz) +
                     tzz * (tzz * tzz + 3 * tzx * tzx + 3 * tyz * tyz) + 6 * txy * tyz * tzx;
  const double X = n2 * (n2 * n2 - 3 * u2) + 4.5 * (uTu - tr3);
  const double c3 = io * io / (24 * 3.14159265358979323846);
  d[0] = n0 * io + (n2 * n1 - uq) * io * io + 2 * X * c3 * io;
  d[1] = n1 * io + 3 * (n2 * n2 - u2) * c3;
  d[2] = n2 * io;
  d[3] = -sfmt_log1m(n3);
  d[4] = -qx * io + (9 * tux - 6 * n2 * ux) * c3;
  d[5] = -qy * io + (9 * tuy - 6 * n2 * uy) * c3;
  d[6] = -qz * io + (9 * tuz - 6 * n2 * uz) * c3;
  d[7] = -ux * io;
  d[8] = -uy * io;
  d[9] = -uz * io;
  const double h = 4.5 * c3;
  d[10] = h * (ux * ux - 3 * (txx * txx + txy * txy + tzx * tzx));
  d[11] = h * (uy * uy - 3 * (txy * txy + tyy * tyy + tyz * tyz));
  d[12] = h * (uz * uz - 3 * (tzx * tzx + tyz * tyz + tzz * tzz));
  d[13] = h * (2 * ux * uy - 6 * (txx * txy + txy * tyy + tzx * tyz));
  d[14] = h * (2 * uy * uz - 6 * (txy * tzx + tyy * tyz + tyz * tzz));
  d[15] = h * (2 * uz * ux - 6 * (txx * tzx + txy * tyz + tzx * tzz));
}

// Homogeneous fluid: k = 0 weighted densities of a uniform density n
inline void sfmt_uniform_fields(double n, const SFMTConst& c, double (&v)[16], double (&w0)[4]) {
  double r[7];
  sfmt_radial(0.0, c, r);
  for (int i = 0; i < 16; ++i) v[i] = 0.0;
  for (int i = 0; i < 4; ++i) {
    w0[i] = r[i];
    v[i] = r[i] * n;
  }
}

}  // namespace deft
