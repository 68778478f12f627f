// Contact-density estimators of a converged density (the post-processing the
// contact-paper drivers run next to the density, papers/contact/figs/
// inner-sphere.cpp:220-233): n0, gSigmaS, YuWu_S, nA, gSigmaA and the Gross
// estimate (src/haskell/WhiteBear.hs:68-142, src/ContactDensity.cpp:362-400).
// FP64 throughout; the convolutions go through the plan's FFT seam.
#pragma once

#include "passes.hpp"

namespace deft {

enum SpecKernel : int {
  SK_COPY = 0,         // 1
  SK_STEP = 1,         // w3 of radius r (fm_weights)
  SK_SHELL = 2,        // w2 of radius r
  SK_VSHELL = 3,       // i wv(r) k_j
  SK_SHELL_PRIME = 4,  // s 4 pi (-sin kr/k - r cos kr), -8 pi r at k = 0 (src/haskell/FMT.hs:66-68)
  SK_VSHELL_PRIME = 5  // i s 4 pi r sin(kr)/k k_j, 0 at k = 0 (src/haskell/FMT.hs:70-72)
};

// out = (accumulate ? out : 0) + coef K(k) in over the valid half spectrum
static __global__ __launch_bounds__(kThreads) void k_spec_apply(GridArgs g, const cx<double>* __restrict__ in,
                                                                cx<double>* __restrict__ out, int kind, int comp,
                                                                double r, double coef, int accumulate) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long e = (long long)blockIdx.x * kThreads + threadIdx.x; e < nk; e += (long long)gridDim.x * kThreads) {
    const int kz = (int)(e % g.Nkz);
    const long long row = e / g.Nkz;
    const int y = (int)(row % g.Ny), x = (int)(row / g.Ny);
    const long long j = row * g.ldk + kz;
    const double mx = wrapf(x, g.Nx), my = wrapf(y, g.Ny), mz = kz;
    const double kv[3] = {mx * g.b[0] + my * g.b[3] + mz * g.b[6], mx * g.b[1] + my * g.b[4] + mz * g.b[7],
                          mx * g.b[2] + my * g.b[5] + mz * g.b[8]};
    const double k = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
    double kre = 0.0, kim = 0.0;  // kernel value
    if (kind == SK_COPY) {
      kre = 1.0;
    } else if (kind <= SK_VSHELL) {
      double w3, w2, wv;
      fm_weights(k, r, g.six_dr2, w3, w2, wv);
      if (kind == SK_STEP)
        kre = w3;
      else if (kind == SK_SHELL)
        kre = w2;
      else
        kim = wv * kv[comp];
    } else if (k > 0.0) {
      const double kr = k * r;
      double sn, cs;
      sincos(kr, &sn, &cs);
      const double s = exp(-g.six_dr2 * k * k);
      if (kind == SK_SHELL_PRIME)
        kre = s * 4.0 * M_PI * (-sn / k - r * cs);
      else
        kim = s * 4.0 * M_PI * r * sn / k * kv[comp];
    } else if (kind == SK_SHELL_PRIME) {
      kre = -8.0 * M_PI * r;
    }
    const cx<double> a = in[j];
    cx<double> v = {coef * (kre * a.re - kim * a.im), coef * (kre * a.im + kim * a.re)};
    if (accumulate) v = v + out[j];
    out[j] = v;
  }
}

// Local parts.  w: n3, n2, n2v_x/y/z (N each); out: n0, gSigmaS, YuWu_S, nA,
// (gSigmaA later), Gross, gSigmaS in the C++ form; abcd: the eight fields A, B, C_xyz, D_xyz whose
// convolutions make the gSigmaA numerator.  mark2: 0 White-Bear, else the
// Mark II g3 prefactor s3 (3/2 or 1).
static __global__ __launch_bounds__(kThreads) void k_contact_point(long long N, const double* __restrict__ w,
                                                                   const double* __restrict__ shell2,
                                                                   const double* __restrict__ step2, double R,
                                                                   double s3, double* __restrict__ out,
                                                                   double* __restrict__ abcd) {
  const double pi = M_PI;
  const double fpR = 4 * pi * R, fpR2 = 4 * pi * R * R;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads) {
    const double n3 = w[i], n2 = w[N + i], vx = w[2 * N + i], vy = w[3 * N + i], vz = w[4 * N + i];
    const double om = 1.0 - n3, io = 1.0 / om, L = log1m(n3);
    double g2, g2p, g3, g3p;
    if (s3 > 0.0) {
      markii_of(n3, io, L, s3, g2, g2p, g3, g3p);
    } else {
      f3_of(n3, om, io, L, g3, g3p);
      g2 = io;
      g2p = io * io;
    }
    const double n0 = n2 / fpR2, n1 = n2 / fpR;
    const double v2 = vx * vx + vy * vy + vz * vz;
    const double c36 = 1.0 / (36.0 * pi);
    const double vtt = n2 * (n2 * n2 - 3.0 * v2);
    // partials with n0, n1, n2, n3, n1v, n2v independent
    const double dn0 = -L;
    const double dn1 = g2 * n2;
    const double dn2 = g2 * n1 + g3 * (3.0 * n2 * n2 - 3.0 * v2) * c36;
    const double dn3 = n0 * io + g2p * (n1 * n2 - v2 / fpR) + g3p * vtt * c36;
    const double v[3] = {vx, vy, vz};
    double dn1v[3], dn2v[3], d2v_dot_v = 0.0, d1v_dot_v = 0.0;
    for (int c = 0; c < 3; ++c) {
      dn1v[c] = -g2 * v[c];
      dn2v[c] = -g2 * v[c] / fpR - g3 * 6.0 * n2 * v[c] * c36;
      d2v_dot_v += dn2v[c] * v[c];
      d1v_dot_v += dn1v[c] * v[c];
    }
    const double dAdR = dn3 * n2 + 2.0 * n2 * dn2 / R + n0 * dn1 + (2.0 / R) * d2v_dot_v + 3.0 * d1v_dot_v / (fpR * R);
    const double zeta2 = (R / 3.0) * n2;
    const double zeta = 1.0 - v2 / (n2 * n2);
    const double ghs = io * (1.0 + 0.5 * (io * zeta2) * zeta * (3.0 + io * zeta2));
    const double eta = step2[i] / 8.0;
    const double oe = 1.0 - eta;
    out[i] = n0;
    out[N + i] = dAdR / (n0 * n0 * 4.0 * pi * (2.0 * R) * (2.0 * R));
    out[2 * N + i] = ghs * zeta * zeta;
    out[3 * N + i] = shell2[i] / (4.0 * pi * (2.0 * R) * (2.0 * R));
    out[5 * N + i] = (1.0 - 0.5 * eta) / (oe * oe * oe);
    {
      // gSigmaS in the C++ by-hand form (n2' = -2 n2/R for every component;
      // the vector partials are c1 n2v and c2 n2v)
      const double n2p = -2.0 * n2 / R;
      const double c1 = -g2, c2 = -g2 / fpR - g3 * 6.0 * n2 * c36;
      double a = n2 * (dn3 - dn1 / fpR2 - dn0 / (2.0 * pi * R * R * R)) - n2p * (dn2 + dn1 / fpR + dn0 / fpR2);
      for (int c = 0; c < 3; ++c) a -= v[c] * (c2 * n2p + c1 * n2p / fpR + c1 * v[c] / fpR2);
      out[6 * N + i] = a / (4.0 * pi * (2.0 * R) * (2.0 * R)) / (n0 * n0);
    }
    abcd[i] = dn3 - dn1 / fpR2 - dn0 / (2.0 * pi * R * R * R);
    abcd[N + i] = dn2 + dn1 / fpR + dn0 / fpR2;
    for (int c = 0; c < 3; ++c) {
      abcd[(2 + c) * N + i] = dn2v[c] + dn1v[c] / fpR;
      abcd[(5 + c) * N + i] = dn1v[c] / fpR2;
    }
  }
}

static __global__ __launch_bounds__(kThreads) void k_divide(long long N, const double* __restrict__ num,
                                                            const double* __restrict__ den, double* __restrict__ out) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads)
    out[i] = num[i] / den[i];
}

}  // namespace deft
