// Kernels of the square-well liquid's attraction (sw.hpp) on the generic
// line passes: pointwise work on FP64 fields of N values, k-space products on
// half spectra at the FFT precision T.  All kernel products carry 1/N, as the
// generic weight products do, so raw forward and inverse transforms give the
// continuous convolutions.
#pragma once

#include "passes.hpp"
#include "sw.hpp"

namespace deft {

// |k| and k of flat half-spectrum element i; idx is its padded position
__device__ __forceinline__ double sw_kvec(const GridArgs& g, long long i, long long& idx, double (&kv)[3]) {
  const int kz = (int)(i % g.Nkz);
  const long long r = i / g.Nkz;
  const int y = (int)(r % g.Ny), x = (int)(r / g.Ny);
  idx = r * g.ldk + kz;
  const double mx = wrapf(x, g.Nx), my = wrapf(y, g.Ny), mz = kz;
  kv[0] = mx * g.b[0] + my * g.b[3] + mz * g.b[6];
  kv[1] = mx * g.b[1] + my * g.b[4] + mz * g.b[7];
  kv[2] = mx * g.b[2] + my * g.b[5] + mz * g.b[8];
  return sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
}

// The gSigmaA kernels at k (times 1/N): shell w2(R), -shell'(R) (the kernel
// of B), the vector radial factors vshell'(R) and vshell(R) (the kernels are
// i times them times k_j), and the denominator's shell w2(2R)
// (src/haskell/FMT.hs:21-81, the same forms as the contact estimators).
struct SWGKernels {
  double shell, mshellp, vsp, wv, shell2;
};
__device__ __forceinline__ SWGKernels sw_gkernels(const GridArgs& g, double k, double R) {
  double w3, w2, wv, w3d, w2d, wvd;
  fm_weights(k, R, g.six_dr2, w3, w2, wv);
  fm_weights(k, 2.0 * R, g.six_dr2, w3d, w2d, wvd);
  SWGKernels K;
  const double fp = 4.0 * 3.14159265358979323846;
  if (k > 0.0) {
    double sn, cs;
    sincos(k * R, &sn, &cs);
    const double s = exp(-g.six_dr2 * k * k);
    K.mshellp = -s * fp * (-sn / k - R * cs);
    K.vsp = s * fp * R * sn / k;
  } else {
    K.mshellp = 2.0 * fp * R;  // -shell'(0) = 8 pi R
    K.vsp = 0.0;
  }
  K.shell = w2 * g.invN;
  K.mshellp *= g.invN;
  K.vsp *= g.invN;
  K.wv = wv * g.invN;
  K.shell2 = w2d * g.invN;
  return K;
}

// the eight numerator fields a (8 fields) from the White-Bear weighted densities w (5 fields)
static __global__ __launch_bounds__(kThreads) void k_sw_afields(long long N, double R, const double* __restrict__ w,
                                                                double* __restrict__ a) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads) {
    double v[8];
    sw_afields(w[i], w[N + i], w[2 * N + i], w[3 * N + i], w[4 * N + i], R, v);
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j * N + i] = v[j];
  }
}

// in place over the eight a-field spectra: field 0 <- the numerator
// shell(A) - shell'(B) + vshell'.C + vshell.D, field 1 <- shell_2R(n) from n~
template <typename T>
__global__ __launch_bounds__(kThreads) void kg_sw_num(GridArgs g, SWConst c, cx<T>* __restrict__ spec,
                                                      const cx<T>* __restrict__ nspec) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    long long idx;
    double kv[3];
    const double k = sw_kvec(g, i, idx, kv);
    const SWGKernels K = sw_gkernels(g, k, c.R);
    cx<T> s[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = spec[j * g.ks + idx];
    double re = K.shell * (double)s[0].re + K.mshellp * (double)s[1].re;
    double im = K.shell * (double)s[0].im + K.mshellp * (double)s[1].im;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      // i x (re + i im) = -x im + i x re
      const double xc = K.vsp * kv[d], xd = K.wv * kv[d];
      re += -xc * (double)s[2 + d].im - xd * (double)s[5 + d].im;
      im += xc * (double)s[2 + d].re + xd * (double)s[5 + d].re;
    }
    const cx<T> nv = nspec[idx];
    spec[idx] = {(T)re, (T)im};
    spec[g.ks + idx] = {(T)(K.shell2 * (double)nv.re), (T)(K.shell2 * (double)nv.im)};
  }
}

// density of the plan's rule (0 on hard walls)
template <typename T>
__device__ __forceinline__ double sw_density(const GridArgs& g, double x, T ve) {
  return isinf(ve) ? 0.0 : exp(-x / g.kT);
}

// g = num/den (0 where the density or den vanishes) and f_i = n P_i(g) (5 fields)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_sw_point_e(GridArgs g, const double* __restrict__ x,
                                                         const T* __restrict__ vext, const double* __restrict__ nd,
                                                         double* __restrict__ gout, double* __restrict__ f) {
  const long long N = g.N;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads) {
    const double n = sw_density(g, x[i], vext[i]);
    const double den = nd[N + i];
    const double gg = (n != 0.0 && den != 0.0) ? nd[i] / den : 0.0;
    double P[5], dP[5];
    sw_P(gg, P, dP);
    gout[i] = gg;
#pragma unroll
    for (int j = 0; j < 5; ++j) f[j * N + i] = n * P[j];
  }
}

// in place: field 0 <- sum_i Xi_i field_i
template <typename T>
__global__ __launch_bounds__(kThreads) void kg_sw_xi_sum(GridArgs g, SWConst c, cx<T>* __restrict__ spec) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    long long idx;
    double kv[3];
    const double k = sw_kvec(g, i, idx, kv);
    double xi[5];
    sw_xi(k, c, xi);
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const cx<T> s = spec[j * g.ks + idx];
      re += xi[j] * (double)s.re;
      im += xi[j] * (double)s.im;
    }
    spec[idx] = {(T)(re * g.invN), (T)(im * g.invN)};
  }
}

// out_i <- Xi_i n~ (5 fields)
template <typename T>
__global__ __launch_bounds__(kThreads) void kg_sw_xi_split(GridArgs g, SWConst c, const cx<T>* __restrict__ nspec,
                                                           cx<T>* __restrict__ out) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    long long idx;
    double kv[3];
    const double k = sw_kvec(g, i, idx, kv);
    double xi[5];
    sw_xi(k, c, xi);
    const cx<T> s = nspec[idx];
#pragma unroll
    for (int j = 0; j < 5; ++j)
      out[j * g.ks + idx] = {(T)(xi[j] * g.invN * (double)s.re), (T)(xi[j] * g.invN * (double)s.im)};
  }
}

// Energy partial sums (eps/2) n c into SLOT_SW
template <typename T>
__global__ __launch_bounds__(kThreads) void k_sw_energy(GridArgs g, SWConst c, const double* __restrict__ x,
                                                        const T* __restrict__ vext, const double* __restrict__ cf,
                                                        double* __restrict__ partial, int pstride) {
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < g.N; i += (long long)gridDim.x * kThreads)
    acc[0] += 0.5 * c.eps * sw_density(g, x[i], vext[i]) * cf[i];
  block_sum_store<1>(acc, partial, SLOT_SW, pstride, blockIdx.x);
}

// the direct part of dF_sw/dn into swg, and u = B/den, t = B g/den with
// B = (eps/2) n sum_i (Xi_i * n) P_i'(g)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_sw_point_g(GridArgs g, SWConst c, const double* __restrict__ x,
                                                         const T* __restrict__ vext, const double* __restrict__ gf,
                                                         const double* __restrict__ nd, const double* __restrict__ cf,
                                                         const double* __restrict__ kn, double* __restrict__ swg,
                                                         double* __restrict__ ut) {
  const long long N = g.N;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads) {
    const double n = sw_density(g, x[i], vext[i]);
    double P[5], dP[5];
    sw_P(gf[i], P, dP);
    double sp = 0.0, sd = 0.0;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const double kj = kn[j * N + i];
      sp += P[j] * kj;
      sd += dP[j] * kj;
    }
    swg[i] = 0.5 * c.eps * (cf[i] + sp);
    const double B = 0.5 * c.eps * n * sd;
    const double den = nd[N + i];
    const double u = (n != 0.0 && den != 0.0) ? B / den : 0.0;
    ut[i] = u;
    ut[N + i] = u * gf[i];
  }
}

// from u~ (field 0) and t~ (field 1): fields 0..7 <- conj(K_j) u~ (the
// adjoint numerator kernels), field 8 <- -shell_2R t~
template <typename T>
__global__ __launch_bounds__(kThreads) void kg_sw_adj(GridArgs g, SWConst c, cx<T>* __restrict__ spec) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    long long idx;
    double kv[3];
    const double k = sw_kvec(g, i, idx, kv);
    const SWGKernels K = sw_gkernels(g, k, c.R);
    const cx<T> uv = spec[idx], tv = spec[g.ks + idx];
    const double ur = (double)uv.re, ui = (double)uv.im;
    spec[idx] = {(T)(K.shell * ur), (T)(K.shell * ui)};
    spec[g.ks + idx] = {(T)(K.mshellp * ur), (T)(K.mshellp * ui)};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      // conj(i x) (ur + i ui) = -i x (ur + i ui) = x ui - i x ur
      const double xc = K.vsp * kv[d], xd = K.wv * kv[d];
      spec[(2 + d) * g.ks + idx] = {(T)(xc * ui), (T)(-xc * ur)};
      spec[(5 + d) * g.ks + idx] = {(T)(xd * ui), (T)(-xd * ur)};
    }
    spec[8 * g.ks + idx] = {(T)(-K.shell2 * (double)tv.re), (T)(-K.shell2 * (double)tv.im)};
  }
}

// r_m = sum_j (d a_j / d w_m) q_j, m = 0..4, by five dual-number passes
static __global__ __launch_bounds__(kThreads) void k_sw_jac(long long N, double R, const double* __restrict__ w,
                                                            const double* __restrict__ q, double* __restrict__ r) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads) {
    double wv[5], qv[8];
#pragma unroll
    for (int m = 0; m < 5; ++m) wv[m] = w[m * N + i];
#pragma unroll
    for (int j = 0; j < 8; ++j) qv[j] = q[j * N + i];
    for (int m = 0; m < 5; ++m) {
      Dual in[5];
#pragma unroll
      for (int t = 0; t < 5; ++t) in[t] = {wv[t], t == m ? 1.0 : 0.0};
      Dual a[8];
      sw_afields(in[0], in[1], in[2], in[3], in[4], R, a);
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 8; ++j) s += a[j].d * qv[j];
      r[m * N + i] = s;
    }
  }
}

// a += b over a half spectrum
template <typename T>
__global__ __launch_bounds__(kThreads) void kg_spec_add(GridArgs g, cx<T>* __restrict__ a, const cx<T>* __restrict__ b) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    const int kz = (int)(i % g.Nkz);
    const long long idx = (i / g.Nkz) * g.ldk + kz;
    a[idx] = a[idx] + b[idx];
  }
}

// a += b over N values
static __global__ __launch_bounds__(kThreads) void k_sw_add(long long N, double* __restrict__ a,
                                                            const double* __restrict__ b) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads)
    a[i] += b[i];
}

}  // namespace deft
