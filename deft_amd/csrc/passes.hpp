// HBM passes of the White-Bear energy / gradient evaluation on one grid.
//
// Layouts (deft src/Grid.h:37-42, FFTW r2c layout used by src/Grid.cpp:470-478):
//   real field      r[x][y][z],  index (x*Ny + y)*Nz + z,        N  = Nx*Ny*Nz
//   half spectrum   s[x][y][kz], index (x*Ny + y)*Nkz + kz,      Nk = Nx*Ny*(Nz/2+1)
// A 3-D transform is three line passes, each staged through LDS:
//   z pass: 2P contiguous real lines -> P complex FFTs of length Nz (two real
//           lines packed as re/im), split into two half spectra
//   y pass: C consecutive kz columns of one x plane, FFT along y (stride Nkz)
//   x pass: C consecutive (y,kz) columns, FFT along x (stride Ny*Nkz)
// The fundamental-measure multiply is fused into the x pass (forward x FFT,
// five weight products, five inverse x FFTs), Phi and its partials into the
// z pass (five c2r lines -> FP64 pointwise -> five r2c lines), and the
// gradient chain rule into the final c2r z pass.  Energies and CG dot
// products are FP64 per-block partial sums, reduced in a fixed order.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_core.hpp"
#include "wb_point.hpp"

namespace deft {

constexpr int kThreads = 256;

// reduction slots of the per-block partial-sum buffer
enum Slot : int {
  SLOT_IDEAL = 0,  // sum of -n (Veff + kT ln nQ + kT)
  SLOT_EXT = 1,    // sum of Vext n
  SLOT_MUN = 2,    // sum of -mu n
  SLOT_PHI1 = 3,
  SLOT_PHI2 = 4,
  SLOT_PHI3 = 5,
  SLOT_PGG = 6,    // pgrad . grad
  SLOT_PGO = 7,    // pgrad . oldgrad
  SLOT_GD = 8,     // grad . direction
  SLOT_NORM = 9,   // generic dot for tests
  NUM_SLOTS = 10
};

struct GridArgs {
  int Nx, Ny, Nz, Nkz;
  long long N, Nk;
  double b[9];  // reciprocal lattice vectors as rows: k = mx*b0 + my*b1 + mz*b2
  double R, six_dr2, invN;
  double kT, mu, logNQ, dV;
  WBConst wb;
};

// line pairs per z-pass block and columns per y/x-pass block
constexpr int zpairs(int L) { return (1024 / L) > 0 ? (1024 / L) : 1; }
constexpr int ccols(int L) { return (2048 / L) > 16 ? 16 : ((2048 / L) < 1 ? 1 : (2048 / L)); }

// Block-wide FP64 sums of NV values, stored to partial[(slot0+i)*stride + blockIdx.x].
template <int NV>
__device__ __forceinline__ void block_sum_store(double (&v)[NV], double* __restrict__ partial, int slot0,
                                                int stride) {
  __shared__ double red[NV][kThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = v[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) red[i][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) s += red[threadIdx.x][w];
    partial[(long long)(slot0 + threadIdx.x) * stride + blockIdx.x] = s;
  }
}

static __device__ __forceinline__ int wrapf(int m, int n) { return (m > n / 2) ? m - n : m; }

// k-vector length and weights (scaled by 1/N) of spectral element (x, y, kz).
template <typename T>
__device__ __forceinline__ void element_weights(const GridArgs& g, int x, int y, int kz, T& w3, T& w2, T (&wk)[3]) {
  const double mx = wrapf(x, g.Nx), my = wrapf(y, g.Ny), mz = kz;
  const double kx = mx * g.b[0] + my * g.b[3] + mz * g.b[6];
  const double ky = mx * g.b[1] + my * g.b[4] + mz * g.b[7];
  const double kz_ = mx * g.b[2] + my * g.b[5] + mz * g.b[8];
  const double k = sqrt(kx * kx + ky * ky + kz_ * kz_);
  double a3, a2, av;
  fm_weights(k, g.R, g.six_dr2, a3, a2, av);
  w3 = (T)(a3 * g.invN);
  w2 = (T)(a2 * g.invN);
  const double s = av * g.invN;
  wk[0] = (T)(s * kx);
  wk[1] = (T)(s * ky);
  wk[2] = (T)(s * kz_);
}

// Split P packed complex FFTs Z = FFT(a + i b) into the half spectra of a and b
// and store lines [line0, line0+2P) of `dst` (Nkz = L/2+1 per line).
template <typename T, int L, int P>
__device__ __forceinline__ void store_r2c_halves(const cx<T>* Z, cx<T>* __restrict__ dst, long long line0,
                                                 long long lines) {
  constexpr int NKZ = L / 2 + 1;
  for (int e = threadIdx.x; e < 2 * P * NKZ; e += kThreads) {
    const int l = e / NKZ, k = e - l * NKZ;
    if (line0 + l < lines) {
      const cx<T>* zl = Z + (l >> 1) * L;
      const cx<T> a = zl[k];
      const cx<T> b = cconj(zl[(L - k) % L]);
      dst[(line0 + l) * NKZ + k] = (l & 1) ? scale(mul_mi(a - b), (T)0.5) : scale(a + b, (T)0.5);
    }
  }
}

// Load lines [line0, line0+2P) of half spectra and build the P packed complex
// spectra Z = A + i B (Hermitian completion along z; the imaginary parts of the
// kz = 0 and kz = Nz/2 elements are dropped, i.e. c2r returns the real part).
// raw needs P*(L+2) elements, Z needs P*L.
template <typename T, int L, int P>
__device__ __forceinline__ void load_c2r_packed(const cx<T>* __restrict__ src, long long line0, long long lines,
                                                cx<T>* raw, cx<T>* Z) {
  constexpr int NKZ = L / 2 + 1;
  for (int e = threadIdx.x; e < 2 * P * NKZ; e += kThreads) {
    const int l = e / NKZ;
    cx<T> v = {(T)0, (T)0};
    if (line0 + l < lines) v = src[line0 * NKZ + e];
    raw[e] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < P * L; e += kThreads) {
    const int p = e / L, z = e - p * L;
    cx<T> a, b;
    if (z <= L / 2) {
      a = raw[(2 * p) * NKZ + z];
      b = raw[(2 * p + 1) * NKZ + z];
      if (z == 0 || z == L / 2) {
        a.im = (T)0;
        b.im = (T)0;
      }
    } else {
      a = cconj(raw[(2 * p) * NKZ + (L - z)]);
      b = cconj(raw[(2 * p + 1) * NKZ + (L - z)]);
    }
    Z[e] = {a.re - b.im, a.im + b.re};
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// z pass 1: (optional CG moves) Veff -> n = exp(-Veff/kT) -> r2c along z.
// Also accumulates the local energy terms (ideal gas, Vext n, -mu n).
// ---------------------------------------------------------------------------
template <typename T, int L>
__global__ __launch_bounds__(kThreads) void k_zfwd_density(GridArgs g, double* __restrict__ x,
                                                           const double* __restrict__ d, double m1, double m2,
                                                           int nmoves, const float* __restrict__ vext,
                                                           cx<T>* __restrict__ out, double* __restrict__ partial,
                                                           int pstride, const cx<T>* __restrict__ tw) {
  constexpr int P = zpairs(L);
  __shared__ cx<T> A[P * L];
  __shared__ cx<T> B[P * L];
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)blockIdx.x * 2 * P;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int e = threadIdx.x; e < 2 * P * L; e += kThreads) {
    const int l = e / L, z = e - l * L;
    double nv = 0.0;
    if (line0 + l < lines) {
      const long long idx = (line0 + l) * L + z;
      double xv = x[idx];
      if (nmoves > 0) {
        const double dv = d[idx];
        xv = __dadd_rn(xv, __dmul_rn(m1, dv));
        if (nmoves > 1) xv = __dadd_rn(xv, __dmul_rn(m2, dv));
        x[idx] = xv;
      }
      const float ve = vext[idx];
      if (isfinite(ve)) {
        nv = exp(-xv / g.kT);
        acc[0] += -nv * (xv + g.kT * g.logNQ + g.kT);
        acc[1] += (double)ve * nv;
        acc[2] += -g.mu * nv;
      }
    }
    cx<T>* c = &A[(l >> 1) * L + z];
    if (l & 1)
      c->im = (T)nv;
    else
      c->re = (T)nv;
  }
  __syncthreads();
  const cx<T>* Z = fft_lds<T, L, L, false>(A, B, P, tw);
  store_r2c_halves<T, L, P>(Z, out, line0, lines);
  block_sum_store<3>(acc, partial, SLOT_IDEAL, pstride);
}

// Plain r2c along z of a real FP64 field (FFT seam and tests).
template <typename T, int L>
__global__ __launch_bounds__(kThreads) void k_zfwd_plain(GridArgs g, const double* __restrict__ in,
                                                         cx<T>* __restrict__ out, const cx<T>* __restrict__ tw) {
  constexpr int P = zpairs(L);
  __shared__ cx<T> A[P * L];
  __shared__ cx<T> B[P * L];
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)blockIdx.x * 2 * P;
  for (int e = threadIdx.x; e < 2 * P * L; e += kThreads) {
    const int l = e / L, z = e - l * L;
    T v = (T)0;
    if (line0 + l < lines) v = (T)in[(line0 + l) * L + z];
    cx<T>* c = &A[(l >> 1) * L + z];
    if (l & 1)
      c->im = v;
    else
      c->re = v;
  }
  __syncthreads();
  const cx<T>* Z = fft_lds<T, L, L, false>(A, B, P, tw);
  store_r2c_halves<T, L, P>(Z, out, line0, lines);
}

// Plain c2r along z, out = scale * real part, for blockIdx.y fields.
template <typename T, int L>
__global__ __launch_bounds__(kThreads) void k_zinv_plain(GridArgs g, const cx<T>* __restrict__ in,
                                                         long long fstride_in, double* __restrict__ out,
                                                         long long fstride_out, double scl,
                                                         const cx<T>* __restrict__ tw) {
  constexpr int P = zpairs(L);
  __shared__ cx<T> A[P * (L + 2)];
  __shared__ cx<T> B[P * (L + 2)];
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)blockIdx.x * 2 * P;
  load_c2r_packed<T, L, P>(in + blockIdx.y * fstride_in, line0, lines, A, B);
  const cx<T>* R = fft_lds<T, L, L, true>(B, A, P, tw);
  double* o = out + blockIdx.y * fstride_out;
  for (int e = threadIdx.x; e < 2 * P * L; e += kThreads) {
    const int l = e / L, z = e - l * L;
    if (line0 + l < lines) {
      const cx<T> v = R[(l >> 1) * L + z];
      o[(line0 + l) * L + z] = scl * (double)((l & 1) ? v.im : v.re);
    }
  }
}

// ---------------------------------------------------------------------------
// y pass: FFT along y of blockIdx.y fields (in place).
// ---------------------------------------------------------------------------
template <typename T, int L, bool INV>
__global__ __launch_bounds__(kThreads) void k_ypass(GridArgs g, cx<T>* __restrict__ data, long long fstride,
                                                    const cx<T>* __restrict__ tw) {
  constexpr int C = ccols(L);
  constexpr int SS = L + 1;
  __shared__ cx<T> A[C * SS];
  __shared__ cx<T> B[C * SS];
  const int nch = (g.Nkz + C - 1) / C;
  const int x = blockIdx.x / nch, ch = blockIdx.x - x * nch;
  const int kz0 = ch * C;
  const int ncol = min(C, g.Nkz - kz0);
  cx<T>* base = data + blockIdx.y * fstride + (long long)x * L * g.Nkz + kz0;
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int y = e / C, c = e - y * C;
    cx<T> v = {(T)0, (T)0};
    if (c < ncol) v = base[(long long)y * g.Nkz + c];
    A[c * SS + y] = v;
  }
  __syncthreads();
  const cx<T>* R = fft_lds<T, L, SS, INV>(A, B, ncol, tw);
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int y = e / C, c = e - y * C;
    if (c < ncol) base[(long long)y * g.Nkz + c] = R[c * SS + y];
  }
}

// Plain FFT along x of blockIdx.y fields, in place, times `scl`.
template <typename T, int L, bool INV>
__global__ __launch_bounds__(kThreads) void k_xpass_plain(GridArgs g, cx<T>* __restrict__ data, long long fstride,
                                                          double scl, const cx<T>* __restrict__ tw) {
  constexpr int C = ccols(L);
  constexpr int SS = L + 1;
  __shared__ cx<T> A[C * SS];
  __shared__ cx<T> B[C * SS];
  const long long ncols = (long long)g.Ny * g.Nkz;
  const long long col0 = (long long)blockIdx.x * C;
  const int ncol = (int)min((long long)C, ncols - col0);
  cx<T>* base = data + blockIdx.y * fstride + col0;
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int xx = e / C, c = e - xx * C;
    cx<T> v = {(T)0, (T)0};
    if (c < ncol) v = base[(long long)xx * ncols + c];
    A[c * SS + xx] = v;
  }
  __syncthreads();
  const cx<T>* R = fft_lds<T, L, SS, INV>(A, B, ncol, tw);
  const T s = (T)scl;
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int xx = e / C, c = e - xx * C;
    if (c < ncol) base[(long long)xx * ncols + c] = scale(R[c * SS + xx], s);
  }
}

// ---------------------------------------------------------------------------
// x pass with the weight multiply: n~ -> n3, n2, n2v_x, n2v_y, n2v_z spectra.
// ---------------------------------------------------------------------------
template <typename T, int L>
__global__ __launch_bounds__(kThreads) void k_xconv_fwd(GridArgs g, const cx<T>* __restrict__ in,
                                                        cx<T>* __restrict__ out, long long fstride,
                                                        const cx<T>* __restrict__ tw) {
  constexpr int C = ccols(L);
  constexpr int SS = L + 1;
  constexpr int PT = (C * L + kThreads - 1) / kThreads;
  __shared__ cx<T> A[C * SS];
  __shared__ cx<T> B[C * SS];
  __shared__ cx<T> S2[C * SS];
  const long long ncols = (long long)g.Ny * g.Nkz;
  const long long col0 = (long long)blockIdx.x * C;
  const int ncol = (int)min((long long)C, ncols - col0);
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int xx = e / C, c = e - xx * C;
    cx<T> v = {(T)0, (T)0};
    if (c < ncol) v = in[(long long)xx * ncols + col0 + c];
    A[c * SS + xx] = v;
  }
  __syncthreads();
  cx<T>* S = fft_lds<T, L, SS, false>(A, B, ncol, tw);
  cx<T>* W1 = (S == A) ? B : A;
  T w3[PT], w2[PT], wk[PT][3];
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const int e = threadIdx.x + i * kThreads;
    const int xx = e / C, c = e - xx * C;
    w3[i] = w2[i] = wk[i][0] = wk[i][1] = wk[i][2] = (T)0;
    if (e < C * L && c < ncol) {
      const long long col = col0 + c;
      const int y = (int)(col / g.Nkz), kz = (int)(col - (long long)y * g.Nkz);
      element_weights<T>(g, xx, y, kz, w3[i], w2[i], wk[i]);
    }
  }
#pragma unroll 1
  for (int f = 0; f < 5; ++f) {
#pragma unroll
    for (int i = 0; i < PT; ++i) {
      const int e = threadIdx.x + i * kThreads;
      if (e < C * L) {
        const int xx = e / C, c = e - xx * C;
        const cx<T> s = S[c * SS + xx];
        cx<T> r;
        if (f == 0)
          r = scale(s, w3[i]);
        else if (f == 1)
          r = scale(s, w2[i]);
        else
          r = mul_i(scale(s, wk[i][f - 2]));
        W1[c * SS + xx] = r;
      }
    }
    __syncthreads();
    const cx<T>* R = fft_lds<T, L, SS, true>(W1, S2, ncol, tw);
    cx<T>* o = out + f * fstride + col0;
    for (int e = threadIdx.x; e < C * L; e += kThreads) {
      const int xx = e / C, c = e - xx * C;
      if (c < ncol) o[(long long)xx * ncols + c] = R[c * SS + xx];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// x pass of the back-convolution: sum_a w_a(-k) FFT_x(P_a) -> inverse FFT_x.
// Even weights keep their sign, odd (vector) weights flip it (deft
// src/Functional.h:407-417).
// ---------------------------------------------------------------------------
template <typename T, int L>
__global__ __launch_bounds__(kThreads) void k_xconv_bwd(GridArgs g, const cx<T>* __restrict__ in, long long fstride,
                                                        cx<T>* __restrict__ out, const cx<T>* __restrict__ tw) {
  constexpr int C = ccols(L);
  constexpr int SS = L + 1;
  constexpr int PT = (C * L + kThreads - 1) / kThreads;
  __shared__ cx<T> A[C * SS];
  __shared__ cx<T> B[C * SS];
  const long long ncols = (long long)g.Ny * g.Nkz;
  const long long col0 = (long long)blockIdx.x * C;
  const int ncol = (int)min((long long)C, ncols - col0);
  T w3[PT], w2[PT], wk[PT][3];
  cx<T> acc[PT];
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const int e = threadIdx.x + i * kThreads;
    const int xx = e / C, c = e - xx * C;
    w3[i] = w2[i] = wk[i][0] = wk[i][1] = wk[i][2] = (T)0;
    acc[i] = {(T)0, (T)0};
    if (e < C * L && c < ncol) {
      const long long col = col0 + c;
      const int y = (int)(col / g.Nkz), kz = (int)(col - (long long)y * g.Nkz);
      element_weights<T>(g, xx, y, kz, w3[i], w2[i], wk[i]);
    }
  }
#pragma unroll 1
  for (int f = 0; f < 5; ++f) {
    const cx<T>* src = in + f * fstride + col0;
    for (int e = threadIdx.x; e < C * L; e += kThreads) {
      const int xx = e / C, c = e - xx * C;
      cx<T> v = {(T)0, (T)0};
      if (c < ncol) v = src[(long long)xx * ncols + c];
      A[c * SS + xx] = v;
    }
    __syncthreads();
    const cx<T>* S = fft_lds<T, L, SS, false>(A, B, ncol, tw);
#pragma unroll
    for (int i = 0; i < PT; ++i) {
      const int e = threadIdx.x + i * kThreads;
      if (e < C * L) {
        const int xx = e / C, c = e - xx * C;
        const cx<T> s = S[c * SS + xx];
        if (f == 0)
          acc[i] = acc[i] + scale(s, w3[i]);
        else if (f == 1)
          acc[i] = acc[i] + scale(s, w2[i]);
        else
          acc[i] = acc[i] + mul_mi(scale(s, wk[i][f - 2]));
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const int e = threadIdx.x + i * kThreads;
    if (e < C * L) {
      const int xx = e / C, c = e - xx * C;
      A[c * SS + xx] = acc[i];
    }
  }
  __syncthreads();
  const cx<T>* R = fft_lds<T, L, SS, true>(A, B, ncol, tw);
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int xx = e / C, c = e - xx * C;
    if (c < ncol) out[(long long)xx * ncols + col0 + c] = R[c * SS + xx];
  }
}

// ---------------------------------------------------------------------------
// z pass 2: c2r of the five weighted densities -> FP64 Phi (energy partial
// sums) and, with GRAD, the five partials -> r2c back into the same buffers.
// ---------------------------------------------------------------------------
template <typename T, int L, bool GRAD>
__global__ __launch_bounds__(kThreads) void k_zphi(GridArgs g, cx<T>* __restrict__ fields, long long fstride,
                                                   double* __restrict__ partial, int pstride,
                                                   const cx<T>* __restrict__ tw) {
  constexpr int P = zpairs(L);
  constexpr int NPT = 2 * P * L;
  constexpr int PT = (NPT + kThreads - 1) / kThreads;
  __shared__ cx<T> A[P * (L + 2)];
  __shared__ cx<T> B[P * (L + 2)];
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)blockIdx.x * 2 * P;
  T val[5][PT];
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    load_c2r_packed<T, L, P>(fields + f * fstride, line0, lines, A, B);
    const cx<T>* R = fft_lds<T, L, L, true>(B, A, P, tw);
#pragma unroll
    for (int i = 0; i < PT; ++i) {
      const int e = threadIdx.x + i * kThreads;
      val[f][i] = (T)0;
      if (e < NPT) {
        const int l = e / L, z = e - l * L;
        const cx<T> v = R[(l >> 1) * L + z];
        val[f][i] = (l & 1) ? v.im : v.re;
      }
    }
    __syncthreads();
  }
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const int e = threadIdx.x + i * kThreads;
    const bool valid = (e < NPT) && (line0 + e / L < lines);
    double p1 = 0.0, p2 = 0.0, p3 = 0.0;
    if constexpr (GRAD) {
      double dd[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (valid)
        wb_phi_partials((double)val[0][i], (double)val[1][i], (double)val[2][i], (double)val[3][i],
                        (double)val[4][i], g.wb, p1, p2, p3, dd);
#pragma unroll
      for (int f = 0; f < 5; ++f) val[f][i] = (T)dd[f];
    } else {
      if (valid)
        wb_phi((double)val[0][i], (double)val[1][i], (double)val[2][i], (double)val[3][i], (double)val[4][i], g.wb,
               p1, p2, p3);
    }
    acc[0] += p1;
    acc[1] += p2;
    acc[2] += p3;
  }
  if constexpr (GRAD) {
#pragma unroll
    for (int f = 0; f < 5; ++f) {
#pragma unroll
      for (int i = 0; i < PT; ++i) {
        const int e = threadIdx.x + i * kThreads;
        if (e < NPT) {
          const int l = e / L, z = e - l * L;
          cx<T>* c = &A[(l >> 1) * L + z];
          if (l & 1)
            c->im = val[f][i];
          else
            c->re = val[f][i];
        }
      }
      __syncthreads();
      const cx<T>* Z = fft_lds<T, L, L, false>(A, B, P, tw);
      store_r2c_halves<T, L, P>(Z, fields + f * fstride, line0, lines);
      __syncthreads();
    }
  }
  block_sum_store<3>(acc, partial, SLOT_PHI1, pstride);
}

// ---------------------------------------------------------------------------
// z pass 3: c2r of the back-convolution -> dF/dn -> grad and pgrad (FP64),
// with pgrad.grad and pgrad.oldgrad partial sums.
// ---------------------------------------------------------------------------
template <typename T, int L>
__global__ __launch_bounds__(kThreads) void k_zgrad(GridArgs g, const cx<T>* __restrict__ accs,
                                                    const double* __restrict__ x, const float* __restrict__ vext,
                                                    const double* __restrict__ oldgrad, int have_old, int precond,
                                                    double* __restrict__ grad, double* __restrict__ pgrad,
                                                    double* __restrict__ partial, int pstride,
                                                    const cx<T>* __restrict__ tw) {
  constexpr int P = zpairs(L);
  __shared__ cx<T> A[P * (L + 2)];
  __shared__ cx<T> B[P * (L + 2)];
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)blockIdx.x * 2 * P;
  load_c2r_packed<T, L, P>(accs, line0, lines, A, B);
  const cx<T>* R = fft_lds<T, L, L, true>(B, A, P, tw);
  double dots[2] = {0.0, 0.0};
  for (int e = threadIdx.x; e < 2 * P * L; e += kThreads) {
    const int l = e / L, z = e - l * L;
    if (line0 + l < lines) {
      const long long idx = (line0 + l) * L + z;
      const cx<T> v = R[(l >> 1) * L + z];
      const double conv = (double)((l & 1) ? v.im : v.re);
      const double xv = x[idx];
      const float ve = vext[idx];
      double gr = 0.0, pg = 0.0;
      if (isfinite(ve)) {
        const double n = exp(-xv / g.kT);
        const double local = -xv - g.kT * g.logNQ + (double)ve - g.mu;
        const double dfdn = (g.kT * conv + local) * g.dV;
        gr = -dfdn * n / g.kT;
        pg = precond ? -dfdn / g.kT : gr;
      }
      grad[idx] = gr;
      pgrad[idx] = pg;
      dots[0] += pg * gr;
      if (have_old) dots[1] += pg * oldgrad[idx];
    }
  }
  block_sum_store<2>(dots, partial, SLOT_PGG, pstride);
}

// ---------------------------------------------------------------------------
// CG vector kernels (FP64)
// ---------------------------------------------------------------------------

// direction = -pg + beta*direction (or -pg), oldgrad = grad, partial grad.direction
static __global__ __launch_bounds__(kThreads) void k_cg_update(long long N, const double* __restrict__ pg,
                                                        const double* __restrict__ gr, double* __restrict__ dir,
                                                        double* __restrict__ oldgrad, double beta, int have_dir,
                                                        double* __restrict__ partial, int pstride) {
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads) {
    const double gi = gr[i];
    double di = -pg[i];
    if (have_dir) di = __dadd_rn(di, __dmul_rn(beta, dir[i]));
    dir[i] = di;
    oldgrad[i] = gi;
    acc[0] += gi * di;
  }
  block_sum_store<1>(acc, partial, SLOT_GD, pstride);
}

// direction = -grad, partial grad.direction
static __global__ __launch_bounds__(kThreads) void k_cg_reset(long long N, const double* __restrict__ gr,
                                                       double* __restrict__ dir, double* __restrict__ partial,
                                                       int pstride) {
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads) {
    const double gi = gr[i];
    dir[i] = -gi;
    acc[0] += gi * -gi;
  }
  block_sum_store<1>(acc, partial, SLOT_GD, pstride);
}

// x = x + s*d, rounded like the host expression x += s*d
static __global__ __launch_bounds__(kThreads) void k_axpy(long long N, double* __restrict__ x, const double* __restrict__ d,
                                                   double s) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads)
    x[i] = __dadd_rn(x[i], __dmul_rn(s, d[i]));
}

// partial a.b into SLOT_NORM
static __global__ __launch_bounds__(kThreads) void k_dot(long long N, const double* __restrict__ a,
                                                  const double* __restrict__ b, double* __restrict__ partial,
                                                  int pstride) {
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads)
    acc[0] += a[i] * b[i];
  block_sum_store<1>(acc, partial, SLOT_NORM, pstride);
}

// n = exp(-Veff/kT), 0 on hard walls
static __global__ __launch_bounds__(kThreads) void k_density(long long N, const double* __restrict__ x,
                                                             const float* __restrict__ vext, double kT,
                                                             double* __restrict__ n) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads)
    n[i] = isfinite(vext[i]) ? exp(-x[i] / kT) : 0.0;
}

// out += a (accumulating into caller buffers, as deft's grad does)
static __global__ __launch_bounds__(kThreads) void k_accumulate(long long N, double* __restrict__ out,
                                                         const double* __restrict__ a) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads)
    out[i] += a[i];
}

// out[slot] = sum_{b < nb} partial[slot*pstride + b] for slots [slot0, slot0 + gridDim.x)
static __global__ __launch_bounds__(kThreads) void k_reduce(const double* __restrict__ partial, int pstride, int slot0,
                                                     int nb, double* __restrict__ out) {
  const int slot = slot0 + blockIdx.x;
  const double* p = partial + (long long)slot * pstride;
  double acc[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += kThreads) acc[0] += p[b];
  __shared__ double red[kThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double s = acc[0];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) red[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
    out[slot] = t;
  }
}

}  // namespace deft
