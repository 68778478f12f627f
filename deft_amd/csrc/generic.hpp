// Passes for grid edges without fused kernels: any length from 1 to
// kMaxGenericLength whose prime factors are at most kMaxDirectRadix, as
// FFTW takes any N in the reference (src/GridDescription.cpp:26-30,
// src/Grid.cpp:470-478): 65 = 1 + int(6.4/0.1 + e/100), the 1 x 1 x 3601
// planar-wall lines of papers/contact/figs/walls.cpp:91-93.
//
// The fused passes of passes.hpp are compiled per edge; here the edge and the
// radix schedule of every axis are kernel arguments (LineSched), so the radix
// order is data the autotuner can choose.  An evaluation is split into line
// FFT passes over whole 3-D fields and pointwise passes:
//   density (moves, n, ideal/external/mu sums) -> r2c z, c2c y, c2c x
//   -> weight multiply (1 -> NW spectra) -> c2c x, c2c y, c2r z (NW fields)
//   -> Phi (+ partials) -> [grad: r2c z, c2c y, c2c x of NW fields
//   -> adjoint weight multiply-add -> c2c x, c2c y, c2r z -> chain rule]
// Axes of length 1 skip their c2c pass.  Real fields between the passes are
// FP64 (the values the FFT precision T produced, widened); every FFT reads
// them back at T.  Partial sums use the elementwise block count, one block
// per slot entry, in the same slots as the fused passes.
#pragma once

#include <algorithm>
#include <vector>

#include "passes.hpp"
#include "sfmt.hpp"
#include "sw_kernels.hpp"

namespace deft {

constexpr int kMaxGenericLength = 4096;
constexpr int kMaxStages = 12;  // 4096 = 2^12

// The radix schedule of one axis: stage s combines sub-transforms of length
// rad[0]*...*rad[s-1] into length rad[0]*...*rad[s] (Stockham order).
struct LineSched {
  int L;   // transform length
  int SS;  // LDS stride between sequences (skewed, odd)
  int nst; // stages (0 for L = 1)
  int rad[kMaxStages];
};

template <typename T>
constexpr int rt_lds_stride(int L) {
  return (L + (L >> lds_skew_shift<T>())) | 1;
}

// LDS capacity (complex values per buffer) of the two length buckets.
template <typename T>
constexpr int gen_cap(bool large) {
  return ((rt_lds_stride<T>(large ? kMaxGenericLength : 1024) + 7) / 8) * 8;
}
inline bool gen_large(int L) { return L > 1024; }

// --- host: radix schedules ---------------------------------------------------

// prime factors of n, ascending (none for n = 1)
inline std::vector<int> prime_factors(int n) {
  std::vector<int> f;
  for (int p = 2; p * p <= n; ++p)
    while (n % p == 0) {
      f.push_back(p);
      n /= p;
    }
  if (n > 1) f.push_back(n);
  return f;
}

// Edges the generic passes take: 1..kMaxGenericLength, prime factors <= kMaxDirectRadix.
inline bool generic_length_ok(int n) {
  if (n < 1 || n > kMaxGenericLength) return false;
  const std::vector<int> f = prime_factors(n);
  return f.empty() || f.back() <= kMaxDirectRadix;
}

// A schedule is valid for L when its radices (each >= 2, prime factors within
// kMaxDirectRadix) multiply to L in at most kMaxStages stages.
inline bool sched_valid(int L, const int *rad, int n) {
  if (n < 0 || n > kMaxStages) return false;
  long long prod = 1;
  for (int i = 0; i < n; ++i) {
    if (rad[i] < 2 || !generic_length_ok(rad[i])) return false;
    const std::vector<int> f = prime_factors(rad[i]);
    if (f.size() > 1 && rad[i] != 4 && rad[i] != 8) return false;  // composite radices: 4 and 8 only
    prod *= rad[i];
    if (prod > L) return false;
  }
  return prod == L;
}

inline LineSched make_sched(int L, const std::vector<int> &rad) {
  LineSched s{};
  s.L = L;
  s.nst = (int)rad.size();
  for (int i = 0; i < s.nst; ++i) s.rad[i] = rad[i];
  return s;
}

// Candidate schedules of L for the autotuner; the first is the default (the
// fused kernels' compile-time order: radix 8, then 4, 2, 3, 5, larger primes).
inline std::vector<std::vector<int>> radix_candidates(int L) {
  std::vector<std::vector<int>> out;
  if (L <= 1) {
    out.push_back({});
    return out;
  }
  const std::vector<int> f = prime_factors(L);
  int twos = 0;
  std::vector<int> odd;
  for (int p : f) {
    if (p == 2)
      ++twos;
    else
      odd.push_back(p);
  }
  auto pow2_part = [&](bool prefer8) {
    std::vector<int> r;
    int t = twos;
    if (prefer8) {
      while (t >= 3) r.push_back(8), t -= 3;
    }
    while (t >= 2) r.push_back(4), t -= 2;
    if (t == 1) r.push_back(2);
    return r;
  };
  auto add = [&](std::vector<int> r) {
    if ((int)r.size() > kMaxStages) return;
    for (const auto &o : out)
      if (o == r) return;
    out.push_back(std::move(r));
  };
  for (bool prefer8 : {true, false}) {
    std::vector<int> a = pow2_part(prefer8);
    a.insert(a.end(), odd.begin(), odd.end());  // 8/4/2 first, then ascending odd primes
    add(a);
    std::vector<int> b(a.rbegin(), a.rend());    // largest radices first
    add(b);
    std::vector<int> c(odd.begin(), odd.end());  // odd primes first
    const std::vector<int> t2 = pow2_part(prefer8);
    c.insert(c.end(), t2.begin(), t2.end());
    add(c);
  }
  return out;
}

// --- runtime Stockham stages -------------------------------------------------

// n / d for 0 <= n < 2^22 and a run-time d >= 1 without the integer division
// sequence: a float estimate (relative error below 2^-23, so off by at most
// one) and one correction each way.
struct FastDiv {
  int d;
  float inv;
};
__device__ __forceinline__ FastDiv fast_div(int d) { return FastDiv{d, 1.0f / (float)d}; }
__device__ __forceinline__ int fdiv(int n, const FastDiv& f) {
  int q = (int)((float)n * f.inv);
  if (q * f.d > n) --q;
  if ((q + 1) * f.d <= n) ++q;
  return q;
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void stage_rt(const cx<T>* __restrict__ src, cx<T>* __restrict__ dst, int nseq,
                                         const cx<T>* __restrict__ tw, int L, int SS, int NS) {
  const int M = L / R;
  const int TWS = L / (NS * R);
  const int total = nseq * M;
  const FastDiv dM = fast_div(M), dNS = fast_div(NS);
  for (int b = threadIdx.x; b < total; b += kBlockThreads) {
    const int s = fdiv(b, dM);
    const int j = b - s * M;
    const int jn = fdiv(j, dNS);
    const int k = j - jn * NS;
    const cx<T>* in = src + s * SS;
    cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[skew<T>(j + r * M)];
    if (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const cx<T> w = tw[k * r * TWS];
        v[r] = INV ? cmulc(v[r], w) : cmul(v[r], w);
      }
    }
    dft<T, R, INV>(v);
    cx<T>* out = dst + s * SS;
    const int o0 = jn * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out[skew<T>(o0 + r * NS)] = v[r];
  }
}

// Radix R without a register butterfly: out[q] = sum_r w_R^(r q) twiddled in[r].
template <typename T, bool INV>
__device__ __forceinline__ void stage_direct_rt(const cx<T>* __restrict__ src, cx<T>* __restrict__ dst, int nseq,
                                                const cx<T>* __restrict__ tw, int L, int SS, int R, int NS) {
  const int M = L / R;
  const int TWS = L / (NS * R);
  const int total = nseq * L;
  for (int e = threadIdx.x; e < total; e += kBlockThreads) {
    const int s = e / L;
    const int rem = e - s * L;
    const int q = rem / M;
    const int j = rem - q * M;
    const int k = j % NS;
    const cx<T>* in = src + s * SS;
    cx<T> acc = in[skew<T>(j)];
    int wq = q;
    for (int r = 1; r < R; ++r) {
      cx<T> v = in[skew<T>(j + r * M)];
      if (NS > 1) {
        const cx<T> w = tw[k * r * TWS];
        v = INV ? cmulc(v, w) : cmul(v, w);
      }
      const cx<T> w = tw[wq * M];
      acc = acc + (INV ? cmulc(v, w) : cmul(v, w));
      wq += q;
      if (wq >= R) wq -= R;
    }
    dst[s * SS + skew<T>((j / NS) * NS * R + k + q * NS)] = acc;
  }
}

// Transforms nseq sequences in `a` (stride s.SS) with schedule s; `b` is
// scratch of the same size; returns the buffer holding the result.  The data
// in `a` must be visible to the block.
template <typename T, bool INV>
__device__ __forceinline__ cx<T>* fft_lds_rt(cx<T>* a, cx<T>* b, int nseq, const cx<T>* __restrict__ tw,
                                             const LineSched& s) {
  int NS = 1;
  for (int st = 0; st < s.nst; ++st) {
    const int R = s.rad[st];
    switch (R) {
      case 2:
        stage_rt<T, 2, INV>(a, b, nseq, tw, s.L, s.SS, NS);
        break;
      case 3:
        stage_rt<T, 3, INV>(a, b, nseq, tw, s.L, s.SS, NS);
        break;
      case 4:
        stage_rt<T, 4, INV>(a, b, nseq, tw, s.L, s.SS, NS);
        break;
      case 5:
        stage_rt<T, 5, INV>(a, b, nseq, tw, s.L, s.SS, NS);
        break;
      case 8:
        stage_rt<T, 8, INV>(a, b, nseq, tw, s.L, s.SS, NS);
        break;
      default:
        stage_direct_rt<T, INV>(a, b, nseq, tw, s.L, s.SS, R, NS);
        break;
    }
    __syncthreads();
    cx<T>* t = a;
    a = b;
    b = t;
    NS *= R;
  }
  return a;
}

// --- line passes ---------------------------------------------------------------

// r2c along z of field blockIdx.y of `in` (real FP64, fields fs_in apart) into
// half spectrum blockIdx.y of `out`, times scl.  P line pairs per block.
template <typename T, int CAP>
__global__ __launch_bounds__(kThreads) void kg_zr2c(GridArgs g, LineSched s, int P, const double* __restrict__ in,
                                                    long long fs_in, cx<T>* __restrict__ out, double scl,
                                                    const cx<T>* __restrict__ tw) {
  __shared__ cx<T> A[CAP];
  __shared__ cx<T> B[CAP];
  const int L = s.L, SS = s.SS;
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)blockIdx.x * 2 * P;
  const double* src = in + blockIdx.y * fs_in;
  for (int e = threadIdx.x; e < 2 * P * L; e += kThreads) {
    const int l = e / L, z = e - l * L;
    const T v = line0 + l < lines ? (T)src[(line0 + l) * L + z] : (T)0;
    cx<T>* c = &A[(l >> 1) * SS + skew<T>(z)];
    if (l & 1)
      c->im = v;
    else
      c->re = v;
  }
  __syncthreads();
  const cx<T>* Z = fft_lds_rt<T, false>(A, B, P, tw, s);
  const int NKZ = L / 2 + 1;
  const T h = (T)(0.5 * scl);
  cx<T>* dst = out + blockIdx.y * g.ks;
  for (int e = threadIdx.x; e < 2 * P * NKZ; e += kThreads) {
    const int l = e / NKZ, k = e - l * NKZ;
    if (line0 + l < lines) {
      const cx<T>* zl = Z + (l >> 1) * SS;
      const cx<T> a = zl[skew<T>(k)];
      const cx<T> b = cconj(zl[skew<T>((L - k) % L)]);
      dst[(line0 + l) * g.ldk + k] = (l & 1) ? scale(mul_mi(a - b), h) : scale(a + b, h);
    }
  }
}

// c2r along z of half spectrum blockIdx.y of `in` into real field blockIdx.y
// of `out` (fields fs_out apart), times scl (the real part, as FFTW's c2r).
template <typename T, int CAP>
__global__ __launch_bounds__(kThreads) void kg_zc2r(GridArgs g, LineSched s, int P, const cx<T>* __restrict__ in,
                                                    double* __restrict__ out, long long fs_out, double scl,
                                                    const cx<T>* __restrict__ tw) {
  __shared__ cx<T> A[CAP];
  __shared__ cx<T> B[CAP];
  const int L = s.L, SS = s.SS;
  const int NKZ = L / 2 + 1;
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)blockIdx.x * 2 * P;
  const cx<T>* srcs = in + blockIdx.y * g.ks;
  for (int e = threadIdx.x; e < 2 * P * NKZ; e += kThreads) {
    const int l = e / NKZ, k = e - l * NKZ;
    A[e] = load_or_zero(srcs, (line0 + l) * g.ldk + k, 0, line0 + l < lines);
  }
  __syncthreads();
  // two half spectra -> one packed complex line Z = A + i B (Hermitian completion)
  for (int e = threadIdx.x; e < P * L; e += kThreads) {
    const int p = e / L, z = e - p * L;
    cx<T> a, b;
    if (2 * z <= L) {
      a = A[(2 * p) * NKZ + z];
      b = A[(2 * p + 1) * NKZ + z];
      if (z == 0 || 2 * z == L) {
        a.im = (T)0;
        b.im = (T)0;
      }
    } else {
      a = cconj(A[(2 * p) * NKZ + (L - z)]);
      b = cconj(A[(2 * p + 1) * NKZ + (L - z)]);
    }
    B[p * SS + skew<T>(z)] = {a.re - b.im, a.im + b.re};
  }
  __syncthreads();
  const cx<T>* R = fft_lds_rt<T, true>(B, A, P, tw, s);
  double* o = out + blockIdx.y * fs_out;
  for (int e = threadIdx.x; e < 2 * P * L; e += kThreads) {
    const int l = e / L, z = e - l * L;
    if (line0 + l < lines) {
      const cx<T> v = R[(l >> 1) * SS + skew<T>(z)];
      o[(line0 + l) * L + z] = scl * (double)((l & 1) ? v.im : v.re);
    }
  }
}

// c2c along y (axis 1: columns (x, kz), stride ldk) or x (axis 0: columns
// (y, kz), stride Ny*ldk) of half spectrum blockIdx.y of `data`, in place; C
// columns per block, consecutive kz in consecutive slots.
template <typename T, int CAP, bool INV>
__global__ __launch_bounds__(kThreads) void kg_col(GridArgs g, LineSched s, int C, int axis, cx<T>* __restrict__ data,
                                                   const cx<T>* __restrict__ tw) {
  __shared__ cx<T> A[CAP];
  __shared__ cx<T> B[CAP];
  const int L = s.L, SS = s.SS;
  const long long ncols = (long long)(axis == 1 ? g.Nx : g.Ny) * g.Nkz;
  const long long col0 = (long long)blockIdx.x * C;
  const int ncol = (int)min((long long)C, ncols - col0);
  const long long stride = axis == 1 ? (long long)g.ldk : (long long)g.Ny * g.ldk;
  const long long outer = axis == 1 ? (long long)g.Ny * g.ldk : (long long)g.ldk;
  cx<T>* base = data + blockIdx.y * g.ks;
  // each column's offset once (one 64-bit division per column, not per
  // element), and a fixed thread -> (column, first row) mapping: the rows of
  // a column are kThreads / C apart
  __shared__ long long coff[kThreads];
  if (threadIdx.x < ncol) {
    const long long v = col0 + threadIdx.x;
    const long long o = v / g.Nkz;
    coff[threadIdx.x] = o * outer + (v - o * g.Nkz);
  }
  __syncthreads();
  const int rows = kThreads / C;  // rows handled at once
  const int c = threadIdx.x % C, i0 = threadIdx.x / C;
  const bool act = i0 < rows && c < ncol;
  const long long co = act ? coff[c] : 0;
  if (act)
    for (int i = i0; i < L; i += rows) A[c * SS + skew<T>(i)] = base[co + i * stride];
  __syncthreads();
  const cx<T>* R = fft_lds_rt<T, INV>(A, B, ncol, tw, s);
  if (act)
    for (int i = i0; i < L; i += rows) base[co + i * stride] = R[c * SS + skew<T>(i)];
}

// --- pointwise passes (gridDim.x = the plan's elementwise block count) -----------

// CG moves, n = exp(-Veff/kT) into `n` (0 on hard walls), and the local energy
// sums; the arithmetic of zfwd_density_body.
template <typename T>
__global__ __launch_bounds__(kThreads) void kg_density(GridArgs g, double* __restrict__ x, const double* __restrict__ d,
                                                       double m1, double m2, int nmoves, const T* __restrict__ vext,
                                                       double* __restrict__ n, double* __restrict__ partial,
                                                       int pstride) {
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < g.N; i += (long long)gridDim.x * kThreads) {
    double xv = x[i];
    if (nmoves > 0) {
      const double dv = d[i];
      xv = __dadd_rn(xv, __dmul_rn(m1, dv));
      if (nmoves > 1) xv = __dadd_rn(xv, __dmul_rn(m2, dv));
      x[i] = xv;
    }
    const T ve = vext[i];
    double nv = 0.0;
    if (!isinf(ve)) {
      nv = exp(-xv / g.kT);
      acc[0] += -nv * (xv + g.kT * g.logNQ + g.kT);
      if (!isnan(ve)) acc[1] += (double)ve * nv;
      acc[2] += -g.mu * nv;
    }
    n[i] = nv;
  }
  block_sum_store<3>(acc, partial, SLOT_IDEAL, pstride, blockIdx.x);
}

// Weight product of field f for one spectral element (w3, w2, i w2v_j, tensor).
template <typename T>
__device__ __forceinline__ cx<T> weight_one(int f, const cx<T>& s, T w3, T w2, const T (&wk)[3], const T (&wt)[6]) {
  switch (f) {
    case 0:
      return scale(s, w3);
    case 1:
      return scale(s, w2);
    case 2:
    case 3:
    case 4:
      return mul_i(scale(s, wk[f - 2]));
    default:
      return scale(s, wt[f - 5]);
  }
}

template <typename T, int NW>
__device__ __forceinline__ void spectral_element(const GridArgs& g, long long i, long long& idx, T& w3, T& w2,
                                                 T (&wk)[3], T (&wt)[6]) {
  const int kz = (int)(i % g.Nkz);
  const long long r = i / g.Nkz;
  const int y = (int)(r % g.Ny), x = (int)(r / g.Ny);
  idx = r * g.ldk + kz;
  element_weights<T>(g, x, y, kz, w3, w2, wk);
#pragma unroll
  for (int j = 0; j < 6; ++j) wt[j] = (T)0;
  if constexpr (NW == 11) element_tensor_weights<T>(g, x, y, kz, w3, w2, wt);
}

// spectrum `in` -> NW weighted spectra out + f*ks
template <typename T, int NW>
__global__ __launch_bounds__(kThreads) void kg_wmul_fwd(GridArgs g, const cx<T>* __restrict__ in,
                                                        cx<T>* __restrict__ out) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    long long idx;
    T w3, w2, wk[3], wt[6];
    spectral_element<T, NW>(g, i, idx, w3, w2, wk, wt);
    const cx<T> s = in[idx];
#pragma unroll
    for (int f = 0; f < NW; ++f) out[f * g.ks + idx] = weight_one<T>(f, s, w3, w2, wk, wt);
  }
}

// sum over f of w_f(-k) in[f] (odd vector weights flip sign) -> out
template <typename T, int NW>
__global__ __launch_bounds__(kThreads) void kg_wmul_bwd(GridArgs g, const cx<T>* __restrict__ in,
                                                        cx<T>* __restrict__ out) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    long long idx;
    T w3, w2, wk[3], wt[6];
    spectral_element<T, NW>(g, i, idx, w3, w2, wk, wt);
    cx<T> acc = {(T)0, (T)0};
#pragma unroll
    for (int f = 0; f < NW; ++f) {
      const cx<T> p = weight_one<T>(f, in[f * g.ks + idx], w3, w2, wk, wt);
      acc = (f < 2 || f >= 5) ? acc + p : acc - p;
    }
    out[idx] = acc;
  }
}

// --- soft FMT (sfmt.hpp): sixteen weighted fields from a per-plan table ----

// The table: the seven radial functions of |k| (w3, w2, w1, w0, W2v, W1v, W2t)
// times 1/N for every valid spectral element, 8 doubles each, rows padded to ldk.
static __global__ __launch_bounds__(kThreads) void k_sfmt_table(GridArgs g, SFMTConst c, double* __restrict__ tab) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    const int kz = (int)(i % g.Nkz);
    const long long r = i / g.Nkz;
    const int y = (int)(r % g.Ny), x = (int)(r / g.Ny);
    const double mx = wrapf(x, g.Nx), my = wrapf(y, g.Ny), mz = kz;
    const double kx = mx * g.b[0] + my * g.b[3] + mz * g.b[6];
    const double ky = mx * g.b[1] + my * g.b[4] + mz * g.b[7];
    const double kz_ = mx * g.b[2] + my * g.b[5] + mz * g.b[8];
    double rad[7];
    sfmt_radial(sqrt(kx * kx + ky * ky + kz_ * kz_), c, rad);
    double* t = tab + 8 * (r * g.ldk + kz);
#pragma unroll
    for (int j = 0; j < 7; ++j) t[j] = rad[j] * g.invN;
    t[7] = 0.0;
  }
}

// The sixteen weights of one spectral element: scalars (w3, w2, w1, w0), the
// vector radial factors times k_j (the weights are i k_j W), and the tensor
// weights (W2t (1 - 3 kh_i^2) on the diagonal, 3 W2t kh_i kh_j off it).
template <typename T>
__device__ __forceinline__ void sfmt_element(const GridArgs& g, long long i, long long& idx, T (&wt)[16]) {
  const int kz = (int)(i % g.Nkz);
  const long long r = i / g.Nkz;
  const int y = (int)(r % g.Ny), x = (int)(r / g.Ny);
  idx = r * g.ldk + kz;
  const double mx = wrapf(x, g.Nx), my = wrapf(y, g.Ny), mz = kz;
  const double k[3] = {mx * g.b[0] + my * g.b[3] + mz * g.b[6], mx * g.b[1] + my * g.b[4] + mz * g.b[7],
                       mx * g.b[2] + my * g.b[5] + mz * g.b[8]};
  const double* t = g.stab + 8 * idx;
  const double kk = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
  const double ik2 = kk > 0.0 ? 1.0 / kk : 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) wt[j] = (T)t[j];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    wt[4 + j] = (T)(t[4] * k[j]);
    wt[7 + j] = (T)(t[5] * k[j]);
  }
  const double w2t = t[6];
  wt[10] = (T)(w2t * (1.0 - 3.0 * k[0] * k[0] * ik2));
  wt[11] = (T)(w2t * (1.0 - 3.0 * k[1] * k[1] * ik2));
  wt[12] = (T)(w2t * (1.0 - 3.0 * k[2] * k[2] * ik2));
  wt[13] = (T)(w2t * 3.0 * k[0] * k[1] * ik2);
  wt[14] = (T)(w2t * 3.0 * k[1] * k[2] * ik2);
  wt[15] = (T)(w2t * 3.0 * k[2] * k[0] * ik2);
}

// vector fields (4..9) carry i k_j: odd under k -> -k
__device__ __forceinline__ constexpr bool sfmt_odd(int f) { return f >= 4 && f < 10; }

template <typename T>
__global__ __launch_bounds__(kThreads) void kg_wmul_fwd_sfmt(GridArgs g, const cx<T>* __restrict__ in,
                                                             cx<T>* __restrict__ out) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    long long idx;
    T wt[16];
    sfmt_element<T>(g, i, idx, wt);
    const cx<T> s = in[idx];
#pragma unroll
    for (int f = 0; f < 16; ++f) out[f * g.ks + idx] = sfmt_odd(f) ? mul_i(scale(s, wt[f])) : scale(s, wt[f]);
  }
}

// sum over f of w_f(-k) in[f] -> out
template <typename T>
__global__ __launch_bounds__(kThreads) void kg_wmul_bwd_sfmt(GridArgs g, const cx<T>* __restrict__ in,
                                                             cx<T>* __restrict__ out) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    long long idx;
    T wt[16];
    sfmt_element<T>(g, i, idx, wt);
    cx<T> acc = {(T)0, (T)0};
#pragma unroll
    for (int f = 0; f < 16; ++f) {
      const cx<T> s = in[f * g.ks + idx];
      if (sfmt_odd(f))
        acc = acc - mul_i(scale(s, wt[f]));
      else
        acc = acc + scale(s, wt[f]);
    }
    out[idx] = acc;
  }
}

// Phi sums of the NW weighted densities in w (fields N apart) and, with GRAD,
// their partials written over them (the arithmetic of zphi_body; FN = 4 is
// the soft-FMT Phi of sfmt.hpp).
template <typename T, bool GRAD, int FN>
__global__ __launch_bounds__(kThreads) void kg_phi(GridArgs g, double* __restrict__ w, double* __restrict__ partial,
                                                   int pstride) {
  constexpr int NW = FN == 4 ? 16 : (FN == 3 ? 11 : 5);
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < g.N; i += (long long)gridDim.x * kThreads) {
    double v[NW];
#pragma unroll
    for (int f = 0; f < NW; ++f) v[f] = w[f * g.N + i];
    double p1 = 0.0, p2 = 0.0, p3 = 0.0;
    if constexpr (NW == 16) {
      sfmt_phi(v, p1, p2, p3);
      if constexpr (GRAD) {
        double dd[16];
        sfmt_phi_partials(v, dd);
#pragma unroll
        for (int f = 0; f < 16; ++f) w[f * g.N + i] = (double)(T)dd[f];
      }
    } else if constexpr (NW == 11) {
      wbt_phi(v, g.wb, p1, p2, p3);
      if constexpr (GRAD) {
        double dd[11], q1, q2, q3;
        wbt_phi_partials(v, g.wb, q1, q2, q3, dd);
#pragma unroll
        for (int f = 0; f < 11; ++f) w[f * g.N + i] = (double)(T)dd[f];
      }
    } else {
      wb_phi_k<FN == 1>(v[0], v[1], v[2], v[3], v[4], g.wb, p1, p2, p3);
      if constexpr (GRAD) {
        double dd[5], q1, q2, q3;
        wb_phi_partials_k<FN == 1>(v[0], v[1], v[2], v[3], v[4], g.wb, q1, q2, q3, dd);
#pragma unroll
        for (int f = 0; f < 5; ++f) w[f * g.N + i] = (double)(T)dd[f];
      }
    }
    acc[0] += p1;
    acc[1] += p2;
    acc[2] += p3;
  }
  block_sum_store<3>(acc, partial, SLOT_PHI1, pstride, blockIdx.x);
}

// dF/dn chain rule of zgrad_body from the back-convolution `conv` (real),
// plus `extra` (dF/dn of another term, e.g. the square-well attraction) where given.
template <typename T>
__global__ __launch_bounds__(kThreads) void kg_grad(GridArgs g, const double* __restrict__ conv,
                                                    const double* __restrict__ x, const T* __restrict__ vext,
                                                    const double* __restrict__ oldgrad, int have_old, int precond,
                                                    double* __restrict__ grad, double* __restrict__ pgrad,
                                                    double* __restrict__ partial, int pstride,
                                                    const double* __restrict__ extra) {
  double dots[2] = {0.0, 0.0};
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < g.N; i += (long long)gridDim.x * kThreads) {
    const double cv = conv[i];
    const double xv = x[i];
    const T ve = vext[i];
    double gr = 0.0, pg = 0.0;
    if (isfinite(ve)) {
      const double n = exp(-xv / g.kT);
      const double local = -xv - g.kT * g.logNQ + (double)ve - g.mu + (extra ? extra[i] : 0.0);
      const double dfdn = (g.kT * cv + local) * g.dV;
      gr = -dfdn * n / g.kT;
      pg = point_pgrad(precond, dfdn, n, xv, gr, g);
    }
    grad[i] = gr;
    pgrad[i] = pg;
    dots[0] += pg * gr;
    if (have_old) dots[1] += pg * oldgrad[i];
  }
  block_sum_store<2>(dots, partial, SLOT_PGG, pstride, blockIdx.x);
}

}  // namespace deft
