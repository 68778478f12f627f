// HBM passes of the White-Bear energy / gradient evaluation on one grid.
//
// Layouts (deft src/Grid.h:37-42, FFTW r2c layout used by src/Grid.cpp:470-478):
//   real field      r[x][y][z],  index (x*Ny + y)*Nz + z,        N  = Nx*Ny*Nz
//   half spectrum   s[x][y][kz], index (x*Ny + y)*ldk + kz,      kz < Nkz = Nz/2+1
// The device keeps half-spectrum rows padded to ldk = Nkz rounded up to 16
// complex values, so every row starts on a 128-byte line; the padding is never
// read or written.  (The FFT seam of the C ABI converts to the compact layout.)
// A 3-D transform is three line passes, each staged through LDS:
//   z pass: 2P contiguous real lines -> P complex FFTs of length Nz (two real
//           lines packed as re/im), split into two half spectra
//   y pass: C consecutive kz columns of one x plane, FFT along y (stride ldk);
//           the columns past the last full chunk of C (the Nyquist column for
//           Nz = 128) are gathered from C different (x, kz) pairs per block
//   x pass: C consecutive valid (y,kz) columns, FFT along x (stride Ny*ldk)
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

constexpr int kThreads = kBlockThreads;

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
  SLOT_SW = 9,     // square-well liquid: (eps/2) n (Xi * n P(g))
  NUM_SLOTS = 10
};

struct GridArgs {
  const double* wtab;  // FP64 weight table: (w3/N, w2/N) per spectral element, rows padded to ldk
  const float* wtab32;  // the same table rounded to FP32, for the FP32 passes
  const double* stab;  // soft FMT: the seven radial weights /N per spectral element (8 doubles), rows padded to ldk
  int Nx, Ny, Nz, Nkz;
  int ldk;        // padded half-spectrum row stride (complex values)
  long long N;    // real points
  long long ks;   // complex values per stored half spectrum, Nx*Ny*ldk
  double b[9];    // reciprocal lattice vectors as rows: k = mx*b0 + my*b1 + mz*b2
  double R, six_dr2, invN;
  double kT, mu, logNQ, dV;
  double sk_floor;  // structure-factor preconditioner: densities below keep grad/n
  WBConst wb;
};

constexpr int padded_row(int nkz) { return (nkz + 15) / 16 * 16; }

// Line pairs per z-pass block (a block holds about ZB complex values) and
// columns per y/x-pass block (a tile holds about TB bytes); both are chosen per
// plan by the autotuner among the instantiated values.
constexpr int kZBlocks[2] = {512, 1024};
constexpr int kTileBytes[2] = {8192, 16384};
constexpr int zpairs(int L, int ZB = 1024) { return (ZB / L) > 0 ? (ZB / L) : 1; }
// (a y/x tile holds C columns of L complex values, about 16 KiB)
template <typename T>
constexpr int ccols(int L, int TB) {
  const int c = TB / (L * 2 * (int)sizeof(T));
  return c > 16 ? 16 : (c < 1 ? 1 : (c >= 8 ? (c >= 16 ? 16 : 8) : (c >= 4 ? 4 : (c >= 2 ? 2 : 1))));
}
static_assert(256 % ccols<float>(20, 16384) == 0 && ccols<double>(128, 8192) == 4, "column tiles");

// Block-wide FP64 sums of NV values, stored to partial[(slot0+i)*stride + bx].
// The first NWV waves of the block hold the values (any further waves' are
// not counted); every thread of the block must call it.
template <int NV, int NWV = kThreads / 64>
__device__ __forceinline__ void block_sum_store(double (&v)[NV], double* __restrict__ partial, int slot0,
                                                int stride, int bx) {
  __shared__ double red[NV][NWV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (wave < NWV) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = v[i];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
      if (lane == 0) red[i][wave] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NWV; ++w) s += red[threadIdx.x][w];
    partial[(long long)(slot0 + threadIdx.x) * stride + bx] = s;
  }
}

static __device__ __forceinline__ int wrapf(int m, int n) { return (m > n / 2) ? m - n : m; }

// Blocks are dealt to the 8 XCDs round robin (block b runs on XCD b % 8), and
// each XCD has its own L2.  An x-pass tile of 8 KiB reads 64 of every 128
// bytes of a row; the next tile reads the other half.  This mapping gives
// tiles 2k and 2k+1 to blocks 16q + r and 16q + r + 8, which share an XCD, so
// the second half of each line is an L2 hit instead of a second HBM fetch.
// The last (nb mod 16) blocks keep their index.
constexpr int kXcds = 8;
static __device__ __forceinline__ int xcd_pair_tile(int b, int nb) {
  const int full = nb / (2 * kXcds) * (2 * kXcds);
  if (b >= full) return b;
  const int q = b / (2 * kXcds), r = b - q * 2 * kXcds;
  return q * 2 * kXcds + 2 * (r % kXcds) + r / kXcds;
}

// p[i] if ok, else 0; the load itself always goes to a valid address (safe)
// (a conditionally assigned complex<double> otherwise lands in scratch)
template <typename T>
__device__ __forceinline__ cx<T> load_or_zero(const cx<T>* __restrict__ p, long long i, long long safe, bool ok) {
  const cx<T> v = p[ok ? i : safe];
  return {ok ? v.re : (T)0, ok ? v.im : (T)0};
}

// Offset of valid column v (= y*Nkz + kz) inside one x plane.
static __device__ __forceinline__ long long col_offset(const GridArgs& g, long long v) {
  const long long y = v / g.Nkz;
  return y * g.ldk + (v - y * g.Nkz);
}

// Component d of the k-vector (mx, my, mz) . b in R, with contraction off so
// that its rounding does not depend on the kernel it is inlined into (the
// x-pass variants recompute it).
template <typename R>
__device__ __forceinline__ R kvec_component(const GridArgs& g, R mx, R my, R mz, int d) {
#pragma clang fp contract(off)
  return mx * (R)g.b[d] + my * (R)g.b[3 + d] + mz * (R)g.b[6 + d];
}

// Weights (scaled by 1/N) of spectral element (x, y, kz): w3, w2 and the
// vector weight's i-coefficient times k_j, which is -w3 k_j.  w3 and w2 come
// from the plan's table, computed once in FP64 (FP64 reads it as is, FP32 its
// FP32 rounding): evaluated per element, their sines, cosines and
// exponentials had dominated the x passes' VALU time in both precisions.
template <typename T>
__device__ __forceinline__ void element_weights(const GridArgs& g, int x, int y, int kz, T& w3, T& w2, T (&wk)[3]) {
  if constexpr (sizeof(T) <= 4) {
    const float mx = (float)wrapf(x, g.Nx), my = (float)wrapf(y, g.Ny), mz = (float)kz;
    const float kx = kvec_component<float>(g, mx, my, mz, 0);
    const float ky = kvec_component<float>(g, mx, my, mz, 1);
    const float kz_ = kvec_component<float>(g, mx, my, mz, 2);
    const long long idx = ((long long)x * g.Ny + y) * g.ldk + kz;
    const float2 w = reinterpret_cast<const float2*>(g.wtab32)[idx];
    w3 = w.x;
    w2 = w.y;
    wk[0] = -w3 * kx;
    wk[1] = -w3 * ky;
    wk[2] = -w3 * kz_;
  } else {
    const double mx = wrapf(x, g.Nx), my = wrapf(y, g.Ny), mz = kz;
    const double kx = kvec_component<double>(g, mx, my, mz, 0);
    const double ky = kvec_component<double>(g, mx, my, mz, 1);
    const double kz_ = kvec_component<double>(g, mx, my, mz, 2);
    const long long idx = ((long long)x * g.Ny + y) * g.ldk + kz;
    const double2 w = reinterpret_cast<const double2*>(g.wtab)[idx];
    const double a3 = w.x;
    w3 = (T)a3;
    w2 = (T)w.y;
    wk[0] = (T)(-a3 * kx);
    wk[1] = (T)(-a3 * ky);
    wk[2] = (T)(-a3 * kz_);
  }
}

// FP64 weight table (w3/N, w2/N) of every valid spectral element, built once per
// plan, and its FP32 rounding (either output may be null).
static __global__ __launch_bounds__(kThreads) void k_weight_table(GridArgs g, double* __restrict__ tab,
                                                                  float* __restrict__ tab32) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    const int kz = (int)(i % g.Nkz);
    const long long r = i / g.Nkz;
    const int y = (int)(r % g.Ny), x = (int)(r / g.Ny);
    const double mx = wrapf(x, g.Nx), my = wrapf(y, g.Ny), mz = kz;
    const double kx = mx * g.b[0] + my * g.b[3] + mz * g.b[6];
    const double ky = mx * g.b[1] + my * g.b[4] + mz * g.b[7];
    const double kz_ = mx * g.b[2] + my * g.b[5] + mz * g.b[8];
    double a3, a2, av;
    fm_weights(sqrt(kx * kx + ky * ky + kz_ * kz_), g.R, g.six_dr2, a3, a2, av);
    const long long j = r * g.ldk + kz;  // padded rows, (w3, w2) interleaved
    if (tab) {
      tab[2 * j] = a3 * g.invN;
      tab[2 * j + 1] = a2 * g.invN;
    }
    if (tab32) {
      tab32[2 * j] = (float)(a3 * g.invN);
      tab32[2 * j + 1] = (float)(a2 * g.invN);
    }
  }
}

// Split P packed complex FFTs Z = FFT(a + i b) into the half spectra of a and b
// and store lines [line0, line0+2P) of `dst` (Nkz = L/2+1 values per line).
template <typename T, int L, int P>
__device__ __forceinline__ void store_r2c_halves(const cx<T>* Z, cx<T>* __restrict__ dst, long long line0,
                                                 long long lines, int ldk) {
  constexpr int NKZ = L / 2 + 1;
  constexpr int SS = lds_stride<T, L>();
  for (int e = threadIdx.x; e < 2 * P * NKZ; e += kThreads) {
    const int l = e / NKZ, k = e - l * NKZ;
    if (line0 + l < lines) {
      const cx<T>* zl = Z + (l >> 1) * SS;
      const cx<T> a = zl[skew<T>(k)];
      const cx<T> b = cconj(zl[skew<T>((L - k) % L)]);
      dst[(line0 + l) * ldk + k] = (l & 1) ? scale(mul_mi(a - b), (T)0.5) : scale(a + b, (T)0.5);
    }
  }
}

// Load lines [line0, line0+2P) of half spectra and build the P packed complex
// spectra Z = A + i B (Hermitian completion along z; the imaginary parts of the
// kz = 0 and, for even Nz, kz = Nz/2 elements are dropped, i.e. c2r returns the
// real part).
// Both buffers need zbuf_size<T, L, P>() elements.
template <typename T, int L, int P>
constexpr int zbuf_size() {
  return P * ((L + 2) > lds_stride<T, L>() ? (L + 2) : lds_stride<T, L>());
}

template <typename T, int L, int P>
__device__ __forceinline__ void pack_c2r(const cx<T>* raw, cx<T>* Z);

template <typename T, int L, int P>
__device__ __forceinline__ void load_c2r_packed(const cx<T>* __restrict__ src, long long line0, long long lines,
                                                int ldk, cx<T>* raw, cx<T>* Z) {
  constexpr int NKZ = L / 2 + 1;
  for (int e = threadIdx.x; e < 2 * P * NKZ; e += kThreads) {
    const int l = e / NKZ, k = e - l * NKZ;
    raw[e] = load_or_zero(src, (line0 + l) * ldk + k, 0, line0 + l < lines);
  }
  __syncthreads();
  pack_c2r<T, L, P>(raw, Z);
}

// The second half of load_c2r_packed: the raw half-spectrum rows in `raw`
// (visible to the block) -> packed complex lines in Z, then a barrier.
template <typename T, int L, int P>
__device__ __forceinline__ void pack_c2r(const cx<T>* raw, cx<T>* Z) {
  constexpr int NKZ = L / 2 + 1;
  constexpr int SS = lds_stride<T, L>();
  for (int e = threadIdx.x; e < P * L; e += kThreads) {
    const int p = e / L, z = e - p * L;
    cx<T> a, b;
    if (z <= L / 2) {
      a = raw[(2 * p) * NKZ + z];
      b = raw[(2 * p + 1) * NKZ + z];
      if (z == 0 || (L % 2 == 0 && z == L / 2)) {
        a.im = (T)0;
        b.im = (T)0;
      }
    } else {
      a = cconj(raw[(2 * p) * NKZ + (L - z)]);
      b = cconj(raw[(2 * p + 1) * NKZ + (L - z)]);
    }
    Z[p * SS + skew<T>(z)] = {a.re - b.im, a.im + b.re};
  }
  __syncthreads();
}

// pack_c2r for nf fields at once: field q's raw rows at raw + q * 2P(L/2+1),
// its packed lines as sequences q*P .. q*P+P-1 of Z; one barrier.
template <typename T, int L, int P>
__device__ __forceinline__ void pack_c2r_fields(const cx<T>* raw, cx<T>* Z, int nf) {
  constexpr int NKZ = L / 2 + 1;
  constexpr int SS = lds_stride<T, L>();
  for (int e = threadIdx.x; e < nf * P * L; e += kThreads) {
    const int pq = e / L, z = e - pq * L;
    const int q = pq / P, p = pq - q * P;
    const cx<T>* r = raw + q * (2 * P * NKZ);
    cx<T> a, b;
    if (z <= L / 2) {
      a = r[(2 * p) * NKZ + z];
      b = r[(2 * p + 1) * NKZ + z];
      if (z == 0 || (L % 2 == 0 && z == L / 2)) {
        a.im = (T)0;
        b.im = (T)0;
      }
    } else {
      a = cconj(r[(2 * p) * NKZ + (L - z)]);
      b = cconj(r[(2 * p + 1) * NKZ + (L - z)]);
    }
    Z[pq * SS + skew<T>(z)] = {a.re - b.im, a.im + b.re};
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// z pass 1: (optional CG moves) Veff -> n = exp(-Veff/kT) -> r2c along z.
// Also accumulates the local energy terms (ideal gas, Vext n, -mu n).
// ---------------------------------------------------------------------------
template <typename T, int L, int ZB>
__device__ __forceinline__ void zfwd_density_body(int bx, const GridArgs& g, double* __restrict__ x,
                                                  const double* __restrict__ d, double m1, double m2, int nmoves,
                                                  const T* __restrict__ vext, cx<T>* __restrict__ out,
                                                  double* __restrict__ partial, int pstride,
                                                  const cx<T>* __restrict__ tw) {
  constexpr int P = zpairs(L, ZB);
  constexpr int SS = lds_stride<T, L>();
  __shared__ cx<T> A[P * SS];
  __shared__ cx<T> B[P * SS];
  const cx<T>* twl = lds_twiddles<T, L>(tw);
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)bx * 2 * P;
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
      const T ve = vext[idx];
      // +inf: hard wall, n = 0; NaN: constrained point (the old-API
      // constrain(mask, f), src/Functional.cpp:759-838): its density counts in
      // the energy, without external potential, and its gradient is zero
      if (!isinf(ve)) {
        nv = exp(-xv / g.kT);
        acc[0] += -nv * (xv + g.kT * g.logNQ + g.kT);
        if (!isnan(ve)) acc[1] += (double)ve * nv;
        acc[2] += -g.mu * nv;
      }
    }
    cx<T>* c = &A[(l >> 1) * SS + skew<T>(z)];
    if (l & 1)
      c->im = (T)nv;
    else
      c->re = (T)nv;
  }
  __syncthreads();
  const cx<T>* Z = fft_lds<T, L, SS, false>(A, B, P, twl);
  store_r2c_halves<T, L, P>(Z, out, line0, lines, g.ldk);
  block_sum_store<3>(acc, partial, SLOT_IDEAL, pstride, bx);
}

template <typename T, int L, int ZB>
__global__ __launch_bounds__(kThreads) void k_zfwd_density(GridArgs g, double* __restrict__ x,
                                                           const double* __restrict__ d, double m1, double m2,
                                                           int nmoves, const T* __restrict__ vext,
                                                           cx<T>* __restrict__ out, double* __restrict__ partial,
                                                           int pstride, const cx<T>* __restrict__ tw) {
  zfwd_density_body<T, L, ZB>(blockIdx.x, g, x, d, m1, m2, nmoves, vext, out, partial, pstride, tw);
}

// Plain r2c along z of a real FP64 field (FFT seam and tests).
template <typename T, int L, int ZB>
__global__ __launch_bounds__(kThreads) void k_zfwd_plain(GridArgs g, const double* __restrict__ in,
                                                         cx<T>* __restrict__ out, const cx<T>* __restrict__ tw) {
  constexpr int P = zpairs(L, ZB);
  constexpr int SS = lds_stride<T, L>();
  __shared__ cx<T> A[P * SS];
  __shared__ cx<T> B[P * SS];
  const cx<T>* twl = lds_twiddles<T, L>(tw);
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)blockIdx.x * 2 * P;
  for (int e = threadIdx.x; e < 2 * P * L; e += kThreads) {
    const int l = e / L, z = e - l * L;
    T v = (T)0;
    if (line0 + l < lines) v = (T)in[(line0 + l) * L + z];
    cx<T>* c = &A[(l >> 1) * SS + skew<T>(z)];
    if (l & 1)
      c->im = v;
    else
      c->re = v;
  }
  __syncthreads();
  const cx<T>* Z = fft_lds<T, L, SS, false>(A, B, P, twl);
  store_r2c_halves<T, L, P>(Z, out, line0, lines, g.ldk);
}

// Plain c2r along z, out = scale * real part, for blockIdx.y fields.
template <typename T, int L, int ZB>
__global__ __launch_bounds__(kThreads) void k_zinv_plain(GridArgs g, const cx<T>* __restrict__ in,
                                                         double* __restrict__ out, long long fstride_out, double scl,
                                                         const cx<T>* __restrict__ tw) {
  constexpr int P = zpairs(L, ZB);
  constexpr int SS = lds_stride<T, L>();
  __shared__ cx<T> A[zbuf_size<T, L, P>()];
  __shared__ cx<T> B[zbuf_size<T, L, P>()];
  const cx<T>* twl = lds_twiddles<T, L>(tw);
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)blockIdx.x * 2 * P;
  load_c2r_packed<T, L, P>(in + blockIdx.y * g.ks, line0, lines, g.ldk, A, B);
  const cx<T>* R = fft_lds<T, L, SS, true>(B, A, P, twl);
  double* o = out + blockIdx.y * fstride_out;
  for (int e = threadIdx.x; e < 2 * P * L; e += kThreads) {
    const int l = e / L, z = e - l * L;
    if (line0 + l < lines) {
      const cx<T> v = R[(l >> 1) * SS + skew<T>(z)];
      o[(line0 + l) * L + z] = scl * (double)((l & 1) ? v.im : v.re);
    }
  }
}

// ---------------------------------------------------------------------------
// y pass: FFT along y of NF consecutive fields (in place, field = blockIdx.y).  Blocks below
// Nx*(Nkz/C) take C consecutive kz columns of one x plane (128-byte aligned
// rows); the remaining blocks take the columns past the last full chunk, C
// (x, kz) pairs per block.
// ---------------------------------------------------------------------------
inline int y_main_blocks(int Nx, int Nkz, int C) { return Nx * (Nkz / C); }
inline int y_tail_blocks(int Nx, int Nkz, int C) {
  const int ntail = Nkz - (Nkz / C) * C;
  return (Nx * ntail + C - 1) / C;
}

template <typename T, int L, bool INV, int TB>
__device__ __forceinline__ void ypass_body(int bx, int by, const GridArgs& g, cx<T>* __restrict__ data,
                                           const cx<T>* __restrict__ tw) {
  constexpr int C = ccols<T>(L, TB);
  constexpr int SS = lds_stride<T, L>();
  __shared__ cx<T> A[C * SS];
  __shared__ cx<T> B[C * SS];
  const cx<T>* twl = lds_twiddles<T, L>(tw);
  const int nfull = g.Nkz / C;
  const int ntail = g.Nkz - nfull * C;
  const int nmain = g.Nx * nfull;
  const long long plane = (long long)L * g.ldk;
  int nseq;
  // per-sequence base offsets (x plane + kz) for this thread's column slot
  const int myc = threadIdx.x % C;  // kThreads is a multiple of C
  long long off_c = 0;
  if (bx < nmain) {
    const int x = bx / nfull, ch = bx - x * nfull;
    nseq = C;
    off_c = (long long)x * plane + ch * C + myc;
  } else {
    const int q0 = (bx - nmain) * C;
    nseq = min(C, g.Nx * ntail - q0);
    const int q = q0 + myc;
    if (myc < nseq) off_c = (long long)(q / ntail) * plane + nfull * C + (q % ntail);
  }
  cx<T>* base = data + by * g.ks;
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int y = e / C;  // e % C == myc
    A[myc * SS + skew<T>(y)] = load_or_zero(base, off_c + (long long)y * g.ldk, (long long)y * g.ldk, myc < nseq);
  }
  __syncthreads();
  const cx<T>* R = fft_lds<T, L, SS, INV>(A, B, nseq, twl);
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int y = e / C;
    if (myc < nseq) base[off_c + (long long)y * g.ldk] = R[myc * SS + skew<T>(y)];
  }
}

template <typename T, int L, bool INV, int NF, int TB>
__global__ __launch_bounds__(kThreads) void k_ypass(GridArgs g, cx<T>* __restrict__ data,
                                                    const cx<T>* __restrict__ tw) {
  ypass_body<T, L, INV, TB>(blockIdx.x, blockIdx.y, g, data, tw);  // gridDim.y == NF
}

// Plain FFT along x of one field, in place, times `scl`.
template <typename T, int L, bool INV, int TB>
__global__ __launch_bounds__(kThreads) void k_xpass_plain(GridArgs g, cx<T>* __restrict__ data, double scl,
                                                          const cx<T>* __restrict__ tw) {
  constexpr int C = ccols<T>(L, TB);
  constexpr int SS = lds_stride<T, L>();
  __shared__ cx<T> A[C * SS];
  __shared__ cx<T> B[C * SS];
  const cx<T>* twl = lds_twiddles<T, L>(tw);
  const long long ncols = (long long)g.Ny * g.Nkz;
  const long long plane = (long long)g.Ny * g.ldk;
  const long long col0 = (long long)blockIdx.x * C;
  const int ncol = (int)min((long long)C, ncols - col0);
  const int myc = threadIdx.x % C;
  const long long off = myc < ncol ? col_offset(g, col0 + myc) : 0;
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int xx = e / C;
    A[myc * SS + skew<T>(xx)] = load_or_zero(data, (long long)xx * plane + off, 0, myc < ncol);
  }
  __syncthreads();
  const cx<T>* R = fft_lds<T, L, SS, INV>(A, B, ncol, twl);
  const T s = (T)scl;
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int xx = e / C;
    if (myc < ncol) data[(long long)xx * plane + off] = scale(R[myc * SS + skew<T>(xx)], s);
  }
}

// Column tile of an x-pass block, aligned like the y pass: blocks below
// Ny*(Nkz/C) take C consecutive kz columns of one y row (row-aligned, so a tile
// reads whole 64- or 128-byte segments); the remaining blocks gather the
// columns past the last full chunk of each row, C (y, kz) pairs per block.
// Returns this thread's column (y, kz), whether it exists (`ok`), and the
// number of columns of the block.  The block count equals ceil(Ny*Nkz/C).
template <int C>
__device__ __forceinline__ int x_tile_column(const GridArgs& g, int bx, int myc, int& y, int& kz, bool& ok) {
  const int nfull = g.Nkz / C;
  const int ntail = g.Nkz - nfull * C;
  const int nmain = g.Ny * nfull;
  if (bx < nmain) {
    y = bx / nfull;
    kz = (bx - y * nfull) * C + myc;
    ok = true;
    return C;
  }
  const int q0 = (bx - nmain) * C;
  const int ncol = min(C, g.Ny * ntail - q0);
  ok = myc < ncol;
  const int q = q0 + (ok ? myc : 0);
  y = q / ntail;
  kz = nfull * C + (q - y * ntail);
  return ncol;
}

// Weight products of field f (0: w3, 1: w2, 2..4: i w2v_j) for a thread's elements.
template <typename T, int PT>
__device__ __forceinline__ cx<T> weight_mul(int f, int i, const cx<T>& s, const T (&w3)[PT], const T (&w2)[PT],
                                            const T (&wk)[PT][3], const T (&wt)[PT][6]) {
  switch (f) {
    case 0:
      return scale(s, w3[i]);
    case 1:
      return scale(s, w2[i]);
    case 2:
      return mul_i(scale(s, wk[i][0]));
    case 3:
      return mul_i(scale(s, wk[i][1]));
    case 4:
      return mul_i(scale(s, wk[i][2]));
    default:
      return scale(s, wt[i][f - 5]);
  }
}

// The six shell-tensor weights (xx, yy, zz, xy, yz, zx; scaled by 1/N) of
// spectral element (x, y, kz) given its scaled w3 and w2.
template <typename T>
__device__ __forceinline__ void element_tensor_weights(const GridArgs& g, int x, int y, int kz, T w3, T w2,
                                                       T (&wt)[6]) {
  const double mx = wrapf(x, g.Nx), my = wrapf(y, g.Ny), mz = kz;
  const T kx = (T)(mx * g.b[0] + my * g.b[3] + mz * g.b[6]);
  const T ky = (T)(mx * g.b[1] + my * g.b[4] + mz * g.b[7]);
  const T kz_ = (T)(mx * g.b[2] + my * g.b[5] + mz * g.b[8]);
  const T k = sqrt(kx * kx + ky * ky + kz_ * kz_);
  if (k == (T)0) {
#pragma unroll
    for (int j = 0; j < 6; ++j) wt[j] = (T)0;
    return;
  }
  const T a = tensor_amplitude<T>(k, (T)g.R, (T)g.six_dr2, w3, w2, (T)g.invN);
  const T ik = (T)1 / k;
  const T hx = kx * ik, hy = ky * ik, hz = kz_ * ik;
  const T third = (T)(1.0 / 3.0);
  wt[0] = a * (hx * hx - third);
  wt[1] = a * (hy * hy - third);
  wt[2] = a * (hz * hz - third);
  wt[3] = a * (hx * hy);
  wt[4] = a * (hy * hz);
  wt[5] = a * (hz * hx);
}

// ---------------------------------------------------------------------------
// x pass with the weight multiply: n~ -> n3, n2, n2v_x, n2v_y, n2v_z spectra.
// ---------------------------------------------------------------------------
template <typename T, int L, int TB, int NW = 5>
__device__ __forceinline__ void xconv_fwd_body(int bx, const GridArgs& g, const cx<T>* __restrict__ in,
                                               cx<T>* __restrict__ out, const cx<T>* __restrict__ tw) {
  constexpr int C = ccols<T>(L, TB);
  constexpr int SS = lds_stride<T, L>();
  constexpr int PT = (C * L + kThreads - 1) / kThreads;
  // FP64: two weighted fields per inverse-FFT round (twice the butterflies
  // per stage, half the barriers; 3% faster at 128^3).  FP32 and the eleven
  // tensor fields take one field per round: the doubled buffers cost FP32 a
  // third of its resident blocks and 14%.  Up to L = 128 the transformed
  // input stays in registers between the rounds, so two buffers suffice;
  // beyond, those registers (and the doubled buffers) cost resident blocks,
  // so the input stays in LDS and a third buffer serves the inverse FFTs.
  constexpr bool kSReg = L <= 128;
  constexpr int FG = (NW == 5 && sizeof(T) == 8 && kSReg) ? 2 : 1;
  __shared__ cx<T> A[FG * C * SS];
  __shared__ cx<T> B[FG * C * SS];
  __shared__ cx<T> S2[kSReg ? 1 : C * SS];
  const cx<T>* twl = lds_twiddles<T, L>(tw);
  const long long plane = (long long)g.Ny * g.ldk;
  const int myc = threadIdx.x % C;
  int cy, ckz;
  bool cok;
  const int ncol = x_tile_column<C>(g, bx, myc, cy, ckz, cok);  // cok == (myc < ncol)
  (void)cok;
  const long long off = (long long)cy * g.ldk + ckz;
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int xx = e / C;
    A[myc * SS + skew<T>(xx)] = load_or_zero(in, (long long)xx * plane + off, 0, myc < ncol);
  }
  __syncthreads();
  cx<T>* S = fft_lds<T, L, SS, false>(A, B, ncol, twl);
  cx<T>* W1 = kSReg ? A : (S == A ? B : A);
  cx<T>* W2 = kSReg ? B : S2;
  T w3[PT], w2[PT], wk[PT][3], wt[PT][6];
  cx<T> sv[kSReg ? PT : 1];
  {
    const int y = cy, kz = ckz;
#pragma unroll
    for (int i = 0; i < PT; ++i) {
      const int e = threadIdx.x + i * kThreads;
      w3[i] = w2[i] = wk[i][0] = wk[i][1] = wk[i][2] = (T)0;
#pragma unroll
      for (int j = 0; j < 6; ++j) wt[i][j] = (T)0;
      if constexpr (kSReg) {
        sv[i] = {(T)0, (T)0};
        if (e < C * L) sv[i] = S[myc * SS + skew<T>(e / C)];
      }
      if (e < C * L && myc < ncol) {
        element_weights<T>(g, e / C, y, kz, w3[i], w2[i], wk[i]);
        if constexpr (NW == 11) element_tensor_weights<T>(g, e / C, y, kz, w3[i], w2[i], wt[i]);
      }
    }
  }
  if constexpr (kSReg) __syncthreads();  // S is read before the rounds overwrite both buffers
#pragma unroll
  for (int f0 = 0; f0 < NW; f0 += FG) {
    const int nf = NW - f0 < FG ? NW - f0 : FG;
#pragma unroll
    for (int q = 0; q < FG; ++q) {
      if (f0 + q < NW) {
#pragma unroll
        for (int i = 0; i < PT; ++i) {
          const int e = threadIdx.x + i * kThreads;
          if (e < C * L) {
            const int j = skew<T>(e / C);
            const cx<T> v = kSReg ? sv[kSReg ? i : 0] : S[myc * SS + j];
            W1[(q * C + myc) * SS + j] = weight_mul<T, PT>(f0 + q, i, v, w3, w2, wk, wt);
          }
        }
      }
    }
    __syncthreads();
    const cx<T>* R = fft_lds<T, L, SS, true>(W1, W2, FG == 1 ? ncol : nf * C, twl);
    for (int q = 0; q < nf; ++q) {
      cx<T>* o = out + (f0 + q) * g.ks;
      for (int e = threadIdx.x; e < C * L; e += kThreads) {
        const int xx = e / C;
        if (myc < ncol) o[(long long)xx * plane + off] = R[(q * C + myc) * SS + skew<T>(xx)];
      }
    }
    __syncthreads();
  }
}

template <typename T, int L, int TB, int NW = 5>
__global__ __launch_bounds__(kThreads) void k_xconv_fwd(GridArgs g, const cx<T>* __restrict__ in,
                                                        cx<T>* __restrict__ out, const cx<T>* __restrict__ tw) {
  xconv_fwd_body<T, L, TB, NW>(xcd_pair_tile(blockIdx.x, gridDim.x), g, in, out, tw);
}

// ---------------------------------------------------------------------------
// x pass of the back-convolution: sum_a w_a(-k) FFT_x(P_a) -> inverse FFT_x.
// Even weights keep their sign, odd (vector) weights flip it (deft
// src/Functional.h:407-417).
// ---------------------------------------------------------------------------
template <typename T, int L, int TB, int NW = 5>
__device__ __forceinline__ void xconv_bwd_body(int bx, const GridArgs& g, const cx<T>* __restrict__ in,
                                               cx<T>* __restrict__ out, const cx<T>* __restrict__ tw) {
  constexpr int C = ccols<T>(L, TB);
  constexpr int SS = lds_stride<T, L>();
  constexpr int PT = (C * L + kThreads - 1) / kThreads;
  __shared__ cx<T> A[C * SS];
  __shared__ cx<T> B[C * SS];
  const cx<T>* twl = lds_twiddles<T, L>(tw);
  const long long plane = (long long)g.Ny * g.ldk;
  const int myc = threadIdx.x % C;
  int cy, ckz;
  bool cok;
  const int ncol = x_tile_column<C>(g, bx, myc, cy, ckz, cok);  // cok == (myc < ncol)
  (void)cok;
  const long long off = (long long)cy * g.ldk + ckz;
  // FP32: the next field's columns are loaded into registers while the
  // current field is transformed and accumulated (two fields' loads in
  // flight; 10% faster at 128^3).  FP64 loads each field in turn: there the
  // extra registers halve the occupancy and the pass ran 10% slower.
  constexpr bool kPrefetch = sizeof(T) == 4;
  cx<T> pre[PT];
  auto fetch = [&](int f) {
    const cx<T>* src = in + f * g.ks;
#pragma unroll
    for (int i = 0; i < PT; ++i) {
      const int e = threadIdx.x + i * kThreads;  // e % C == myc
      pre[i] = load_or_zero(src, (long long)(e / C) * plane + off, 0, e < C * L && myc < ncol);
    }
  };
  // the first field's loads go out before the weight-table reads, so their
  // latencies overlap
  if constexpr (kPrefetch) fetch(0);
  T w3[PT], w2[PT], wk[PT][3], wt[PT][6];
  cx<T> acc[PT];
  {
    const int y = cy, kz = ckz;
#pragma unroll
    for (int i = 0; i < PT; ++i) {
      const int e = threadIdx.x + i * kThreads;
      w3[i] = w2[i] = wk[i][0] = wk[i][1] = wk[i][2] = (T)0;
#pragma unroll
      for (int j = 0; j < 6; ++j) wt[i][j] = (T)0;
      acc[i] = {(T)0, (T)0};
      if (e < C * L && myc < ncol) {
        element_weights<T>(g, e / C, y, kz, w3[i], w2[i], wk[i]);
        if constexpr (NW == 11) element_tensor_weights<T>(g, e / C, y, kz, w3[i], w2[i], wt[i]);
      }
    }
  }
#pragma unroll
  for (int f = 0; f < NW; ++f) {
    if constexpr (kPrefetch) {
#pragma unroll
      for (int i = 0; i < PT; ++i) {
        const int e = threadIdx.x + i * kThreads;
        if (e < C * L) A[myc * SS + skew<T>(e / C)] = pre[i];
      }
      __syncthreads();
      if (f + 1 < NW) fetch(f + 1);
    } else {
      const cx<T>* src = in + f * g.ks;
      for (int e = threadIdx.x; e < C * L; e += kThreads) {
        const int xx = e / C;
        A[myc * SS + skew<T>(xx)] = load_or_zero(src, (long long)xx * plane + off, 0, myc < ncol);
      }
      __syncthreads();
    }
    const cx<T>* S = fft_lds<T, L, SS, false>(A, B, ncol, twl);
#pragma unroll
    for (int i = 0; i < PT; ++i) {
      const int e = threadIdx.x + i * kThreads;
      if (e < C * L) {
        const cx<T> p = weight_mul<T, PT>(f, i, S[myc * SS + skew<T>(e / C)], w3, w2, wk, wt);
        // even kernels (w3, w2, tensor) are their own adjoints, odd ones (vector) flip sign
        acc[i] = (f < 2 || f >= 5) ? acc[i] + p : acc[i] - p;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const int e = threadIdx.x + i * kThreads;
    if (e < C * L) A[myc * SS + skew<T>(e / C)] = acc[i];
  }
  __syncthreads();
  const cx<T>* R = fft_lds<T, L, SS, true>(A, B, ncol, twl);
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int xx = e / C;
    if (myc < ncol) out[(long long)xx * plane + off] = R[myc * SS + skew<T>(xx)];
  }
}

template <typename T, int L, int TB, int NW = 5>
__global__ __launch_bounds__(kThreads) void k_xconv_bwd(GridArgs g, const cx<T>* __restrict__ in,
                                                        cx<T>* __restrict__ out, const cx<T>* __restrict__ tw) {
  xconv_bwd_body<T, L, TB, NW>(xcd_pair_tile(blockIdx.x, gridDim.x), g, in, out, tw);
}

// ---------------------------------------------------------------------------
// z pass 2: c2r of the five weighted densities -> FP64 Phi (energy partial
// sums) and, with GRAD, the five partials -> r2c back into the same buffers.
// ---------------------------------------------------------------------------
// FN: 0 White-Bear, 1 Mark II (either form), 3 tensor White-Bear (11 fields)
template <typename T, int L, bool GRAD, int ZB, int FN = 0>
__device__ __forceinline__ void zphi_body(int bx, const GridArgs& g, cx<T>* __restrict__ fields,
                                          double* __restrict__ partial, int pstride, const cx<T>* __restrict__ tw) {
  constexpr int P = zpairs(L, ZB);
  constexpr int NPT = 2 * P * L;
  constexpr int PT = (NPT + kThreads - 1) / kThreads;
  constexpr int SS = lds_stride<T, L>();
  constexpr int NW = FN == 3 ? 11 : 5;
  // Two fields per FFT round (the eleven tensor fields one at a time): twice
  // the butterflies per stage and half the barriers of one field per round.
  // At L >= 256 the paired round needs 184 FP64 registers (two waves per
  // SIMD instead of four), which costs more than the barriers it saves.
  constexpr int FG = (FN == 3 || L > 128) ? 1 : 2;
  __shared__ cx<T> A[zbuf_size<T, L, FG * P>()];
  __shared__ cx<T> B[zbuf_size<T, L, FG * P>()];
  // The butterfly loops stride by blockDim.x here: on a compile-time stride
  // the compiler keeps every stage's twiddles in registers across the field
  // loop and the pass no longer fits 128 VGPRs (it spilled 156 bytes in FP32,
  // 244 in FP64).  The FP64 variants read their twiddles from global memory:
  // their LDS copy measured 13% slower at 128^3.
  const cx<T>* twl = sizeof(T) == 4 ? lds_twiddles<T, L>(tw) : tw;
  const int fstride = (int)blockDim.x;
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)bx * 2 * P;
  T val[NW][PT];
  // the next round's half-spectrum rows are loaded into registers while the
  // current round is packed and transformed
  constexpr int NRAW = 2 * P * (L / 2 + 1);
  constexpr int PF = (NRAW + kThreads - 1) / kThreads;
  cx<T> pre[FG][PF];
  auto fetch = [&](int f0) {
#pragma unroll
    for (int q = 0; q < FG; ++q) {
#pragma unroll
      for (int i = 0; i < PF; ++i) {
        const int e = threadIdx.x + i * kThreads;
        const int l = e / (L / 2 + 1), k = e - l * (L / 2 + 1);
        pre[q][i] = {(T)0, (T)0};
        if (f0 + q < NW)  // past the last field nothing is read, not even the base
          pre[q][i] = load_or_zero(fields + (f0 + q) * g.ks, (line0 + l) * g.ldk + k, 0,
                                   e < NRAW && line0 + l < lines);
      }
    }
  };
  fetch(0);
#pragma unroll
  for (int f0 = 0; f0 < NW; f0 += FG) {
    const int nf = NW - f0 < FG ? NW - f0 : FG;
#pragma unroll
    for (int q = 0; q < FG; ++q) {
#pragma unroll
      for (int i = 0; i < PF; ++i) {
        const int e = threadIdx.x + i * kThreads;
        if (q < nf && e < NRAW) A[q * NRAW + e] = pre[q][i];
      }
    }
    __syncthreads();
    if (f0 + FG < NW) fetch(f0 + FG);
    pack_c2r_fields<T, L, P>(A, B, nf);
    const cx<T>* R = fft_lds<T, L, SS, true>(B, A, nf * P, twl, fstride);
#pragma unroll
    for (int q = 0; q < FG; ++q) {
      if (f0 + q < NW) {
#pragma unroll
        for (int i = 0; i < PT; ++i) {
          const int e = threadIdx.x + i * kThreads;
          val[f0 + q][i] = (T)0;
          if (e < NPT) {
            const int l = e / L, z = e - l * L;
            const cx<T> v = R[(q * P + (l >> 1)) * SS + skew<T>(z)];
            val[f0 + q][i] = (l & 1) ? v.im : v.re;
          }
        }
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
    if constexpr (NW == 11) {
      double v[11];
#pragma unroll
      for (int f = 0; f < 11; ++f) v[f] = (double)val[f][i];
      if constexpr (GRAD) {
        double dd[11];
#pragma unroll
        for (int f = 0; f < 11; ++f) dd[f] = 0.0;
        if (valid) {
          double q1, q2, q3;
          wbt_phi(v, g.wb, p1, p2, p3);
          wbt_phi_partials(v, g.wb, q1, q2, q3, dd);
        }
#pragma unroll
        for (int f = 0; f < 11; ++f) val[f][i] = (T)dd[f];
      } else {
        if (valid) wbt_phi(v, g.wb, p1, p2, p3);
      }
    } else if constexpr (GRAD) {
      double dd[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (valid) {
        // the energy terms come from wb_phi, as in the energy-only variant
        double q1, q2, q3;
        wb_phi_k<FN == 1>((double)val[0][i], (double)val[1][i], (double)val[2][i], (double)val[3][i],
                          (double)val[4][i], g.wb, p1, p2, p3);
        wb_phi_partials_k<FN == 1>((double)val[0][i], (double)val[1][i], (double)val[2][i], (double)val[3][i],
                                   (double)val[4][i], g.wb, q1, q2, q3, dd);
      }
#pragma unroll
      for (int f = 0; f < 5; ++f) val[f][i] = (T)dd[f];
    } else {
      if (valid)
        wb_phi_k<FN == 1>((double)val[0][i], (double)val[1][i], (double)val[2][i], (double)val[3][i],
                          (double)val[4][i], g.wb, p1, p2, p3);
    }
    acc[0] += p1;
    acc[1] += p2;
    acc[2] += p3;
  }
  if constexpr (GRAD) {
#pragma unroll
    for (int f0 = 0; f0 < NW; f0 += FG) {
      const int nf = NW - f0 < FG ? NW - f0 : FG;
#pragma unroll
      for (int q = 0; q < FG; ++q) {
        if (f0 + q < NW) {
#pragma unroll
          for (int i = 0; i < PT; ++i) {
            const int e = threadIdx.x + i * kThreads;
            if (e < NPT) {
              const int l = e / L, z = e - l * L;
              cx<T>* c = &A[(q * P + (l >> 1)) * SS + skew<T>(z)];
              if (l & 1)
                c->im = val[f0 + q][i];
              else
                c->re = val[f0 + q][i];
            }
          }
        }
      }
      __syncthreads();
      const cx<T>* Z = fft_lds<T, L, SS, false>(A, B, nf * P, twl, fstride);
      for (int q = 0; q < nf; ++q) store_r2c_halves<T, L, P>(Z + q * P * SS, fields + (f0 + q) * g.ks, line0, lines, g.ldk);
      __syncthreads();
    }
  }
  block_sum_store<3>(acc, partial, SLOT_PHI1, pstride, bx);
}

// At most 128 VGPRs (four waves per SIMD): the White-Bear FP64
// energy+gradient pass at 512-value blocks needs 132 otherwise and drops to
// three waves.  The FP64 Mark II variants keep the default budget (held to
// 128 they spill to scratch; at 128^3 the default budget with the field
// prefetch ran 5-12% faster than four waves without it); the FP32 Mark II
// ones measured faster at four waves, spills included.
template <typename T, int ZB, int FN>
constexpr int zphi_min_waves() {
  return (FN == 1 && sizeof(T) == 8) ? 1 : 4;
}
template <typename T, int L, bool GRAD, int ZB, int FN = 0>
__global__ __launch_bounds__(kThreads, (zphi_min_waves<T, ZB, FN>())) void k_zphi(
    GridArgs g, cx<T>* __restrict__ fields, double* __restrict__ partial, int pstride,
    const cx<T>* __restrict__ tw) {
  zphi_body<T, L, GRAD, ZB, FN>(blockIdx.x, g, fields, partial, pstride, tw);
}

// the eleven-field tensor variant keeps the default register budget
template <typename T, int L, bool GRAD, int ZB>
__global__ __launch_bounds__(kThreads) void k_zphi11(GridArgs g, cx<T>* __restrict__ fields,
                                                     double* __restrict__ partial, int pstride,
                                                     const cx<T>* __restrict__ tw) {
  zphi_body<T, L, GRAD, ZB, 3>(blockIdx.x, g, fields, partial, pstride, tw);
}

// ---------------------------------------------------------------------------
// z pass 3: c2r of the back-convolution -> dF/dn -> grad and pgrad (FP64),
// with pgrad.grad and pgrad.oldgrad partial sums.  The preconditioned
// gradient is one of (`precond`):
//   PGRAD_NONE 0   pgrad = grad
//   PGRAD_NEW  1   grad/n = -(dF/dn)/kT, the new API's (src/haskell/NewCode.hs:229-242)
//   PGRAD_OLD  2   (dF/dn)/(Veff - kT), the old API's OfEffectivePotential
//                  (src/EffectivePotentialToDensity.cpp:62-71)
//   PGRAD_SK   3   the structure-factor preconditioner: this pass stores
//                  grad/sqrt(n) (0 below sk_floor), which k_sk_finish turns
//                  into n^-1/2 F^-1 S(k) F (grad/sqrt(n)) after the transforms
// ---------------------------------------------------------------------------
enum PgradKind : int { PGRAD_NONE = 0, PGRAD_NEW = 1, PGRAD_OLD = 2, PGRAD_SK = 3 };

// pgrad of one point from dF/dn (finite Vext)
__device__ __forceinline__ double point_pgrad(int precond, double dfdn, double n, double xv, double gr,
                                              const GridArgs& g) {
  switch (precond) {
    case PGRAD_NEW:
      return -dfdn / g.kT;
    case PGRAD_OLD:
      return dfdn / (xv - g.kT);
    case PGRAD_SK:
      return n >= g.sk_floor ? -dfdn * sqrt(n) / g.kT : 0.0;
    default:
      return gr;
  }
}

template <typename T, int L, int ZB>
__device__ __forceinline__ void zgrad_body(int bx, const GridArgs& g, const cx<T>* __restrict__ accs,
                                           const double* __restrict__ x, const T* __restrict__ vext,
                                           const double* __restrict__ oldgrad, int have_old, int precond,
                                           double* __restrict__ grad, double* __restrict__ pgrad,
                                           double* __restrict__ partial, int pstride,
                                           const cx<T>* __restrict__ tw) {
  constexpr int P = zpairs(L, ZB);
  constexpr int SS = lds_stride<T, L>();
  __shared__ cx<T> A[zbuf_size<T, L, P>()];
  __shared__ cx<T> B[zbuf_size<T, L, P>()];
  const cx<T>* twl = lds_twiddles<T, L>(tw);
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)bx * 2 * P;
  load_c2r_packed<T, L, P>(accs, line0, lines, g.ldk, A, B);
  const cx<T>* R = fft_lds<T, L, SS, true>(B, A, P, twl);
  double dots[2] = {0.0, 0.0};
  for (int e = threadIdx.x; e < 2 * P * L; e += kThreads) {
    const int l = e / L, z = e - l * L;
    if (line0 + l < lines) {
      const long long idx = (line0 + l) * L + z;
      const cx<T> v = R[(l >> 1) * SS + skew<T>(z)];
      const double conv = (double)((l & 1) ? v.im : v.re);
      const double xv = x[idx];
      const T ve = vext[idx];
      double gr = 0.0, pg = 0.0;
      if (isfinite(ve)) {
        const double n = exp(-xv / g.kT);
        const double local = -xv - g.kT * g.logNQ + (double)ve - g.mu;
        const double dfdn = (g.kT * conv + local) * g.dV;
        gr = -dfdn * n / g.kT;
        pg = point_pgrad(precond, dfdn, n, xv, gr, g);
      }
      grad[idx] = gr;
      pgrad[idx] = pg;
      dots[0] += pg * gr;
      if (have_old) dots[1] += pg * oldgrad[idx];
    }
  }
  block_sum_store<2>(dots, partial, SLOT_PGG, pstride, bx);
}
template <typename T, int L, int ZB>
__global__ __launch_bounds__(kThreads) void k_zgrad(GridArgs g, const cx<T>* __restrict__ accs,
                                                    const double* __restrict__ x, const T* __restrict__ vext,
                                                    const double* __restrict__ oldgrad, int have_old, int precond,
                                                    double* __restrict__ grad, double* __restrict__ pgrad,
                                                    double* __restrict__ partial, int pstride,
                                                    const cx<T>* __restrict__ tw) {
  zgrad_body<T, L, ZB>(blockIdx.x, g, accs, x, vext, oldgrad, have_old, precond, grad, pgrad, partial, pstride, tw);
}

// ---------------------------------------------------------------------------
// Structure-factor preconditioner (PGRAD_SK): pg = n^-1/2 F^-1 S(k) F n^-1/2 g
// on points with n >= sk_floor, grad/n on the dilute ones, 0 on hard-wall and
// constrained points.  S(k) = (kT/n_b)/(kT/n_b + H_ex(k)) is the uniform
// fluid's structure factor, H_ex(k) = kT [P33 w3^2 + 2 P32 w3 w2 + P22 w2^2
// + cv k^2 w3^2] with Phi's second derivatives at the bulk weighted
// densities; in the uniform fluid the operator is (dV/kT) times the inverse
// Hessian over Veff, and with S = 1 it is the new API's grad/n.
// ---------------------------------------------------------------------------

// S(k)/N on the padded half spectrum from the plan's weight table (w/N)
static __global__ __launch_bounds__(kThreads) void k_sk_table(GridArgs g, double P33, double P32, double P22,
                                                              double cv, double nb, double clamp,
                                                              double* __restrict__ out) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  const double N = (double)g.N;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    const int kz = (int)(i % g.Nkz);
    const long long r = i / g.Nkz;
    const int y = (int)(r % g.Ny), x = (int)(r / g.Ny);
    const double mx = wrapf(x, g.Nx), my = wrapf(y, g.Ny), mz = kz;
    const double kx = mx * g.b[0] + my * g.b[3] + mz * g.b[6];
    const double ky = mx * g.b[1] + my * g.b[4] + mz * g.b[7];
    const double kz_ = mx * g.b[2] + my * g.b[5] + mz * g.b[8];
    const long long idx = r * g.ldk + kz;
    const double2 w = reinterpret_cast<const double2*>(g.wtab)[idx];
    const double a3 = w.x * N, a2 = w.y * N;
    const double hex = g.kT * (P33 * a3 * a3 + 2.0 * P32 * a3 * a2 + P22 * a2 * a2 +
                               cv * (kx * kx + ky * ky + kz_ * kz_) * a3 * a3);
    double S = (g.kT / nb) / (g.kT / nb + hex);
    if (!(S > 0.0) || !isfinite(S)) S = clamp;
    S = fmin(fmax(S, 1.0 / clamp), clamp);
    out[idx] = S * g.invN;
  }
}

// spectrum *= S(k)/N
template <typename T>
__global__ __launch_bounds__(kThreads) void k_spec_scale(GridArgs g, cx<T>* __restrict__ spec,
                                                         const double* __restrict__ sk) {
  const long long nk = (long long)g.Nx * g.Ny * g.Nkz;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nk; i += (long long)gridDim.x * kThreads) {
    const long long r = i / g.Nkz;
    const long long idx = r * g.ldk + (i - r * g.Nkz);
    spec[idx] = scale(spec[idx], (T)sk[idx]);
  }
}

// pg = conv/sqrt(n) (dense), grad/n (dilute), 0 (wall, constrained), with the
// pgrad.grad and pgrad.oldgrad partial sums; gridDim.x = the block count the
// evaluation's reduction reads
template <typename T>
__global__ __launch_bounds__(kThreads) void k_sk_finish(GridArgs g, const double* __restrict__ x,
                                                        const T* __restrict__ vext, const double* __restrict__ grad,
                                                        const double* __restrict__ oldgrad, int have_old,
                                                        double* __restrict__ pgrad, double* __restrict__ partial,
                                                        int pstride) {
  double dots[2] = {0.0, 0.0};
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < g.N; i += (long long)gridDim.x * kThreads) {
    const T ve = vext[i];
    const double gr = grad[i];
    double pg = 0.0;
    if (isfinite(ve)) {
      const double n = exp(-x[i] / g.kT);
      if (n >= g.sk_floor)
        pg = pgrad[i] / sqrt(n);
      else if (n > 0.0)
        pg = gr / n;
    }
    pgrad[i] = pg;
    dots[0] += pg * gr;
    if (have_old) dots[1] += pg * oldgrad[i];
  }
  block_sum_store<2>(dots, partial, SLOT_PGG, pstride, blockIdx.x);
}

// The same preconditioner fused into the passes of a fixed-edge plan (the
// operations per element are those of the split path above, so the results
// are identical; no real-space round trip between the passes):
//   k_zgrad_sk     c2r of the back-convolution -> grad (stored) and
//                  grad/sqrt(n) (0 below sk_floor) -> r2c along z, in place
//   y pass         forward, one spectrum
//   k_xsk          FFT along x, times S(k)/N, inverse FFT along x
//   y pass         inverse, one spectrum
//   k_zsk_finish   c2r -> pg as k_sk_finish, with the pgrad.grad and
//                  pgrad.oldgrad partial sums
template <typename T, int L, int ZB>
__device__ __forceinline__ void zgrad_sk_body(int bx, const GridArgs& g, cx<T>* __restrict__ accs,
                                              const double* __restrict__ x, const T* __restrict__ vext,
                                              double* __restrict__ grad, const cx<T>* __restrict__ tw) {
  constexpr int P = zpairs(L, ZB);
  constexpr int NPT = 2 * P * L;
  constexpr int PT = (NPT + kThreads - 1) / kThreads;
  constexpr int SS = lds_stride<T, L>();
  __shared__ cx<T> A[zbuf_size<T, L, P>()];
  __shared__ cx<T> B[zbuf_size<T, L, P>()];
  const cx<T>* twl = lds_twiddles<T, L>(tw);
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)bx * 2 * P;
  load_c2r_packed<T, L, P>(accs, line0, lines, g.ldk, A, B);
  cx<T>* R = fft_lds<T, L, SS, true>(B, A, P, twl);
  T val[PT];
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const int e = threadIdx.x + i * kThreads;
    val[i] = (T)0;
    if (e < NPT) {
      const int l = e / L, z = e - l * L;
      if (line0 + l < lines) {
        const long long idx = (line0 + l) * L + z;
        const cx<T> v = R[(l >> 1) * SS + skew<T>(z)];
        const double conv = (double)((l & 1) ? v.im : v.re);
        const double xv = x[idx];
        const T ve = vext[idx];
        double gr = 0.0, pg = 0.0;
        if (isfinite(ve)) {
          const double n = exp(-xv / g.kT);
          const double local = -xv - g.kT * g.logNQ + (double)ve - g.mu;
          const double dfdn = (g.kT * conv + local) * g.dV;
          gr = -dfdn * n / g.kT;
          pg = point_pgrad(PGRAD_SK, dfdn, n, xv, gr, g);
        }
        grad[idx] = gr;
        val[i] = (T)pg;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const int e = threadIdx.x + i * kThreads;
    if (e < NPT) {
      const int l = e / L, z = e - l * L;
      cx<T>* c = &R[(l >> 1) * SS + skew<T>(z)];
      if (l & 1)
        c->im = val[i];
      else
        c->re = val[i];
    }
  }
  __syncthreads();
  const cx<T>* Z = fft_lds<T, L, SS, false>(R, R == A ? B : A, P, twl);
  store_r2c_halves<T, L, P>(Z, accs, line0, lines, g.ldk);
}
template <typename T, int L, int ZB>
__global__ __launch_bounds__(kThreads) void k_zgrad_sk(GridArgs g, cx<T>* __restrict__ accs,
                                                       const double* __restrict__ x, const T* __restrict__ vext,
                                                       double* __restrict__ grad, const cx<T>* __restrict__ tw) {
  zgrad_sk_body<T, L, ZB>(blockIdx.x, g, accs, x, vext, grad, tw);
}

template <typename T, int L, int TB>
__device__ __forceinline__ void xsk_body(int bx, const GridArgs& g, cx<T>* __restrict__ data,
                                         const double* __restrict__ sk, const cx<T>* __restrict__ tw) {
  constexpr int C = ccols<T>(L, TB);
  constexpr int SS = lds_stride<T, L>();
  __shared__ cx<T> A[C * SS];
  __shared__ cx<T> B[C * SS];
  const cx<T>* twl = lds_twiddles<T, L>(tw);
  const long long plane = (long long)g.Ny * g.ldk;
  const int myc = threadIdx.x % C;
  int cy, ckz;
  bool cok;
  const int ncol = x_tile_column<C>(g, bx, myc, cy, ckz, cok);  // cok == (myc < ncol)
  const long long off = (long long)cy * g.ldk + ckz;
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int xx = e / C;
    A[myc * SS + skew<T>(xx)] = load_or_zero(data, (long long)xx * plane + off, 0, cok);
  }
  __syncthreads();
  cx<T>* S = fft_lds<T, L, SS, false>(A, B, ncol, twl);
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int xx = e / C;
    if (cok) {
      cx<T>* c = &S[myc * SS + skew<T>(xx)];
      *c = scale(*c, (T)sk[(long long)xx * plane + off]);
    }
  }
  __syncthreads();
  const cx<T>* R = fft_lds<T, L, SS, true>(S, S == A ? B : A, ncol, twl);
  for (int e = threadIdx.x; e < C * L; e += kThreads) {
    const int xx = e / C;
    if (cok) data[(long long)xx * plane + off] = R[myc * SS + skew<T>(xx)];
  }
}
template <typename T, int L, int TB>
__global__ __launch_bounds__(kThreads) void k_xsk(GridArgs g, cx<T>* __restrict__ data, const double* __restrict__ sk,
                                                  const cx<T>* __restrict__ tw) {
  xsk_body<T, L, TB>(xcd_pair_tile(blockIdx.x, gridDim.x), g, data, sk, tw);
}

template <typename T, int L, int ZB>
__device__ __forceinline__ void zsk_finish_body(int bx, const GridArgs& g, const cx<T>* __restrict__ spec,
                                                const double* __restrict__ x, const T* __restrict__ vext,
                                                const double* __restrict__ grad,
                                                const double* __restrict__ oldgrad, int have_old,
                                                double* __restrict__ pgrad, double* __restrict__ partial,
                                                int pstride, const cx<T>* __restrict__ tw) {
  constexpr int P = zpairs(L, ZB);
  constexpr int SS = lds_stride<T, L>();
  __shared__ cx<T> A[zbuf_size<T, L, P>()];
  __shared__ cx<T> B[zbuf_size<T, L, P>()];
  const cx<T>* twl = lds_twiddles<T, L>(tw);
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)bx * 2 * P;
  load_c2r_packed<T, L, P>(spec, line0, lines, g.ldk, A, B);
  const cx<T>* R = fft_lds<T, L, SS, true>(B, A, P, twl);
  double dots[2] = {0.0, 0.0};
  for (int e = threadIdx.x; e < 2 * P * L; e += kThreads) {
    const int l = e / L, z = e - l * L;
    if (line0 + l < lines) {
      const long long idx = (line0 + l) * L + z;
      const cx<T> v = R[(l >> 1) * SS + skew<T>(z)];
      const T ve = vext[idx];
      const double gr = grad[idx];
      double pg = 0.0;
      if (isfinite(ve)) {
        const double n = exp(-x[idx] / g.kT);
        if (n >= g.sk_floor)
          pg = (double)((l & 1) ? v.im : v.re) / sqrt(n);
        else if (n > 0.0)
          pg = gr / n;
      }
      pgrad[idx] = pg;
      dots[0] += pg * gr;
      if (have_old) dots[1] += pg * oldgrad[idx];
    }
  }
  block_sum_store<2>(dots, partial, SLOT_PGG, pstride, bx);
}
template <typename T, int L, int ZB>
__global__ __launch_bounds__(kThreads) void k_zsk_finish(GridArgs g, const cx<T>* __restrict__ spec,
                                                         const double* __restrict__ x, const T* __restrict__ vext,
                                                         const double* __restrict__ grad,
                                                         const double* __restrict__ oldgrad, int have_old,
                                                         double* __restrict__ pgrad, double* __restrict__ partial,
                                                         int pstride, const cx<T>* __restrict__ tw) {
  zsk_finish_body<T, L, ZB>(blockIdx.x, g, spec, x, vext, grad, oldgrad, have_old, pgrad, partial, pstride, tw);
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
  block_sum_store<1>(acc, partial, SLOT_GD, pstride, blockIdx.x);
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
  block_sum_store<1>(acc, partial, SLOT_GD, pstride, blockIdx.x);
}

// x = x + s*d, rounded like the host expression x += s*d
static __global__ __launch_bounds__(kThreads) void k_fill(long long N, double* __restrict__ x, double v) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads)
    x[i] = v;
}

static __global__ __launch_bounds__(kThreads) void k_axpy(long long N, double* __restrict__ x, const double* __restrict__ d,
                                                   double s) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads)
    x[i] = __dadd_rn(x[i], __dmul_rn(s, d[i]));
}

// n = exp(-Veff/kT), 0 on hard walls
static __global__ __launch_bounds__(kThreads) void k_density(long long N, const double* __restrict__ x,
                                                             const float* __restrict__ vext, double kT,
                                                             double* __restrict__ n) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads)
    n[i] = isinf(vext[i]) ? 0.0 : exp(-x[i] / kT);
}

// out += a (accumulating into caller buffers, as deft's grad does)
static __global__ __launch_bounds__(kThreads) void k_accumulate(long long N, double* __restrict__ out,
                                                         const double* __restrict__ a) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads)
    out[i] += a[i];
}

// out[slot] = sum_{b < nb} partial[slot*pstride + b], in a fixed order
__device__ __forceinline__ void reduce_body(int slot, const double* __restrict__ partial, int pstride, int nb,
                                            double* __restrict__ out) {
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
    __threadfence_system();  // `out` may be host-mapped: visible before any later flag
  }
}

// Raises the plan's completion flag (host-mapped) to seq once all earlier work
// of the stream is done; the host waits on it instead of synchronising.
static __global__ void k_signal(unsigned long long* __restrict__ flag, unsigned long long seq) {
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// out[slot] for slots [slot0, slot0 + gridDim.x)
static __global__ __launch_bounds__(kThreads) void k_reduce(const double* __restrict__ partial, int pstride, int slot0,
                                                     int nb, double* __restrict__ out) {
  reduce_body(slot0 + blockIdx.x, partial, pstride, nb, out);
}

// k_reduce that also raises the completion flag: each block, after its sum
// has reached the host (reduce_body's system fence), counts itself in `cnt`;
// the last one re-arms the counter and stores seq to the host-mapped flag,
// so the host needs no separate flag kernel.
static __global__ __launch_bounds__(kThreads) void k_reduce_flag(const double* __restrict__ partial, int pstride,
                                                          int slot0, int nb, double* __restrict__ out,
                                                          unsigned* __restrict__ cnt,
                                                          unsigned long long* __restrict__ flag,
                                                          unsigned long long seq) {
  reduce_body(slot0 + blockIdx.x, partial, pstride, nb, out);
  if (threadIdx.x == 0) {
    const unsigned prev = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == gridDim.x - 1) {
      __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

}  // namespace deft
