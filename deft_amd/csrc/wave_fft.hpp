// Wave-parallel z Phi pass for power-of-two z edges (16..512).
//
// k_zphi (passes.hpp) transforms all of a block's lines together: every
// Stockham stage is a block-wide step between two __syncthreads(), and at
// L = 128 a radix-8 stage over the block's few packed lines keeps one or two
// of its four waves busy.  Here each wave owns one weighted-density field: it
// loads that field's half-spectrum rows, runs the c2r transforms of its packed
// lines with every lane holding eight elements, exchanges elements between
// stages through its own slice of LDS (one wave's LDS instructions execute in
// issue order, so the exchanges need no block barrier), and after the Phi step
// runs the r2c transforms of the five partials the same way.  Two block
// barriers per group of lines remain, around the Phi step, which reads all
// five fields of a point.
//
// The stages are those of fft_lds (same radix order, butterflies, twiddle
// indices and packing), and the Phi step gives every point and every partial
// sum to the thread k_zphi gives it, with the same block reduction, so the two
// variants compute the same operations on the same values.
#pragma once

#include "passes.hpp"

namespace deft {

// Orders the calling wave's memory operations before the call against those
// after it (the compiler may not move them across); used between the LDS
// writes of one stage and the reads of the next by other lanes of the wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lengths the wave transforms take: eight elements per lane, L/8 lanes per
// sequence, 64/(L/8) sequences per wave.
constexpr bool wave_fft_length(int L) { return L >= 16 && L <= 512 && (L & (L - 1)) == 0; }

// Radix of the stage that combines sub-transforms of length NS (the order
// fft_lds takes for powers of two: eights, then one 4 or 2).
template <int L, int NS>
constexpr int wave_radix() {
  return (L / NS) % 8 == 0 ? 8 : ((L / NS) % 4 == 0 ? 4 : 2);
}

// Butterflies of one radix-R stage.  Lane t of a sequence holds butterflies
// j = t + (L/8) i for i < 8/R; v[i R + r] is element j + r L/R of the stage
// input.  Same twiddle indices and butterflies as stockham_stage.
template <typename T, int L, int R, int NS, bool INV>
__device__ __forceinline__ void wave_butterflies(cx<T> (&v)[8], int t, const cx<T>* __restrict__ tw) {
  constexpr int LPS = L / 8;
  constexpr int TWS = L / (NS * R);
#pragma unroll
  for (int i = 0; i < 8 / R; ++i) {
    if constexpr (NS > 1) {
      const int k = (t + LPS * i) % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const cx<T> w = tw[k * r * TWS];
        v[i * R + r] = INV ? cmulc(v[i * R + r], w) : cmul(v[i * R + r], w);
      }
    }
    dft<T, R, INV>(&v[i * R]);
  }
}

template <typename T, int L, int R>
__device__ __forceinline__ void wave_read(cx<T> (&v)[8], const cx<T>* seq, int t) {
  constexpr int LPS = L / 8, M = L / R;
#pragma unroll
  for (int i = 0; i < 8 / R; ++i)
#pragma unroll
    for (int r = 0; r < R; ++r) v[i * R + r] = seq[skew<T>(t + LPS * i + r * M)];
}

template <typename T, int L, int R, int NS>
__device__ __forceinline__ void wave_write(const cx<T> (&v)[8], cx<T>* seq, int t) {
  constexpr int LPS = L / 8;
#pragma unroll
  for (int i = 0; i < 8 / R; ++i) {
    const int j = t + LPS * i;
    const int o0 = (j / NS) * NS * R + j % NS;
#pragma unroll
    for (int r = 0; r < R; ++r) seq[skew<T>(o0 + r * NS)] = v[i * R + r];
  }
}

// Stages NS, NS*R, ... of the transform of `seq` (skewed, in place): each
// stage reads its inputs, transforms them in registers and writes them back
// after the wave's reads are done.
template <typename T, int L, int NS, bool INV>
__device__ __forceinline__ void wave_stages(cx<T>* seq, int t, const cx<T>* __restrict__ tw) {
  if constexpr (NS < L) {
    constexpr int R = wave_radix<L, NS>();
    cx<T> v[8];
    wave_read<T, L, R>(v, seq, t);
    wave_butterflies<T, L, R, NS, INV>(v, t, tw);
    wave_lds_sync();
    wave_write<T, L, R, NS>(v, seq, t);
    wave_lds_sync();
    wave_stages<T, L, NS * R, INV>(seq, t, tw);
  }
}

// LDS slice of one field: the packed lines (stride SS, skewed) and, before the
// first stage, the raw half-spectrum rows of the group's lines.
template <typename T, int L>
constexpr int wave_slice() {
  constexpr int SPW = 64 / (L / 8);
  constexpr int a = SPW * lds_stride<T, L>(), b = 2 * SPW * (L / 2 + 1);
  return a > b ? a : b;
}

constexpr int kWaveFields = 5;
constexpr int kWaveThreads = kWaveFields * 64;

// White-Bear energy (and with GRAD the five partials, r2c'd back in place).
// Block b covers lines [2Pb, 2P(b+1)) with P = zpairs(L, 512), as k_zphi does
// at 512-value blocks: 2*SPW lines, 1024 points.  (At 1024-value blocks two
// such groups per block need more than 128 VGPRs and spill, so this variant
// is instantiated for 512 only.)
template <typename T, int L, bool GRAD>
__global__ __launch_bounds__(kWaveThreads) void k_zphi_wave(GridArgs g, cx<T>* __restrict__ fields,
                                                            double* __restrict__ partial, int pstride,
                                                            const cx<T>* __restrict__ tw) {
  static_assert(wave_fft_length(L), "the wave transforms take power-of-two lengths 16..512");
  constexpr int LPS = L / 8, SPW = 64 / LPS, NKZ = L / 2 + 1;
  constexpr int SS = lds_stride<T, L>();
  constexpr int REG = wave_slice<T, L>();
  constexpr int GL = 2 * SPW;  // lines per block
  static_assert(GL == 2 * zpairs(L, 512) && GL * L == 4 * kThreads, "a block is 1024 points");
  __shared__ cx<T> S[kWaveFields * REG];
  __shared__ cx<T> twl[L];
  for (int i = threadIdx.x; i < L; i += kWaveThreads) twl[i] = tw[i];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sq = lane / LPS, t = lane - sq * LPS;
  cx<T>* reg = S + wave * REG;
  cx<T>* seq = reg + sq * SS;
  cx<T>* fld = fields + wave * g.ks;
  const long long lines = (long long)g.Nx * g.Ny;
  const long long line0 = (long long)blockIdx.x * GL;
  for (int e = lane; e < GL * NKZ; e += 64) {
    const int l = e / NKZ, k = e - l * NKZ;
    reg[e] = load_or_zero(fld, (line0 + l) * g.ldk + k, 0, line0 + l < lines);
  }
  __syncthreads();  // the twiddles (and this wave's rows)
  {
    // first stage, its inputs packed from the raw rows as pack_c2r_fields does:
    // Z = A + i B for lines 2 sq and 2 sq + 1 (Hermitian completion along z)
    cx<T> v[8];
    const cx<T>* ra = reg + (2 * sq) * NKZ;
    const cx<T>* rb = ra + NKZ;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int z = t + r * LPS;
      cx<T> a, b;
      if (z <= L / 2) {
        a = ra[z];
        b = rb[z];
        if (z == 0 || z == L / 2) {
          a.im = (T)0;
          b.im = (T)0;
        }
      } else {
        a = cconj(ra[L - z]);
        b = cconj(rb[L - z]);
      }
      v[r] = {a.re - b.im, a.im + b.re};
    }
    wave_butterflies<T, L, 8, 1, true>(v, t, twl);
    wave_lds_sync();
    wave_write<T, L, 8, 1>(v, seq, t);
    wave_lds_sync();
  }
  wave_stages<T, L, 8, true>(seq, t, twl);
  __syncthreads();
  double acc[3] = {0.0, 0.0, 0.0};
  if (threadIdx.x < kThreads) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = threadIdx.x + i * kThreads;
      const int l = e / L, z = e - l * L;
      const bool odd = l & 1;
      const int at = (l >> 1) * SS + skew<T>(z);
      double v[kWaveFields];
#pragma unroll
      for (int f = 0; f < kWaveFields; ++f) {
        const cx<T> c = S[f * REG + at];
        v[f] = (double)(odd ? c.im : c.re);
      }
      double p1 = 0.0, p2 = 0.0, p3 = 0.0;
      const bool valid = line0 + l < lines;
      if constexpr (GRAD) {
        double dd[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (valid) {
          double q1, q2, q3;
          wb_phi_k<false>(v[0], v[1], v[2], v[3], v[4], g.wb, p1, p2, p3);
          wb_phi_partials_k<false>(v[0], v[1], v[2], v[3], v[4], g.wb, q1, q2, q3, dd);
        }
#pragma unroll
        for (int f = 0; f < kWaveFields; ++f) {
          cx<T>* c = &S[f * REG + at];
          if (odd)
            c->im = (T)dd[f];
          else
            c->re = (T)dd[f];
        }
      } else {
        if (valid) wb_phi_k<false>(v[0], v[1], v[2], v[3], v[4], g.wb, p1, p2, p3);
      }
      acc[0] += p1;
      acc[1] += p2;
      acc[2] += p3;
    }
  }
  if constexpr (GRAD) {
    __syncthreads();
    wave_stages<T, L, 1, false>(seq, t, twl);
    // split the packed spectra into the two lines' half spectra, as store_r2c_halves
    for (int e = lane; e < GL * NKZ; e += 64) {
      const int l = e / NKZ, k = e - l * NKZ;
      if (line0 + l < lines) {
        const cx<T>* zl = reg + (l >> 1) * SS;
        const cx<T> a = zl[skew<T>(k)];
        const cx<T> b = cconj(zl[skew<T>((L - k) % L)]);
        fld[(line0 + l) * g.ldk + k] = (l & 1) ? scale(mul_mi(a - b), (T)0.5) : scale(a + b, (T)0.5);
      }
    }
  }
  block_sum_store<3, kThreads / 64>(acc, partial, SLOT_PHI1, pstride, blockIdx.x);
}

// The vector weight -w3 k_d of one element (d = 0, 1, 2), the expression
// element_weights evaluates, for a caller that keeps only w3 and w2.
template <typename T>
__device__ __forceinline__ T element_wk(const GridArgs& g, int x, int y, int kz, T w3, int d) {
  if constexpr (sizeof(T) <= 4) {
    const float mx = (float)wrapf(x, g.Nx), my = (float)wrapf(y, g.Ny), mz = (float)kz;
    return -w3 * kvec_component<float>(g, mx, my, mz, d);
  } else {
    const double mx = wrapf(x, g.Nx), my = wrapf(y, g.Ny), mz = kz;
    return (T)(-w3 * kvec_component<double>(g, mx, my, mz, d));
  }
}

// Forward x pass with the weight multiply (the five White-Bear fields), one
// wave per tile of SPW = 64/(L/8) columns: the wave loads its columns into its
// LDS slice, transforms them, keeps each lane's eight spectral elements and
// their w3, w2 in registers, and for every field multiplies, runs the inverse
// transform (its first stage straight from registers) and stores.  Four waves
// per block take four consecutive tiles (16 columns at L = 128, so a block
// reads whole 128-byte rows).  Same stages and weights as k_xconv_fwd; the
// compiler contracts some products into FMAs differently in the two kernels
// (the weights' k-vector and, here, the products that feed the first inverse
// stage in registers), so the results agree to rounding, not bit for bit.
template <typename T, int L>
__global__ __launch_bounds__(kThreads) void k_xconv_fwd_wave(GridArgs g, const cx<T>* __restrict__ in,
                                                              cx<T>* __restrict__ out,
                                                              const cx<T>* __restrict__ tw) {
  static_assert(wave_fft_length(L), "the wave transforms take power-of-two lengths 16..512");
  constexpr int LPS = L / 8, SPW = 64 / LPS, NWV = kThreads / 64;
  constexpr int SS = lds_stride<T, L>();
  __shared__ cx<T> S[NWV * SPW * SS];
  __shared__ cx<T> twl[L];
  for (int i = threadIdx.x; i < L; i += kThreads) twl[i] = tw[i];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long ntiles = ((long long)g.Ny * g.Nkz + SPW - 1) / SPW;
  const int tile = blockIdx.x * NWV + wave;
  if (tile >= ntiles) return;  // no block barrier follows
  cx<T>* slice = S + wave * SPW * SS;
  const long long plane = (long long)g.Ny * g.ldk;
  // column lc = lane % SPW for the loads and stores (64 % SPW == 0, so every
  // element e = lane + 64 i of the tile belongs to it)
  const int lc = lane % SPW;
  int ly, lkz;
  bool lok;
  x_tile_column<SPW>(g, tile, lc, ly, lkz, lok);
  const long long loff = (long long)ly * g.ldk + lkz;
  for (int e = lane; e < SPW * L; e += 64)
    slice[lc * SS + skew<T>(e / SPW)] = load_or_zero(in, (long long)(e / SPW) * plane + loff, 0, lok);
  wave_lds_sync();
  // column sq = lane / (L/8) for the transforms; the lane holds its elements
  // kx = t + r L/8
  const int sq = lane / LPS, t = lane - sq * LPS;
  cx<T>* seq = slice + sq * SS;
  wave_stages<T, L, 1, false>(seq, t, twl);
  int cy, ckz;
  bool cok;
  x_tile_column<SPW>(g, tile, sq, cy, ckz, cok);  // (cy, ckz) valid also where !cok
  cx<T> sv[8];
  T w3[8], w2[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    sv[r] = seq[skew<T>(t + r * LPS)];
    T wk[3];
    element_weights<T>(g, t + r * LPS, cy, ckz, w3[r], w2[r], wk);
  }
  wave_lds_sync();
  // the n2 field first: its weights w2 are then dead for the other four
#pragma unroll
  for (int fo = 0; fo < 5; ++fo) {
    const int f = fo == 0 ? 1 : (fo == 1 ? 0 : fo);
    cx<T> v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      if (f == 0)
        v[r] = scale(sv[r], w3[r]);
      else if (f == 1)
        v[r] = scale(sv[r], w2[r]);
      else
        v[r] = mul_i(scale(sv[r], element_wk<T>(g, t + r * LPS, cy, ckz, w3[r], f - 2)));
    }
    wave_butterflies<T, L, 8, 1, true>(v, t, twl);
    wave_lds_sync();
    wave_write<T, L, 8, 1>(v, seq, t);
    wave_lds_sync();
    wave_stages<T, L, 8, true>(seq, t, twl);
    cx<T>* o = out + f * g.ks;
    if (lok)
      for (int e = lane; e < SPW * L; e += 64) o[(long long)(e / SPW) * plane + loff] = slice[lc * SS + skew<T>(e / SPW)];
    wave_lds_sync();  // the next field's first stage overwrites the slice
  }
}

}  // namespace deft
