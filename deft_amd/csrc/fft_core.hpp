// LDS-resident mixed-radix Stockham FFT building blocks for gfx950.
//
// A workgroup holds `nseq` complex sequences of length L in LDS (sequence
// stride SS >= L, padded to spread LDS banks) and transforms all of them in
// place with radix-8/4/2/3/5 Stockham stages.  Each stage reads one buffer and
// writes the other, so two buffers ping-pong; the caller learns which one holds
// the result.  Twiddles come from a per-length table tw[m] = exp(-2 pi i m / L)
// computed in FP64 on the host and stored in T.
//
// Forward = exp(-i k x) (FFTW's FFTW_FORWARD sign, used by deft's r2c call in
// src/Grid.cpp:473); inverse = exp(+i k x); neither is normalised here.
#pragma once

#include <hip/hip_runtime.h>

namespace deft {

// Every kernel that runs these stages launches blocks of kBlockThreads
// threads; the stage loops stride by the constant rather than by blockDim.x
// (which the compiler would read from the dispatch packet in every loop).
constexpr int kBlockThreads = 256;

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> scale(cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.re, -a.im}; }
template <typename T>
__device__ __forceinline__ cx<T> mul_i(cx<T> a) { return {-a.im, a.re}; }
template <typename T>
__device__ __forceinline__ cx<T> mul_mi(cx<T> a) { return {a.im, -a.re}; }
// multiply by -i for the forward transform, +i for the inverse
template <typename T, bool INV>
__device__ __forceinline__ cx<T> rot(cx<T> a) { return INV ? mul_i(a) : mul_mi(a); }

// ---------------------------------------------------------------------------
// small DFT kernels on registers
// ---------------------------------------------------------------------------

template <typename T, bool INV>
__device__ __forceinline__ void dft2(cx<T>& a, cx<T>& b) {
  const cx<T> t = a;
  a = t + b;
  b = t - b;
}

template <typename T, bool INV>
__device__ __forceinline__ void dft4(cx<T>& a0, cx<T>& a1, cx<T>& a2, cx<T>& a3) {
  const cx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3;
  const cx<T> t3 = rot<T, INV>(a1 - a3);
  a0 = t0 + t2;
  a2 = t0 - t2;
  a1 = t1 + t3;
  a3 = t1 - t3;
}

template <typename T, bool INV>
__device__ __forceinline__ void dft8(cx<T>* v) {
  cx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<T, INV>(e0, e1, e2, e3);
  dft4<T, INV>(o0, o1, o2, o3);
  const T h = (T)0.70710678118654752440;
  // o1 *= w8, o2 *= w8^2, o3 *= w8^3 with w8 = exp(-+ i pi/4)
  if (INV) {
    o1 = {(o1.re - o1.im) * h, (o1.re + o1.im) * h};
    o3 = {(-o3.re - o3.im) * h, (o3.re - o3.im) * h};
  } else {
    o1 = {(o1.re + o1.im) * h, (o1.im - o1.re) * h};
    o3 = {(o3.im - o3.re) * h, (-o3.re - o3.im) * h};
  }
  o2 = rot<T, INV>(o2);
  v[0] = e0 + o0;
  v[4] = e0 - o0;
  v[1] = e1 + o1;
  v[5] = e1 - o1;
  v[2] = e2 + o2;
  v[6] = e2 - o2;
  v[3] = e3 + o3;
  v[7] = e3 - o3;
}

template <typename T, bool INV>
__device__ __forceinline__ void dft3(cx<T>* v) {
  const T c = (T)-0.5;
  const T s = (T)0.86602540378443864676;
  const cx<T> t1 = v[1] + v[2];
  const cx<T> t2 = v[0] + scale(t1, c);
  const cx<T> u = rot<T, INV>(scale(v[1] - v[2], s));
  v[0] = v[0] + t1;
  v[1] = t2 + u;
  v[2] = t2 - u;
}

template <typename T, bool INV>
__device__ __forceinline__ void dft5(cx<T>* v) {
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
  const cx<T> b1 = v[1] + v[4], b2 = v[2] + v[3];
  const cx<T> d1 = v[1] - v[4], d2 = v[2] - v[3];
  const cx<T> a0 = v[0];
  const cx<T> r1 = a0 + scale(b1, c1) + scale(b2, c2);
  const cx<T> r2 = a0 + scale(b1, c2) + scale(b2, c1);
  const cx<T> i1 = rot<T, INV>(scale(d1, s1) + scale(d2, s2));
  const cx<T> i2 = rot<T, INV>(scale(d1, s2) - scale(d2, s1));
  v[0] = a0 + b1 + b2;
  v[1] = r1 + i1;
  v[4] = r1 - i1;
  v[2] = r2 + i2;
  v[3] = r2 - i2;
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(cx<T>* v) {
  if constexpr (R == 2) {
    dft2<T, INV>(v[0], v[1]);
  } else if constexpr (R == 3) {
    dft3<T, INV>(v);
  } else if constexpr (R == 4) {
    dft4<T, INV>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 5) {
    dft5<T, INV>(v);
  } else {
    static_assert(R == 8, "unsupported radix");
    dft8<T, INV>(v);
  }
}

// ---------------------------------------------------------------------------
// LDS layout: element i of a sequence sits at skew(i) = i + (i >> SK), one pad
// slot per 2^SK elements (SK = 4 for 8-byte FP32 complex, 3 for 16-byte FP64
// complex).  Stockham stages write with strides R*NS (powers of two); the skew
// spreads those writes over all LDS banks.  Sequences are lds_stride<T, L>()
// apart (odd, so column-wise loads of several sequences do not collide either).
// ---------------------------------------------------------------------------
template <typename T>
constexpr int lds_skew_shift() {
  return sizeof(T) <= 4 ? 4 : 3;
}
template <typename T>
__device__ __forceinline__ constexpr int skew(int i) {
  return i + (i >> lds_skew_shift<T>());
}
template <typename T, int L>
constexpr int lds_stride() {
  return (L + (L >> lds_skew_shift<T>())) | 1;
}

// ---------------------------------------------------------------------------
// Stockham stages
// ---------------------------------------------------------------------------

// Radices 2..8 have register butterflies; any other prime factor p of L is one
// direct-DFT stage of radix p (O(p) per output, for odd grid edges such as
// 201 = 3 * 67 that the reference's FFTW plans take in stride).
constexpr int kMaxDirectRadix = 512;

constexpr int smallest_factor(int n) {
  for (int f = 2; f * f <= n; ++f)
    if (n % f == 0) return f;
  return n;
}

constexpr int pick_radix(int rem) {
  return (rem % 8 == 0)   ? 8
         : (rem % 4 == 0) ? 4
         : (rem % 2 == 0) ? 2
         : (rem % 3 == 0) ? 3
         : (rem % 5 == 0) ? 5
                          : smallest_factor(rem);
}

template <int L, int NS = 1>
constexpr int num_stages() {
  if constexpr (NS >= L) {
    return 0;
  } else {
    return 1 + num_stages<L, NS * pick_radix(L / NS)>();
  }
}

constexpr bool supported_length(int L) {
  if (L < 2) return false;
  for (int r = L; r > 1;) {
    const int f = pick_radix(r);
    if (f > kMaxDirectRadix) return false;
    r /= f;
  }
  return true;
}

// One radix-R stage: sub-transforms of length NS are combined into length NS*R.
template <typename T, int L, int SS, int R, int NS, bool INV>
__device__ __forceinline__ void stockham_stage(const cx<T>* __restrict__ src, cx<T>* __restrict__ dst, int nseq,
                                               const cx<T>* __restrict__ tw, int stride) {
  constexpr int M = L / R;
  constexpr int TWS = L / (NS * R);
  const int total = nseq * M;
  for (int b = threadIdx.x; b < total; b += stride) {
    const int s = b / M;
    const int j = b - s * M;
    const int k = j % NS;
    const cx<T>* in = src + s * SS;
    cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[skew<T>(j + r * M)];
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const cx<T> w = tw[k * r * TWS];
        v[r] = INV ? cmulc(v[r], w) : cmul(v[r], w);
      }
    }
    dft<T, R, INV>(v);
    cx<T>* out = dst + s * SS;
    const int o0 = (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out[skew<T>(o0 + r * NS)] = v[r];
  }
}

// Radix-R stage for a prime R without a register butterfly: one output per
// thread iteration, out[q] = sum_r w_R^(r q) * twiddled in[r] with w_R = tw[L/R]
// (exponents reduced mod R), accumulated in T.  Same data movement as
// stockham_stage.
template <typename T, int L, int SS, int R, int NS, bool INV>
__device__ __forceinline__ void stockham_stage_direct(const cx<T>* __restrict__ src, cx<T>* __restrict__ dst,
                                                      int nseq, const cx<T>* __restrict__ tw, int stride) {
  constexpr int M = L / R;
  constexpr int TWS = L / (NS * R);
  const int total = nseq * L;
  for (int e = threadIdx.x; e < total; e += stride) {
    const int s = e / L;
    const int rem = e - s * L;
    const int q = rem / M;
    const int j = rem - q * M;
    const int k = j % NS;
    const cx<T>* in = src + s * SS;
    cx<T> acc = in[skew<T>(j)];
    int wq = q;  // (r q) mod R
    for (int r = 1; r < R; ++r) {
      cx<T> v = in[skew<T>(j + r * M)];
      if constexpr (NS > 1) {
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

template <typename T, int L, int SS, int NS, bool INV>
__device__ __forceinline__ void stockham_run(cx<T>* a, cx<T>* b, int nseq, const cx<T>* __restrict__ tw,
                                             int stride) {
  if constexpr (NS < L) {
    constexpr int R = pick_radix(L / NS);
    if constexpr (R == 2 || R == 3 || R == 4 || R == 5 || R == 8)
      stockham_stage<T, L, SS, R, NS, INV>(a, b, nseq, tw, stride);
    else
      stockham_stage_direct<T, L, SS, R, NS, INV>(a, b, nseq, tw, stride);
    __syncthreads();
    stockham_run<T, L, SS, NS * R, INV>(b, a, nseq, tw, stride);
  }
}

// Copies the length-L twiddle table into LDS (one static array per T and L)
// and returns it; the caller's next __syncthreads() makes it visible.  The
// stages of a pass then read their twiddles from LDS instead of from global
// memory, whose latency every stage would otherwise wait out.
template <typename T, int L>
__device__ __forceinline__ const cx<T>* lds_twiddles(const cx<T>* __restrict__ tw) {
  __shared__ cx<T> t[L];
  for (int i = threadIdx.x; i < L; i += kBlockThreads) t[i] = tw[i];
  return t;
}

// Transforms the nseq sequences held in `a` (stride SS, elements at skew(i));
// `b` is scratch of the same size.  Returns the buffer holding the result.  All threads of the block
// must call it; the data in `a` must be visible (a __syncthreads() before).
// `stride` is the loop stride over butterflies, kBlockThreads unless the
// caller passes blockDim.x to keep the compiler from scheduling on a known
// trip count (see zphi_body).
template <typename T, int L, int SS, bool INV>
__device__ __forceinline__ cx<T>* fft_lds(cx<T>* a, cx<T>* b, int nseq, const cx<T>* __restrict__ tw,
                                          int stride = kBlockThreads) {
  static_assert(supported_length(L), "FFT length has a prime factor above kMaxDirectRadix");
  stockham_run<T, L, SS, 1, INV>(a, b, nseq, tw, stride);
  return (num_stages<L>() % 2 == 0) ? a : b;
}

}  // namespace deft
