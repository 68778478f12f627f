// Batched launches of the evaluation passes: one launch covers every grid of a
// batch whose pass axis has the same length and tuning.  Block b of a launch
// belongs to the grid i with off[i] <= b < off[i+1] and runs that grid's block
// b - off[i] through the same device body as the single-grid kernel, so every
// grid gets bit-for-bit the result of its own single launch.
//
// `sel` lets one launch layout serve every round (the graph-captured rounds
// launch every pass over every grid of the batch): a block returns at once
// unless its grid's request kind (0 none, 1 energy, 2 energy + gradient)
// matches: sel 0 runs all, 1 any request, 2 gradient requests, 3 energy-only
// requests.  The test is per block, before any barrier.
#pragma once

#include "passes.hpp"

namespace deft {

// Everything one grid's evaluation passes need (one entry per grid and round).
template <typename T>
struct BatchArgs {
  GridArgs g;
  double* x;
  const double* d;
  double m1, m2;
  int nmoves;
  int have_old, precond;
  int nzb;  // z-pass blocks (reduction length)
  const T* vext;  // FP32 for FP32 passes, FP64 for FP64 passes
  cx<T>* spec0;  // the density spectrum / back-convolution
  cx<T>* spec1;  // the five weighted-density spectra
  double* partial;
  int pstride;
  const cx<T>* twx;
  const cx<T>* twy;
  const cx<T>* twz;
  const double* og;
  double* grad;
  double* pgrad;
  double* dscal;
  const double* sktab;  // S(k)/N of the structure-factor preconditioner (gradient requests that use it)
  int kind;  // 0 no request this round, 1 energy, 2 energy + gradient
};

__device__ __forceinline__ bool batch_skip(int kind, int sel) {
  return (sel == 1 && kind == 0) || (sel == 2 && kind != 2) || (sel == 3 && kind != 1);
}

// One launch group: n grids, block offsets off[0..n] (off[0] = 0) and the
// index of each grid's BatchArgs entry.
struct BatchGroup {
  const int* off;
  const int* idx;
  int n;
};

__device__ __forceinline__ int batch_find(const BatchGroup& grp, int b) {
  int lo = 0, hi = grp.n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (grp.off[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <typename T, int L, int ZB>
__global__ __launch_bounds__(kThreads) void kb_zfwd_density(BatchGroup grp, const BatchArgs<T>* __restrict__ args,
                                                            int sel) {
  const int i = batch_find(grp, blockIdx.x);
  const BatchArgs<T>& a = args[grp.idx[i]];
  if (batch_skip(a.kind, sel)) return;
  zfwd_density_body<T, L, ZB>(blockIdx.x - grp.off[i], a.g, a.x, a.d, a.m1, a.m2, a.nmoves, a.vext, a.spec0,
                              a.partial, a.pstride, a.twz);
}

// field set: 0 = spec0, 1 = spec1 (gridDim.y fields of it)
template <typename T, int L, bool INV, int TB>
__global__ __launch_bounds__(kThreads) void kb_ypass(BatchGroup grp, const BatchArgs<T>* __restrict__ args,
                                                     int which, int sel) {
  const int i = batch_find(grp, blockIdx.x);
  const BatchArgs<T>& a = args[grp.idx[i]];
  if (batch_skip(a.kind, sel)) return;
  ypass_body<T, L, INV, TB>(blockIdx.x - grp.off[i], blockIdx.y, a.g, which ? a.spec1 : a.spec0, a.twy);
}

// (x passes: the XCD pairing of the single-grid kernels is applied to the
// batch-wide block index, so the tiles a block gets can belong to a
// neighbouring grid; each tile is computed exactly as in the single launch)
template <typename T, int L, int TB>
__global__ __launch_bounds__(kThreads) void kb_xconv_fwd(BatchGroup grp, const BatchArgs<T>* __restrict__ args,
                                                         int sel) {
  const int b = xcd_pair_tile(blockIdx.x, gridDim.x);
  const int i = batch_find(grp, b);
  const BatchArgs<T>& a = args[grp.idx[i]];
  if (batch_skip(a.kind, sel)) return;
  xconv_fwd_body<T, L, TB>(b - grp.off[i], a.g, a.spec0, a.spec1, a.twx);
}

template <typename T, int L, int TB>
__global__ __launch_bounds__(kThreads) void kb_xconv_bwd(BatchGroup grp, const BatchArgs<T>* __restrict__ args,
                                                         int sel) {
  const int b = xcd_pair_tile(blockIdx.x, gridDim.x);
  const int i = batch_find(grp, b);
  const BatchArgs<T>& a = args[grp.idx[i]];
  if (batch_skip(a.kind, sel)) return;
  xconv_bwd_body<T, L, TB>(b - grp.off[i], a.g, a.spec1, a.spec0, a.twx);
}

template <typename T, int L, bool GRAD, int ZB>
__global__ __launch_bounds__(kThreads) void kb_zphi(BatchGroup grp, const BatchArgs<T>* __restrict__ args, int sel) {
  const int i = batch_find(grp, blockIdx.x);
  const BatchArgs<T>& a = args[grp.idx[i]];
  if (batch_skip(a.kind, sel)) return;
  zphi_body<T, L, GRAD, ZB>(blockIdx.x - grp.off[i], a.g, a.spec1, a.partial, a.pstride, a.twz);
}

template <typename T, int L, int ZB>
__global__ __launch_bounds__(kThreads) void kb_zgrad(BatchGroup grp, const BatchArgs<T>* __restrict__ args, int sel) {
  const int i = batch_find(grp, blockIdx.x);
  const BatchArgs<T>& a = args[grp.idx[i]];
  if (batch_skip(a.kind, sel)) return;
  zgrad_body<T, L, ZB>(blockIdx.x - grp.off[i], a.g, a.spec0, a.x, a.vext, a.og, a.have_old, a.precond, a.grad,
                       a.pgrad, a.partial, a.pstride, a.twz);
}

// The structure-factor preconditioner's passes (gradient requests with
// precond == PGRAD_SK): the fused z gradient pass, then after the one-field y
// pass the x pass with S(k)/N, and after the inverse y pass the finishing z pass.
template <typename T, int L, int ZB>
__global__ __launch_bounds__(kThreads) void kb_zgrad_sk(BatchGroup grp, const BatchArgs<T>* __restrict__ args) {
  const int i = batch_find(grp, blockIdx.x);
  const BatchArgs<T>& a = args[grp.idx[i]];
  zgrad_sk_body<T, L, ZB>(blockIdx.x - grp.off[i], a.g, a.spec0, a.x, a.vext, a.grad, a.twz);
}

template <typename T, int L, int TB>
__global__ __launch_bounds__(kThreads) void kb_xsk(BatchGroup grp, const BatchArgs<T>* __restrict__ args) {
  const int b = xcd_pair_tile(blockIdx.x, gridDim.x);
  const int i = batch_find(grp, b);
  const BatchArgs<T>& a = args[grp.idx[i]];
  xsk_body<T, L, TB>(b - grp.off[i], a.g, a.spec0, a.sktab, a.twx);
}

template <typename T, int L, int ZB>
__global__ __launch_bounds__(kThreads) void kb_zsk_finish(BatchGroup grp, const BatchArgs<T>* __restrict__ args) {
  const int i = batch_find(grp, blockIdx.x);
  const BatchArgs<T>& a = args[grp.idx[i]];
  zsk_finish_body<T, L, ZB>(blockIdx.x - grp.off[i], a.g, a.spec0, a.x, a.vext, a.grad, a.og, a.have_old, a.pgrad,
                            a.partial, a.pstride, a.twz);
}

// nslots blocks per grid: block b reduces slot b % nslots of grid b / nslots
// (with sel != 0: nothing for grids without a request, the six energy slots
// for energy-only requests)
template <typename T>
__global__ __launch_bounds__(kThreads) void kb_reduce(BatchGroup grp, const BatchArgs<T>* __restrict__ args,
                                                      int nslots, int sel) {
  const int i = blockIdx.x / nslots;
  const BatchArgs<T>& a = args[grp.idx[i]];
  const int s = (int)(blockIdx.x - i * nslots);
  if (sel != 0 && (a.kind == 0 || (a.kind == 1 && s >= 6))) return;
  reduce_body(SLOT_IDEAL + s, a.partial, a.pstride, a.nzb, a.dscal);
}

// One grid's conjugate-gradient update inside a round: direction = -pg +
// beta * direction (or -pg without a direction; -grad when `reset`), oldgrad =
// grad, and the grad.direction partial sums, exactly as the single-grid
// k_cg_update / k_cg_reset with eblocks blocks.
struct CgArgs {
  long long N;
  const double* pg;
  const double* g;
  double* d;
  double* og;
  double beta;
  int have_dir, reset;
  double* partial;
  int pstride, eblocks;
  double* dscal;
};

static __global__ __launch_bounds__(kThreads) void kb_cg(BatchGroup grp, const CgArgs* __restrict__ args) {
  const int i = batch_find(grp, blockIdx.x);
  const CgArgs& a = args[grp.idx[i]];
  const int bx = blockIdx.x - grp.off[i];
  double acc[1] = {0.0};
  const long long stride = (long long)a.eblocks * kThreads;
  if (a.reset) {
    for (long long k = (long long)bx * kThreads + threadIdx.x; k < a.N; k += stride) {
      const double gi = a.g[k];
      a.d[k] = -gi;
      acc[0] += gi * -gi;
    }
  } else {
    for (long long k = (long long)bx * kThreads + threadIdx.x; k < a.N; k += stride) {
      const double gi = a.g[k];
      double di = -a.pg[k];
      if (a.have_dir) di = __dadd_rn(di, __dmul_rn(a.beta, a.d[k]));
      a.d[k] = di;
      a.og[k] = gi;
      acc[0] += gi * di;
    }
  }
  block_sum_store<1>(acc, a.partial, SLOT_GD, a.pstride, bx);
}

// block r reduces request r's grad.direction sums
static __global__ __launch_bounds__(kThreads) void kb_cg_reduce(const CgArgs* __restrict__ args) {
  const CgArgs& a = args[blockIdx.x];
  reduce_body(SLOT_GD, a.partial, a.pstride, a.eblocks, a.dscal);
}

}  // namespace deft
