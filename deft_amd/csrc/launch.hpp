// Kernel launch sequences of one energy / energy+gradient evaluation and of
// the FFT seam, templated on the FFT precision T.  Included by exactly one
// translation unit per T, which instantiates Launch<T>.
#pragma once

#include <type_traits>

#include "plan.hpp"

namespace deft {

#define DEFT_HIP_CHECK(expr)                                                                     \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return set_error(DEFT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// The grid edges the FFT kernels are instantiated for.
template <typename Fn>
inline bool with_length(int L, Fn &&fn) {
  switch (L) {
#define DEFT_CASE(n)                        \
  case n:                                   \
    fn(std::integral_constant<int, n>{});   \
    return true;
    DEFT_CASE(16)
    DEFT_CASE(20)
    DEFT_CASE(32)
    DEFT_CASE(48)
    DEFT_CASE(64)
    DEFT_CASE(96)
    DEFT_CASE(128)
    DEFT_CASE(256)
#undef DEFT_CASE
    default:
      return false;
  }
}

inline int y_blocks(const deft_plan_s *p) {
  const int C = ccols(p->Ny);
  return y_main_blocks(p->Nx, p->Nkz, C) + y_tail_blocks(p->Nx, p->Nkz, C);
}
inline int x_blocks(const deft_plan_s *p) {
  const long long ncols = (long long)p->Ny * p->Nkz;
  return (int)((ncols + ccols(p->Nx) - 1) / ccols(p->Nx));
}

// Applies all pending moves but the last two with axpy; returns the last two
// for fusion into the density pass.
inline void take_moves(deft_plan_s *p, double &m1, double &m2, int &nm) {
  const int first = p->npend > 2 ? p->npend - 2 : 0;
  for (int i = 0; i < first; ++i) {
    ProfScope ps(p, 2 * K_AXPY, 24.0 * p->N);
    hipLaunchKernelGGL(k_axpy, dim3(p->eblocks), dim3(kThreads), 0, p->stream, p->N, p->x, p->d, p->pend[i]);
  }
  nm = p->npend - first;
  m1 = nm > 0 ? p->pend[first] : 0.0;
  m2 = nm > 1 ? p->pend[first + 1] : 0.0;
  p->npend = 0;
}

template <typename T>
static void launch_density_to_weighted(deft_plan_s *p, const GridArgs &g) {
  hipStream_t st = p->stream;
  const double cb = (double)sizeof(cx<T>) * p->Nk;  // bytes of one half spectrum
  cx<T> *S0 = spec_ptr<T>(p, 0), *S1 = spec_ptr<T>(p, 1);
  double m1, m2;
  int nm;
  take_moves(p, m1, m2, nm);
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    ProfScope ps(p, kid<T>(K_ZFWD_DENSITY), (nm > 0 ? 28.0 : 12.0) * p->N + cb);
    hipLaunchKernelGGL((k_zfwd_density<T, L>), dim3(p->zblocks), dim3(kThreads), 0, st, g, p->x, p->d, m1, m2, nm,
                       p->vext, S0, p->partial, p->pstride, tw_ptr<T>(p, 2));
  });
  with_length(p->Ny, [&](auto LY) {
    constexpr int L = decltype(LY)::value;
    ProfScope ps(p, kid<T>(K_YPASS1), 2 * cb);
    hipLaunchKernelGGL((k_ypass<T, L, false, 1>), dim3(y_blocks(p)), dim3(kThreads), 0, st, g, S0,
                       tw_ptr<T>(p, 1));
  });
  with_length(p->Nx, [&](auto LX) {
    constexpr int L = decltype(LX)::value;
    ProfScope ps(p, kid<T>(K_XCONV_FWD), 6 * cb);
    hipLaunchKernelGGL((k_xconv_fwd<T, L>), dim3(x_blocks(p)), dim3(kThreads), 0, st, g, S0, S1,
                       tw_ptr<T>(p, 0));
  });
  with_length(p->Ny, [&](auto LY) {
    constexpr int L = decltype(LY)::value;
    ProfScope ps(p, kid<T>(K_YPASS5), 10 * cb);
    hipLaunchKernelGGL((k_ypass<T, L, true, 5>), dim3(y_blocks(p), 5), dim3(kThreads), 0, st, g, S1,
                       tw_ptr<T>(p, 1));
  });
}

template <typename T>
int Launch<T>::evaluate(deft_plan_s *p, bool grad, bool have_old, bool precond) {
  const GridArgs g = grid_args(p);
  hipStream_t st = p->stream;
  const double cb = (double)sizeof(cx<T>) * p->Nk;
  cx<T> *S0 = spec_ptr<T>(p, 0), *S1 = spec_ptr<T>(p, 1);
  launch_density_to_weighted<T>(p, g);
  if (!grad) {
    with_length(p->Nz, [&](auto LZ) {
      constexpr int L = decltype(LZ)::value;
      ProfScope ps(p, kid<T>(K_ZPHI_E), 5 * cb);
      hipLaunchKernelGGL((k_zphi<T, L, false>), dim3(p->zblocks), dim3(kThreads), 0, st, g, S1, p->partial,
                         p->pstride, tw_ptr<T>(p, 2));
    });
    ProfScope ps(p, 2 * K_REDUCE, 48.0 * p->zblocks);
    hipLaunchKernelGGL(k_reduce, dim3(6), dim3(kThreads), 0, st, p->partial, p->pstride, (int)SLOT_IDEAL,
                       p->zblocks, p->dscal);
  } else {
    with_length(p->Nz, [&](auto LZ) {
      constexpr int L = decltype(LZ)::value;
      ProfScope ps(p, kid<T>(K_ZPHI_EG), 10 * cb);
      hipLaunchKernelGGL((k_zphi<T, L, true>), dim3(p->zblocks), dim3(kThreads), 0, st, g, S1, p->partial,
                         p->pstride, tw_ptr<T>(p, 2));
    });
    with_length(p->Ny, [&](auto LY) {
      constexpr int L = decltype(LY)::value;
      ProfScope ps(p, kid<T>(K_YPASS5), 10 * cb);
      hipLaunchKernelGGL((k_ypass<T, L, false, 5>), dim3(y_blocks(p), 5), dim3(kThreads), 0, st, g, S1,
                         tw_ptr<T>(p, 1));
    });
    with_length(p->Nx, [&](auto LX) {
      constexpr int L = decltype(LX)::value;
      ProfScope ps(p, kid<T>(K_XCONV_BWD), 6 * cb);
      hipLaunchKernelGGL((k_xconv_bwd<T, L>), dim3(x_blocks(p)), dim3(kThreads), 0, st, g, S1, S0,
                         tw_ptr<T>(p, 0));
    });
    with_length(p->Ny, [&](auto LY) {
      constexpr int L = decltype(LY)::value;
      ProfScope ps(p, kid<T>(K_YPASS1), 2 * cb);
      hipLaunchKernelGGL((k_ypass<T, L, true, 1>), dim3(y_blocks(p)), dim3(kThreads), 0, st, g, S0,
                         tw_ptr<T>(p, 1));
    });
    with_length(p->Nz, [&](auto LZ) {
      constexpr int L = decltype(LZ)::value;
      ProfScope ps(p, kid<T>(K_ZGRAD), cb + (have_old ? 36.0 : 28.0) * p->N);
      hipLaunchKernelGGL((k_zgrad<T, L>), dim3(p->zblocks), dim3(kThreads), 0, st, g, S0, p->x, p->vext, p->og,
                         have_old ? 1 : 0, precond ? 1 : 0, p->g, p->pg, p->partial, p->pstride, tw_ptr<T>(p, 2));
    });
    ProfScope ps(p, 2 * K_REDUCE, 64.0 * p->zblocks);
    hipLaunchKernelGGL(k_reduce, dim3(8), dim3(kThreads), 0, st, p->partial, p->pstride, (int)SLOT_IDEAL,
                       p->zblocks, p->dscal);
  }
  DEFT_HIP_CHECK(hipGetLastError());
  return plan_sync(p);
}

template <typename T>
int Launch<T>::fft_forward(deft_plan_s *p, const double *in_dev, double scl) {
  const GridArgs g = grid_args(p);
  hipStream_t st = p->stream;
  cx<T> *S0 = spec_ptr<T>(p, 0);
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    hipLaunchKernelGGL((k_zfwd_plain<T, L>), dim3(p->zblocks), dim3(kThreads), 0, st, g, in_dev, S0,
                       tw_ptr<T>(p, 2));
  });
  with_length(p->Ny, [&](auto LY) {
    constexpr int L = decltype(LY)::value;
    hipLaunchKernelGGL((k_ypass<T, L, false, 1>), dim3(y_blocks(p)), dim3(kThreads), 0, st, g, S0,
                       tw_ptr<T>(p, 1));
  });
  with_length(p->Nx, [&](auto LX) {
    constexpr int L = decltype(LX)::value;
    hipLaunchKernelGGL((k_xpass_plain<T, L, false>), dim3(x_blocks(p)), dim3(kThreads), 0, st, g, S0, scl,
                       tw_ptr<T>(p, 0));
  });
  DEFT_HIP_CHECK(hipGetLastError());
  return DEFT_OK;
}

template <typename T>
int Launch<T>::fft_inverse(deft_plan_s *p, double *out_dev, double scl) {
  const GridArgs g = grid_args(p);
  hipStream_t st = p->stream;
  cx<T> *S0 = spec_ptr<T>(p, 0);
  with_length(p->Nx, [&](auto LX) {
    constexpr int L = decltype(LX)::value;
    hipLaunchKernelGGL((k_xpass_plain<T, L, true>), dim3(x_blocks(p)), dim3(kThreads), 0, st, g, S0, scl,
                       tw_ptr<T>(p, 0));
  });
  with_length(p->Ny, [&](auto LY) {
    constexpr int L = decltype(LY)::value;
    hipLaunchKernelGGL((k_ypass<T, L, true, 1>), dim3(y_blocks(p)), dim3(kThreads), 0, st, g, S0,
                       tw_ptr<T>(p, 1));
  });
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    hipLaunchKernelGGL((k_zinv_plain<T, L>), dim3(p->zblocks, 1), dim3(kThreads), 0, st, g, S0, out_dev,
                       p->N, 1.0, tw_ptr<T>(p, 2));
  });
  DEFT_HIP_CHECK(hipGetLastError());
  return DEFT_OK;
}

template <typename T>
int Launch<T>::weighted_densities(deft_plan_s *p, double *out_dev) {
  const GridArgs g = grid_args(p);
  hipStream_t st = p->stream;
  cx<T> *S1 = spec_ptr<T>(p, 1);
  launch_density_to_weighted<T>(p, g);
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    hipLaunchKernelGGL((k_zinv_plain<T, L>), dim3(p->zblocks, 5), dim3(kThreads), 0, st, g, S1, out_dev,
                       p->N, 1.0, tw_ptr<T>(p, 2));
  });
  DEFT_HIP_CHECK(hipGetLastError());
  return DEFT_OK;
}

}  // namespace deft
