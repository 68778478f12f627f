// Launches of the tensor White-Bear passes (eleven weighted fields), kept out of
// the main launch translation units so the two compile in parallel.
#pragma once

#include "launch.hpp"

namespace deft {

template <typename T>
void TensorLaunch<T>::xconv_fwd(deft_plan_s *p, const GridArgs &g) {
  const double cb = (double)sizeof(cx<T>) * p->Nk;
  with_length(p->Nx, [&](auto LX) {
    constexpr int L = decltype(LX)::value;
    ProfScope ps(p, kid<T>(K_XCONV_FWD), 12 * cb);
    hipLaunchKernelGGL((k_xconv_fwd<T, L, kTileBytes[0], 11>), dim3(x_blocks<T>(p, kTileBytes[0])), dim3(kThreads), 0,
                       p->stream, g, spec_ptr<T>(p, 0), spec_ptr<T>(p, 1), tw_ptr<T>(p, 0));
  });
}

template <typename T>
void TensorLaunch<T>::xconv_bwd(deft_plan_s *p, const GridArgs &g) {
  const double cb = (double)sizeof(cx<T>) * p->Nk;
  with_length(p->Nx, [&](auto LX) {
    constexpr int L = decltype(LX)::value;
    ProfScope ps(p, kid<T>(K_XCONV_BWD), 12 * cb);
    hipLaunchKernelGGL((k_xconv_bwd<T, L, kTileBytes[0], 11>), dim3(x_blocks<T>(p, kTileBytes[0])), dim3(kThreads), 0,
                       p->stream, g, spec_ptr<T>(p, 1), spec_ptr<T>(p, 0), tw_ptr<T>(p, 0));
  });
}

template <typename T>
void TensorLaunch<T>::zphi(deft_plan_s *p, const GridArgs &g, bool grad, int zb) {
  const double cb = (double)sizeof(cx<T>) * p->Nk;
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    with_zb(zb, [&](auto ZBc) {
      constexpr int ZB = decltype(ZBc)::value;
      ProfScope ps(p, kid<T>(grad ? K_ZPHI_EG : K_ZPHI_E), (grad ? 22 : 11) * cb);
      if (grad)
        hipLaunchKernelGGL((k_zphi<T, L, true, ZB, 11>), dim3(zblocks_for(p, ZB)), dim3(kThreads), 0, p->stream, g,
                           spec_ptr<T>(p, 1), p->partial, p->pstride, tw_ptr<T>(p, 2));
      else
        hipLaunchKernelGGL((k_zphi<T, L, false, ZB, 11>), dim3(zblocks_for(p, ZB)), dim3(kThreads), 0, p->stream, g,
                           spec_ptr<T>(p, 1), p->partial, p->pstride, tw_ptr<T>(p, 2));
    });
  });
}

}  // namespace deft
