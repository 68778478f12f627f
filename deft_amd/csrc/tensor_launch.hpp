// Launches of the passes of the other functionals (tensor White-Bear's eleven
// weighted fields, Mark II's Phi pass), kept out of the main launch
// translation units so they compile in parallel.
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
  const bool tensor = p->functional == DEFT_WHITE_BEAR_TENSOR;
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    with_zb(zb, [&](auto ZBc) {
      constexpr int ZB = decltype(ZBc)::value;
      const int nw = tensor ? 11 : 5;
      ProfScope ps(p, kid<T>(grad ? K_ZPHI_EG : K_ZPHI_E), (grad ? 2 : 1) * nw * cb);
      const dim3 grid(zblocks_for(p, ZB));
      if (tensor && grad)
        hipLaunchKernelGGL((k_zphi11<T, L, true, ZB>), grid, dim3(kThreads), 0, p->stream, g, spec_ptr<T>(p, 1),
                           p->partial, p->pstride, tw_ptr<T>(p, 2));
      else if (tensor)
        hipLaunchKernelGGL((k_zphi11<T, L, false, ZB>), grid, dim3(kThreads), 0, p->stream, g, spec_ptr<T>(p, 1),
                           p->partial, p->pstride, tw_ptr<T>(p, 2));
      else if (grad)
        hipLaunchKernelGGL((k_zphi<T, L, true, ZB, 1>), grid, dim3(kThreads), 0, p->stream, g, spec_ptr<T>(p, 1),
                           p->partial, p->pstride, tw_ptr<T>(p, 2));
      else
        hipLaunchKernelGGL((k_zphi<T, L, false, ZB, 1>), grid, dim3(kThreads), 0, p->stream, g, spec_ptr<T>(p, 1),
                           p->partial, p->pstride, tw_ptr<T>(p, 2));
    });
  });
}

}  // namespace deft
