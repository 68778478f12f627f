// Kernel launch sequences of one energy / energy+gradient evaluation and of
// the FFT seam, templated on the FFT precision T, plus the autotuner of the
// pass variants.  Included by exactly one translation unit per T, which
// instantiates Launch<T>.
#pragma once

#include <cstring>
#include <map>
#include <tuple>
#include <type_traits>
#include <vector>

#include "batch_kernels.hpp"
#include "plan.hpp"
#include "wave_fft.hpp"

namespace deft {

#define DEFT_HIP_CHECK(expr)                                                                     \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return set_error(DEFT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TRY_LAUNCH(expr)             \
  do {                               \
    const int rc_ = (expr);          \
    if (rc_ != DEFT_OK) return rc_;  \
  } while (0)

template <int V>
using ic = std::integral_constant<int, V>;

// The grid edges the FFT kernels are instantiated for.
template <typename Fn>
inline bool with_length(int L, Fn &&fn) {
  switch (L) {
#define DEFT_CASE(n) \
  case n:            \
    fn(ic<n>{});     \
    return true;
    DEFT_CASE(16)
    DEFT_CASE(20)
    DEFT_CASE(21)
    DEFT_CASE(32)
    DEFT_CASE(48)
    DEFT_CASE(64)
    DEFT_CASE(96)
    DEFT_CASE(128)
    DEFT_CASE(201)
    DEFT_CASE(256)
#undef DEFT_CASE
    default:
      return false;
  }
}

// tunable variants: z-pass block size and y/x tile bytes
template <typename Fn>
inline void with_zb(int zb, Fn &&fn) {
  if (zb == kZBlocks[0])
    fn(ic<kZBlocks[0]>{});
  else
    fn(ic<kZBlocks[1]>{});
}
template <typename Fn>
inline void with_tb(int tb, Fn &&fn) {
  if (tb == kTileBytes[0])
    fn(ic<kTileBytes[0]>{});
  else
    fn(ic<kTileBytes[1]>{});
}

template <typename T>
inline int y_blocks(const deft_plan_s *p, int tb) {
  const int C = ccols<T>(p->Ny, tb);
  return y_main_blocks(p->Nx, p->Nkz, C) + y_tail_blocks(p->Nx, p->Nkz, C);
}
template <typename T>
inline int x_blocks(const deft_plan_s *p, int tb) {
  const long long ncols = (long long)p->Ny * p->Nkz;
  const int C = ccols<T>(p->Nx, tb);
  return (int)((ncols + C - 1) / C);
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

// --- single launches (each dispatches on the grid edge and the tuning) -----

template <typename T>
static void launch_ypass(deft_plan_s *p, const GridArgs &g, cx<T> *data, bool inv, int nf, int tb) {
  const double cb = (double)sizeof(cx<T>) * p->Nk;
  const int fam = nf == 1 ? (inv ? K_YPASS1_INV : K_YPASS1) : (inv ? K_YPASS5 : K_YPASS5_FWD);
  with_length(p->Ny, [&](auto LY) {
    constexpr int L = decltype(LY)::value;
    with_tb(tb, [&](auto TBc) {
      constexpr int TB = decltype(TBc)::value;
      ProfScope ps(p, kid<T>(fam), 2.0 * nf * cb);
      const dim3 grid(y_blocks<T>(p, TB), nf);
      if (inv && nf == 1)
        hipLaunchKernelGGL((k_ypass<T, L, true, 1, TB>), grid, dim3(kThreads), 0, p->stream, g, data, tw_ptr<T>(p, 1));
      else if (inv)
        hipLaunchKernelGGL((k_ypass<T, L, true, 5, TB>), grid, dim3(kThreads), 0, p->stream, g, data, tw_ptr<T>(p, 1));
      else if (nf == 1)
        hipLaunchKernelGGL((k_ypass<T, L, false, 1, TB>), grid, dim3(kThreads), 0, p->stream, g, data, tw_ptr<T>(p, 1));
      else
        hipLaunchKernelGGL((k_ypass<T, L, false, 5, TB>), grid, dim3(kThreads), 0, p->stream, g, data, tw_ptr<T>(p, 1));
    });
  });
}

// Plans with another functional than White-Bear run their own Phi pass
// (Mark II: compile-time factors; tensor White-Bear: eleven weighted fields,
// whose x passes use the first tile size only).  Those launches are compiled
// in their own translation units (tensor_f32.hip, tensor_f64.hip).
inline int plan_fields(const deft_plan_s *p) {
  return p->functional == DEFT_SFMT ? 16 : (p->functional == DEFT_WHITE_BEAR_TENSOR ? 11 : 5);
}

template <typename T>
struct TensorLaunch {
  static void xconv_fwd(deft_plan_s *p, const GridArgs &g);
  static void xconv_bwd(deft_plan_s *p, const GridArgs &g);
  static void zphi(deft_plan_s *p, const GridArgs &g, bool grad, int zb);
};

template <typename T>
static void launch_xconv_fwd(deft_plan_s *p, const GridArgs &g, int tb) {
  const double cb = (double)sizeof(cx<T>) * p->Nk;
  if (plan_fields(p) == 11) {
    TensorLaunch<T>::xconv_fwd(p, g);
    return;
  }
  const bool wave = p->xw[prec_index<T>()] != 0;
  with_length(p->Nx, [&](auto LX) {
    constexpr int L = decltype(LX)::value;
    if constexpr (wave_fft_length(L)) {
      if (wave) {
        constexpr int SPW = 64 / (L / 8);
        const long long tiles = ((long long)p->Ny * p->Nkz + SPW - 1) / SPW;
        const int blocks = (int)((tiles + kThreads / 64 - 1) / (kThreads / 64));
        ProfScope ps(p, kid<T>(K_XCONV_FWD), 6 * cb);
        hipLaunchKernelGGL((k_xconv_fwd_wave<T, L>), dim3(blocks), dim3(kThreads), 0, p->stream, g,
                           spec_ptr<T>(p, 0), spec_ptr<T>(p, 1), tw_ptr<T>(p, 0));
        return;
      }
    }
    with_tb(tb, [&](auto TBc) {
      constexpr int TB = decltype(TBc)::value;
      ProfScope ps(p, kid<T>(K_XCONV_FWD), 6 * cb);
      hipLaunchKernelGGL((k_xconv_fwd<T, L, TB>), dim3(x_blocks<T>(p, TB)), dim3(kThreads), 0, p->stream, g,
                         spec_ptr<T>(p, 0), spec_ptr<T>(p, 1), tw_ptr<T>(p, 0));
    });
  });
}

template <typename T>
static void launch_xconv_bwd(deft_plan_s *p, const GridArgs &g, int tb) {
  const double cb = (double)sizeof(cx<T>) * p->Nk;
  if (plan_fields(p) == 11) {
    TensorLaunch<T>::xconv_bwd(p, g);
    return;
  }
  with_length(p->Nx, [&](auto LX) {
    constexpr int L = decltype(LX)::value;
    with_tb(tb, [&](auto TBc) {
      constexpr int TB = decltype(TBc)::value;
      ProfScope ps(p, kid<T>(K_XCONV_BWD), 6 * cb);
      hipLaunchKernelGGL((k_xconv_bwd<T, L, TB>), dim3(x_blocks<T>(p, TB)), dim3(kThreads), 0, p->stream, g,
                         spec_ptr<T>(p, 1), spec_ptr<T>(p, 0), tw_ptr<T>(p, 0));
    });
  });
}

template <typename T>
static void launch_zphi(deft_plan_s *p, const GridArgs &g, bool grad, int zb) {
  const double cb = (double)sizeof(cx<T>) * p->Nk;
  if (p->functional != DEFT_WHITE_BEAR) {
    TensorLaunch<T>::zphi(p, g, grad, zb);
    return;
  }
  const bool wave = p->zw[prec_index<T>()] != 0;
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    with_zb(zb, [&](auto ZBc) {
      constexpr int ZB = decltype(ZBc)::value;
      if constexpr (wave_fft_length(L) && ZB == 512) {
        if (wave) {
          const dim3 grid(zblocks_for(p, ZB)), blk(kWaveThreads);
          if (grad) {
            ProfScope ps(p, kid<T>(K_ZPHI_EG), 10 * cb);
            hipLaunchKernelGGL((k_zphi_wave<T, L, true>), grid, blk, 0, p->stream, g, spec_ptr<T>(p, 1),
                               p->partial, p->pstride, tw_ptr<T>(p, 2));
          } else {
            ProfScope ps(p, kid<T>(K_ZPHI_E), 5 * cb);
            hipLaunchKernelGGL((k_zphi_wave<T, L, false>), grid, blk, 0, p->stream, g, spec_ptr<T>(p, 1),
                               p->partial, p->pstride, tw_ptr<T>(p, 2));
          }
          return;
        }
      }
      if (grad) {
        ProfScope ps(p, kid<T>(K_ZPHI_EG), 10 * cb);
        hipLaunchKernelGGL((k_zphi<T, L, true, ZB>), dim3(zblocks_for(p, ZB)), dim3(kThreads), 0, p->stream, g,
                           spec_ptr<T>(p, 1), p->partial, p->pstride, tw_ptr<T>(p, 2));
      } else {
        ProfScope ps(p, kid<T>(K_ZPHI_E), 5 * cb);
        hipLaunchKernelGGL((k_zphi<T, L, false, ZB>), dim3(zblocks_for(p, ZB)), dim3(kThreads), 0, p->stream, g,
                           spec_ptr<T>(p, 1), p->partial, p->pstride, tw_ptr<T>(p, 2));
      }
    });
  });
}

template <typename T>
static void launch_density_to_weighted(deft_plan_s *p, const GridArgs &g) {
  const int ti = prec_index<T>();
  const double cb = (double)sizeof(cx<T>) * p->Nk;  // bytes of one half spectrum
  double m1, m2;
  int nm;
  take_moves(p, m1, m2, nm);
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    with_zb(p->tz[ti], [&](auto ZBc) {
      constexpr int ZB = decltype(ZBc)::value;
      ProfScope ps(p, kid<T>(K_ZFWD_DENSITY), (nm > 0 ? 24.0 : 8.0) * p->N + (double)sizeof(T) * p->N + cb);
      hipLaunchKernelGGL((k_zfwd_density<T, L, ZB>), dim3(zblocks_for(p, ZB)), dim3(kThreads), 0, p->stream, g, p->x,
                         p->d, m1, m2, nm, vext_ptr<T>(p), spec_ptr<T>(p, 0), p->partial, p->pstride,
                         tw_ptr<T>(p, 2));
    });
  });
  launch_ypass<T>(p, g, spec_ptr<T>(p, 0), false, 1, p->ty[ti]);
  launch_xconv_fwd<T>(p, g, p->tx[ti]);
  launch_ypass<T>(p, g, spec_ptr<T>(p, 1), true, plan_fields(p), p->ty[ti]);
}

// --- edges without fused kernels (generic.hpp) --------------------------------

template <typename T>
inline LineSched line_sched(const deft_plan_s *p, int axis) {
  LineSched s = p->sched[prec_index<T>()][axis];
  s.SS = rt_lds_stride<T>(s.L);
  return s;
}

template <typename T, typename Fn>
inline void with_cap(int L, Fn &&fn) {
  if (gen_large(L))
    fn(ic<gen_cap<T>(true)>{});
  else
    fn(ic<gen_cap<T>(false)>{});
}

inline int kblocks_for(const deft_plan_s *p) {
  return (int)std::min<long long>(4096, (p->Nk + kThreads - 1) / kThreads);
}

// line pairs per z block, columns per y/x block
inline int gen_pairs(const LineSched &s, int cap) { return std::max(1, cap / std::max(s.SS, s.L + 2)); }
inline int gen_cols(const LineSched &s, int cap) { return std::max(1, std::min(kThreads, cap / s.SS)); }

template <typename T, bool INV>
static void gen_col(deft_plan_s *p, const GridArgs &g, int axis, cx<T> *data, int nf) {
  const LineSched s = line_sched<T>(p, axis);
  if (s.L == 1) return;
  with_cap<T>(s.L, [&](auto CAPc) {
    constexpr int CAP = decltype(CAPc)::value;
    const int C = gen_cols(s, CAP);
    const long long ncols = (long long)(axis == 1 ? p->Nx : p->Ny) * p->Nkz;
    ProfScope ps(p, kid<T>(K_GEN_LINE), 2.0 * nf * sizeof(cx<T>) * p->Nk);
    hipLaunchKernelGGL((kg_col<T, CAP, INV>), dim3((unsigned)((ncols + C - 1) / C), nf), dim3(kThreads), 0, p->stream,
                       g, s, C, axis, data, tw_ptr<T>(p, axis));
  });
}

// nf real FP64 fields (fs_in apart) -> half spectra out + f*ks, times scl
template <typename T>
static void gen_forward(deft_plan_s *p, const GridArgs &g, const double *in, long long fs_in, int nf, cx<T> *out,
                        double scl) {
  const LineSched s = line_sched<T>(p, 2);
  with_cap<T>(s.L, [&](auto CAPc) {
    constexpr int CAP = decltype(CAPc)::value;
    const int P = gen_pairs(s, CAP);
    const long long lines = (long long)p->Nx * p->Ny;
    ProfScope ps(p, kid<T>(K_GEN_LINE), nf * (8.0 * p->N + sizeof(cx<T>) * p->Nk));
    hipLaunchKernelGGL((kg_zr2c<T, CAP>), dim3((unsigned)((lines + 2 * P - 1) / (2 * P)), nf), dim3(kThreads), 0,
                       p->stream, g, s, P, in, fs_in, out, scl, tw_ptr<T>(p, 2));
  });
  gen_col<T, false>(p, g, 1, out, nf);
  gen_col<T, false>(p, g, 0, out, nf);
}

// half spectra in + f*ks (consumed) -> nf real FP64 fields (fs_out apart), times scl
template <typename T>
static void gen_inverse(deft_plan_s *p, const GridArgs &g, cx<T> *in, int nf, double *out, long long fs_out,
                        double scl) {
  gen_col<T, true>(p, g, 0, in, nf);
  gen_col<T, true>(p, g, 1, in, nf);
  const LineSched s = line_sched<T>(p, 2);
  with_cap<T>(s.L, [&](auto CAPc) {
    constexpr int CAP = decltype(CAPc)::value;
    const int P = gen_pairs(s, CAP);
    const long long lines = (long long)p->Nx * p->Ny;
    ProfScope ps(p, kid<T>(K_GEN_LINE), nf * (8.0 * p->N + sizeof(cx<T>) * p->Nk));
    hipLaunchKernelGGL((kg_zc2r<T, CAP>), dim3((unsigned)((lines + 2 * P - 1) / (2 * P)), nf), dim3(kThreads), 0,
                       p->stream, g, s, P, in, out, fs_out, scl, tw_ptr<T>(p, 2));
  });
}

template <typename Fn>
inline void with_nw(int nw, Fn &&fn) {
  if (nw == 11)
    fn(ic<11>{});
  else
    fn(ic<5>{});
}

// density -> the plan's weighted densities in gwork (nw fields), moves applied
template <typename T>
static void gen_density_to_weighted(deft_plan_s *p, const GridArgs &g, double *w) {
  double m1, m2;
  int nm;
  take_moves(p, m1, m2, nm);
  const int nw = plan_fields(p);
  {
    ProfScope ps(p, kid<T>(K_GEN_POINT), (nm > 0 ? 32.0 : 16.0) * p->N + (double)sizeof(T) * p->N);
    hipLaunchKernelGGL(kg_density<T>, dim3(p->eblocks), dim3(kThreads), 0, p->stream, g, p->x, p->d, m1, m2, nm,
                       vext_ptr<T>(p), p->gwork, p->partial, p->pstride);
  }
  gen_forward<T>(p, g, p->gwork, p->N, 1, spec_ptr<T>(p, 0), 1.0);
  if (p->functional == DEFT_SFMT) {
    ProfScope ps(p, kid<T>(K_GEN_POINT), (1.0 + nw) * sizeof(cx<T>) * p->Nk + 64.0 * p->Nk);
    hipLaunchKernelGGL((kg_wmul_fwd_sfmt<T>), dim3(kblocks_for(p)), dim3(kThreads), 0, p->stream, g,
                       spec_ptr<T>(p, 0), spec_ptr<T>(p, 1));
  } else {
    with_nw(nw, [&](auto NWc) {
      constexpr int NW = decltype(NWc)::value;
      ProfScope ps(p, kid<T>(K_GEN_POINT), (1.0 + NW) * sizeof(cx<T>) * p->Nk);
      hipLaunchKernelGGL((kg_wmul_fwd<T, NW>), dim3(kblocks_for(p)), dim3(kThreads), 0, p->stream, g,
                         spec_ptr<T>(p, 0), spec_ptr<T>(p, 1));
    });
  }
  gen_inverse<T>(p, g, spec_ptr<T>(p, 1), nw, w, p->N, 1.0);
}

template <typename T>
static int sk_apply(deft_plan_s *p, const GridArgs &g, int nb, bool have_old);

// The square-well liquid's attraction (sw_kernels.hpp), with the White-Bear
// weighted densities in w and n~ in spectrum 0 (as gen_density_to_weighted
// leaves them): its energy partial sums (SLOT_SW) and, with grad, its dF/dn in
// the plan's swg field.  Works in spectra 1..9 (spectrum 0 last) and in the
// plan's 26 real SW fields:
//   a/q (8) | num, den | f/kn (5) | g | c | swg | u, t | r (5) | scratch
template <typename T>
static void sw_enqueue(deft_plan_s *p, const GridArgs &g, const double *w, bool grad) {
  const long long N = p->N;
  double *B = p->swbuf;
  double *A8 = B, *ND = B + 8 * N, *F5 = B + 10 * N, *G = B + 15 * N, *C = B + 16 * N, *SWG = B + 17 * N,
         *UT = B + 18 * N, *R5 = B + 20 * N, *TMP = B + 25 * N;
  const SWConst c = make_sw_const(p->R, p->sw_eps, p->sw_lambda);
  const dim3 eb(p->eblocks), tb(kThreads), kb(kblocks_for(p));
  const double cb = (double)sizeof(cx<T>) * p->Nk;
  const int pk = kid<T>(K_GEN_POINT);
  {
    ProfScope ps(p, pk, 104.0 * N);
    hipLaunchKernelGGL(k_sw_afields, eb, tb, 0, p->stream, N, p->R, w, A8);
  }
  gen_forward<T>(p, g, A8, N, 8, spec_ptr<T>(p, 1), 1.0);
  {
    ProfScope ps(p, pk, 11.0 * cb);
    hipLaunchKernelGGL((kg_sw_num<T>), kb, tb, 0, p->stream, g, c, spec_ptr<T>(p, 1), spec_ptr<T>(p, 0));
  }
  gen_inverse<T>(p, g, spec_ptr<T>(p, 1), 2, ND, N, 1.0);
  {
    ProfScope ps(p, pk, 64.0 * N + sizeof(T) * N);
    hipLaunchKernelGGL((k_sw_point_e<T>), eb, tb, 0, p->stream, g, p->x, vext_ptr<T>(p), ND, G, F5);
  }
  gen_forward<T>(p, g, F5, N, 5, spec_ptr<T>(p, 1), 1.0);
  {
    ProfScope ps(p, pk, 6.0 * cb);
    hipLaunchKernelGGL((kg_sw_xi_sum<T>), kb, tb, 0, p->stream, g, c, spec_ptr<T>(p, 1));
  }
  gen_inverse<T>(p, g, spec_ptr<T>(p, 1), 1, C, N, 1.0);
  {
    ProfScope ps(p, pk, 16.0 * N + sizeof(T) * N);
    hipLaunchKernelGGL((k_sw_energy<T>), eb, tb, 0, p->stream, g, c, p->x, vext_ptr<T>(p), C, p->partial, p->pstride);
  }
  if (!grad) return;
  {
    ProfScope ps(p, pk, 6.0 * cb);
    hipLaunchKernelGGL((kg_sw_xi_split<T>), kb, tb, 0, p->stream, g, c, spec_ptr<T>(p, 0), spec_ptr<T>(p, 1));
  }
  gen_inverse<T>(p, g, spec_ptr<T>(p, 1), 5, F5, N, 1.0);
  {
    ProfScope ps(p, pk, 104.0 * N + sizeof(T) * N);
    hipLaunchKernelGGL((k_sw_point_g<T>), eb, tb, 0, p->stream, g, c, p->x, vext_ptr<T>(p), G, ND, C, F5, SWG, UT);
  }
  gen_forward<T>(p, g, UT, N, 2, spec_ptr<T>(p, 1), 1.0);
  {
    ProfScope ps(p, pk, 11.0 * cb);
    hipLaunchKernelGGL((kg_sw_adj<T>), kb, tb, 0, p->stream, g, c, spec_ptr<T>(p, 1));
  }
  gen_inverse<T>(p, g, spec_ptr<T>(p, 1), 8, A8, N, 1.0);
  {
    ProfScope ps(p, pk, 144.0 * N);
    hipLaunchKernelGGL(k_sw_jac, eb, tb, 0, p->stream, N, p->R, w, A8, R5);
  }
  gen_forward<T>(p, g, R5, N, 5, spec_ptr<T>(p, 1), 1.0);
  {
    ProfScope ps(p, pk, 9.0 * cb);
    hipLaunchKernelGGL((kg_wmul_bwd<T, 5>), kb, tb, 0, p->stream, g, spec_ptr<T>(p, 1), spec_ptr<T>(p, 0));
    hipLaunchKernelGGL((kg_spec_add<T>), kb, tb, 0, p->stream, g, spec_ptr<T>(p, 0), spec_ptr<T>(p, 9));
  }
  gen_inverse<T>(p, g, spec_ptr<T>(p, 0), 1, TMP, N, 1.0);
  {
    ProfScope ps(p, pk, 24.0 * N);
    hipLaunchKernelGGL(k_sw_add, eb, tb, 0, p->stream, N, SWG, TMP);
  }
}

template <typename T>
static int gen_enqueue_evaluation(deft_plan_s *p, bool grad, bool have_old, int precond, int *nb) {
  const GridArgs g = grid_args(p);
  const int nw = plan_fields(p);
  double *w = p->gwork;
  gen_density_to_weighted<T>(p, g, w);
  const bool sw = p->functional == DEFT_SW_LIQUID;
  if (sw) sw_enqueue<T>(p, g, w, grad);
  const int fn = (p->functional == DEFT_WHITE_BEAR || sw) ? 0 : (nw == 16 ? 4 : (nw == 11 ? 3 : 1));
  {
    ProfScope ps(p, kid<T>(K_GEN_POINT), (grad ? 16.0 : 8.0) * nw * p->N);
    auto phi = [&](auto FNc) {
      constexpr int FN = decltype(FNc)::value;
      if (grad)
        hipLaunchKernelGGL((kg_phi<T, true, FN>), dim3(p->eblocks), dim3(kThreads), 0, p->stream, g, w, p->partial,
                           p->pstride);
      else
        hipLaunchKernelGGL((kg_phi<T, false, FN>), dim3(p->eblocks), dim3(kThreads), 0, p->stream, g, w, p->partial,
                           p->pstride);
    };
    if (fn == 0)
      phi(ic<0>{});
    else if (fn == 1)
      phi(ic<1>{});
    else if (fn == 3)
      phi(ic<3>{});
    else
      phi(ic<4>{});
  }
  if (grad) {
    gen_forward<T>(p, g, w, p->N, nw, spec_ptr<T>(p, 1), 1.0);
    if (p->functional == DEFT_SFMT) {
      ProfScope ps(p, kid<T>(K_GEN_POINT), (1.0 + nw) * sizeof(cx<T>) * p->Nk + 64.0 * p->Nk);
      hipLaunchKernelGGL((kg_wmul_bwd_sfmt<T>), dim3(kblocks_for(p)), dim3(kThreads), 0, p->stream, g,
                         spec_ptr<T>(p, 1), spec_ptr<T>(p, 0));
    } else {
      with_nw(nw, [&](auto NWc) {
        constexpr int NW = decltype(NWc)::value;
        ProfScope ps(p, kid<T>(K_GEN_POINT), (1.0 + NW) * sizeof(cx<T>) * p->Nk);
        hipLaunchKernelGGL((kg_wmul_bwd<T, NW>), dim3(kblocks_for(p)), dim3(kThreads), 0, p->stream, g,
                           spec_ptr<T>(p, 1), spec_ptr<T>(p, 0));
      });
    }
    gen_inverse<T>(p, g, spec_ptr<T>(p, 0), 1, w, p->N, 1.0);
    ProfScope ps(p, kid<T>(K_GEN_POINT), (have_old ? 48.0 : 40.0) * p->N + (double)sizeof(T) * p->N);
    hipLaunchKernelGGL(kg_grad<T>, dim3(p->eblocks), dim3(kThreads), 0, p->stream, g, w, p->x, vext_ptr<T>(p), p->og,
                       have_old ? 1 : 0, precond, p->g, p->pg, p->partial, p->pstride,
                       sw ? (const double *)(p->swbuf + 17 * p->N) : (const double *)nullptr);
  }
  if (grad && precond == PGRAD_SK) TRY_LAUNCH(sk_apply<T>(p, g, p->eblocks, have_old));
  *nb = p->eblocks;
  DEFT_HIP_CHECK(hipGetLastError());
  return DEFT_OK;
}

// Times the candidate radix schedules of every axis longer than 1 (the
// forward and inverse line passes of that axis over the plan's weighted
// fields) and keeps the fastest.
template <typename T>
static int gen_autotune(deft_plan_s *p);

// The structure-factor preconditioner's transforms after a gradient pass that
// stored grad/sqrt(n) in p->pg (PGRAD_SK): pg <- F^-1 [S(k)/N F pg], then
// k_sk_finish; nb blocks write the pgrad.grad partial sums.
template <typename T>
static int sk_apply(deft_plan_s *p, const GridArgs &g, int nb, bool have_old) {
  if (!p->sktab) return set_error(DEFT_ERR_STATE, "structure-factor table missing");
  TRY_LAUNCH(Launch<T>::fft_forward(p, p->pg, 1.0));
  {
    ProfScope ps(p, kid<T>(K_FFT_SEAM), 2.0 * sizeof(cx<T>) * p->Nk + 8.0 * p->Nk);
    hipLaunchKernelGGL(k_spec_scale<T>, dim3(kblocks_for(p)), dim3(kThreads), 0, p->stream, g, spec_ptr<T>(p, 0),
                       p->sktab);
  }
  TRY_LAUNCH(Launch<T>::fft_inverse(p, p->pg, 1.0));
  ProfScope ps(p, kid<T>(K_SK_FINISH), (have_old ? 48.0 : 40.0) * p->N + (double)sizeof(T) * p->N);
  hipLaunchKernelGGL(k_sk_finish<T>, dim3(nb), dim3(kThreads), 0, p->stream, g, p->x, vext_ptr<T>(p), p->g, p->og,
                     have_old ? 1 : 0, p->pg, p->partial, p->pstride);
  DEFT_HIP_CHECK(hipGetLastError());
  return DEFT_OK;
}

// The structure-factor preconditioner fused into the passes of a fixed-edge
// plan (k_zgrad_sk, the y passes, k_xsk, k_zsk_finish), after the
// back-convolution's inverse y pass; nzb z blocks write the partial sums.
template <typename T>
static int sk_apply_fused(deft_plan_s *p, const GridArgs &g, int zb, int nzb, bool have_old) {
  if (!p->sktab) return set_error(DEFT_ERR_STATE, "structure-factor table missing");
  const int ti = prec_index<T>();
  const double cb = (double)sizeof(cx<T>) * p->Nk;
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    with_zb(zb, [&](auto ZBc) {
      constexpr int ZB = decltype(ZBc)::value;
      ProfScope ps(p, kid<T>(K_ZGRAD), 2.0 * cb + 24.0 * p->N + (double)sizeof(T) * p->N);
      hipLaunchKernelGGL((k_zgrad_sk<T, L, ZB>), dim3(nzb), dim3(kThreads), 0, p->stream, g, spec_ptr<T>(p, 0), p->x,
                         vext_ptr<T>(p), p->g, tw_ptr<T>(p, 2));
    });
  });
  launch_ypass<T>(p, g, spec_ptr<T>(p, 0), false, 1, p->ty[ti]);
  with_length(p->Nx, [&](auto LX) {
    constexpr int L = decltype(LX)::value;
    with_tb(p->tx[ti], [&](auto TBc) {
      constexpr int TB = decltype(TBc)::value;
      ProfScope ps(p, kid<T>(K_FFT_SEAM), 2.0 * cb + 8.0 * p->Nk);
      hipLaunchKernelGGL((k_xsk<T, L, TB>), dim3(x_blocks<T>(p, TB)), dim3(kThreads), 0, p->stream, g,
                         spec_ptr<T>(p, 0), p->sktab, tw_ptr<T>(p, 0));
    });
  });
  launch_ypass<T>(p, g, spec_ptr<T>(p, 0), true, 1, p->ty[ti]);
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    with_zb(zb, [&](auto ZBc) {
      constexpr int ZB = decltype(ZBc)::value;
      ProfScope ps(p, kid<T>(K_SK_FINISH), cb + (have_old ? 40.0 : 32.0) * p->N + (double)sizeof(T) * p->N);
      hipLaunchKernelGGL((k_zsk_finish<T, L, ZB>), dim3(nzb), dim3(kThreads), 0, p->stream, g, spec_ptr<T>(p, 0),
                         p->x, vext_ptr<T>(p), p->g, p->og, have_old ? 1 : 0, p->pg, p->partial, p->pstride,
                         tw_ptr<T>(p, 2));
    });
  });
  DEFT_HIP_CHECK(hipGetLastError());
  return DEFT_OK;
}

template <typename T>
int Launch<T>::evaluate(deft_plan_s *p, bool grad, bool have_old, int precond) {
  int nzb = 0;
  TRY_LAUNCH(enqueue_evaluation(p, grad, have_old, precond, &nzb));
  // every pass wrote its partial sums from nzb blocks
  if (p->functional == DEFT_SW_LIQUID) TRY_LAUNCH(plan_reduce_wait(p, SLOT_SW, 1, nzb));
  return plan_reduce_wait(p, SLOT_IDEAL, grad ? 8 : 6, nzb);
}

template <typename T>
int Launch<T>::enqueue_evaluation(deft_plan_s *p, bool grad, bool have_old, int precond, int *nzb_out) {
  if (p->generic) return gen_enqueue_evaluation<T>(p, grad, have_old, precond, nzb_out);
  const GridArgs g = grid_args(p);
  const int ti = prec_index<T>();
  const int zb = p->tz[ti];
  const int nzb = zblocks_for(p, zb);
  *nzb_out = nzb;
  launch_density_to_weighted<T>(p, g);
  launch_zphi<T>(p, g, grad, zb);
  if (grad) {
    launch_ypass<T>(p, g, spec_ptr<T>(p, 1), false, plan_fields(p), p->ty[ti]);
    launch_xconv_bwd<T>(p, g, p->tx[ti]);
    launch_ypass<T>(p, g, spec_ptr<T>(p, 0), true, 1, p->ty[ti]);
    if (precond == PGRAD_SK) {
      TRY_LAUNCH(sk_apply_fused<T>(p, g, zb, nzb, have_old));
    } else {
      with_length(p->Nz, [&](auto LZ) {
        constexpr int L = decltype(LZ)::value;
        with_zb(zb, [&](auto ZBc) {
          constexpr int ZB = decltype(ZBc)::value;
          ProfScope ps(p, kid<T>(K_ZGRAD),
                       (double)sizeof(cx<T>) * p->Nk + (have_old ? 32.0 : 24.0) * p->N + (double)sizeof(T) * p->N);
          hipLaunchKernelGGL((k_zgrad<T, L, ZB>), dim3(nzb), dim3(kThreads), 0, p->stream, g, spec_ptr<T>(p, 0),
                             p->x, vext_ptr<T>(p), p->og, have_old ? 1 : 0, precond, p->g, p->pg, p->partial,
                             p->pstride, tw_ptr<T>(p, 2));
        });
      });
    }
  }
  DEFT_HIP_CHECK(hipGetLastError());
  return DEFT_OK;
}

// --- batched evaluation ----------------------------------------------------

namespace batch_detail {

struct Group {
  int L = 0, tune = 0, flag = 0;
  std::vector<int> members;  // request indices
  std::vector<int> off;      // block offsets, size members + 1
  size_t off_at = 0, idx_at = 0;
  int total() const { return off.back(); }
};

// groups of the requests accepted by `pred`, keyed by (axis length, tuning, flag)
template <typename KeyFn, typename BlocksFn, typename PredFn>
std::vector<Group> make_groups(int n, KeyFn key, BlocksFn blocks, PredFn pred) {
  std::map<std::tuple<int, int, int>, Group> m;
  for (int i = 0; i < n; ++i) {
    if (!pred(i)) continue;
    const auto k = key(i);
    Group &g = m[k];
    g.L = std::get<0>(k);
    g.tune = std::get<1>(k);
    g.flag = std::get<2>(k);
    if (g.off.empty()) g.off.push_back(0);
    g.members.push_back(i);
    g.off.push_back(g.off.back() + blocks(i));
  }
  std::vector<Group> out;
  for (auto &kv : m) out.push_back(std::move(kv.second));
  return out;
}

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace batch_detail

template <typename T>
int Launch<T>::evaluate_batch(hipStream_t st, BatchStage &S, const BatchReq *reqs, int n, long long *launches) {
  using namespace batch_detail;
  if (n <= 0) return DEFT_OK;
  const int ti = prec_index<T>();
  std::vector<BatchArgs<T>> args(n);
  for (int i = 0; i < n; ++i) {
    deft_plan_s *p = reqs[i].p;
    BatchArgs<T> &a = args[i];
    take_moves(p, a.m1, a.m2, a.nmoves);
    a.g = grid_args(p);
    a.x = p->x;
    a.d = p->d;
    a.have_old = reqs[i].have_old ? 1 : 0;
    a.precond = reqs[i].precond;
    a.nzb = zblocks_for(p, p->tz[ti]);
    a.vext = vext_ptr<T>(p);
    a.spec0 = spec_ptr<T>(p, 0);
    a.spec1 = spec_ptr<T>(p, 1);
    a.partial = p->partial;
    a.pstride = p->pstride;
    a.twx = tw_ptr<T>(p, 0);
    a.twy = tw_ptr<T>(p, 1);
    a.twz = tw_ptr<T>(p, 2);
    a.og = p->og;
    a.grad = p->g;
    a.pgrad = p->pg;
    a.dscal = p->dscal;
    a.sktab = p->sktab;
    a.kind = reqs[i].grad ? 2 : 1;
    if (reqs[i].grad && reqs[i].precond == PGRAD_SK && !(p->sk_ok && p->sktab))
      return set_error(DEFT_ERR_STATE, "structure-factor table missing");
  }
  auto P = [&](int i) { return reqs[i].p; };
  auto all = [&](int) { return true; };
  auto gr = [&](int i) { return reqs[i].grad; };
  auto gr_plain = [&](int i) { return reqs[i].grad && reqs[i].precond != PGRAD_SK; };
  auto gr_sk = [&](int i) { return reqs[i].grad && reqs[i].precond == PGRAD_SK; };
  auto zkey = [&](int i) { return std::make_tuple(P(i)->Nz, P(i)->tz[ti], 0); };
  auto zkey_g = [&](int i) { return std::make_tuple(P(i)->Nz, P(i)->tz[ti], reqs[i].grad ? 1 : 0); };
  auto ykey = [&](int i) { return std::make_tuple(P(i)->Ny, P(i)->ty[ti], 0); };
  auto xkey = [&](int i) { return std::make_tuple(P(i)->Nx, P(i)->tx[ti], 0); };
  auto zbl = [&](int i) { return zblocks_for(P(i), P(i)->tz[ti]); };
  auto ybl = [&](int i) { return y_blocks<T>(P(i), P(i)->ty[ti]); };
  auto xbl = [&](int i) { return x_blocks<T>(P(i), P(i)->tx[ti]); };
  auto rkey = [&](int i) { return std::make_tuple(0, 0, reqs[i].grad ? 8 : 6); };
  std::vector<Group> g_dens = make_groups(n, zkey, zbl, all);
  std::vector<Group> g_y1 = make_groups(n, ykey, ybl, all);
  std::vector<Group> g_xf = make_groups(n, xkey, xbl, all);
  std::vector<Group> g_phi = make_groups(n, zkey_g, zbl, all);
  std::vector<Group> g_yg = make_groups(n, ykey, ybl, gr);
  std::vector<Group> g_xb = make_groups(n, xkey, xbl, gr);
  std::vector<Group> g_zg = make_groups(n, zkey, zbl, gr_plain);
  std::vector<Group> g_zsk = make_groups(n, zkey, zbl, gr_sk);
  std::vector<Group> g_ysk = make_groups(n, ykey, ybl, gr_sk);
  std::vector<Group> g_xsk = make_groups(n, xkey, xbl, gr_sk);
  std::vector<Group> g_red = make_groups(n, rkey, [&](int i) { return reqs[i].grad ? 8 : 6; }, all);
  std::vector<Group> *stages[] = {&g_dens, &g_y1, &g_xf, &g_phi, &g_yg, &g_xb, &g_zg, &g_zsk, &g_ysk, &g_xsk, &g_red};
  // staging layout: args, then offsets and member indices of every group
  size_t at = align_up(S.used, 256);
  const size_t args_at = at;
  at += sizeof(BatchArgs<T>) * n;
  for (auto *gv : stages)
    for (Group &g : *gv) {
      at = align_up(at, 16);
      g.off_at = at;
      at += sizeof(int) * g.off.size();
      g.idx_at = at;
      at += sizeof(int) * g.members.size();
    }
  if (at > S.cap) return set_error(DEFT_ERR_ALLOC, "batch staging buffer too small");
  std::memcpy(S.host + args_at, args.data(), sizeof(BatchArgs<T>) * n);
  for (auto *gv : stages)
    for (Group &g : *gv) {
      std::memcpy(S.host + g.off_at, g.off.data(), sizeof(int) * g.off.size());
      std::memcpy(S.host + g.idx_at, g.members.data(), sizeof(int) * g.members.size());
    }
  DEFT_HIP_CHECK(hipMemcpyAsync(S.dev + S.used, S.host + S.used, at - S.used, hipMemcpyHostToDevice, st));
  S.used = at;
  const BatchArgs<T> *dargs = reinterpret_cast<const BatchArgs<T> *>(S.dev + args_at);
  auto grp = [&](const Group &g) {
    return BatchGroup{reinterpret_cast<const int *>(S.dev + g.off_at), reinterpret_cast<const int *>(S.dev + g.idx_at),
                      (int)g.members.size()};
  };
  long long nl = 0;
  auto zpass = [&](const std::vector<Group> &gs, auto launch) {
    for (const Group &g : gs)
      with_length(g.L, [&](auto LZ) {
        with_zb(g.tune, [&](auto ZBc) {
          launch(LZ, ZBc, g);
          ++nl;
        });
      });
  };
  auto tpass = [&](const std::vector<Group> &gs, auto launch) {
    for (const Group &g : gs)
      with_length(g.L, [&](auto LA) {
        with_tb(g.tune, [&](auto TBc) {
          launch(LA, TBc, g);
          ++nl;
        });
      });
  };
  const dim3 blk(kThreads);
  zpass(g_dens, [&](auto LZ, auto ZBc, const Group &g) {
    constexpr int L = decltype(LZ)::value, ZB = decltype(ZBc)::value;
    hipLaunchKernelGGL((kb_zfwd_density<T, L, ZB>), dim3(g.total()), blk, 0, st, grp(g), dargs, 0);
  });
  tpass(g_y1, [&](auto LY, auto TBc, const Group &g) {
    constexpr int L = decltype(LY)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_ypass<T, L, false, TB>), dim3(g.total(), 1), blk, 0, st, grp(g), dargs, 0, 0);
  });
  tpass(g_xf, [&](auto LX, auto TBc, const Group &g) {
    constexpr int L = decltype(LX)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_xconv_fwd<T, L, TB>), dim3(g.total()), blk, 0, st, grp(g), dargs, 0);
  });
  tpass(g_y1, [&](auto LY, auto TBc, const Group &g) {  // same grouping: inverse y of the 5 fields
    constexpr int L = decltype(LY)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_ypass<T, L, true, TB>), dim3(g.total(), 5), blk, 0, st, grp(g), dargs, 1, 0);
  });
  zpass(g_phi, [&](auto LZ, auto ZBc, const Group &g) {
    constexpr int L = decltype(LZ)::value, ZB = decltype(ZBc)::value;
    if (g.flag)
      hipLaunchKernelGGL((kb_zphi<T, L, true, ZB>), dim3(g.total()), blk, 0, st, grp(g), dargs, 0);
    else
      hipLaunchKernelGGL((kb_zphi<T, L, false, ZB>), dim3(g.total()), blk, 0, st, grp(g), dargs, 0);
  });
  tpass(g_yg, [&](auto LY, auto TBc, const Group &g) {
    constexpr int L = decltype(LY)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_ypass<T, L, false, TB>), dim3(g.total(), 5), blk, 0, st, grp(g), dargs, 1, 0);
  });
  tpass(g_xb, [&](auto LX, auto TBc, const Group &g) {
    constexpr int L = decltype(LX)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_xconv_bwd<T, L, TB>), dim3(g.total()), blk, 0, st, grp(g), dargs, 0);
  });
  tpass(g_yg, [&](auto LY, auto TBc, const Group &g) {
    constexpr int L = decltype(LY)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_ypass<T, L, true, TB>), dim3(g.total(), 1), blk, 0, st, grp(g), dargs, 0, 0);
  });
  zpass(g_zg, [&](auto LZ, auto ZBc, const Group &g) {
    constexpr int L = decltype(LZ)::value, ZB = decltype(ZBc)::value;
    hipLaunchKernelGGL((kb_zgrad<T, L, ZB>), dim3(g.total()), blk, 0, st, grp(g), dargs, 0);
  });
  // structure-factor preconditioner (sk_apply_fused's passes, batched)
  zpass(g_zsk, [&](auto LZ, auto ZBc, const Group &g) {
    constexpr int L = decltype(LZ)::value, ZB = decltype(ZBc)::value;
    hipLaunchKernelGGL((kb_zgrad_sk<T, L, ZB>), dim3(g.total()), blk, 0, st, grp(g), dargs);
  });
  tpass(g_ysk, [&](auto LY, auto TBc, const Group &g) {
    constexpr int L = decltype(LY)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_ypass<T, L, false, TB>), dim3(g.total(), 1), blk, 0, st, grp(g), dargs, 0, 0);
  });
  tpass(g_xsk, [&](auto LX, auto TBc, const Group &g) {
    constexpr int L = decltype(LX)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_xsk<T, L, TB>), dim3(g.total()), blk, 0, st, grp(g), dargs);
  });
  tpass(g_ysk, [&](auto LY, auto TBc, const Group &g) {
    constexpr int L = decltype(LY)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_ypass<T, L, true, TB>), dim3(g.total(), 1), blk, 0, st, grp(g), dargs, 0, 0);
  });
  zpass(g_zsk, [&](auto LZ, auto ZBc, const Group &g) {
    constexpr int L = decltype(LZ)::value, ZB = decltype(ZBc)::value;
    hipLaunchKernelGGL((kb_zsk_finish<T, L, ZB>), dim3(g.total()), blk, 0, st, grp(g), dargs);
  });
  for (const Group &g : g_red) {
    hipLaunchKernelGGL(kb_reduce<T>, dim3(g.total()), blk, 0, st, grp(g), dargs, g.flag, 0);
    ++nl;
  }
  DEFT_HIP_CHECK(hipGetLastError());
  if (launches) *launches += nl;
  return DEFT_OK;
}

template <typename T>
static void fill_batch_args(BatchArgs<T> &a, deft_plan_s *p, int ti) {
  a.g = grid_args(p);
  a.x = p->x;
  a.d = p->d;
  a.nzb = zblocks_for(p, p->tz[ti]);
  a.vext = vext_ptr<T>(p);
  a.spec0 = spec_ptr<T>(p, 0);
  a.spec1 = spec_ptr<T>(p, 1);
  a.partial = p->partial;
  a.pstride = p->pstride;
  a.twx = tw_ptr<T>(p, 0);
  a.twy = tw_ptr<T>(p, 1);
  a.twz = tw_ptr<T>(p, 2);
  a.og = p->og;
  a.grad = p->g;
  a.pgrad = p->pg;
  a.dscal = p->dscal;
}

template <typename T>
int Launch<T>::evaluate_batch_graph(hipStream_t st, BatchStage &S, size_t region, deft_plan_s *const *plans,
                                    int nplans, const BatchReq *reqs, int n, BatchGraph &gr, long long *launches) {
  using namespace batch_detail;
  if (n <= 0) return DEFT_OK;
  const int ti = prec_index<T>();
  std::vector<BatchArgs<T>> args(nplans);
  for (int i = 0; i < nplans; ++i) {
    std::memset(&args[i], 0, sizeof(BatchArgs<T>));
    fill_batch_args<T>(args[i], plans[i], ti);
  }
  for (int r = 0; r < n; ++r) {
    deft_plan_s *p = reqs[r].p;
    BatchArgs<T> &a = args[p->batch_index];
    take_moves(p, a.m1, a.m2, a.nmoves);
    a.have_old = reqs[r].have_old ? 1 : 0;
    a.precond = reqs[r].precond;
    a.kind = reqs[r].grad ? 2 : 1;
    if (reqs[r].grad && reqs[r].precond == PGRAD_SK)
      return set_error(DEFT_ERR_UNSUPPORTED, "graph-captured batch rounds take the new API's grad/n only");
  }
  std::vector<int> sig;
  for (int i = 0; i < nplans; ++i) {
    const deft_plan_s *p = plans[i];
    sig.insert(sig.end(), {p->Nx, p->Ny, p->Nz, p->tz[ti], p->ty[ti], p->tx[ti]});
  }
  const size_t args_bytes = sizeof(BatchArgs<T>) * nplans;
  if (gr.exec && (sig != gr.sig || gr.at != region)) {
    (void)hipGraphExecDestroy(gr.exec);
    (void)hipGraphDestroy(gr.graph);
    gr.exec = nullptr;
    gr.graph = nullptr;
  }
  if (gr.exec) {
    std::memcpy(S.host + region, args.data(), args_bytes);
    DEFT_HIP_CHECK(hipMemcpyAsync(S.dev + region, S.host + region, args_bytes, hipMemcpyHostToDevice, st));
    DEFT_HIP_CHECK(hipGraphLaunch(gr.exec, st));
    if (launches) *launches += gr.kernels;
    return DEFT_OK;
  }
  // first round (or changed shapes): lay out the tables and capture the passes
  auto P = [&](int i) { return plans[i]; };
  auto all = [&](int) { return true; };
  auto zkey = [&](int i) { return std::make_tuple(P(i)->Nz, P(i)->tz[ti], 0); };
  auto ykey = [&](int i) { return std::make_tuple(P(i)->Ny, P(i)->ty[ti], 0); };
  auto xkey = [&](int i) { return std::make_tuple(P(i)->Nx, P(i)->tx[ti], 0); };
  auto zbl = [&](int i) { return zblocks_for(P(i), P(i)->tz[ti]); };
  auto ybl = [&](int i) { return y_blocks<T>(P(i), P(i)->ty[ti]); };
  auto xbl = [&](int i) { return x_blocks<T>(P(i), P(i)->tx[ti]); };
  std::vector<Group> g_z = make_groups(nplans, zkey, zbl, all);
  std::vector<Group> g_y = make_groups(nplans, ykey, ybl, all);
  std::vector<Group> g_x = make_groups(nplans, xkey, xbl, all);
  std::vector<Group> g_red = make_groups(nplans, [](int) { return std::make_tuple(0, 0, 8); },
                                         [](int) { return 8; }, all);
  std::vector<Group> *stages[] = {&g_z, &g_y, &g_x, &g_red};
  size_t at = region + args_bytes;
  for (auto *gv : stages)
    for (Group &g : *gv) {
      at = align_up(at, 16);
      g.off_at = at;
      at += sizeof(int) * g.off.size();
      g.idx_at = at;
      at += sizeof(int) * g.members.size();
    }
  if (at > S.cap) return set_error(DEFT_ERR_ALLOC, "batch staging buffer too small");
  std::memcpy(S.host + region, args.data(), args_bytes);
  for (auto *gv : stages)
    for (Group &g : *gv) {
      std::memcpy(S.host + g.off_at, g.off.data(), sizeof(int) * g.off.size());
      std::memcpy(S.host + g.idx_at, g.members.data(), sizeof(int) * g.members.size());
    }
  DEFT_HIP_CHECK(hipMemcpyAsync(S.dev + region, S.host + region, at - region, hipMemcpyHostToDevice, st));
  const BatchArgs<T> *dargs = reinterpret_cast<const BatchArgs<T> *>(S.dev + region);
  auto grp = [&](const Group &g) {
    return BatchGroup{reinterpret_cast<const int *>(S.dev + g.off_at), reinterpret_cast<const int *>(S.dev + g.idx_at),
                      (int)g.members.size()};
  };
  long long nl = 0;
  auto zpass = [&](auto launch) {
    for (const Group &g : g_z)
      with_length(g.L, [&](auto LZ) {
        with_zb(g.tune, [&](auto ZBc) {
          launch(LZ, ZBc, g);
          ++nl;
        });
      });
  };
  auto tpass = [&](const std::vector<Group> &gs, auto launch) {
    for (const Group &g : gs)
      with_length(g.L, [&](auto LA) {
        with_tb(g.tune, [&](auto TBc) {
          launch(LA, TBc, g);
          ++nl;
        });
      });
  };
  const dim3 blk(kThreads);
  DEFT_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  // sel: 1 every request, 2 gradient requests, 3 energy-only requests
  zpass([&](auto LZ, auto ZBc, const Group &g) {
    constexpr int L = decltype(LZ)::value, ZB = decltype(ZBc)::value;
    hipLaunchKernelGGL((kb_zfwd_density<T, L, ZB>), dim3(g.total()), blk, 0, st, grp(g), dargs, 1);
  });
  tpass(g_y, [&](auto LY, auto TBc, const Group &g) {
    constexpr int L = decltype(LY)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_ypass<T, L, false, TB>), dim3(g.total(), 1), blk, 0, st, grp(g), dargs, 0, 1);
  });
  tpass(g_x, [&](auto LX, auto TBc, const Group &g) {
    constexpr int L = decltype(LX)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_xconv_fwd<T, L, TB>), dim3(g.total()), blk, 0, st, grp(g), dargs, 1);
  });
  tpass(g_y, [&](auto LY, auto TBc, const Group &g) {
    constexpr int L = decltype(LY)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_ypass<T, L, true, TB>), dim3(g.total(), 5), blk, 0, st, grp(g), dargs, 1, 1);
  });
  zpass([&](auto LZ, auto ZBc, const Group &g) {
    constexpr int L = decltype(LZ)::value, ZB = decltype(ZBc)::value;
    hipLaunchKernelGGL((kb_zphi<T, L, false, ZB>), dim3(g.total()), blk, 0, st, grp(g), dargs, 3);
  });
  zpass([&](auto LZ, auto ZBc, const Group &g) {
    constexpr int L = decltype(LZ)::value, ZB = decltype(ZBc)::value;
    hipLaunchKernelGGL((kb_zphi<T, L, true, ZB>), dim3(g.total()), blk, 0, st, grp(g), dargs, 2);
  });
  tpass(g_y, [&](auto LY, auto TBc, const Group &g) {
    constexpr int L = decltype(LY)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_ypass<T, L, false, TB>), dim3(g.total(), 5), blk, 0, st, grp(g), dargs, 1, 2);
  });
  tpass(g_x, [&](auto LX, auto TBc, const Group &g) {
    constexpr int L = decltype(LX)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_xconv_bwd<T, L, TB>), dim3(g.total()), blk, 0, st, grp(g), dargs, 2);
  });
  tpass(g_y, [&](auto LY, auto TBc, const Group &g) {
    constexpr int L = decltype(LY)::value, TB = decltype(TBc)::value;
    hipLaunchKernelGGL((kb_ypass<T, L, true, TB>), dim3(g.total(), 1), blk, 0, st, grp(g), dargs, 0, 2);
  });
  zpass([&](auto LZ, auto ZBc, const Group &g) {
    constexpr int L = decltype(LZ)::value, ZB = decltype(ZBc)::value;
    hipLaunchKernelGGL((kb_zgrad<T, L, ZB>), dim3(g.total()), blk, 0, st, grp(g), dargs, 2);
  });
  for (const Group &g : g_red) {
    hipLaunchKernelGGL(kb_reduce<T>, dim3(g.total()), blk, 0, st, grp(g), dargs, 8, 1);
    ++nl;
  }
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(st, &graph);
  if (ec != hipSuccess) return set_error(DEFT_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(ec));
  hipGraphExec_t exec = nullptr;
  if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return set_error(DEFT_ERR_HIP, "batch graph instantiation failed");
  }
  gr.graph = graph;
  gr.exec = exec;
  gr.at = region;
  gr.args_bytes = args_bytes;
  gr.sig = sig;
  gr.kernels = nl;
  DEFT_HIP_CHECK(hipGraphLaunch(exec, st));
  if (launches) *launches += nl;
  return DEFT_OK;
}

template <typename T>
int Launch<T>::fft_forward(deft_plan_s *p, const double *in_dev, double scl) {
  const GridArgs g = grid_args(p);
  const int ti = prec_index<T>();
  if (p->generic) {
    gen_forward<T>(p, g, in_dev, p->N, 1, spec_ptr<T>(p, 0), scl);
    DEFT_HIP_CHECK(hipGetLastError());
    return DEFT_OK;
  }
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    with_zb(p->tz[ti], [&](auto ZBc) {
      constexpr int ZB = decltype(ZBc)::value;
      ProfScope ps(p, kid<T>(K_FFT_SEAM), 8.0 * p->N + (double)sizeof(cx<T>) * p->Nk);
      hipLaunchKernelGGL((k_zfwd_plain<T, L, ZB>), dim3(zblocks_for(p, ZB)), dim3(kThreads), 0, p->stream, g, in_dev,
                         spec_ptr<T>(p, 0), tw_ptr<T>(p, 2));
    });
  });
  launch_ypass<T>(p, g, spec_ptr<T>(p, 0), false, 1, p->ty[ti]);
  with_length(p->Nx, [&](auto LX) {
    constexpr int L = decltype(LX)::value;
    with_tb(p->tx[ti], [&](auto TBc) {
      constexpr int TB = decltype(TBc)::value;
      ProfScope ps(p, kid<T>(K_FFT_SEAM), 2.0 * sizeof(cx<T>) * p->Nk);
      hipLaunchKernelGGL((k_xpass_plain<T, L, false, TB>), dim3(x_blocks<T>(p, TB)), dim3(kThreads), 0, p->stream, g,
                         spec_ptr<T>(p, 0), scl, tw_ptr<T>(p, 0));
    });
  });
  DEFT_HIP_CHECK(hipGetLastError());
  return DEFT_OK;
}

template <typename T>
int Launch<T>::fft_inverse(deft_plan_s *p, double *out_dev, double scl) {
  const GridArgs g = grid_args(p);
  const int ti = prec_index<T>();
  if (p->generic) {
    gen_inverse<T>(p, g, spec_ptr<T>(p, 0), 1, out_dev, p->N, scl);
    DEFT_HIP_CHECK(hipGetLastError());
    return DEFT_OK;
  }
  with_length(p->Nx, [&](auto LX) {
    constexpr int L = decltype(LX)::value;
    with_tb(p->tx[ti], [&](auto TBc) {
      constexpr int TB = decltype(TBc)::value;
      ProfScope ps(p, kid<T>(K_FFT_SEAM), 2.0 * sizeof(cx<T>) * p->Nk);
      hipLaunchKernelGGL((k_xpass_plain<T, L, true, TB>), dim3(x_blocks<T>(p, TB)), dim3(kThreads), 0, p->stream, g,
                         spec_ptr<T>(p, 0), scl, tw_ptr<T>(p, 0));
    });
  });
  launch_ypass<T>(p, g, spec_ptr<T>(p, 0), true, 1, p->ty[ti]);
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    with_zb(p->tz[ti], [&](auto ZBc) {
      constexpr int ZB = decltype(ZBc)::value;
      ProfScope ps(p, kid<T>(K_FFT_SEAM), 8.0 * p->N + (double)sizeof(cx<T>) * p->Nk);
      hipLaunchKernelGGL((k_zinv_plain<T, L, ZB>), dim3(zblocks_for(p, ZB), 1), dim3(kThreads), 0, p->stream, g,
                         spec_ptr<T>(p, 0), out_dev, p->N, 1.0, tw_ptr<T>(p, 2));
    });
  });
  DEFT_HIP_CHECK(hipGetLastError());
  return DEFT_OK;
}

template <typename T>
int Launch<T>::weighted_densities(deft_plan_s *p, double *out_dev) {
  const GridArgs g = grid_args(p);
  const int ti = prec_index<T>();
  if (p->generic) {
    gen_density_to_weighted<T>(p, g, out_dev);
    DEFT_HIP_CHECK(hipGetLastError());
    return DEFT_OK;
  }
  launch_density_to_weighted<T>(p, g);
  with_length(p->Nz, [&](auto LZ) {
    constexpr int L = decltype(LZ)::value;
    with_zb(p->tz[ti], [&](auto ZBc) {
      constexpr int ZB = decltype(ZBc)::value;
      hipLaunchKernelGGL((k_zinv_plain<T, L, ZB>), dim3(zblocks_for(p, ZB), plan_fields(p)), dim3(kThreads), 0, p->stream, g,
                         spec_ptr<T>(p, 1), out_dev, p->N, 1.0, tw_ptr<T>(p, 2));
    });
  });
  DEFT_HIP_CHECK(hipGetLastError());
  return DEFT_OK;
}

// Mean event time (ms) of `reps` launches issued by fn() on the plan's stream,
// after two warm-up launches.
template <typename Fn>
static float time_launches(deft_plan_s *p, Fn &&fn, int reps = 5) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return 1e30f;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return 1e30f;
  }
  fn();
  fn();
  (void)hipEventRecord(a, p->stream);
  for (int i = 0; i < reps; ++i) fn();
  (void)hipEventRecord(b, p->stream);
  float ms = 1e30f;
  if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) ms = 1e30f;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

// Times every instantiated variant of the z, y and x passes on this grid
// (the heaviest kernel of each pass) and keeps the fastest; spectra and
// partial sums are overwritten, Veff and the CG vectors are not touched.
template <typename T>
int Launch<T>::autotune(deft_plan_s *p) {
  if (p->generic) return gen_autotune<T>(p);
  const GridArgs g = grid_args(p);
  const int ti = prec_index<T>();
  const bool was_prof = p->prof;
  p->prof = false;
  float best = 1e30f;
  p->zw[ti] = p->xw[ti] = 0;
  for (int zb : kZBlocks) {
    const float t = time_launches(p, [&] { launch_zphi<T>(p, g, true, zb); });
    if (t < best) {
      best = t;
      p->tz[ti] = zb;
    }
  }
  best = 1e30f;
  for (int tb : kTileBytes) {
    const float t = time_launches(p, [&] { launch_ypass<T>(p, g, spec_ptr<T>(p, 1), false, 5, tb); });
    if (t < best) {
      best = t;
      p->ty[ti] = tb;
    }
  }
  best = 1e30f;
  for (int tb : kTileBytes) {
    const float t = time_launches(p, [&] {
      launch_xconv_fwd<T>(p, g, tb);
      launch_xconv_bwd<T>(p, g, tb);
    });
    if (t < best) {
      best = t;
      p->tx[ti] = tb;
    }
  }
  // The one-per-wave Phi and forward x passes are chosen by timing whole
  // energy+gradient evaluations: launched back to back on their own they time
  // up to 10% faster than the block passes, inside an evaluation they did not
  // (at 128^3 FP64 the x pass measured 35.5 against 35.0 us).
  const bool zok = p->functional == DEFT_WHITE_BEAR && wave_fft_length(p->Nz);
  const bool xok = plan_fields(p) == 5 && wave_fft_length(p->Nx);
  if (zok || xok) {
    const int tz0 = p->tz[ti];
    int bz = 0, bx = 0;
    best = 1e30f;
    for (int zw = 0; zw <= (zok ? 1 : 0); ++zw) {
      for (int xw = 0; xw <= (xok ? 1 : 0); ++xw) {
        p->zw[ti] = zw;
        p->xw[ti] = xw;
        p->tz[ti] = zw ? 512 : tz0;  // the wave Phi pass exists at 512-value blocks
        const float t = time_evaluation(p);
        if (t < best) {
          best = t;
          bz = zw;
          bx = xw;
        }
      }
    }
    p->zw[ti] = bz;
    p->xw[ti] = bx;
    p->tz[ti] = bz ? 512 : tz0;
  }
  p->prof = was_prof;
  DEFT_HIP_CHECK(hipGetLastError());
  return plan_sync(p);
}

template <typename T>
float Launch<T>::time_evaluation(deft_plan_s *p) {
  const bool was_prof = p->prof;
  p->prof = false;
  int nzb = 0, rc = DEFT_OK;
  const float ms = time_launches(p, [&] {
    if (rc == DEFT_OK) rc = enqueue_evaluation(p, true, false, PGRAD_NEW, &nzb);
  });
  p->prof = was_prof;
  return rc == DEFT_OK ? ms : 1e30f;
}

template <typename T>
static int gen_autotune(deft_plan_s *p) {
  const GridArgs g = grid_args(p);
  const bool was_prof = p->prof;
  p->prof = false;
  const int nw = plan_fields(p);
  for (int axis = 0; axis < 3; ++axis) {
    const int L = axis == 0 ? p->Nx : (axis == 1 ? p->Ny : p->Nz);
    if (L == 1) continue;
    float best = 1e30f;
    LineSched &cur = p->sched[prec_index<T>()][axis];
    LineSched keep = cur;
    for (const std::vector<int> &cand : radix_candidates(L)) {
      cur = make_sched(L, cand);
      const float t = time_launches(p, [&] {
        if (axis == 2) {
          gen_forward<T>(p, g, p->gwork, p->N, nw, spec_ptr<T>(p, 1), 1.0);
          gen_inverse<T>(p, g, spec_ptr<T>(p, 1), nw, p->gwork, p->N, 1.0);
        } else {
          gen_col<T, false>(p, g, axis, spec_ptr<T>(p, 1), nw);
          gen_col<T, true>(p, g, axis, spec_ptr<T>(p, 1), nw);
        }
      });
      if (t < best) {
        best = t;
        keep = cur;
      }
    }
    cur = keep;
  }
  p->prof = was_prof;
  DEFT_HIP_CHECK(hipGetLastError());
  return plan_sync(p);
}

}  // namespace deft
