// Device-resident state of one White-Bear fluid grid (shared by the launch
// units and the C ABI implementation).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/deft_amd.h"
#include "generic.hpp"

namespace deft {

// Kernel families timed by the per-plan profiler (HIP events around each
// launch); each family is counted separately for FP32 and FP64 FFTs and for
// each transform direction, so one entry is one kernel instance for a given
// grid.
enum KernelId : int {
  K_ZFWD_DENSITY = 0,
  K_YPASS1,      // forward y pass of the density spectrum
  K_YPASS5,      // inverse y pass of the five weighted-density spectra
  K_XCONV_FWD,
  K_ZPHI_E,
  K_ZPHI_EG,
  K_XCONV_BWD,
  K_ZGRAD,
  K_CG_UPDATE,
  K_CG_RESET,
  K_AXPY,
  K_REDUCE,
  K_FFT_SEAM,
  K_YPASS5_FWD,  // forward y pass of the five partial-derivative spectra
  K_YPASS1_INV,  // inverse y pass of the back-convolution
  K_GEN_LINE,    // line FFT passes of edges without fused kernels
  K_GEN_POINT,   // pointwise passes of edges without fused kernels
  K_SK_FINISH,   // last pass of the structure-factor preconditioner
  NUM_FAMILIES
};
constexpr int NUM_KERNELS = 2 * NUM_FAMILIES;  // entry 2*family + (FP64 ? 1 : 0)

template <typename T>
constexpr int kid(int family) {
  return 2 * family + (sizeof(T) == sizeof(double) ? 1 : 0);
}

inline const char *kernel_name(int id) {
  static const char *names[NUM_KERNELS] = {
      "zfwd_density_f32", "zfwd_density_f64", "ypass1_fwd_f32", "ypass1_fwd_f64", "ypass5_inv_f32", "ypass5_inv_f64",
      "xconv_fwd_f32",    "xconv_fwd_f64",    "zphi_energy_f32", "zphi_energy_f64", "zphi_grad_f32", "zphi_grad_f64",
      "xconv_bwd_f32",    "xconv_bwd_f64",    "zgrad_f32", "zgrad_f64", "cg_update", "cg_update_",
      "cg_reset", "cg_reset_", "axpy", "axpy_", "reduce", "reduce_", "fft_seam_f32", "fft_seam_f64",
      "ypass5_fwd_f32", "ypass5_fwd_f64", "ypass1_inv_f32", "ypass1_inv_f64", "gen_line_f32", "gen_line_f64",
      "gen_point_f32", "gen_point_f64", "sk_finish_f32", "sk_finish_f64"};
  return (id >= 0 && id < NUM_KERNELS) ? names[id] : "?";
}

struct KernelStat {
  long long launches = 0;
  double ms = 0.0;
  double bytes = 0.0;  // algorithmic bytes moved by all timed launches
};

struct PendingEvent {
  hipEvent_t start, stop;
  int id;
  double bytes;
};

}  // namespace deft

struct deft_batch_s;

struct deft_plan_s {
  int Nx = 0, Ny = 0, Nz = 0, Nkz = 0;
  int ldk = 0;        // padded half-spectrum row stride on the device
  long long N = 0, Nk = 0;
  long long ks = 0;   // stored complex values per half spectrum, Nx*Ny*ldk
  double cell[9] = {0};
  double recip[9] = {0};  // rows b1, b2, b3 with a_i . b_j = 2 pi delta_ij
  double volume = 0, dV = 0;
  double R = 1, kT = 1, mu = 0;
  int functional = DEFT_WHITE_BEAR;  // a deft_functional_kind
  int prec = DEFT_FFT_FP32;
  int device = 0;
  hipStream_t stream = nullptr;

  // FP64 vectors (N each): Veff (the minimization variable), CG direction,
  // gradient, preconditioned gradient, previous gradient
  double *x = nullptr, *d = nullptr, *g = nullptr, *pg = nullptr, *og = nullptr;
  // Vext at both FFT precisions: the FP32 passes read the rounded copy, the
  // FP64 passes the caller's values (+inf marks hard walls, NaN constrained points)
  float *vext = nullptr;
  double *vext64 = nullptr;
  void *spec = nullptr;    // spec_fields half spectra of ks complex values (sized for FP64)
  int spec_fields = 6;     // 6, or 12 once the tensor functional was chosen
  int batch_index = -1;    // position in its batch
  double *partial = nullptr;
  int pstride = 0;         // partial-sum slot stride (>= blocks of any reducing kernel)
  unsigned long long *hflag = nullptr;  // completion flag, host-mapped pinned
  unsigned long long *dflag = nullptr;  // its device address
  unsigned long long flag_seq = 0;
  double *hscal = nullptr;  // NUM_SLOTS reduced sums in host-mapped pinned memory
  double *dscal = nullptr;  // device address of hscal (written by k_reduce)
  unsigned *dcount = nullptr;  // arrival counter of k_reduce_flag, zero between launches
  void *twbuf = nullptr;
  deft::cx<float> *twf[3] = {nullptr, nullptr, nullptr};
  deft::cx<double> *twd[3] = {nullptr, nullptr, nullptr};
  double *wtab = nullptr;  // FP64 weight table, (w3, w2) per element of the padded half spectrum (2 ks doubles)
  float *wtab32 = nullptr;  // its FP32 rounding (2 ks floats)
  double *tmp = nullptr;   // scratch for host-seam calls (grows on demand)
  size_t tmp_bytes = 0;

  // moves x += s*d not yet applied (flushed into the next density pass)
  double pend[8] = {0};
  int npend = 0;

  deft_energy_terms last_terms{};

  // profiler: events recorded around launches while `prof` is set
  bool prof = false;
  std::vector<deft::PendingEvent> pending;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> event_pool;
  size_t pool_used = 0;
  deft::KernelStat kstat[deft::NUM_KERNELS];
  int eblocks = 0;  // blocks of the elementwise FP64 kernels
  deft_batch_s *batch = nullptr;  // evaluations go through this batch's rounds
  hipStream_t own_stream = nullptr;  // the plan's stream while it is in a batch
  // tuning per FFT precision (index 0 FP32, 1 FP64): z-pass block size (complex
  // values), y-pass and x-pass tile bytes; set by deft_plan_autotune
  int tz[2] = {1024, 512};  // (FP64 at 1024 exceeds the Phi pass register budget)
  int ty[2] = {16384, 16384};
  int tx[2] = {8192, 8192};
  // White-Bear Phi pass: 0 the block-wide FFT stages (k_zphi), 1 one field per
  // wave (k_zphi_wave, power-of-two z edges); both give the same values
  int zw[2] = {0, 0};
  // forward x pass (five fields): 0 block-wide FFT stages (k_xconv_fwd), 1 one
  // tile of columns per wave (k_xconv_fwd_wave, power-of-two x edges)
  int xw[2] = {0, 0};
  // edges without fused kernels (generic.hpp): line-pass radix schedules per
  // FFT precision and axis (SS is filled at launch) and the real work fields
  bool generic = false;
  deft::LineSched sched[2][3] = {};
  double *gwork = nullptr;  // gwork_fields real FP64 fields of N values
  int gwork_fields = 0;
  // structure-factor preconditioner: S(k)/N on the padded half spectrum (ks
  // doubles), valid for the current mu, functional and radius while sk_ok
  double *sktab = nullptr;
  bool sk_ok = false;
  double sk_nb = 0.0, sk_floor = 0.0;
  // evaluation times of both engines from the last autotune (-1: not timed)
  double engine_fused_ms = -1.0, engine_generic_ms = -1.0;
  // soft FMT (DEFT_SFMT): sigma, epsilon, Xi and the radial weight table
  // (8 doubles per element of the padded half spectrum), valid while stab_ok
  double sfmt_sigma = 1.0, sfmt_epsilon = 1.0, sfmt_xi = 0.1;
  double *stab = nullptr;
  bool stab_ok = false;
  // square-well liquid (DEFT_SW_LIQUID): well depth, range lambda (in units of
  // sigma = 2R) and its 26 real work fields of N values
  double sw_eps = 1.0, sw_lambda = 1.3;
  double *swbuf = nullptr;
};

namespace deft {

inline double water_nQ(double kT) {
  // nQ = (m kT / 2 pi)^(3/2), m = water mass in atomic units (deft src/IdealGas.cpp:54-59)
  const double mass = 18.01528 * 1822.8885;
  return std::pow(mass * kT / (2 * M_PI), 1.5);
}

inline bool det3(const double *c, double &det) {
  det = c[0] * (c[4] * c[8] - c[5] * c[7]) - c[1] * (c[3] * c[8] - c[5] * c[6]) + c[2] * (c[3] * c[7] - c[4] * c[6]);
  return std::isfinite(det) && det != 0.0;
}

// rows b_j of 2 pi inv(cell)^T, so that a_i . b_j = 2 pi delta_ij
inline void reciprocal(const double *c, double det, double *b) {
  double inv[9];
  inv[0] = (c[4] * c[8] - c[5] * c[7]) / det;
  inv[1] = (c[2] * c[7] - c[1] * c[8]) / det;
  inv[2] = (c[1] * c[5] - c[2] * c[4]) / det;
  inv[3] = (c[5] * c[6] - c[3] * c[8]) / det;
  inv[4] = (c[0] * c[8] - c[2] * c[6]) / det;
  inv[5] = (c[2] * c[3] - c[0] * c[5]) / det;
  inv[6] = (c[3] * c[7] - c[4] * c[6]) / det;
  inv[7] = (c[1] * c[6] - c[0] * c[7]) / det;
  inv[8] = (c[0] * c[4] - c[1] * c[3]) / det;
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) b[j * 3 + k] = 2 * M_PI * inv[k * 3 + j];
}

inline GridArgs grid_args(const deft_plan_s *p) {
  GridArgs a{};
  a.wtab = p->wtab;
  a.wtab32 = p->wtab32;
  a.stab = p->stab;
  a.Nx = p->Nx;
  a.Ny = p->Ny;
  a.Nz = p->Nz;
  a.Nkz = p->Nkz;
  a.ldk = p->ldk;
  a.N = p->N;
  a.ks = p->ks;
  for (int i = 0; i < 9; ++i) a.b[i] = p->recip[i];
  a.R = p->R;
  const double dr = std::cbrt(p->dV);
  a.six_dr2 = 6.0 * dr * dr;
  a.invN = 1.0 / (double)p->N;
  a.kT = p->kT;
  a.mu = p->mu;
  a.logNQ = std::log(water_nQ(p->kT));
  a.dV = p->dV;
  a.sk_floor = p->sk_floor;
  a.wb = make_wb_const(p->R, (p->functional == DEFT_SFMT || p->functional == DEFT_SW_LIQUID) ? DEFT_WHITE_BEAR
                                                                                               : p->functional);
  return a;
}

template <typename T>
inline cx<T> *spec_ptr(deft_plan_s *p, int field) {
  return reinterpret_cast<cx<T> *>(p->spec) + (long long)field * p->ks;
}

template <typename T>
inline const T *vext_ptr(const deft_plan_s *p);
template <>
inline const float *vext_ptr<float>(const deft_plan_s *p) {
  return p->vext;
}
template <>
inline const double *vext_ptr<double>(const deft_plan_s *p) {
  return p->vext64;
}

template <typename T>
inline cx<T> *tw_ptr(deft_plan_s *p, int axis);
template <>
inline cx<float> *tw_ptr<float>(deft_plan_s *p, int axis) {
  return p->twf[axis];
}
template <>
inline cx<double> *tw_ptr<double>(deft_plan_s *p, int axis) {
  return p->twd[axis];
}

int set_error(int code, const std::string &msg);

template <typename T>
constexpr int prec_index() {
  return sizeof(T) == sizeof(double) ? 1 : 0;
}

// blocks of a z pass whose blocks hold about zb complex values
inline int zblocks_for(const deft_plan_s *p, int zb) {
  const long long lines = (long long)p->Nx * p->Ny;
  const int pairs = zpairs(p->Nz, zb);
  return (int)((lines + 2 * pairs - 1) / (2 * pairs));
}

// Records HIP events around one launch when the plan's profiler is on.
struct ProfScope {
  deft_plan_s *p;
  int id;
  double bytes;
  size_t slot = 0;
  bool on;
  ProfScope(deft_plan_s *p_, int id_, double bytes_) : p(p_), id(id_), bytes(bytes_), on(p_->prof) {
    if (!on) return;
    if (p->pool_used == p->event_pool.size()) {
      hipEvent_t a = nullptr, b = nullptr;
      if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        on = false;
        return;
      }
      p->event_pool.push_back({a, b});
    }
    slot = p->pool_used++;
    (void)hipEventRecord(p->event_pool[slot].first, p->stream);
  }
  ~ProfScope() {
    if (!on) return;
    (void)hipEventRecord(p->event_pool[slot].second, p->stream);
    p->pending.push_back({p->event_pool[slot].first, p->event_pool[slot].second, id, bytes});
  }
};

// Synchronises the plan's stream and folds finished profiler events into the stats.
inline int plan_sync(deft_plan_s *p) {
  const hipError_t e = hipStreamSynchronize(p->stream);
  if (e != hipSuccess) return set_error(DEFT_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  for (const PendingEvent &pe : p->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pe.start, pe.stop) == hipSuccess) {
      p->kstat[pe.id].launches += 1;
      p->kstat[pe.id].ms += ms;
      p->kstat[pe.id].bytes += pe.bytes;
    }
  }
  p->pending.clear();
  p->pool_used = 0;
  return DEFT_OK;
}

// One grid's evaluation request in a batched round.
struct BatchReq {
  deft_plan_s *p;
  bool grad, have_old;
  int precond;  // PgradKind
  int cg = 0;  // 0: an evaluation; 1: CG update, 2: CG reset (then grad.direction)
  double beta = 0.0;
  int have_dir = 0;
};

// Pinned host staging and its device copy for one round's argument tables.
struct BatchStage {
  char *host = nullptr;
  char *dev = nullptr;
  size_t cap = 0;
  size_t used = 0;
};

// One precision's graph of a batch round: every pass launched over every grid
// of the batch (blocks of grids without a matching request return at once),
// so one captured graph serves every round.  The staging region at `at` holds
// the grids' BatchArgs (rewritten every round) followed by the group tables
// (written once); `sig` is the per-plan shape and tuning the graph was
// captured for.
struct BatchGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t at = 0, args_bytes = 0;
  std::vector<int> sig;
  long long kernels = 0;
};

// Waits until the work queued so far on the plan's stream has finished: a
// one-thread kernel raises a host-mapped flag and the host spins on it (a
// stream synchronisation leaves the GPU idle for ~30 us per wait).  After
// kSpinLimitUs without the flag the host falls back to a blocking
// synchronisation, which also reports any error; with profiling on, or in a
// batch, the blocking path is used throughout.
constexpr long long kSpinLimitUs = 20000;
// Spins until *hflag >= want: 1 = flag seen, 0 = spin limit reached (the
// caller then synchronises the stream).
inline int flag_spin_wait(const unsigned long long *hflag, unsigned long long want) {
  const auto t0 = std::chrono::steady_clock::now();
  for (long long n = 0;; ++n) {
    if (__atomic_load_n(hflag, __ATOMIC_ACQUIRE) >= want) return 1;
    if ((n & 1023) == 1023 &&
        std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count() >
            kSpinLimitUs)
      return 0;
    __builtin_ia32_pause();
  }
}

// The wait on any stream with its own flag pair: a one-thread kernel raises
// the flag; as flag_spin_wait, or negative on a launch error.
inline int flag_wait(hipStream_t st, unsigned long long *hflag, unsigned long long *dflag,
                     unsigned long long &seq) {
  const unsigned long long want = ++seq;
  hipLaunchKernelGGL(k_signal, dim3(1), dim3(64), 0, st, dflag, want);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DEFT_ERR_HIP, std::string("k_signal: ") + hipGetErrorString(e)), -1;
  return flag_spin_wait(hflag, want);
}

// Reduces slots [slot0, slot0 + nslots) of nb blocks' partial sums into
// p->hscal and waits for them; when the wait spins, the reduction's last block
// raises the flag itself (one launch instead of two).
inline int plan_reduce_wait(deft_plan_s *p, int slot0, int nslots, int nb) {
  const bool spin = !(p->prof || p->batch || !p->hflag || !p->dcount);
  const unsigned long long want = spin ? ++p->flag_seq : 0;
  {
    ProfScope ps(p, 2 * K_REDUCE, 8.0 * nslots * nb);
    if (spin)
      hipLaunchKernelGGL(k_reduce_flag, dim3(nslots), dim3(kThreads), 0, p->stream, p->partial, p->pstride, slot0, nb,
                         p->dscal, p->dcount, p->dflag, want);
    else
      hipLaunchKernelGGL(k_reduce, dim3(nslots), dim3(kThreads), 0, p->stream, p->partial, p->pstride, slot0, nb,
                         p->dscal);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DEFT_ERR_HIP, std::string("k_reduce: ") + hipGetErrorString(e));
  if (!spin) return plan_sync(p);
  return flag_spin_wait(p->hflag, want) == 1 ? DEFT_OK : plan_sync(p);
}

// Kernel launch sequences, instantiated for T = float and T = double in
// separate translation units.
template <typename T>
struct Launch {
  // The same evaluation for several grids at once (all at precision T): the
  // passes are launched once per group of grids with equal axis length and
  // tuning; enqueued on `st` without synchronising.  launches += kernels run.
  static int evaluate_batch(hipStream_t st, BatchStage &stage, const BatchReq *reqs, int n, long long *launches);
  // The same through a graph over all `nplans` plans of the batch (plans[i]
  // has batch_index i), captured on first use into `gr` and replayed after.
  static int evaluate_batch_graph(hipStream_t st, BatchStage &stage, size_t region, deft_plan_s *const *plans,
                                  int nplans, const BatchReq *reqs, int n, BatchGraph &gr, long long *launches);
  // Energy (grad=false) or energy + grad + pgrad (grad=true) of the resident
  // Veff after applying pending moves; reduced sums land in p->hscal.
  static int evaluate(deft_plan_s *p, bool grad, bool have_old, int precond);
  // The same passes enqueued on p->stream without the reduction and the wait;
  // *nzb = blocks whose partial sums the passes wrote (slots SLOT_IDEAL..).
  static int enqueue_evaluation(deft_plan_s *p, bool grad, bool have_old, int precond, int *nzb);
  // times the tunable kernel variants on this grid and keeps the fastest
  static int autotune(deft_plan_s *p);
  // milliseconds of one energy + gradient evaluation (grad/n) with the plan's
  // current engine and tuning; the gradient buffers are overwritten
  static float time_evaluation(deft_plan_s *p);
  // fft seam: real FP64 (device, N) -> spectrum in field 0 (scaled by scl)
  static int fft_forward(deft_plan_s *p, const double *in_dev, double scl);
  // ifft seam: spectrum in field 0 -> real FP64 (device, N) times scl
  static int fft_inverse(deft_plan_s *p, double *out_dev, double scl);
  // five weighted densities of the resident Veff into out_dev (5N doubles)
  static int weighted_densities(deft_plan_s *p, double *out_dev);
};

extern template struct Launch<float>;
extern template struct Launch<double>;

}  // namespace deft
