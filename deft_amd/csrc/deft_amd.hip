// C ABI of the MI355X White-Bear functional: plans, the functional and FFT
// seams, the bulk path, and the device-resident conjugate-gradient minimizer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <ucontext.h>

#include "contact.hpp"
#include "launch.hpp"
#include "minimizer.hpp"

using namespace deft;

namespace deft {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

}  // namespace deft

#define TRY(expr)                    \
  do {                               \
    const int rc_ = (expr);          \
    if (rc_ != DEFT_OK) return rc_;  \
  } while (0)

namespace {

constexpr int kVersion = 2;

// edges with fused per-edge kernels; any other supported edge runs the generic passes
bool fused_ok(int n) {
  return with_length(n, [](auto) {});
}

// room for the generic passes' real work fields (plan_fields(p) of N doubles)
int ensure_gwork(deft_plan_s *p) {
  const int nf = plan_fields(p);
  if (!p->generic || p->gwork_fields >= nf) return DEFT_OK;
  double *w = nullptr;
  if (p->stream && hipStreamSynchronize(p->stream) != hipSuccess)
    return set_error(DEFT_ERR_HIP, "stream synchronisation failed");
  if (hipMalloc(&w, sizeof(double) * (size_t)nf * (size_t)p->N) != hipSuccess)
    return set_error(DEFT_ERR_ALLOC, "hipMalloc of the generic work fields failed");
  (void)hipFree(p->gwork);
  p->gwork = w;
  p->gwork_fields = nf;
  return DEFT_OK;
}

// Soft FMT: the radial weight table of the plan's cell, kT and softness
// (k_sfmt_table), built whenever one of them is set.
int plan_build_sfmt_table(deft_plan_s *p) {
  if (!p->stab) {
    if (p->stream && hipStreamSynchronize(p->stream) != hipSuccess)
      return set_error(DEFT_ERR_HIP, "stream synchronisation failed");
    if (hipMalloc(&p->stab, 8 * sizeof(double) * (size_t)p->ks) != hipSuccess) {
      p->stab = nullptr;
      return set_error(DEFT_ERR_ALLOC, "hipMalloc of the soft-FMT weight table failed");
    }
  }
  const SFMTConst c = make_sfmt_const(p->sfmt_sigma, p->sfmt_epsilon, p->kT, p->sfmt_xi);
  const GridArgs g = grid_args(p);
  const long long nk = (long long)p->Nx * p->Ny * p->Nkz;
  const int blocks = (int)std::min<long long>(4096, (nk + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_sfmt_table, dim3(blocks), dim3(kThreads), 0, p->stream, g, c, p->stab);
  DEFT_HIP_CHECK(hipGetLastError());
  DEFT_HIP_CHECK(hipStreamSynchronize(p->stream));
  p->stab_ok = true;
  return DEFT_OK;
}

// Uniform soft-FMT fluid at density n: kT Phi + kT n (ln(n/nQ) - 1) and its d/dn
// (the k = 0 weights; HomogeneousSFMTFluid of src/haskell/SFMT.hs:39-41 at mu = 0).
void sfmt_bulk(const SFMTConst &c, double n, double kT, double *f, double *d) {
  double v[16], w0[4], p1, p2, p3, dd[16];
  sfmt_uniform_fields(n, c, v, w0);
  sfmt_phi(v, p1, p2, p3);
  sfmt_phi_partials(v, dd);
  const double lg = std::log(n / water_nQ(kT));
  *f = kT * (p1 + p2 + p3) + kT * n * (lg - 1.0);
  *d = kT * (dd[0] * w0[0] + dd[1] * w0[1] + dd[2] * w0[2] + dd[3] * w0[3]) + kT * lg;
}

int batch_evaluate(deft_plan_s *p, int prec, bool grad, bool have_old, int precond);
int batch_cg(deft_plan_s *p, int reset, double beta, int have_dir);
int plan_ensure_sk(deft_plan_s *p);

int evaluate_at(deft_plan_s *p, int prec, bool grad, bool have_old, int precond) {
  if (grad && precond == PGRAD_SK) {
    // batched plans build their table when the batch is created or the
    // preconditioner is chosen (deft_minimize_set_preconditioner), never
    // inside the batch's rounds
    if (p->batch) {
      if (!p->sk_ok)
        return set_error(DEFT_ERR_STATE, "structure-factor table missing (choose the preconditioner after the last mu change)");
    } else {
      TRY(plan_ensure_sk(p));
    }
  }
  if (p->batch) return batch_evaluate(p, prec, grad, have_old, precond);
  if (prec == DEFT_FFT_FP64) return Launch<double>::evaluate(p, grad, have_old, precond);
  return Launch<float>::evaluate(p, grad, have_old, precond);
}

int evaluate(deft_plan_s *p, bool grad, bool have_old, int precond) {
  return evaluate_at(p, p->prec, grad, have_old, precond);
}

// Per-term energies from the reduced sums, summed in a fixed order.
deft_energy_terms terms_from_sums(const deft_plan_s *p) {
  const double *s = p->hscal;
  deft_energy_terms t;
  t.kTphi1 = p->kT * s[SLOT_PHI1] * p->dV;
  t.kTphi2 = p->kT * s[SLOT_PHI2] * p->dV;
  t.kTphi3 = p->kT * s[SLOT_PHI3] * p->dV;
  t.ideal = s[SLOT_IDEAL] * p->dV;
  t.int_V_n = s[SLOT_EXT] * p->dV;
  t.minus_mu_N = s[SLOT_MUN] * p->dV;
  t.sw = p->functional == DEFT_SW_LIQUID ? s[SLOT_SW] * p->dV : 0.0;
  t.total = t.kTphi1 + t.kTphi2 + t.kTphi3 + t.ideal + t.int_V_n + t.minus_mu_N + t.sw;
  return t;
}

int ensure_tmp(deft_plan_s *p, size_t bytes) {
  if (p->tmp_bytes >= bytes) return DEFT_OK;
  if (p->tmp) (void)hipFree(p->tmp);
  p->tmp = nullptr;
  p->tmp_bytes = 0;
  if (hipMalloc(&p->tmp, bytes) != hipSuccess) return set_error(DEFT_ERR_ALLOC, "hipMalloc of scratch failed");
  p->tmp_bytes = bytes;
  return DEFT_OK;
}

int flush_moves(deft_plan_s *p) {
  for (int i = 0; i < p->npend; ++i) {
    ProfScope ps(p, 2 * K_AXPY, 24.0 * p->N);
    hipLaunchKernelGGL(k_axpy, dim3(p->eblocks), dim3(kThreads), 0, p->stream, p->N, p->x, p->d, p->pend[i]);
  }
  p->npend = 0;
  DEFT_HIP_CHECK(hipGetLastError());
  return DEFT_OK;
}

void queue_move(deft_plan_s *p, double s) {
  if (p->npend == 8) {
    ProfScope ps(p, 2 * K_AXPY, 24.0 * p->N);
    hipLaunchKernelGGL(k_axpy, dim3(p->eblocks), dim3(kThreads), 0, p->stream, p->N, p->x, p->d, p->pend[0]);
    for (int i = 1; i < 8; ++i) p->pend[i - 1] = p->pend[i];
    p->npend = 7;
  }
  p->pend[p->npend++] = s;
}

int reduce_slot(deft_plan_s *p, int slot, int nb, double *out) {
  TRY(plan_reduce_wait(p, slot, 1, nb));
  *out = p->hscal[slot];
  return DEFT_OK;
}

void free_plan(deft_plan_s *p) {
  if (!p) return;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (auto &ev : p->event_pool) {
    (void)hipEventDestroy(ev.first);
    (void)hipEventDestroy(ev.second);
  }
  (void)hipFree(p->x);
  (void)hipFree(p->d);
  (void)hipFree(p->g);
  (void)hipFree(p->pg);
  (void)hipFree(p->og);
  (void)hipFree(p->vext);
  (void)hipFree(p->vext64);
  (void)hipFree(p->spec);
  (void)hipFree(p->partial);
  (void)hipFree(p->dcount);
  (void)hipFree(p->twbuf);
  (void)hipFree(p->tmp);
  (void)hipFree(p->wtab);
  (void)hipFree(p->wtab32);
  (void)hipFree(p->gwork);
  (void)hipFree(p->sktab);
  (void)hipFree(p->stab);
  (void)hipFree(p->swbuf);
  if (p->hscal) (void)hipHostFree(p->hscal);
  if (p->hflag) (void)hipHostFree(p->hflag);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

}  // namespace

// ===========================================================================
// minimizers: the control logic is minimizer.hpp's, on the plan's device
// vectors through DeviceOps
// ===========================================================================
// new API Minimize (src/new/Minimize.h:13-199) with the precision ladder
struct deft_minimize_s : deft::MinState {
  deft_plan_s *p = nullptr;
};

// old API ConjugateGradient / PreconditionedConjugateGradient with
// QuadraticLineMinimizer, Precision and MaxIter (src/ConjugateGradient.cpp:59-171)
struct deft_cg_s : deft::OldCgState {
  deft_plan_s *p = nullptr;
};

// ===========================================================================
// batches: several plans whose evaluations run in lockstep rounds; every
// minimizer of a batch runs deft's algorithm unchanged in a thread of its own,
// each evaluation it asks for joins the current round, and the round is
// launched (Launch<T>::evaluate_batch, one launch per pass and group of equal
// axis lengths) once every running minimizer has posted a request
// ===========================================================================
struct deft_batch_s {
  std::vector<deft_plan_s *> plans;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::condition_variable cv;
  int active = 0;                  // minimizer threads still running
  std::vector<BatchReq> posted[2]; // this round's requests: FP32, FP64
  std::vector<BatchReq> posted_cg; // this round's CG updates
  int nposted = 0;
  long long round = 0;
  int round_rc = DEFT_OK;
  std::string round_err;
  BatchStage stage;
  long long rounds_run = 0, launches = 0, requests = 0;
  bool graphs = false;  // rounds replay one captured graph per precision
  BatchGraph graph[2];  // FP32, FP64
  // The minimizers run as fibers (user-level contexts) of the thread that
  // called deft_batch_minimize: a fiber posts its request and switches back to
  // the scheduler, which runs the round once every live fiber has posted.
  ucontext_t sched;
  struct Fiber *current = nullptr;
  // host-mapped completion flag of the batch stream (see flag_wait)
  unsigned long long *hflag = nullptr, *dflag = nullptr, flag_seq = 0;
};

struct Fiber {
  ucontext_t ctx;
  std::vector<char> stack;
  deft_minimize_s *m = nullptr;
  deft_batch_s *b = nullptr;
  int rc = DEFT_OK;
  std::string err;
  bool done = false;
};

namespace {

// The plan's device vectors as the minimizers' backend (minimizer.hpp).
struct DeviceOps {
  deft_plan_s *p;
  int plan_prec() const { return p->prec; }
  int energy(int prec, double *E) {
    TRY(evaluate_at(p, prec, false, false, PGRAD_NONE));
    p->last_terms = terms_from_sums(p);
    *E = p->last_terms.total;
    return DEFT_OK;
  }
  int gradient(int prec, bool have_old, int kind, double *E, double *pgg, double *pgo) {
    TRY(evaluate_at(p, prec, true, have_old, kind));
    p->last_terms = terms_from_sums(p);
    *E = p->last_terms.total;
    *pgg = p->hscal[SLOT_PGG];
    *pgo = p->hscal[SLOT_PGO];
    return DEFT_OK;
  }
  // direction = -pg + beta direction (or -pg), oldgrad = grad; grad.direction
  int cg_update(double beta, bool have_dir, double *gd) {
    if (p->batch) {
      // batched plans: the update joins a round with the other grids' updates
      TRY(batch_cg(p, 0, beta, have_dir ? 1 : 0));
      *gd = p->hscal[SLOT_GD];
      return DEFT_OK;
    }
    {
      ProfScope ps(p, 2 * K_CG_UPDATE, 40.0 * p->N);
      hipLaunchKernelGGL(k_cg_update, dim3(p->eblocks), dim3(kThreads), 0, p->stream, p->N, p->pg, p->g, p->d, p->og,
                         beta, have_dir ? 1 : 0, p->partial, p->pstride);
    }
    return reduce_slot(p, SLOT_GD, p->eblocks, gd);
  }
  // direction = -grad; grad.direction
  int cg_reset(double *gd) {
    if (p->batch) {
      TRY(batch_cg(p, 1, 0.0, 0));
      *gd = p->hscal[SLOT_GD];
      return DEFT_OK;
    }
    {
      ProfScope ps(p, 2 * K_CG_RESET, 16.0 * p->N);
      hipLaunchKernelGGL(k_cg_reset, dim3(p->eblocks), dim3(kThreads), 0, p->stream, p->N, p->g, p->d, p->partial,
                         p->pstride);
    }
    return reduce_slot(p, SLOT_GD, p->eblocks, gd);
  }
  void move(double s) { queue_move(p, s); }
  double newton_scale() const { return p->kT / p->dV; }
  // the structure-factor pgrad: plans whose mu has a bulk density
  bool sk_available() {
    // batched plans: built when the batch was created; graph-captured rounds
    // run grad/n only
    if (p->batch) return p->sk_ok && !p->batch->graphs;
    if (plan_ensure_sk(p) == DEFT_OK) return true;
    g_last_error.clear();
    return false;
  }
};

int improve_energy_mixed(deft_minimize_s *m, int verbose, int *more) {
  DeviceOps ops{m->p};
  return deft::improve_energy_mixed(*m, ops, verbose, more);
}

int m_energy(deft_minimize_s *m, double *E) {
  DeviceOps ops{m->p};
  return min_energy(*m, ops, E);
}

// Launches the posted requests as one round and wakes the waiting minimizers;
// the caller holds b->mu.
// The round's CG updates: one launch over all of them and one reduction
// launch (tables at `at` in the staging buffer).
int launch_cg_batch(deft_batch_s *b, size_t at) {
  const std::vector<BatchReq> &q = b->posted_cg;
  const int n = (int)q.size();
  std::vector<CgArgs> args(n);
  std::vector<int> off(n + 1, 0), idx(n);
  for (int r = 0; r < n; ++r) {
    deft_plan_s *p = q[r].p;
    args[r] = CgArgs{p->N, p->pg, p->g, p->d, p->og, q[r].beta, q[r].have_dir, q[r].cg == 2 ? 1 : 0,
                     p->partial, p->pstride, p->eblocks, p->dscal};
    off[r + 1] = off[r] + p->eblocks;
    idx[r] = r;
  }
  at = (at + 255) / 256 * 256;
  const size_t a_at = at, o_at = a_at + sizeof(CgArgs) * n, i_at = o_at + sizeof(int) * (n + 1);
  const size_t end = i_at + sizeof(int) * n;
  if (end > b->stage.cap) return set_error(DEFT_ERR_ALLOC, "batch staging buffer too small");
  std::memcpy(b->stage.host + a_at, args.data(), sizeof(CgArgs) * n);
  std::memcpy(b->stage.host + o_at, off.data(), sizeof(int) * (n + 1));
  std::memcpy(b->stage.host + i_at, idx.data(), sizeof(int) * n);
  DEFT_HIP_CHECK(hipMemcpyAsync(b->stage.dev + a_at, b->stage.host + a_at, end - a_at, hipMemcpyHostToDevice,
                                b->stream));
  const CgArgs *dargs = reinterpret_cast<const CgArgs *>(b->stage.dev + a_at);
  const BatchGroup grp{reinterpret_cast<const int *>(b->stage.dev + o_at),
                       reinterpret_cast<const int *>(b->stage.dev + i_at), n};
  hipLaunchKernelGGL(kb_cg, dim3(off[n]), dim3(kThreads), 0, b->stream, grp, dargs);
  hipLaunchKernelGGL(kb_cg_reduce, dim3(n), dim3(kThreads), 0, b->stream, dargs);
  DEFT_HIP_CHECK(hipGetLastError());
  b->launches += 2;
  if (!b->graphs) b->stage.used = end;
  return DEFT_OK;
}

int run_round(deft_batch_s *b) {
  b->stage.used = 0;
  int rc = DEFT_OK;
  const int np = (int)b->plans.size();
  // staging: graph tables for FP32 and FP64 in the first two thirds, the CG
  // tables in the last (grouped rounds append everything from the start)
  if (!b->posted_cg.empty()) rc = launch_cg_batch(b, b->graphs ? 2 * (b->stage.cap / 3) : 0);
  if (rc != DEFT_OK) {
  } else if (b->graphs) {
    if (!b->posted[0].empty())
      rc = Launch<float>::evaluate_batch_graph(b->stream, b->stage, 0, b->plans.data(), np, b->posted[0].data(),
                                               (int)b->posted[0].size(), b->graph[0], &b->launches);
    if (rc == DEFT_OK && !b->posted[1].empty())
      rc = Launch<double>::evaluate_batch_graph(b->stream, b->stage, (b->stage.cap / 3 + 255) / 256 * 256,
                                                b->plans.data(), np, b->posted[1].data(), (int)b->posted[1].size(),
                                                b->graph[1], &b->launches);
  } else {
    if (!b->posted[0].empty())
      rc = Launch<float>::evaluate_batch(b->stream, b->stage, b->posted[0].data(), (int)b->posted[0].size(),
                                         &b->launches);
    if (rc == DEFT_OK && !b->posted[1].empty())
      rc = Launch<double>::evaluate_batch(b->stream, b->stage, b->posted[1].data(), (int)b->posted[1].size(),
                                          &b->launches);
  }
  if (rc == DEFT_OK) {
    const int w = b->hflag ? flag_wait(b->stream, b->hflag, b->dflag, b->flag_seq) : 0;
    if (w < 0) {
      rc = DEFT_ERR_HIP;
    } else if (w == 0) {
      const hipError_t e = hipStreamSynchronize(b->stream);
      if (e != hipSuccess) rc = set_error(DEFT_ERR_HIP, std::string("batch round: ") + hipGetErrorString(e));
    }
  }
  b->round_rc = rc;
  b->round_err = rc == DEFT_OK ? std::string() : g_last_error;
  b->requests += b->nposted;
  b->posted[0].clear();
  b->posted[1].clear();
  b->posted_cg.clear();
  b->nposted = 0;
  b->round++;
  b->rounds_run++;
  return rc;
}

// Posts a request of a minimizer thread and returns once its round has run.
int batch_post(deft_plan_s *p, const BatchReq &req, int prec) {
  deft_batch_s *b = p->batch;
  if (!b->current) return set_error(DEFT_ERR_STATE, "batched plans evaluate inside deft_batch_minimize only");
  if (req.cg)
    b->posted_cg.push_back(req);
  else
    b->posted[prec == DEFT_FFT_FP64 ? 1 : 0].push_back(req);
  b->nposted++;
  Fiber *f = b->current;
  if (swapcontext(&f->ctx, &b->sched) != 0) return set_error(DEFT_ERR_STATE, "fiber switch failed");
  // resumed by the scheduler after the round ran
  if (b->round_rc != DEFT_OK) return set_error(b->round_rc, b->round_err);
  return DEFT_OK;
}

int batch_evaluate(deft_plan_s *p, int prec, bool grad, bool have_old, int precond) {
  return batch_post(p, BatchReq{p, grad, have_old, precond}, prec);
}

// CG update (or reset) of a batched plan in the next round; grad.direction lands in hscal
int batch_cg(deft_plan_s *p, int reset, double beta, int have_dir) {
  BatchReq r{p, false, false, PGRAD_NONE};
  r.cg = reset ? 2 : 1;
  r.beta = beta;
  r.have_dir = have_dir;
  return batch_post(p, r, DEFT_FFT_FP64);
}

thread_local Fiber *g_fiber_start = nullptr;

// Body of a minimizer fiber; returning resumes the scheduler (uc_link).
void fiber_main() {
  Fiber *f = g_fiber_start;
  int more = 1, rc = DEFT_OK;
  while (more && rc == DEFT_OK) rc = improve_energy_mixed(f->m, 0, &more);
  f->rc = rc;
  if (rc != DEFT_OK) f->err = g_last_error;
  f->done = true;
  f->b->active--;
}

// Copies between the compact half-spectrum layout of the seam and the padded
// rows kept on the device.
template <typename T>
void compact_from_padded(const deft_plan_s *p, const T *pad, double *out) {
  const long long rows = (long long)p->Nx * p->Ny;
  for (long long r = 0; r < rows; ++r)
    for (int k = 0; k < 2 * p->Nkz; ++k) out[r * 2 * p->Nkz + k] = (double)pad[r * 2 * p->ldk + k];
}

template <typename T>
void padded_from_compact(const deft_plan_s *p, const double *in, T *pad) {
  const long long rows = (long long)p->Nx * p->Ny;
  for (long long r = 0; r < rows; ++r) {
    for (int k = 0; k < 2 * p->Nkz; ++k) pad[r * 2 * p->ldk + k] = (T)in[r * 2 * p->Nkz + k];
    for (int k = 2 * p->Nkz; k < 2 * p->ldk; ++k) pad[r * 2 * p->ldk + k] = (T)0;
  }
}

int check_plan(deft_plan p) {
  if (!p) return set_error(DEFT_ERR_ARG, "null plan");
  return DEFT_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *deft_last_error(void) { return g_last_error.c_str(); }

int deft_version(void) { return kVersion; }

int deft_supported_length(int n) { return (n > 0 && (fused_ok(n) || generic_length_ok(n))) ? 1 : 0; }

int deft_fused_length(int n) { return (n > 0 && fused_ok(n)) ? 1 : 0; }

int deft_device_count(int *count) {
  if (!count) return set_error(DEFT_ERR_ARG, "null count");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c;
  return DEFT_OK;
}

int deft_plan_create(int nx, int ny, int nz, const double cell[9], double R, double kT, double mu,
                     int fft_precision, deft_plan *out) {
  if (!out || !cell) return set_error(DEFT_ERR_ARG, "null argument");
  *out = nullptr;
  if (!deft_supported_length(nx) || !deft_supported_length(ny) || !deft_supported_length(nz))
    return set_error(DEFT_ERR_UNSUPPORTED, "grid edges must be 1..4096 with no prime factor above 512");
  double det;
  if (!det3(cell, det)) return set_error(DEFT_ERR_ARG, "cell matrix is singular or not finite");
  if (!(R > 0) || !(kT > 0) || !std::isfinite(R) || !std::isfinite(kT) || !std::isfinite(mu))
    return set_error(DEFT_ERR_ARG, "R and kT must be positive and finite, mu finite");
  if (fft_precision != DEFT_FFT_FP32 && fft_precision != DEFT_FFT_FP64)
    return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(DEFT_ERR_HIP, "no HIP device");

  deft_plan_s *p = new deft_plan_s();
  p->Nx = nx;
  p->Ny = ny;
  p->Nz = nz;
  p->Nkz = nz / 2 + 1;
  p->N = (long long)nx * ny * nz;
  p->Nk = (long long)nx * ny * p->Nkz;
  p->ldk = padded_row(p->Nkz);
  p->ks = (long long)nx * ny * p->ldk;
  std::memcpy(p->cell, cell, sizeof(p->cell));
  reciprocal(cell, det, p->recip);
  p->volume = std::fabs(det);
  p->dV = p->volume / (double)p->N;
  p->R = R;
  p->kT = kT;
  p->mu = mu;
  p->prec = fft_precision;
  p->generic = !(fused_ok(nx) && fused_ok(ny) && fused_ok(nz));
  {
    const int lens[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) p->sched[0][a] = p->sched[1][a] = make_sched(lens[a], radix_candidates(lens[a])[0]);
  }
  p->eblocks = (int)std::min<long long>(1024, (p->N + kThreads - 1) / kThreads);
  p->pstride = std::max(std::max(zblocks_for(p, kZBlocks[0]), zblocks_for(p, kZBlocks[1])), p->eblocks);

  auto fail = [&](int code, const char *what) {
    free_plan(p);
    return set_error(code, what);
  };
  if (hipGetDevice(&p->device) != hipSuccess) return fail(DEFT_ERR_HIP, "hipGetDevice failed");
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(DEFT_ERR_HIP, "hipStreamCreate failed");
  const size_t nb = sizeof(double) * (size_t)p->N;
  if (hipMalloc(&p->x, nb) != hipSuccess || hipMalloc(&p->d, nb) != hipSuccess || hipMalloc(&p->g, nb) != hipSuccess ||
      hipMalloc(&p->pg, nb) != hipSuccess || hipMalloc(&p->og, nb) != hipSuccess ||
      hipMalloc(&p->vext, sizeof(float) * (size_t)p->N) != hipSuccess ||
      hipMalloc(&p->vext64, sizeof(double) * (size_t)p->N) != hipSuccess ||
      hipMalloc(&p->spec, 6 * sizeof(cx<double>) * (size_t)p->ks) != hipSuccess ||
      hipMalloc(&p->partial, sizeof(double) * NUM_SLOTS * (size_t)p->pstride) != hipSuccess ||
      hipMalloc(&p->wtab, 2 * sizeof(double) * (size_t)p->ks) != hipSuccess ||
      hipMalloc(&p->wtab32, 2 * sizeof(float) * (size_t)p->ks) != hipSuccess ||
      hipMalloc(&p->dcount, sizeof(unsigned)) != hipSuccess)
    return fail(DEFT_ERR_ALLOC, "device allocation failed");
  if (ensure_gwork(p) != DEFT_OK) return fail(DEFT_ERR_ALLOC, "device allocation failed");
  if (hipMemsetAsync(p->dcount, 0, sizeof(unsigned), p->stream) != hipSuccess)
    return fail(DEFT_ERR_HIP, "hipMemsetAsync failed");
  if (hipHostMalloc((void **)&p->hflag, sizeof(unsigned long long), hipHostMallocMapped | hipHostMallocCoherent) !=
          hipSuccess ||
      hipHostGetDevicePointer((void **)&p->dflag, p->hflag, 0) != hipSuccess)
    return fail(DEFT_ERR_ALLOC, "pinned allocation failed");
  *p->hflag = 0;
  // reduced sums land directly in host-mapped pinned memory: no copy per evaluation
  if (hipHostMalloc(&p->hscal, sizeof(double) * NUM_SLOTS, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
    return fail(DEFT_ERR_ALLOC, "pinned allocation failed");
  if (hipHostGetDevicePointer((void **)&p->dscal, p->hscal, 0) != hipSuccess)
    return fail(DEFT_ERR_HIP, "hipHostGetDevicePointer failed");
  // twiddle tables exp(-2 pi i m / L), FP64 on the host
  const int lens[3] = {nx, ny, nz};
  size_t ntw = 0;
  for (int a = 0; a < 3; ++a) ntw += lens[a];
  if (hipMalloc(&p->twbuf, ntw * (sizeof(cx<float>) + sizeof(cx<double>))) != hipSuccess)
    return fail(DEFT_ERR_ALLOC, "twiddle allocation failed");
  std::vector<cx<float>> twf(ntw);
  std::vector<cx<double>> twd(ntw);
  size_t off = 0;
  for (int a = 0; a < 3; ++a) {
    const int L = lens[a];
    for (int m = 0; m < L; ++m) {
      const double ang = 2 * M_PI * (double)m / (double)L;
      twd[off + m] = {std::cos(ang), -std::sin(ang)};
      twf[off + m] = {(float)std::cos(ang), (float)-std::sin(ang)};
    }
    off += L;
  }
  cx<double> *dbase = reinterpret_cast<cx<double> *>(p->twbuf);
  cx<float> *fbase = reinterpret_cast<cx<float> *>(dbase + ntw);
  off = 0;
  for (int a = 0; a < 3; ++a) {
    p->twd[a] = dbase + off;
    p->twf[a] = fbase + off;
    off += lens[a];
  }
  if (hipMemcpy(dbase, twd.data(), ntw * sizeof(cx<double>), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(fbase, twf.data(), ntw * sizeof(cx<float>), hipMemcpyHostToDevice) != hipSuccess)
    return fail(DEFT_ERR_HIP, "twiddle upload failed");
  if (hipMemset(p->x, 0, nb) != hipSuccess || hipMemset(p->d, 0, nb) != hipSuccess ||
      hipMemset(p->og, 0, nb) != hipSuccess || hipMemset(p->g, 0, nb) != hipSuccess ||
      hipMemset(p->pg, 0, nb) != hipSuccess || hipMemset(p->vext, 0, sizeof(float) * (size_t)p->N) != hipSuccess ||
      hipMemset(p->vext64, 0, nb) != hipSuccess)
    return fail(DEFT_ERR_HIP, "initialisation failed");
  if (hipMemset(p->spec, 0, 6 * sizeof(cx<double>) * (size_t)p->ks) != hipSuccess)
    return fail(DEFT_ERR_HIP, "initialisation failed");
  hipLaunchKernelGGL(k_weight_table, dim3(p->eblocks), dim3(kThreads), 0, p->stream, grid_args(p), p->wtab,
                     p->wtab32);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(p->stream) != hipSuccess)
    return fail(DEFT_ERR_HIP, "weight table failed");
  *out = p;
  return DEFT_OK;
}

int deft_plan_destroy(deft_plan p) {
  if (p && p->batch) return set_error(DEFT_ERR_STATE, "the plan is in a batch: destroy the batch first");
  free_plan(p);
  return DEFT_OK;
}

int deft_batch_create(deft_plan *plans, int n, deft_batch *out) {
  if (!out || !plans || n < 1) return set_error(DEFT_ERR_ARG, "need plans and an output");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    if (!plans[i]) return set_error(DEFT_ERR_ARG, "null plan");
    if (plans[i]->batch) return set_error(DEFT_ERR_STATE, "a plan is already in a batch");
    if (plans[i]->functional != DEFT_WHITE_BEAR)
      return set_error(DEFT_ERR_UNSUPPORTED, "batches take White-Bear plans only");
    if (plans[i]->generic)
      return set_error(DEFT_ERR_UNSUPPORTED, "batches take plans whose edges have fused kernels (deft_fused_length)");
    for (int j = 0; j < i; ++j)
      if (plans[j] == plans[i]) return set_error(DEFT_ERR_ARG, "a plan appears twice");
  }
  deft_batch_s *b = new deft_batch_s();
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return set_error(DEFT_ERR_HIP, "stream creation failed");
  }
  b->stage.cap = 131072 + (size_t)n * 8192;
  if (hipHostMalloc((void **)&b->hflag, sizeof(unsigned long long), hipHostMallocMapped | hipHostMallocCoherent) !=
          hipSuccess ||
      hipHostGetDevicePointer((void **)&b->dflag, b->hflag, 0) != hipSuccess) {
    if (b->hflag) (void)hipHostFree(b->hflag);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return set_error(DEFT_ERR_ALLOC, "batch flag allocation failed");
  }
  *b->hflag = 0;
  if (hipHostMalloc((void **)&b->stage.host, b->stage.cap, hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void **)&b->stage.dev, b->stage.cap) != hipSuccess) {
    if (b->stage.host) (void)hipHostFree(b->stage.host);
    if (b->stage.dev) (void)hipFree(b->stage.dev);
    (void)hipHostFree(b->hflag);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return set_error(DEFT_ERR_ALLOC, "batch staging allocation failed");
  }
  // flush and drain every plan before binding any of them, so a failure leaves
  // no plan pointing at the batch
  int rc = DEFT_OK;
  for (int i = 0; i < n && rc == DEFT_OK; ++i) {
    rc = flush_moves(plans[i]);
    // the structure-factor table, where the plan's mu has a bulk density: the
    // gradient confirmation of batched minimizers uses it, as unbatched ones do
    if (rc == DEFT_OK && plan_ensure_sk(plans[i]) != DEFT_OK) g_last_error.clear();
    if (rc == DEFT_OK && hipStreamSynchronize(plans[i]->stream) != hipSuccess)
      rc = set_error(DEFT_ERR_HIP, "hipStreamSynchronize of a plan failed");
  }
  if (rc != DEFT_OK) {
    const std::string err = g_last_error;
    (void)hipHostFree(b->stage.host);
    (void)hipFree(b->stage.dev);
    (void)hipHostFree(b->hflag);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return set_error(rc, err);
  }
  b->plans.reserve(n);
  for (int i = 0; i < n; ++i) {
    deft_plan_s *p = plans[i];
    p->own_stream = p->stream;
    p->stream = b->stream;
    p->batch = b;
    p->batch_index = (int)b->plans.size();
    b->plans.push_back(p);
  }
  *out = b;
  return DEFT_OK;
}

int deft_batch_destroy(deft_batch b) {
  if (!b) return DEFT_OK;
  if (b->active > 0) return set_error(DEFT_ERR_STATE, "the batch is minimizing");
  (void)hipStreamSynchronize(b->stream);
  for (deft_plan_s *p : b->plans) {
    p->stream = p->own_stream;
    p->own_stream = nullptr;
    p->batch = nullptr;
    p->batch_index = -1;
  }
  for (BatchGraph &g : b->graph) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  (void)hipHostFree(b->stage.host);
  (void)hipFree(b->stage.dev);
  (void)hipHostFree(b->hflag);
  (void)hipStreamDestroy(b->stream);
  delete b;
  return DEFT_OK;
}

int deft_batch_set_graphs(deft_batch b, int on) {
  if (!b) return set_error(DEFT_ERR_ARG, "null batch");
  if (b->active > 0) return set_error(DEFT_ERR_STATE, "the batch is minimizing");
  b->graphs = on != 0;
  return DEFT_OK;
}

int deft_batch_stats(deft_batch b, long long *rounds, long long *launches, long long *requests) {
  if (!b) return set_error(DEFT_ERR_ARG, "null batch");
  if (rounds) *rounds = b->rounds_run;
  if (launches) *launches = b->launches;
  if (requests) *requests = b->requests;
  return DEFT_OK;
}

// Variant tuning per engine, then (for a plan whose edges have fused kernels,
// outside batches) the engine itself: one FP32 and one FP64 energy+gradient
// evaluation with the fused passes (their radix order fixed per edge) against
// the generic line passes at their own autotuned radix orders; the faster
// engine stays.  The gradient buffers are saved around the timing.
int deft_plan_autotune(deft_plan p) {
  TRY(check_plan(p));
  TRY(flush_moves(p));
  TRY(plan_sync(p));
  TRY(Launch<float>::autotune(p));
  TRY(Launch<double>::autotune(p));
  p->engine_fused_ms = p->engine_generic_ms = -1.0;
  if (p->generic || p->batch || p->functional == DEFT_WHITE_BEAR_TENSOR) return DEFT_OK;
  const size_t nb = sizeof(double) * (size_t)p->N;
  double *keep = nullptr;
  if (hipMalloc(&keep, 2 * nb) != hipSuccess) return set_error(DEFT_ERR_ALLOC, "hipMalloc of the tuning copy failed");
  auto restore = [&](int rc) {
    if (hipMemcpyAsync(p->g, keep, nb, hipMemcpyDeviceToDevice, p->stream) != hipSuccess ||
        hipMemcpyAsync(p->pg, keep + p->N, nb, hipMemcpyDeviceToDevice, p->stream) != hipSuccess)
      rc = set_error(DEFT_ERR_HIP, "restoring the gradient buffers failed");
    if (plan_sync(p) != DEFT_OK && rc == DEFT_OK) rc = DEFT_ERR_HIP;
    (void)hipFree(keep);
    return rc;
  };
  if (hipMemcpyAsync(keep, p->g, nb, hipMemcpyDeviceToDevice, p->stream) != hipSuccess ||
      hipMemcpyAsync(keep + p->N, p->pg, nb, hipMemcpyDeviceToDevice, p->stream) != hipSuccess) {
    (void)hipFree(keep);
    return set_error(DEFT_ERR_HIP, "saving the gradient buffers failed");
  }
  const float fused = Launch<float>::time_evaluation(p) + Launch<double>::time_evaluation(p);
  LineSched sched_keep[2][3];
  std::memcpy(sched_keep, p->sched, sizeof(sched_keep));
  p->generic = true;
  int rc = ensure_gwork(p);
  if (rc == DEFT_OK) rc = Launch<float>::autotune(p);
  if (rc == DEFT_OK) rc = Launch<double>::autotune(p);
  const float gen = rc == DEFT_OK ? Launch<float>::time_evaluation(p) + Launch<double>::time_evaluation(p) : 1e30f;
  p->engine_fused_ms = fused;
  p->engine_generic_ms = gen;
  if (!(gen < fused)) {
    p->generic = false;
    std::memcpy(p->sched, sched_keep, sizeof(sched_keep));  // the fused kernels' compile-time orders
    (void)hipStreamSynchronize(p->stream);
    (void)hipFree(p->gwork);
    p->gwork = nullptr;
    p->gwork_fields = 0;
    if (rc != DEFT_OK) g_last_error.clear();
    rc = DEFT_OK;
  }
  return restore(rc);
}

int deft_plan_set_engine(deft_plan p, int generic) {
  TRY(check_plan(p));
  if (p->batch) return set_error(DEFT_ERR_STATE, "the plan is in a batch");
  if (!generic && !(fused_ok(p->Nx) && fused_ok(p->Ny) && fused_ok(p->Nz)))
    return set_error(DEFT_ERR_UNSUPPORTED, "this plan's edges have no fused kernels");
  if (!generic && (p->functional == DEFT_SFMT || p->functional == DEFT_SW_LIQUID))
    return set_error(DEFT_ERR_UNSUPPORTED, "soft FMT and the square-well liquid run on the generic passes");
  if (!generic == !p->generic) return DEFT_OK;
  TRY(flush_moves(p));
  TRY(plan_sync(p));
  if (!generic) {
    const int lens[3] = {p->Nx, p->Ny, p->Nz};
    for (int a = 0; a < 3; ++a) p->sched[0][a] = p->sched[1][a] = make_sched(lens[a], radix_candidates(lens[a])[0]);
  }
  p->generic = generic != 0;
  return ensure_gwork(p);
}

int deft_plan_engine(deft_plan p, int *generic, double *fused_ms, double *generic_ms) {
  TRY(check_plan(p));
  if (generic) *generic = p->generic ? 1 : 0;
  if (fused_ms) *fused_ms = p->engine_fused_ms;
  if (generic_ms) *generic_ms = p->engine_generic_ms;
  return DEFT_OK;
}

int deft_plan_tuning(deft_plan p, int fft_precision, int *zblock, int *ytile, int *xtile) {
  TRY(check_plan(p));
  if (fft_precision != DEFT_FFT_FP32 && fft_precision != DEFT_FFT_FP64)
    return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  if (zblock) *zblock = p->tz[fft_precision];
  if (ytile) *ytile = p->ty[fft_precision];
  if (xtile) *xtile = p->tx[fft_precision];
  return DEFT_OK;
}

int deft_plan_set_tuning(deft_plan p, int fft_precision, int zblock, int ytile, int xtile) {
  TRY(check_plan(p));
  if (fft_precision != DEFT_FFT_FP32 && fft_precision != DEFT_FFT_FP64)
    return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  auto in = [](int v, const int (&opts)[2]) { return v == opts[0] || v == opts[1]; };
  if (!in(zblock, kZBlocks) || !in(ytile, kTileBytes) || !in(xtile, kTileBytes))
    return set_error(DEFT_ERR_ARG, "zblock must be 512 or 1024, tiles 8192 or 16384");
  p->tz[fft_precision] = zblock;
  p->ty[fft_precision] = ytile;
  p->tx[fft_precision] = xtile;
  return DEFT_OK;
}

int deft_plan_zphi_variant(deft_plan p, int fft_precision, int *variant) {
  TRY(check_plan(p));
  if (fft_precision != DEFT_FFT_FP32 && fft_precision != DEFT_FFT_FP64)
    return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  if (variant) *variant = p->zw[fft_precision];
  return DEFT_OK;
}

int deft_plan_set_zphi_variant(deft_plan p, int fft_precision, int variant) {
  TRY(check_plan(p));
  if (fft_precision != DEFT_FFT_FP32 && fft_precision != DEFT_FFT_FP64)
    return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  if (variant != 0 && variant != 1) return set_error(DEFT_ERR_ARG, "the Phi pass variant is 0 or 1");
  if (variant == 1 && (p->generic || p->functional != DEFT_WHITE_BEAR || !wave_fft_length(p->Nz)))
    return set_error(DEFT_ERR_ARG,
                     "the one-field-per-wave Phi pass takes fused White-Bear plans with a power-of-two z edge");
  p->zw[fft_precision] = variant;
  return DEFT_OK;
}

int deft_plan_xfwd_variant(deft_plan p, int fft_precision, int *variant) {
  TRY(check_plan(p));
  if (fft_precision != DEFT_FFT_FP32 && fft_precision != DEFT_FFT_FP64)
    return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  if (variant) *variant = p->xw[fft_precision];
  return DEFT_OK;
}

int deft_plan_set_xfwd_variant(deft_plan p, int fft_precision, int variant) {
  TRY(check_plan(p));
  if (fft_precision != DEFT_FFT_FP32 && fft_precision != DEFT_FFT_FP64)
    return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  if (variant != 0 && variant != 1) return set_error(DEFT_ERR_ARG, "the x pass variant is 0 or 1");
  if (variant == 1 && (p->generic || plan_fields(p) != 5 || !wave_fft_length(p->Nx)))
    return set_error(DEFT_ERR_ARG,
                     "the one-tile-per-wave x pass takes fused five-field plans with a power-of-two x edge");
  p->xw[fft_precision] = variant;
  return DEFT_OK;
}

int deft_plan_radix_schedule(deft_plan p, int fft_precision, int axis, int *radices, int capacity, int *nstages) {
  TRY(check_plan(p));
  if (fft_precision != DEFT_FFT_FP32 && fft_precision != DEFT_FFT_FP64)
    return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  if (axis < 0 || axis > 2) return set_error(DEFT_ERR_ARG, "axis must be 0, 1 or 2");
  if (!nstages || (capacity > 0 && !radices)) return set_error(DEFT_ERR_ARG, "null output");
  const LineSched &s = p->sched[fft_precision][axis];
  *nstages = s.nst;
  for (int i = 0; i < s.nst && i < capacity; ++i) radices[i] = s.rad[i];
  return DEFT_OK;
}

int deft_plan_set_radix_schedule(deft_plan p, int fft_precision, int axis, const int *radices, int nstages) {
  TRY(check_plan(p));
  if (fft_precision != DEFT_FFT_FP32 && fft_precision != DEFT_FFT_FP64)
    return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  if (axis < 0 || axis > 2) return set_error(DEFT_ERR_ARG, "axis must be 0, 1 or 2");
  if (!p->generic)
    return set_error(DEFT_ERR_UNSUPPORTED, "the fused kernels of this plan's edges fix their radix order at compile time");
  if (nstages > 0 && !radices) return set_error(DEFT_ERR_ARG, "null radices");
  const int L = axis == 0 ? p->Nx : (axis == 1 ? p->Ny : p->Nz);
  if (!sched_valid(L, radices, nstages))
    return set_error(DEFT_ERR_ARG, "radices must be 4, 8 or primes up to 512 whose product is the edge (at most 12)");
  p->sched[fft_precision][axis] = make_sched(L, std::vector<int>(radices, radices + nstages));
  return DEFT_OK;
}

int deft_plan_set_fft_precision(deft_plan p, int prec) {
  TRY(check_plan(p));
  if (prec != DEFT_FFT_FP32 && prec != DEFT_FFT_FP64) return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  p->prec = prec;
  return DEFT_OK;
}

int deft_plan_set_functional(deft_plan p, int functional) {
  TRY(check_plan(p));
  if (functional != DEFT_WHITE_BEAR && functional != DEFT_WHITE_BEAR_MARK_II &&
      functional != DEFT_WHITE_BEAR_MARK_II_SLOW && functional != DEFT_WHITE_BEAR_TENSOR && functional != DEFT_SFMT &&
      functional != DEFT_SW_LIQUID)
    return set_error(DEFT_ERR_ARG, "unknown functional");
  if (functional != DEFT_WHITE_BEAR && p->batch)
    return set_error(DEFT_ERR_STATE, "plans in a batch keep the White-Bear functional");
  // the weighted-density spectra plus the density spectrum
  const int need = functional == DEFT_SFMT ? 17
                   : functional == DEFT_SW_LIQUID ? 10
                   : (functional == DEFT_WHITE_BEAR_TENSOR ? 12 : 6);
  if (p->spec_fields < need) {
    void *spec = nullptr;
    const size_t bytes = need * sizeof(cx<double>) * (size_t)p->ks;
    if (hipStreamSynchronize(p->stream) != hipSuccess) return set_error(DEFT_ERR_HIP, "stream synchronisation failed");
    if (hipMalloc(&spec, bytes) != hipSuccess) return set_error(DEFT_ERR_ALLOC, "hipMalloc of the tensor spectra failed");
    if (hipMemset(spec, 0, bytes) != hipSuccess) {
      (void)hipFree(spec);
      return set_error(DEFT_ERR_HIP, "initialisation failed");
    }
    (void)hipFree(p->spec);
    p->spec = spec;
    p->spec_fields = need;
  }
  // eleven fields per z line: the smaller z blocks keep the tensor Phi pass at
  // two waves per SIMD
  if (functional == DEFT_WHITE_BEAR_TENSOR) p->tz[0] = p->tz[1] = kZBlocks[0];
  const int was = p->functional;
  const bool was_generic = p->generic;
  p->functional = functional;
  p->sk_ok = false;
  // soft FMT and the square-well liquid run on the generic passes; leaving
  // them restores the fused ones where the edges have them
  const bool gen_only = functional == DEFT_SFMT || functional == DEFT_SW_LIQUID;
  if (gen_only)
    p->generic = true;
  else if (was == DEFT_SFMT || was == DEFT_SW_LIQUID)
    p->generic = !(fused_ok(p->Nx) && fused_ok(p->Ny) && fused_ok(p->Nz));
  int rc = ensure_gwork(p);
  if (rc == DEFT_OK && functional == DEFT_SFMT) rc = plan_build_sfmt_table(p);
  if (rc == DEFT_OK && functional == DEFT_SW_LIQUID && !p->swbuf) {
    if (hipMalloc(&p->swbuf, 26 * sizeof(double) * (size_t)p->N) != hipSuccess) {
      p->swbuf = nullptr;
      rc = set_error(DEFT_ERR_ALLOC, "hipMalloc of the square-well work fields failed");
    }
  }
  if (rc != DEFT_OK) {
    p->functional = was;
    p->generic = was_generic;
  }
  return rc;
}

int deft_plan_set_sfmt(deft_plan p, double sigma, double epsilon, double xi) {
  TRY(check_plan(p));
  auto ok = [](double v) { return std::isfinite(v) && v > 0.0; };
  if (!ok(sigma) || !ok(epsilon) || !ok(xi)) return set_error(DEFT_ERR_ARG, "sigma, epsilon and Xi must be positive");
  p->sfmt_sigma = sigma;
  p->sfmt_epsilon = epsilon;
  p->sfmt_xi = xi;
  p->stab_ok = false;
  p->sk_ok = false;
  if (p->functional == DEFT_SFMT) TRY(plan_build_sfmt_table(p));
  return DEFT_OK;
}

int deft_plan_set_sw(deft_plan p, double epsilon, double lambda) {
  TRY(check_plan(p));
  if (!std::isfinite(epsilon) || !(lambda > 1.0) || !std::isfinite(lambda))
    return set_error(DEFT_ERR_ARG, "epsilon must be finite and lambda above 1");
  p->sw_eps = epsilon;
  p->sw_lambda = lambda;
  p->sk_ok = false;
  return DEFT_OK;
}

int deft_sw_homogeneous(double R, double epsilon, double lambda, double kT, double n, double mu, double *energy,
                        double *d_by_dn) {
  auto ok = [](double v) { return std::isfinite(v) && v > 0.0; };
  if (!ok(R) || !ok(kT) || !ok(n) || !std::isfinite(epsilon) || !(lambda > 1.0) || !std::isfinite(mu))
    return set_error(DEFT_ERR_ARG, "R, kT and n must be positive, lambda above 1, epsilon and mu finite");
  double f, d;
  sw_bulk(make_sw_const(R, epsilon, lambda), n, kT, std::log(water_nQ(kT)), &f, &d);
  if (energy) *energy = f - mu * n;
  if (d_by_dn) *d_by_dn = d - mu;
  return DEFT_OK;
}

int deft_sfmt_homogeneous(double sigma, double epsilon, double xi, double kT, double n, double mu, double *energy,
                          double *d_by_dn) {
  auto ok = [](double v) { return std::isfinite(v) && v > 0.0; };
  if (!ok(sigma) || !ok(epsilon) || !ok(xi) || !ok(kT) || !ok(n) || !std::isfinite(mu))
    return set_error(DEFT_ERR_ARG, "sigma, epsilon, Xi, kT and n must be positive, mu finite");
  double f, d;
  sfmt_bulk(make_sfmt_const(sigma, epsilon, kT, xi), n, kT, &f, &d);
  if (energy) *energy = f - mu * n;
  if (d_by_dn) *d_by_dn = d - mu;
  return DEFT_OK;
}

int deft_plan_set_mu(deft_plan p, double mu) {
  TRY(check_plan(p));
  if (!std::isfinite(mu)) return set_error(DEFT_ERR_ARG, "mu must be finite");
  p->mu = mu;
  p->sk_ok = false;
  return DEFT_OK;
}

int deft_plan_set_vext(deft_plan p, const double *vext) {
  TRY(check_plan(p));
  if (!vext) return set_error(DEFT_ERR_ARG, "null vext");
  std::vector<float> f((size_t)p->N);
  for (long long i = 0; i < p->N; ++i) f[i] = (float)vext[i];
  DEFT_HIP_CHECK(hipMemcpyAsync(p->vext, f.data(), sizeof(float) * (size_t)p->N, hipMemcpyHostToDevice, p->stream));
  DEFT_HIP_CHECK(hipMemcpyAsync(p->vext64, vext, sizeof(double) * (size_t)p->N, hipMemcpyHostToDevice, p->stream));
  TRY(plan_sync(p));
  return DEFT_OK;
}

int deft_plan_set_veff(deft_plan p, const double *veff) {
  TRY(check_plan(p));
  if (!veff) return set_error(DEFT_ERR_ARG, "null veff");
  p->npend = 0;
  DEFT_HIP_CHECK(hipMemcpyAsync(p->x, veff, sizeof(double) * (size_t)p->N, hipMemcpyHostToDevice, p->stream));
  TRY(plan_sync(p));
  return DEFT_OK;
}

int deft_plan_fill_veff(deft_plan p, double value) {
  TRY(check_plan(p));
  p->npend = 0;
  hipLaunchKernelGGL(k_fill, dim3(p->eblocks), dim3(kThreads), 0, p->stream, p->N, p->x, value);
  DEFT_HIP_CHECK(hipGetLastError());
  TRY(plan_sync(p));
  return DEFT_OK;
}

int deft_plan_get_veff(deft_plan p, double *veff) {
  TRY(check_plan(p));
  if (!veff) return set_error(DEFT_ERR_ARG, "null veff");
  TRY(flush_moves(p));
  DEFT_HIP_CHECK(hipMemcpyAsync(veff, p->x, sizeof(double) * (size_t)p->N, hipMemcpyDeviceToHost, p->stream));
  TRY(plan_sync(p));
  return DEFT_OK;
}

int deft_plan_get_density(deft_plan p, double *n) {
  TRY(check_plan(p));
  if (!n) return set_error(DEFT_ERR_ARG, "null output");
  TRY(flush_moves(p));
  TRY(ensure_tmp(p, sizeof(double) * (size_t)p->N));
  hipLaunchKernelGGL(k_density, dim3(p->eblocks), dim3(kThreads), 0, p->stream, p->N, p->x, p->vext, p->kT, p->tmp);
  DEFT_HIP_CHECK(hipGetLastError());
  DEFT_HIP_CHECK(hipMemcpyAsync(n, p->tmp, sizeof(double) * (size_t)p->N, hipMemcpyDeviceToHost, p->stream));
  TRY(plan_sync(p));
  return DEFT_OK;
}

int deft_plan_veff_device_ptr(deft_plan p, double **ptr) {
  TRY(check_plan(p));
  if (!ptr) return set_error(DEFT_ERR_ARG, "null output");
  TRY(flush_moves(p));
  TRY(plan_sync(p));
  *ptr = p->x;
  return DEFT_OK;
}

int deft_plan_sizes(deft_plan p, long long *n_real, long long *n_spectral) {
  TRY(check_plan(p));
  if (n_real) *n_real = p->N;
  if (n_spectral) *n_spectral = p->Nk;
  return DEFT_OK;
}

int deft_functional_energy(deft_plan p, const double *veff, double *energy, deft_energy_terms *terms) {
  TRY(check_plan(p));
  if (veff) TRY(deft_plan_set_veff(p, veff));
  TRY(evaluate(p, false, false, PGRAD_NONE));
  p->last_terms = terms_from_sums(p);
  if (energy) *energy = p->last_terms.total;
  if (terms) *terms = p->last_terms;
  return DEFT_OK;
}

int deft_functional_grad_device(deft_plan p, const double *veff_dev, double *outgrad_dev, double *outpgrad_dev,
                                double *energy) {
  return deft_functional_grad_device_kind(p, veff_dev, outgrad_dev, outpgrad_dev, energy, DEFT_PGRAD_NEW);
}

int deft_functional_grad_device_kind(deft_plan p, const double *veff_dev, double *outgrad_dev, double *outpgrad_dev,
                                     double *energy, int pgrad_kind) {
  TRY(check_plan(p));
  if (!outgrad_dev) return set_error(DEFT_ERR_ARG, "null outgrad");
  if (pgrad_kind != DEFT_PGRAD_NEW && pgrad_kind != DEFT_PGRAD_EFFECTIVE_POTENTIAL &&
      pgrad_kind != DEFT_PGRAD_STRUCTURE_FACTOR)
    return set_error(DEFT_ERR_ARG, "unknown pgrad kind");
  if (veff_dev) {
    p->npend = 0;
    DEFT_HIP_CHECK(hipMemcpyAsync(p->x, veff_dev, sizeof(double) * (size_t)p->N, hipMemcpyDeviceToDevice, p->stream));
  }
  TRY(evaluate(p, true, false, pgrad_kind));
  p->last_terms = terms_from_sums(p);
  if (energy) *energy = p->last_terms.total;
  hipLaunchKernelGGL(k_accumulate, dim3(p->eblocks), dim3(kThreads), 0, p->stream, p->N, outgrad_dev, p->g);
  if (outpgrad_dev)
    hipLaunchKernelGGL(k_accumulate, dim3(p->eblocks), dim3(kThreads), 0, p->stream, p->N, outpgrad_dev, p->pg);
  DEFT_HIP_CHECK(hipGetLastError());
  TRY(plan_sync(p));
  return DEFT_OK;
}

int deft_functional_grad(deft_plan p, const double *veff, double *outgrad, double *outpgrad, double *energy) {
  return deft_functional_grad_kind(p, veff, outgrad, outpgrad, energy, DEFT_PGRAD_NEW);
}

int deft_functional_grad_kind(deft_plan p, const double *veff, double *outgrad, double *outpgrad, double *energy,
                              int pgrad_kind) {
  TRY(check_plan(p));
  if (!outgrad) return set_error(DEFT_ERR_ARG, "null outgrad");
  if (pgrad_kind != DEFT_PGRAD_NEW && pgrad_kind != DEFT_PGRAD_EFFECTIVE_POTENTIAL &&
      pgrad_kind != DEFT_PGRAD_STRUCTURE_FACTOR)
    return set_error(DEFT_ERR_ARG, "unknown pgrad kind");
  if (veff) TRY(deft_plan_set_veff(p, veff));
  TRY(evaluate(p, true, false, pgrad_kind));
  p->last_terms = terms_from_sums(p);
  if (energy) *energy = p->last_terms.total;
  std::vector<double> h((size_t)p->N);
  DEFT_HIP_CHECK(hipMemcpyAsync(h.data(), p->g, sizeof(double) * (size_t)p->N, hipMemcpyDeviceToHost, p->stream));
  TRY(plan_sync(p));
  for (long long i = 0; i < p->N; ++i) outgrad[i] += h[i];
  if (outpgrad) {
    DEFT_HIP_CHECK(hipMemcpyAsync(h.data(), p->pg, sizeof(double) * (size_t)p->N, hipMemcpyDeviceToHost, p->stream));
    TRY(plan_sync(p));
    for (long long i = 0; i < p->N; ++i) outpgrad[i] += h[i];
  }
  return DEFT_OK;
}

int deft_weighted_densities(deft_plan p, double *out5) {
  TRY(check_plan(p));
  if (!out5) return set_error(DEFT_ERR_ARG, "null output");
  const int nw = plan_fields(p);
  TRY(ensure_tmp(p, nw * sizeof(double) * (size_t)p->N));
  if (p->prec == DEFT_FFT_FP64)
    TRY(Launch<double>::weighted_densities(p, p->tmp));
  else
    TRY(Launch<float>::weighted_densities(p, p->tmp));
  DEFT_HIP_CHECK(hipMemcpyAsync(out5, p->tmp, nw * sizeof(double) * (size_t)p->N, hipMemcpyDeviceToHost, p->stream));
  TRY(plan_sync(p));
  return DEFT_OK;
}

// four independent 16-byte loads in flight per thread and iteration
static __global__ __launch_bounds__(kThreads) void k_stream_copy(long long n, const double2 *__restrict__ in,
                                                                   double2 *__restrict__ out) {
  const long long stride = (long long)gridDim.x * kThreads;
  long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const double2 a = in[i], b = in[i + stride], c = in[i + 2 * stride], d = in[i + 3 * stride];
    out[i] = a;
    out[i + stride] = b;
    out[i + 2 * stride] = c;
    out[i + 3 * stride] = d;
  }
  for (; i < n; i += stride) out[i] = in[i];
}

int deft_copy_bandwidth(long long bytes, int reps, double *gbps) {
  if (!gbps || bytes < (1 << 20) || reps < 1) return set_error(DEFT_ERR_ARG, "need >= 1 MiB, reps >= 1 and an output");
  const long long n = bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = DEFT_OK;
  if (hipMalloc(&a, 16 * n) != hipSuccess || hipMalloc(&b, 16 * n) != hipSuccess) {
    rc = set_error(DEFT_ERR_ALLOC, "copy buffers");
  } else if (hipMemset(a, 0, 16 * n) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
             hipEventCreate(&e1) != hipSuccess) {
    rc = set_error(DEFT_ERR_HIP, "copy set-up");
  } else {
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // the best of several launch widths (blocks per CU)
    double best = 0.0;
    for (int per_cu : {2, 4, 8, 16, 32}) {
      const dim3 grid((unsigned)std::min<long long>((long long)cus * per_cu, (n + kThreads - 1) / kThreads));
      hipLaunchKernelGGL(k_stream_copy, grid, dim3(kThreads), 0, nullptr, n, a, b);
      (void)hipEventRecord(e0, nullptr);
      for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_stream_copy, grid, dim3(kThreads), 0, nullptr, n, a, b);
      (void)hipEventRecord(e1, nullptr);
      float ms = 0.f;
      if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || !(ms > 0.f)) {
        rc = set_error(DEFT_ERR_HIP, "copy timing");
        break;
      }
      best = std::max(best, 2.0 * 16.0 * (double)n * reps / (ms * 1e-3) / 1e9);
    }
    if (rc == DEFT_OK) *gbps = best;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

int deft_contact_values(deft_plan p, double *out7) {
  TRY(check_plan(p));
  if (!out7) return set_error(DEFT_ERR_ARG, "null output");
  if (p->functional == DEFT_WHITE_BEAR_TENSOR || p->functional == DEFT_SFMT || p->functional == DEFT_SW_LIQUID)
    return set_error(DEFT_ERR_UNSUPPORTED, "contact values are defined for White-Bear and Mark II");
  TRY(flush_moves(p));
  const size_t N = (size_t)p->N;
  TRY(ensure_tmp(p, 23 * sizeof(double) * N));
  double *w5 = p->tmp, *shell2 = w5 + 5 * N, *step2 = w5 + 6 * N, *abcd = w5 + 7 * N, *num = w5 + 15 * N,
         *out = w5 + 16 * N;
  const GridArgs g = grid_args(p);
  const double R = p->R;
  const dim3 eb(p->eblocks), tb(kThreads);
  const int kblocks = (int)std::min<long long>(4096, (p->Nk + kThreads - 1) / kThreads);
  cx<double> *f0 = spec_ptr<double>(p, 0), *f2 = spec_ptr<double>(p, 2), *f3 = spec_ptr<double>(p, 3);
  TRY(Launch<double>::weighted_densities(p, w5));
  hipLaunchKernelGGL(k_density, eb, tb, 0, p->stream, p->N, p->x, p->vext, p->kT, num);
  TRY(Launch<double>::fft_forward(p, num, 1.0));
  hipLaunchKernelGGL(k_spec_apply, dim3(kblocks), tb, 0, p->stream, g, f0, f2, (int)SK_COPY, 0, R, 1.0, 0);
  const double invN = 1.0 / (double)p->N;
  hipLaunchKernelGGL(k_spec_apply, dim3(kblocks), tb, 0, p->stream, g, f2, f0, (int)SK_SHELL, 0, 2 * R, 1.0, 0);
  TRY(Launch<double>::fft_inverse(p, shell2, invN));
  hipLaunchKernelGGL(k_spec_apply, dim3(kblocks), tb, 0, p->stream, g, f2, f0, (int)SK_STEP, 0, 2 * R, 1.0, 0);
  TRY(Launch<double>::fft_inverse(p, step2, invN));
  const double s3 = p->functional == DEFT_WHITE_BEAR ? 0.0 : (p->functional == DEFT_WHITE_BEAR_MARK_II_SLOW ? 1.0 : 1.5);
  hipLaunchKernelGGL(k_contact_point, eb, tb, 0, p->stream, p->N, w5, shell2, step2, R, s3, out, abcd);
  // gSigmaA numerator: shell(A) - shell'(B) + vshell'.(C) + vshell.(D)
  const int kinds[8] = {SK_SHELL, SK_SHELL_PRIME, SK_VSHELL_PRIME, SK_VSHELL_PRIME, SK_VSHELL_PRIME,
                        SK_VSHELL, SK_VSHELL, SK_VSHELL};
  const int comps[8] = {0, 0, 0, 1, 2, 0, 1, 2};
  const double coefs[8] = {1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
  for (int f = 0; f < 8; ++f) {
    TRY(Launch<double>::fft_forward(p, abcd + f * N, 1.0));
    hipLaunchKernelGGL(k_spec_apply, dim3(kblocks), tb, 0, p->stream, g, f0, f3, kinds[f], comps[f], R, coefs[f],
                       f > 0 ? 1 : 0);
  }
  hipLaunchKernelGGL(k_spec_apply, dim3(kblocks), tb, 0, p->stream, g, f3, f0, (int)SK_COPY, 0, R, 1.0, 0);
  TRY(Launch<double>::fft_inverse(p, num, invN));
  hipLaunchKernelGGL(k_divide, eb, tb, 0, p->stream, p->N, num, shell2, out + 4 * N);
  DEFT_HIP_CHECK(hipGetLastError());
  DEFT_HIP_CHECK(hipMemcpyAsync(out7, out, 7 * sizeof(double) * N, hipMemcpyDeviceToHost, p->stream));
  TRY(plan_sync(p));
  return DEFT_OK;
}

int deft_fft(deft_plan p, const double *in, double *out) {
  TRY(check_plan(p));
  if (!in || !out) return set_error(DEFT_ERR_ARG, "null buffer");
  TRY(ensure_tmp(p, sizeof(double) * (size_t)p->N));
  DEFT_HIP_CHECK(hipMemcpyAsync(p->tmp, in, sizeof(double) * (size_t)p->N, hipMemcpyHostToDevice, p->stream));
  if (p->prec == DEFT_FFT_FP64) {
    TRY(Launch<double>::fft_forward(p, p->tmp, p->dV));
    std::vector<double> h(2 * (size_t)p->ks);
    DEFT_HIP_CHECK(hipMemcpyAsync(h.data(), p->spec, sizeof(cx<double>) * (size_t)p->ks, hipMemcpyDeviceToHost, p->stream));
    TRY(plan_sync(p));
    compact_from_padded(p, h.data(), out);
  } else {
    TRY(Launch<float>::fft_forward(p, p->tmp, p->dV));
    std::vector<float> h(2 * (size_t)p->ks);
    DEFT_HIP_CHECK(hipMemcpyAsync(h.data(), p->spec, sizeof(cx<float>) * (size_t)p->ks, hipMemcpyDeviceToHost, p->stream));
    TRY(plan_sync(p));
    compact_from_padded(p, h.data(), out);
  }
  return DEFT_OK;
}

int deft_ifft(deft_plan p, const double *in, double *out) {
  TRY(check_plan(p));
  if (!in || !out) return set_error(DEFT_ERR_ARG, "null buffer");
  TRY(ensure_tmp(p, sizeof(double) * (size_t)p->N));
  if (p->prec == DEFT_FFT_FP64) {
    std::vector<double> h(2 * (size_t)p->ks);
    padded_from_compact(p, in, h.data());
    DEFT_HIP_CHECK(hipMemcpyAsync(p->spec, h.data(), sizeof(cx<double>) * (size_t)p->ks, hipMemcpyHostToDevice, p->stream));
    TRY(plan_sync(p));
    TRY(Launch<double>::fft_inverse(p, p->tmp, 1.0 / p->volume));
  } else {
    std::vector<float> h(2 * (size_t)p->ks);
    padded_from_compact(p, in, h.data());
    DEFT_HIP_CHECK(hipMemcpyAsync(p->spec, h.data(), sizeof(cx<float>) * (size_t)p->ks, hipMemcpyHostToDevice, p->stream));
    TRY(plan_sync(p));
    TRY(Launch<float>::fft_inverse(p, p->tmp, 1.0 / p->volume));
  }
  DEFT_HIP_CHECK(hipMemcpyAsync(out, p->tmp, sizeof(double) * (size_t)p->N, hipMemcpyDeviceToHost, p->stream));
  TRY(plan_sync(p));
  return DEFT_OK;
}

// --- bulk path (src/equation-of-state.cpp:164-192) ---------------------------
static void bulk_terms(double n, double R, double kT, double *f_ex, double *df_ex, int functional = 0) {
  const WBConst c = make_wb_const(R, functional);
  const double n3 = 4 * M_PI / 3 * R * R * R * n, n2 = 4 * M_PI * R * R * n;
  double p1, p2, p3, d[5];
  wb_phi_partials(n3, n2, 0.0, 0.0, 0.0, c, p1, p2, p3, d);
  *f_ex = kT * (p1 + p2 + p3);
  *df_ex = kT * (d[0] * (4 * M_PI / 3 * R * R * R) + d[1] * (4 * M_PI * R * R));
}

int deft_bulk_chemical_potential(double n, double R, double kT, double *mu) {
  if (!mu || !(n > 0) || !(R > 0) || !(kT > 0)) return set_error(DEFT_ERR_ARG, "bad bulk arguments");
  double f, df;
  bulk_terms(n, R, kT, &f, &df);
  *mu = kT * std::log(n / water_nQ(kT)) + df;
  return DEFT_OK;
}

int deft_bulk_free_energy_density(double n, double R, double kT, double *out) {
  if (!out || !(n > 0) || !(R > 0) || !(kT > 0)) return set_error(DEFT_ERR_ARG, "bad bulk arguments");
  double f, df;
  bulk_terms(n, R, kT, &f, &df);
  *out = kT * n * (std::log(n / water_nQ(kT)) - 1) + f;
  return DEFT_OK;
}

int deft_bulk_state(int functional, double n, double R, double kT, double *f, double *mu, double *pr) {
  if (!f || !mu || !pr || !(n > 0) || !(R > 0) || !(kT > 0)) return set_error(DEFT_ERR_ARG, "bad bulk arguments");
  // (the tensor functional's bulk is White-Bear's: T = 0 and n2v = 0 there)
  if (functional != DEFT_WHITE_BEAR && functional != DEFT_WHITE_BEAR_MARK_II &&
      functional != DEFT_WHITE_BEAR_MARK_II_SLOW && functional != DEFT_WHITE_BEAR_TENSOR)
    return set_error(DEFT_ERR_ARG, "unknown functional");
  double fe, dfe;
  bulk_terms(n, R, kT, &fe, &dfe, functional);
  const double lnn = std::log(n / water_nQ(kT));
  *f = kT * n * (lnn - 1) + fe;
  *mu = kT * lnn + dfe;
  *pr = n * *mu - *f;
  return DEFT_OK;
}

}  // extern "C"

namespace {

// Bulk density of the plan's mu and functional: mu(n) rises monotonically from
// -inf (n -> 0) towards +inf (n3 -> 1); bisection in ln n.
bool bulk_density_of_mu(const deft_plan_s *p, double &nb) {
  const bool soft = p->functional == DEFT_SFMT;
  const SFMTConst sc = make_sfmt_const(p->sfmt_sigma, p->sfmt_epsilon, p->kT, p->sfmt_xi);
  double r0[7];
  sfmt_radial(0.0, sc, r0);  // soft FMT: n3 = w3(0) n
  const double nmax = soft ? 1.0 / r0[0] : 1.0 / (4 * M_PI / 3 * p->R * p->R * p->R);
  const SWConst swc = make_sw_const(p->R, p->sw_eps, p->sw_lambda);
  auto mu_of = [&](double n) {
    if (soft) {
      double f, d;
      sfmt_bulk(sc, n, p->kT, &f, &d);
      return d;
    }
    if (p->functional == DEFT_SW_LIQUID) {
      double f, d;
      sw_bulk(swc, n, p->kT, std::log(water_nQ(p->kT)), &f, &d);
      return d;
    }
    double fe, dfe;
    bulk_terms(n, p->R, p->kT, &fe, &dfe, p->functional);
    return p->kT * std::log(n / water_nQ(p->kT)) + dfe;
  };
  double lo = std::log(1e-30 * nmax), hi = std::log((1 - 1e-9) * nmax);
  if (!(mu_of(std::exp(lo)) < p->mu) || !(mu_of(std::exp(hi)) > p->mu)) return false;
  for (int i = 0; i < 200; ++i) {
    const double mid = 0.5 * (lo + hi);
    (mu_of(std::exp(mid)) < p->mu ? lo : hi) = mid;
  }
  nb = std::exp(0.5 * (lo + hi));
  return std::isfinite(nb) && nb > 0;
}

}  // namespace

namespace {

// The structure-factor table of the plan's current mu, functional and radius:
// Phi's second derivatives at the bulk weighted densities by central
// differences (relative step 1e-5) of the FP64 partials, then S(k)/N per
// spectral element on the device.
int plan_ensure_sk(deft_plan_s *p) {
  if (p->sk_ok) return DEFT_OK;
  if (p->functional == DEFT_SFMT || p->functional == DEFT_SW_LIQUID)
    return set_error(DEFT_ERR_UNSUPPORTED, "the structure-factor preconditioner is built for the White-Bear functionals");
  double nb;
  if (!bulk_density_of_mu(p, nb))
    return set_error(DEFT_ERR_ARG, "no bulk density has this chemical potential (structure-factor preconditioner)");
  const WBConst c = make_wb_const(p->R, p->functional == DEFT_WHITE_BEAR_TENSOR ? DEFT_WHITE_BEAR : p->functional);
  const double n3 = 4 * M_PI / 3 * p->R * p->R * p->R * nb, n2 = 4 * M_PI * p->R * p->R * nb;
  auto part = [&](double a3, double a2, double vx, double *d) {
    double p1, p2, p3;
    wb_phi_partials(a3, a2, vx, 0.0, 0.0, c, p1, p2, p3, d);
  };
  double dp[5], dm[5];
  const double h3 = 1e-5 * n3, h2 = 1e-5 * n2;
  part(n3 + h3, n2, 0.0, dp);
  part(n3 - h3, n2, 0.0, dm);
  const double P33 = (dp[0] - dm[0]) / (2 * h3), P32a = (dp[1] - dm[1]) / (2 * h3);
  part(n3, n2 + h2, 0.0, dp);
  part(n3, n2 - h2, 0.0, dm);
  const double P22 = (dp[1] - dm[1]) / (2 * h2), P32b = (dp[0] - dm[0]) / (2 * h2);
  part(n3, n2, h2, dp);
  part(n3, n2, -h2, dm);
  const double cv = (dp[2] - dm[2]) / (2 * h2);
  if (!p->sktab && hipMalloc(&p->sktab, sizeof(double) * (size_t)p->ks) != hipSuccess) {
    p->sktab = nullptr;
    return set_error(DEFT_ERR_ALLOC, "hipMalloc of the structure-factor table failed");
  }
  const int kb = (int)std::min<long long>(4096, (p->Nk + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_sk_table, dim3(kb), dim3(kThreads), 0, p->stream, grid_args(p), P33, 0.5 * (P32a + P32b), P22,
                     cv, nb, 1e3, p->sktab);
  DEFT_HIP_CHECK(hipGetLastError());
  TRY(plan_sync(p));
  p->sk_nb = nb;
  p->sk_floor = 1e-3 * nb;
  p->sk_ok = true;
  return DEFT_OK;
}

}  // namespace

extern "C" {

int deft_plan_bulk_density(deft_plan p, double *nb) {
  TRY(check_plan(p));
  if (!nb) return set_error(DEFT_ERR_ARG, "null output");
  if (!bulk_density_of_mu(p, *nb)) return set_error(DEFT_ERR_ARG, "no bulk density has this chemical potential");
  return DEFT_OK;
}

int deft_bulk_pressure(double n, double R, double kT, double *pr) {
  if (!pr) return set_error(DEFT_ERR_ARG, "null output");
  double mu, f;
  TRY(deft_bulk_chemical_potential(n, R, kT, &mu));
  TRY(deft_bulk_free_energy_density(n, R, kT, &f));
  *pr = n * mu - f;
  return DEFT_OK;
}

// --- minimizer ----------------------------------------------------------------
int deft_minimize_create(deft_plan p, deft_minimize *out) {
  TRY(check_plan(p));
  if (!out) return set_error(DEFT_ERR_ARG, "null output");
  deft_minimize_s *m = new deft_minimize_s();
  m->p = p;
  *out = m;
  return DEFT_OK;
}

int deft_minimize_destroy(deft_minimize m) {
  delete m;
  return DEFT_OK;
}

int deft_minimize_set_precision(deft_minimize m, double v) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  m->precision = v;
  return DEFT_OK;
}

int deft_minimize_set_relative_precision(deft_minimize m, double v) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  m->relative_precision = v;
  return DEFT_OK;
}

int deft_minimize_set_maxiter(deft_minimize m, int v) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  m->maxiter = v;
  return DEFT_OK;
}

int deft_minimize_set_miniter(deft_minimize m, int v) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  m->miniter = v;
  return DEFT_OK;
}

int deft_minimize_set_preconditioner(deft_minimize m, int pgrad_kind) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  if (pgrad_kind != DEFT_PGRAD_NEW && pgrad_kind != DEFT_PGRAD_STRUCTURE_FACTOR)
    return set_error(DEFT_ERR_ARG, "the new-API minimizer takes DEFT_PGRAD_NEW or DEFT_PGRAD_STRUCTURE_FACTOR");
  if (pgrad_kind == DEFT_PGRAD_STRUCTURE_FACTOR) TRY(plan_ensure_sk(m->p));
  if (pgrad_kind != m->precond_kind) {
    m->precond_kind = pgrad_kind;
    min_invalidate(*m);
    m->have_dir = m->have_old = false;
  }
  return DEFT_OK;
}

int deft_minimize_precondition(deft_minimize m, int v) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  m->use_precond = v != 0;
  min_invalidate(*m);
  return DEFT_OK;
}

int deft_minimize_improve_energy(deft_minimize m, int verbose, int *more) {
  if (!m || !more) return set_error(DEFT_ERR_ARG, "null argument");
  return improve_energy_mixed(m, verbose, more);
}

int deft_batch_minimize(deft_batch b, deft_minimize *mins, int n) {
  if (!b || !mins || n < 1) return set_error(DEFT_ERR_ARG, "need a batch and minimizers");
  for (int i = 0; i < n; ++i) {
    if (!mins[i] || mins[i]->p->batch != b) return set_error(DEFT_ERR_ARG, "every minimizer's plan must be in the batch");
    for (int j = 0; j < i; ++j)
      if (mins[j]->p == mins[i]->p) return set_error(DEFT_ERR_ARG, "two minimizers share a plan");
    if (b->graphs && mins[i]->precond_kind == PGRAD_SK)
      return set_error(DEFT_ERR_UNSUPPORTED, "graph-captured batch rounds take the new API's grad/n only");
  }
  if (b->active > 0) return set_error(DEFT_ERR_STATE, "the batch is already minimizing");
  constexpr size_t kFiberStack = 512 * 1024;
  std::vector<Fiber> fibers(n);
  for (int i = 0; i < n; ++i) {
    Fiber &f = fibers[i];
    f.m = mins[i];
    f.b = b;
    f.stack.resize(kFiberStack);
    if (getcontext(&f.ctx) != 0) return set_error(DEFT_ERR_STATE, "getcontext failed");
    f.ctx.uc_stack.ss_sp = f.stack.data();
    f.ctx.uc_stack.ss_size = f.stack.size();
    f.ctx.uc_link = &b->sched;
    makecontext(&f.ctx, fiber_main, 0);
  }
  b->active = n;
  std::vector<bool> begun(n, false);
  while (b->active > 0) {
    // every live fiber runs until it posts its next request or finishes
    for (int i = 0; i < n; ++i) {
      Fiber &f = fibers[i];
      if (f.done) continue;
      b->current = &f;
      if (!begun[i]) {
        begun[i] = true;
        g_fiber_start = &f;
      }
      if (swapcontext(&b->sched, &f.ctx) != 0) {
        // the fibers cannot be resumed: the batch stops minimizing (their
        // minimizers keep the state of their last completed round)
        b->current = nullptr;
        b->active = 0;
        b->nposted = 0;
        b->posted[0].clear();
        b->posted[1].clear();
        b->posted_cg.clear();
        return set_error(DEFT_ERR_STATE, "fiber switch failed");
      }
    }
    b->current = nullptr;
    if (b->nposted > 0) run_round(b);
  }
  for (int i = 0; i < n; ++i)
    if (fibers[i].rc != DEFT_OK) return set_error(fibers[i].rc, fibers[i].err);
  return DEFT_OK;
}

int deft_minimize_run(deft_minimize m, int max_calls, int *iterations) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  int more = 1, n = 0;
  while (more && n < max_calls) {
    TRY(improve_energy_mixed(m, 0, &more));
    ++n;
  }
  if (iterations) *iterations = n;
  return DEFT_OK;
}

int deft_minimize_energy(deft_minimize m, double *energy) {
  if (!m || !energy) return set_error(DEFT_ERR_ARG, "null argument");
  return m_energy(m, energy);
}

int deft_minimize_stats_get(deft_minimize m, deft_minimize_stats *s) {
  if (!m || !s) return set_error(DEFT_ERR_ARG, "null argument");
  s->iter = m->iter;
  s->num_energy_calcs = m->num_energy;
  s->num_grad_calcs = m->num_grad;
  s->stop_reason = m->stop;
  s->energy = m->E;
  s->delta_energy = m->deltaE;
  s->error_estimate = m->error_estimate;
  s->step = m->step;
  return DEFT_OK;
}

int deft_minimize_set_fused_line_search(deft_minimize m, int mode) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  if (mode < -1 || mode > 1) return set_error(DEFT_ERR_ARG, "mode must be -1, 0 or 1");
  m->fuse_ls = mode;
  return DEFT_OK;
}

int deft_minimize_set_mixed_precision(deft_minimize m, double switch_relative_precision, int polish_grad_precision) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  if (!(switch_relative_precision >= 0)) return set_error(DEFT_ERR_ARG, "switch precision must be >= 0");
  if (polish_grad_precision != DEFT_FFT_FP32 && polish_grad_precision != DEFT_FFT_FP64)
    return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  m->switch_rp = switch_relative_precision;
  m->polish_grad_prec = polish_grad_precision;
  if (switch_relative_precision > 0 && m->phase == 0) {
    m->e_prec = DEFT_FFT_FP32;
    m->g_prec = DEFT_FFT_FP32;
  }
  min_invalidate(*m);
  return DEFT_OK;
}

int deft_minimize_fp32_iterations(deft_minimize m, int *iterations) {
  if (!m || !iterations) return set_error(DEFT_ERR_ARG, "null argument");
  *iterations = m->fp32_iters;
  return DEFT_OK;
}

int deft_minimize_phase_iterations(deft_minimize m, int *fp32_iterations, int *mixed_iterations, int *phase) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  if (fp32_iterations) *fp32_iterations = m->fp32_iters;
  if (mixed_iterations) *mixed_iterations = m->phase == 2 ? m->iter_at_phase2 - m->fp32_iters : m->iter - m->fp32_iters;
  if (phase) *phase = m->phase;
  return DEFT_OK;
}

int deft_plan_set_profiling(deft_plan p, int on) {
  TRY(check_plan(p));
  TRY(plan_sync(p));
  p->prof = on != 0;
  return DEFT_OK;
}

int deft_plan_profile_reset(deft_plan p) {
  TRY(check_plan(p));
  TRY(plan_sync(p));
  for (auto &k : p->kstat) k = KernelStat();
  return DEFT_OK;
}

int deft_kernel_count(void) { return NUM_KERNELS; }

int deft_plan_profile_get(deft_plan p, int kernel, const char **name, long long *launches, double *total_ms,
                          double *total_bytes) {
  TRY(check_plan(p));
  if (kernel < 0 || kernel >= NUM_KERNELS) return set_error(DEFT_ERR_ARG, "kernel index out of range");
  TRY(plan_sync(p));
  if (name) *name = kernel_name(kernel);
  if (launches) *launches = p->kstat[kernel].launches;
  if (total_ms) *total_ms = p->kstat[kernel].ms;
  if (total_bytes) *total_bytes = p->kstat[kernel].bytes;
  return DEFT_OK;
}

int deft_minimize_invalidate(deft_minimize m) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  min_invalidate(*m);
  return DEFT_OK;
}

int deft_minimize_set_gradient_check(deft_minimize m, double factor) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  if (!(factor >= 0)) return set_error(DEFT_ERR_ARG, "factor must be >= 0");
  m->grad_check = factor;
  return DEFT_OK;
}

int deft_minimize_gradient_check_get(deft_minimize m, double *last_grad_error, int *restarts) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  if (last_grad_error) *last_grad_error = m->last_grad_error;
  if (restarts) *restarts = m->grad_restarts;
  return DEFT_OK;
}

// --- old-API minimizer stack -------------------------------------------------
int deft_cg_create(deft_plan p, int preconditioned, double stepsize, deft_cg *out) {
  TRY(check_plan(p));
  if (!out) return set_error(DEFT_ERR_ARG, "null output");
  if (p->batch) return set_error(DEFT_ERR_STATE, "the old-API minimizer runs on plans outside batches");
  if (!std::isfinite(stepsize) || stepsize == 0) return set_error(DEFT_ERR_ARG, "stepsize must be finite and nonzero");
  deft_cg_s *m = new deft_cg_s();
  m->p = p;
  m->preconditioned = preconditioned != 0;
  m->step = m->orig_step = stepsize;
  *out = m;
  return DEFT_OK;
}

int deft_cg_destroy(deft_cg m) {
  delete m;
  return DEFT_OK;
}

int deft_cg_set_precision(deft_cg m, double precision) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  if (!(precision >= 0)) return set_error(DEFT_ERR_ARG, "precision must be >= 0");
  m->precision = precision;
  // PrecisionMinimizer's constructor (src/Precision.cpp:39-46)
  m->error_estimate = std::max(1.0, 2 * precision);
  m->dEdn = 0.0;
  m->log_avg = 0.0;
  m->deltaE = 0.0;
  return DEFT_OK;
}

int deft_cg_set_maxiter(deft_cg m, int maxiter) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  m->maxiter = maxiter;
  return DEFT_OK;
}

int deft_cg_set_fft_precision(deft_cg m, int fft_precision) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  if (fft_precision != -1 && fft_precision != DEFT_FFT_FP32 && fft_precision != DEFT_FFT_FP64)
    return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  m->prec = fft_precision;
  m->E_valid = m->G_valid = false;
  return DEFT_OK;
}

int deft_cg_improve_energy(deft_cg m, int verbose, int *more) {
  if (!m || !more) return set_error(DEFT_ERR_ARG, "null argument");
  DeviceOps ops{m->p};
  return old_improve_energy(*m, ops, verbose, more);
}

int deft_cg_energy(deft_cg m, double *energy) {
  if (!m || !energy) return set_error(DEFT_ERR_ARG, "null argument");
  DeviceOps ops{m->p};
  return old_energy(*m, ops, energy);
}

// grad() / pgrad() of the current point into host buffers (N doubles, written)
static int cg_vector(deft_cg m, double *out, bool pgrad) {
  if (!m || !out) return set_error(DEFT_ERR_ARG, "null argument");
  deft_plan_s *p = m->p;
  const bool want_pg = pgrad && m->preconditioned;
  // the device holds the gradient of the last evaluation; evaluate afresh
  // unless the cached one is of the current point and of the right kind
  DeviceOps ops{p};
  double E, pgg, pgo;
  TRY(ops.gradient(old_prec(*m, p->prec), false, want_pg ? MIN_PGRAD_OLD : MIN_PGRAD_NONE, &E, &pgg, &pgo));
  if (!m->E_valid) {
    m->E = E;
    m->E_valid = true;
  }
  DEFT_HIP_CHECK(hipMemcpyAsync(out, pgrad ? p->pg : p->g, sizeof(double) * (size_t)p->N, hipMemcpyDeviceToHost,
                                p->stream));
  TRY(plan_sync(p));
  return DEFT_OK;
}

int deft_cg_grad(deft_cg m, double *grad) { return cg_vector(m, grad, false); }

int deft_cg_pgrad(deft_cg m, double *pgrad) { return cg_vector(m, pgrad, true); }

int deft_cg_minimize(deft_cg m) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  old_restart(*m);
  return DEFT_OK;
}

int deft_cg_invalidate(deft_cg m) {
  if (!m) return set_error(DEFT_ERR_ARG, "null minimizer");
  m->E_valid = m->G_valid = false;
  return DEFT_OK;
}

int deft_cg_stats_get(deft_cg m, deft_minimize_stats *s) {
  if (!m || !s) return set_error(DEFT_ERR_ARG, "null argument");
  s->iter = m->iter;
  s->num_energy_calcs = m->num_energy;
  s->num_grad_calcs = m->num_grad;
  s->stop_reason = m->stop;
  s->energy = m->E;
  s->delta_energy = m->deltaE;
  s->error_estimate = m->error_estimate;
  s->step = m->step;
  return DEFT_OK;
}

}  // extern "C"
