// deft_functional_grad_noalloc: the energy and gradient of deft_functional_grad_device_kind
// for callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).  The grid comes
// as parameters instead of a plan, and every intermediate buffer the passes
// need -- the six half spectra, the FP64 weight table, twiddles, the FP32 copy
// of Vext, the gradient vectors and the partial sums -- lives in one
// file-scope device array, `scratch`, of kScratchBytes per GPU, held from the
// library's load until the process ends.  The launches are the plan path's
// (Launch<T>::enqueue_evaluation), so the results are that entry point's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "plan.hpp"

using namespace deft;

namespace {

// Grids of at most 64^3 points: the FP64 layout needs about 136 bytes per
// point (six FP64 half spectra 96, weight table 16, grad and pgrad 16, Vext 8)
// plus under 100 KiB of partial sums and twiddles.
constexpr long long kMaxPoints = 64LL * 64 * 64;
constexpr size_t kScratchBytes = 36u << 20;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Device address of `scratch` per device, looked up once per device with that
// device current (each GPU holds its own copy of the array).
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {nullptr};

int scratch_for_device(int dev, unsigned char **out) {
  if (dev < 0 || dev >= kMaxDevices) return set_error(DEFT_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error(DEFT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = static_cast<unsigned char *>(g_scratch_ptr[dev]);
  return DEFT_OK;
}

// exp(-2 pi i m / L) for the three axes, FP64 and FP32 (the plan computes the
// same table on the host)
__global__ __launch_bounds__(kThreads) void k_noalloc_twiddles(int L0, int L1, int L2, cx<double> *__restrict__ twd,
                                                               cx<float> *__restrict__ twf) {
  const int n = L0 + L1 + L2;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const int L = i < L0 ? L0 : (i < L0 + L1 ? L1 : L2);
    const int m = i < L0 ? i : (i < L0 + L1 ? i - L0 : i - L0 - L1);
    const double ang = 2 * M_PI * (double)m / (double)L;
    double s, c;
    sincos(ang, &s, &c);
    twd[i] = {c, -s};
    twf[i] = {(float)c, (float)-s};
  }
}

// Vext copies for the passes: FP32 (rounded) and FP64 (0 when no Vext is given)
__global__ __launch_bounds__(kThreads) void k_noalloc_vext(long long N, const double *__restrict__ in,
                                                           float *__restrict__ out32, double *__restrict__ out64) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long long)gridDim.x * kThreads) {
    const double v = in ? in[i] : 0.0;
    if (out32) out32[i] = (float)v;
    if (out64) out64[i] = v;
  }
}

// The energy from the reduced sums, term by term in the plan's order
// (terms_from_sums), without contraction.
__global__ void k_noalloc_energy(const double *__restrict__ s, double kT, double dV, double *__restrict__ energy) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double t1 = __dmul_rn(__dmul_rn(kT, s[SLOT_PHI1]), dV);
    const double t2 = __dmul_rn(__dmul_rn(kT, s[SLOT_PHI2]), dV);
    const double t3 = __dmul_rn(__dmul_rn(kT, s[SLOT_PHI3]), dV);
    const double ti = __dmul_rn(s[SLOT_IDEAL], dV);
    const double te = __dmul_rn(s[SLOT_EXT], dV);
    const double tm = __dmul_rn(s[SLOT_MUN], dV);
    *energy = __dadd_rn(__dadd_rn(__dadd_rn(__dadd_rn(__dadd_rn(t1, t2), t3), ti), te), tm);
  }
}

// bump allocator over the scratch array
struct Carve {
  unsigned char *base;
  size_t used = 0;
  template <typename P>
  P *take(size_t bytes) {
    used = (used + 255) / 256 * 256;
    P *p = reinterpret_cast<P *>(base + used);
    used += bytes;
    return p;
  }
};

}  // namespace

extern "C" int deft_functional_grad_noalloc(int nx, int ny, int nz, const double cell[9], double R, double kT,
                                            double mu, int fft_precision, int pgrad_kind, const double *veff_dev,
                                            const double *vext_dev, double *outgrad_dev, double *outpgrad_dev,
                                            double *energy_dev, void *stream) {
  if (!cell || !veff_dev || !outgrad_dev) return set_error(DEFT_ERR_ARG, "null argument");
  if (!deft_fused_length(nx) || !deft_fused_length(ny) || !deft_fused_length(nz))
    return set_error(DEFT_ERR_UNSUPPORTED, "grid edges must have fused kernels (deft_fused_length)");
  const long long N = (long long)nx * ny * nz;
  if (N > kMaxPoints)
    return set_error(DEFT_ERR_ARG, "deft_functional_grad_noalloc takes grids of at most 64^3 = 262144 points");
  double det;
  if (!det3(cell, det)) return set_error(DEFT_ERR_ARG, "cell matrix is singular or not finite");
  if (!(R > 0) || !(kT > 0) || !std::isfinite(R) || !std::isfinite(kT) || !std::isfinite(mu))
    return set_error(DEFT_ERR_ARG, "R and kT must be positive and finite, mu finite");
  if (fft_precision != DEFT_FFT_FP32 && fft_precision != DEFT_FFT_FP64)
    return set_error(DEFT_ERR_ARG, "unknown FFT precision");
  if (pgrad_kind != DEFT_PGRAD_NEW && pgrad_kind != DEFT_PGRAD_EFFECTIVE_POTENTIAL)
    return set_error(DEFT_ERR_ARG, "unknown pgrad kind");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(DEFT_ERR_HIP, "no HIP device");
  unsigned char *base = nullptr;
  {
    const int rc = scratch_for_device(dev, &base);
    if (rc != DEFT_OK) return rc;
  }

  // a plan-shaped view of the grid whose buffers point into the scratch array
  // (nothing of it is allocated)
  deft_plan_s v;
  v.Nx = nx;
  v.Ny = ny;
  v.Nz = nz;
  v.Nkz = nz / 2 + 1;
  v.N = N;
  v.Nk = (long long)nx * ny * v.Nkz;
  v.ldk = padded_row(v.Nkz);
  v.ks = (long long)nx * ny * v.ldk;
  for (int i = 0; i < 9; ++i) v.cell[i] = cell[i];
  reciprocal(cell, det, v.recip);
  v.volume = std::fabs(det);
  v.dV = v.volume / (double)N;
  v.R = R;
  v.kT = kT;
  v.mu = mu;
  v.prec = fft_precision;
  v.device = dev;
  v.stream = static_cast<hipStream_t>(stream);
  v.eblocks = (int)std::min<long long>(1024, (N + kThreads - 1) / kThreads);
  v.pstride = std::max(std::max(zblocks_for(&v, kZBlocks[0]), zblocks_for(&v, kZBlocks[1])), v.eblocks);
  const bool f64 = fft_precision == DEFT_FFT_FP64;
  Carve c{base};
  const size_t cxb = f64 ? sizeof(cx<double>) : sizeof(cx<float>);
  v.spec = c.take<void>(6 * cxb * (size_t)v.ks);
  v.g = c.take<double>(sizeof(double) * (size_t)N);
  v.pg = c.take<double>(sizeof(double) * (size_t)N);
  if (f64) {
    v.wtab = c.take<double>(2 * sizeof(double) * (size_t)v.ks);
    v.vext64 = vext_dev ? const_cast<double *>(vext_dev) : c.take<double>(sizeof(double) * (size_t)N);
  } else {
    v.wtab32 = c.take<float>(2 * sizeof(float) * (size_t)v.ks);
    v.vext = c.take<float>(sizeof(float) * (size_t)N);
  }
  v.partial = c.take<double>(sizeof(double) * NUM_SLOTS * (size_t)v.pstride);
  double *sums = c.take<double>(sizeof(double) * 16);
  const int ntw = nx + ny + nz;
  cx<double> *twd = c.take<cx<double>>(sizeof(cx<double>) * ntw);
  cx<float> *twf = c.take<cx<float>>(sizeof(cx<float>) * ntw);
  if (c.used > kScratchBytes) return set_error(DEFT_ERR_ARG, "grid needs more scratch than the library holds");
  v.twd[0] = twd;
  v.twd[1] = twd + nx;
  v.twd[2] = twd + nx + ny;
  v.twf[0] = twf;
  v.twf[1] = twf + nx;
  v.twf[2] = twf + nx + ny;
  // the evaluation reads Veff and never moves it (no pending moves)
  v.x = const_cast<double *>(veff_dev);

  const dim3 eb(v.eblocks), tb(kThreads);
  hipLaunchKernelGGL(k_noalloc_twiddles, dim3((ntw + kThreads - 1) / kThreads), tb, 0, v.stream, nx, ny, nz, twd,
                     twf);
  if (f64) {
    hipLaunchKernelGGL(k_weight_table, eb, tb, 0, v.stream, grid_args(&v), v.wtab, nullptr);
    if (!vext_dev) hipLaunchKernelGGL(k_noalloc_vext, eb, tb, 0, v.stream, N, nullptr, nullptr, v.vext64);
  } else {
    hipLaunchKernelGGL(k_weight_table, eb, tb, 0, v.stream, grid_args(&v), nullptr, v.wtab32);
    hipLaunchKernelGGL(k_noalloc_vext, eb, tb, 0, v.stream, N, vext_dev, v.vext, nullptr);
  }
  {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(DEFT_ERR_HIP, std::string("set-up launch: ") + hipGetErrorString(e));
  }
  int nzb = 0;
  const int rc = f64 ? Launch<double>::enqueue_evaluation(&v, true, false, pgrad_kind, &nzb)
                     : Launch<float>::enqueue_evaluation(&v, true, false, pgrad_kind, &nzb);
  if (rc != DEFT_OK) return rc;
  hipLaunchKernelGGL(k_reduce, dim3(8), tb, 0, v.stream, v.partial, v.pstride, (int)SLOT_IDEAL, nzb, sums);
  if (energy_dev) hipLaunchKernelGGL(k_noalloc_energy, dim3(1), dim3(64), 0, v.stream, sums, kT, v.dV, energy_dev);
  hipLaunchKernelGGL(k_accumulate, eb, tb, 0, v.stream, N, outgrad_dev, v.g);
  if (outpgrad_dev) hipLaunchKernelGGL(k_accumulate, eb, tb, 0, v.stream, N, outpgrad_dev, v.pg);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DEFT_ERR_HIP, std::string("deft_functional_grad_noalloc: ") + hipGetErrorString(e));
  return DEFT_OK;
}
