// Host backend for deft_amd/csrc/minimizer.hpp: the library's minimizer
// control logic, compiled with g++ and run on the quadratic
//   E(x) = x.A.x / 2 - b.x,  grad = A x - b,
// with pgrad = grad (plain), dinv * grad (preconditioned).  Reads from stdin
//   n, A (n x n, row-major), b, x0, dinv
// and a command line:
//   old <preconditioned> <stepsize> <precision> <maxcalls>
//   new <preconditioned> <relative_precision> <maxcalls>
//   ladder <switch_rp> <noise> <grad_check> <precision> <maxcalls>
// "ladder" runs the precision ladder with FP32 energies (E rounded to float)
// and FP64 energies carrying a deterministic roughness of relative amplitude
// `noise` (a stand-in for the round-off of FFT-based energies near convergence).
// Prints one line per iteration, "iter E step stop", then "final stop E".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../deft_amd/csrc/minimizer.hpp"

struct HostOps {
  int n = 0;
  std::vector<double> A, b, x, d, g, pg, og, dinv;
  double noise = 0.0;
  int plan_precision = DEFT_FFT_FP64;
  long long energies = 0, gradients = 0;

  double exact(const std::vector<double> &v) const {
    double e = 0.0;
    for (int i = 0; i < n; ++i) {
      double ai = 0.0;
      for (int j = 0; j < n; ++j) ai += A[(size_t)i * n + j] * v[j];
      e += 0.5 * v[i] * ai - b[i] * v[i];
    }
    return e;
  }
  int plan_prec() const { return plan_precision; }
  int energy(int prec, double *E) {
    ++energies;
    double e = exact(x);
    if (prec == DEFT_FFT_FP32) {
      e = (double)(float)e;
    } else if (noise > 0) {
      // deterministic roughness: a function of the point, amplitude noise |e|
      double h = 0.0;
      for (int i = 0; i < n; ++i) h += std::sin(1e7 * x[i] * (i + 1));
      e += noise * std::fabs(e) * h / n;
    }
    *E = e;
    return DEFT_OK;
  }
  int gradient(int prec, bool have_old, int kind, double *E, double *pgg, double *pgo) {
    ++gradients;
    for (int i = 0; i < n; ++i) {
      double s = -b[i];
      for (int j = 0; j < n; ++j) s += A[(size_t)i * n + j] * x[j];
      g[i] = prec == DEFT_FFT_FP32 ? (double)(float)s : s;
    }
    double a = 0.0, c = 0.0;
    for (int i = 0; i < n; ++i) {
      pg[i] = kind == deft::MIN_PGRAD_NONE ? g[i] : dinv[i] * g[i];
      a += pg[i] * g[i];
      if (have_old) c += pg[i] * og[i];
    }
    *pgg = a;
    *pgo = c;
    return energy(prec, E);
  }
  int cg_update(double beta, bool have_dir, double *gd) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      double di = -pg[i];
      if (have_dir) di = di + beta * d[i];
      d[i] = di;
      og[i] = g[i];
      s += g[i] * di;
    }
    *gd = s;
    return DEFT_OK;
  }
  int cg_reset(double *gd) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      d[i] = -g[i];
      s += g[i] * -g[i];
    }
    *gd = s;
    return DEFT_OK;
  }
  void move(double s) {
    for (int i = 0; i < n; ++i) x[i] = x[i] + s * d[i];
  }
  double newton_scale() const { return 1.0; }
  bool sk_available() const { return false; }
};

static void read_vec(std::vector<double> &v, size_t count) {
  v.resize(count);
  for (size_t i = 0; i < count; ++i)
    if (std::scanf("%lf", &v[i]) != 1) std::exit(3);
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  HostOps ops;
  if (std::scanf("%d", &ops.n) != 1) return 3;
  const int n = ops.n;
  read_vec(ops.A, (size_t)n * n);
  read_vec(ops.b, n);
  read_vec(ops.x, n);
  read_vec(ops.dinv, n);
  ops.d.assign(n, 0.0);
  ops.g.assign(n, 0.0);
  ops.pg.assign(n, 0.0);
  ops.og.assign(n, 0.0);
  const char *mode = argv[1];
  int more = 1, rc = DEFT_OK;
  if (!std::strcmp(mode, "old") && argc >= 6) {
    deft::OldCgState m;
    m.preconditioned = std::atoi(argv[2]) != 0;
    m.step = m.orig_step = std::atof(argv[3]);
    m.precision = std::atof(argv[4]);
    m.error_estimate = std::max(1.0, 2 * m.precision);
    const int maxcalls = std::atoi(argv[5]);
    for (int c = 0; c < maxcalls && more && rc == DEFT_OK; ++c) {
      rc = deft::old_improve_energy(m, ops, 0, &more);
      double E;
      deft::old_energy(m, ops, &E);
      std::printf("%d %.17g %.17g %d\n", m.iter, E, m.step, m.stop);
    }
    double E;
    deft::old_energy(m, ops, &E);
    std::printf("final %d %.17g %d\n", m.stop, E, m.num_energy);
  } else if (!std::strcmp(mode, "new") && argc >= 5) {
    deft::MinState m;
    m.use_precond = std::atoi(argv[2]) != 0;
    m.relative_precision = std::atof(argv[3]);
    const int maxcalls = std::atoi(argv[4]);
    for (int c = 0; c < maxcalls && more && rc == DEFT_OK; ++c) {
      rc = deft::improve_energy(m, ops, 0, &more);
      double E;
      deft::min_energy(m, ops, &E);
      std::printf("%d %.17g %.17g %d\n", m.iter, E, m.step, m.stop);
    }
    double E;
    deft::min_energy(m, ops, &E);
    std::printf("final %d %.17g %d\n", m.stop, E, m.num_energy);
  } else if (!std::strcmp(mode, "ladder") && argc >= 7) {
    deft::MinState m;
    m.use_precond = true;
    m.switch_rp = std::atof(argv[2]);
    ops.noise = std::atof(argv[3]);
    m.grad_check = std::atof(argv[4]);
    m.precision = std::atof(argv[5]);
    m.polish_grad_prec = DEFT_FFT_FP64;
    m.e_prec = m.g_prec = DEFT_FFT_FP32;
    const int maxcalls = std::atoi(argv[6]);
    for (int c = 0; c < maxcalls && more && rc == DEFT_OK; ++c) {
      rc = deft::improve_energy_mixed(m, ops, std::getenv("HARNESS_VERBOSE") ? 1 : 0, &more);
      std::printf("%d %.17g %.17g %d %d\n", m.iter, m.E, m.step, m.stop, m.phase);
    }
    std::printf("final %d %.17g %.17g %d %.17g\n", m.stop, ops.exact(ops.x), m.last_grad_error, m.grad_restarts,
                m.E);
  } else {
    return 2;
  }
  return rc == DEFT_OK ? 0 : 1;
}
