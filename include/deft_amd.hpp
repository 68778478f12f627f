// deft_amd.hpp: C++ host layer over the C ABI of include/deft_amd.h.
//
// Keeps the calling shape of deft's FFTW path so a driver written against it
// (papers/contact/figs/inner-sphere.cpp:177-211, papers/fuzzy-fmt/figs/new-bh-walls.cpp:112-155)
// ports line by line:
//   deft                                          deft_amd
//   Lattice, GridDescription(lat, spacing|nx..)   Lattice, GridDescription   (src/lattice.h:79-112,
//                                                                             src/GridDescription.cpp:6-50)
//   Functional::integral(gd, kT, x)               WhiteBearFluid::integral    (src/Functional.h:125-143)
//   Functional::integralgrad(kT, gd, x, g, pg)    WhiteBearFluid::integralgrad (src/Functional.h:144-149)
//   find_chemical_potential(f, kT, n), pressure   find_chemical_potential, pressure
//                                                                            (src/equation-of-state.cpp:164-192)
//   Minimize(&f).improve_energy(v)                Minimize(&f).improve_energy(v)
//                                                                            (src/new/Minimize.h:13-199)
//   Precision(p, PreconditionedConjugateGradient(f, gd, kT, &x, QuadraticLineMinimizer))
//                                                 the same call               (src/Minimizer.h:16-81,
//                                                                             src/ConjugateGradient.cpp:59-171,
//                                                                             src/Precision.cpp:23-156)
//   radial_plot, radial_plot2, z_plot             the same calls             (papers/contact/figs/
//                                                                             inner-sphere.cpp:78-172,
//                                                                             walls.cpp:55-80)
// The functional is OfEffectivePotential(HardSpheresWBnotensor(R) + IdealGas +
// ChemicalPotential(mu)) with an external potential (src/HardSpheres.cpp:137-159,
// src/EffectivePotentialToDensity.cpp:28-90); x is Veff, stored on the GPU.
// Vectors are std::vector<double> (the reference uses Eigen VectorXd).  Unlike
// the reference, failures throw deft_amd::Error (the reference has no error
// channel besides NaN values and asserts).
#pragma once

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "deft_amd.h"

namespace deft_amd {

struct Error : std::runtime_error {
  int status;
  Error(int st, const std::string &what) : std::runtime_error(what), status(st) {}
};

inline void check(int st) {
  if (st != DEFT_OK) throw Error(st, std::string("deft_amd: ") + deft_last_error());
}

// Lattice vectors a1, a2, a3 (src/lattice.h:79-112).
struct Lattice {
  double a[9];
  Lattice(const double a1[3], const double a2[3], const double a3[3]) {
    for (int i = 0; i < 3; ++i) {
      a[i] = a1[i];
      a[3 + i] = a2[i];
      a[6 + i] = a3[i];
    }
  }
  static Lattice orthorhombic(double lx, double ly, double lz) {
    const double a1[3] = {lx, 0, 0}, a2[3] = {0, ly, 0}, a3[3] = {0, 0, lz};
    return Lattice(a1, a2, a3);
  }
  double length(int i) const {
    return std::sqrt(a[3 * i] * a[3 * i] + a[3 * i + 1] * a[3 * i + 1] + a[3 * i + 2] * a[3 * i + 2]);
  }
  double volume() const {
    return std::fabs(a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]));
  }
};

// GridDescription (src/GridDescription.cpp:6-50).  The spacing constructor uses
// the reference's rule N = 1 + int(e/100 + |a|/spacing) (:26-30) and throws
// (DEFT_ERR_UNSUPPORTED) when that N is not an edge the FFT passes support
// (above 4096, or a prime factor above 512), since a different N would change
// the spacing and dV; at_least() rounds up to the next supported edge for
// callers that accept that.
struct GridDescription {
  Lattice Lat;
  int Nx, Ny, Nz;
  long long NxNyNz;
  double dvolume;

  GridDescription(const Lattice &lat, int nx, int ny, int nz) : Lat(lat), Nx(nx), Ny(ny), Nz(nz) {
    NxNyNz = (long long)nx * ny * nz;
    dvolume = lat.volume() / (double)NxNyNz;
  }
  GridDescription(const Lattice &lat, double spacing)
      : GridDescription(lat, edge(lat.length(0), spacing), edge(lat.length(1), spacing), edge(lat.length(2), spacing)) {}

  // the reference's edge for this lattice vector length and spacing
  static int reference_edge(double a, double spacing) { return 1 + int(std::exp(1.0) / 100 + a / spacing); }
  static int edge(double a, double spacing) {
    const int n = reference_edge(a, spacing);
    if (!deft_supported_length(n))
      throw Error(DEFT_ERR_UNSUPPORTED, "deft_amd: GridDescription(lat, " + std::to_string(spacing) + ") gives N = " +
                                            std::to_string(n) +
                                            ", which the FFT kernels do not support (GridDescription::at_least "
                                            "rounds up to the next supported edge)");
    return n;
  }
  // The spacing constructor's grid with each N rounded up to the next supported edge.
  static GridDescription at_least(const Lattice &lat, double spacing) {
    int n[3];
    for (int i = 0; i < 3; ++i) {
      n[i] = reference_edge(lat.length(i), spacing);
      while (!deft_supported_length(n[i]) && n[i] < 4096) ++n[i];
      if (!deft_supported_length(n[i]))
        throw Error(DEFT_ERR_UNSUPPORTED, "deft_amd: GridDescription::at_least finds no supported edge up to 4096");
    }
    return GridDescription(lat, n[0], n[1], n[2]);
  }
  long long index(int x, int y, int z) const { return ((long long)x * Ny + y) * Nz + z; }
  // Cartesian position of grid point (x, y, z) (src/Grid.cpp:43-52 before Wigner-Seitz wrapping)
  void position(int x, int y, int z, double r[3]) const {
    for (int i = 0; i < 3; ++i)
      r[i] = x * Lat.a[i] / Nx + y * Lat.a[3 + i] / Ny + z * Lat.a[6 + i] / Nz;
  }
};

// Bulk path (src/equation-of-state.cpp:164-192) of ideal gas + White-Bear.
inline double find_chemical_potential(double n, double R = 1.0, double kT = 1.0) {
  double mu;
  check(deft_bulk_chemical_potential(n, R, kT, &mu));
  return mu;
}
inline double pressure(double n, double R = 1.0, double kT = 1.0) {
  double p;
  check(deft_bulk_pressure(n, R, kT, &p));
  return p;
}

// The White-Bear fluid of Veff, resident on the GPU.
class WhiteBearFluid {
 public:
  WhiteBearFluid(const GridDescription &gd, double R, double kT, double mu, int fft_precision = DEFT_FFT_FP32)
      : gd_(gd), kT_(kT) {
    check(deft_plan_create(gd.Nx, gd.Ny, gd.Nz, gd.Lat.a, R, kT, mu, fft_precision, &plan_));
  }
  ~WhiteBearFluid() { deft_plan_destroy(plan_); }
  WhiteBearFluid(const WhiteBearFluid &) = delete;
  WhiteBearFluid &operator=(const WhiteBearFluid &) = delete;

  deft_plan plan() const { return plan_; }
  const GridDescription &description() const { return gd_; }
  void set_mu(double mu) { check(deft_plan_set_mu(plan_, mu)); }
  void set_fft_precision(int p) { check(deft_plan_set_fft_precision(plan_, p)); }
  // DEFT_WHITE_BEAR (HardSpheresNoTensor2), DEFT_WHITE_BEAR_MARK_II (WhiteBearMarkII)
  // or DEFT_WHITE_BEAR_MARK_II_SLOW (HardSpheresWBm2slow)
  void set_functional(int f) { check(deft_plan_set_functional(plan_, f)); }
  // Vext = +inf marks hard walls (n = 0 there)
  void set_Vext(const std::vector<double> &v) {
    size_check(v);
    check(deft_plan_set_vext(plan_, v.data()));
  }
  void set_Veff(const std::vector<double> &v) {
    size_check(v);
    check(deft_plan_set_veff(plan_, v.data()));
  }
  std::vector<double> Veff() const {
    std::vector<double> v((size_t)gd_.NxNyNz);
    check(deft_plan_get_veff(plan_, v.data()));
    return v;
  }
  std::vector<double> density() const {
    std::vector<double> v((size_t)gd_.NxNyNz);
    check(deft_plan_get_density(plan_, v.data()));
    return v;
  }

  // Functional::integral(gd, kT, x): the grand free energy of Veff = x
  double integral(const GridDescription &, double, const std::vector<double> &x) {
    size_check(x);
    double e;
    check(deft_functional_energy(plan_, x.data(), &e, &terms_));
    return e;
  }
  // Functional::integralgrad(kT, gd, x, g, pg): g += dF/dVeff, pg += (dF/dn)/(Veff - kT),
  // the old API's OfEffectivePotential pgrad (src/EffectivePotentialToDensity.cpp:62-71)
  void integralgrad(double, const GridDescription &, const std::vector<double> &x, std::vector<double> *g,
                    std::vector<double> *pg = nullptr) {
    grad_of_kind(x, g, pg, DEFT_PGRAD_EFFECTIVE_POTENTIAL);
  }
  // the new API's energy_grad_and_precond: g += dF/dVeff, pg += grad/n; returns the energy
  double energy_grad_and_precond(const std::vector<double> &x, std::vector<double> *g, std::vector<double> *pg) {
    return grad_of_kind(x, g, pg, DEFT_PGRAD_NEW);
  }
  // the seven contact-paper fields of the resident density: n, n0, gSigmaS, YuWu_S,
  // nA, gSigmaA, Gross (the columns radial_plot writes, inner-sphere.cpp:220-233)
  std::vector<std::vector<double>> contact_fields() const {
    const size_t N = (size_t)gd_.NxNyNz;
    std::vector<double> all(7 * N);
    check(deft_contact_values(plan_, all.data()));
    std::vector<std::vector<double>> out;
    out.push_back(density());
    for (int k = 0; k < 6; ++k) out.emplace_back(all.begin() + k * N, all.begin() + (k + 1) * N);
    return out;
  }
  // per-term energies of the last evaluation (print_summary, src/Functional.cpp:281-316)
  const deft_energy_terms &last_terms() const { return terms_; }
  double kT() const { return kT_; }

 private:
  double grad_of_kind(const std::vector<double> &x, std::vector<double> *g, std::vector<double> *pg, int kind) {
    size_check(x);
    if (!g) throw Error(DEFT_ERR_ARG, "deft_amd: null gradient");
    size_check(*g);
    if (pg) size_check(*pg);
    double e;
    check(deft_functional_grad_kind(plan_, x.data(), g->data(), pg ? pg->data() : nullptr, &e, kind));
    return e;
  }
  void size_check(const std::vector<double> &v) const {
    if ((long long)v.size() != gd_.NxNyNz) throw Error(DEFT_ERR_ARG, "deft_amd: field size does not match the grid");
  }
  GridDescription gd_;
  double kT_;
  deft_plan plan_ = nullptr;
  deft_energy_terms terms_{};
};

// Minimize (src/new/Minimize.h:13-199) on the fluid's resident Veff.
class Minimize {
 public:
  explicit Minimize(WhiteBearFluid *f) { check(deft_minimize_create(f->plan(), &m_)); }
  ~Minimize() { deft_minimize_destroy(m_); }
  Minimize(const Minimize &) = delete;
  Minimize &operator=(const Minimize &) = delete;

  void set_precision(double p) { check(deft_minimize_set_precision(m_, p)); }
  void set_relative_precision(double p) { check(deft_minimize_set_relative_precision(m_, p)); }
  void set_maxiter(int n) { check(deft_minimize_set_maxiter(m_, n)); }
  void set_miniter(int n) { check(deft_minimize_set_miniter(m_, n)); }
  void precondition(bool u) { check(deft_minimize_precondition(m_, u ? 1 : 0)); }
  // FP32 -> FP64 precision ladder (no reference counterpart)
  void set_mixed_precision(double switch_rp, int polish_grad_prec = DEFT_FFT_FP32) {
    check(deft_minimize_set_mixed_precision(m_, switch_rp, polish_grad_prec));
  }
  bool improve_energy(bool verbose = false) {
    int more = 0;
    check(deft_minimize_improve_energy(m_, verbose ? 1 : 0, &more));
    return more != 0;
  }
  double energy() {
    double e;
    check(deft_minimize_energy(m_, &e));
    return e;
  }
  int get_iteration_count() const { return stats().iter; }
  deft_minimize_stats stats() const {
    deft_minimize_stats s;
    check(deft_minimize_stats_get(m_, &s));
    return s;
  }

 private:
  deft_minimize m_ = nullptr;
};

// ---------------------------------------------------------------------------
// old-API minimizers (MinimizerInterface, src/Minimizer.h:16-81)
// ---------------------------------------------------------------------------
// The line minimizer argument of the reference's factories; quadratic is the
// one the hot path uses (src/LineMinimizer.h:8-21).
enum LineMinimizerKind { QuadraticLineMinimizer };

// A MinimizerInterface acting on the fluid's resident Veff.  As in the
// reference the minimizer does not own x: the constructor uploads *x, and the
// minimizer writes the current Veff back into *x whenever improve_energy
// returns false, whenever energy() is asked for, and when it is destroyed (the
// reference's drivers read x after their loop; a copy per iteration would cost
// a device-to-host transfer each time).
class Minimizer {
 public:
  Minimizer(WhiteBearFluid &f, double kT, std::vector<double> *x, bool preconditioned, double stepsize)
      : f_(&f), x_(x) {
    if (std::fabs(kT - f.kT()) > 1e-15 * std::fabs(kT))
      throw Error(DEFT_ERR_ARG, "deft_amd: the minimizer's kT differs from the fluid's");
    if (x) f.set_Veff(*x);
    check(deft_cg_create(f.plan(), preconditioned ? 1 : 0, stepsize, &m_));
  }
  Minimizer(Minimizer &&o) noexcept : m_(o.m_), f_(o.f_), x_(o.x_) { o.m_ = nullptr; }
  Minimizer &operator=(Minimizer &&o) noexcept {
    std::swap(m_, o.m_);
    std::swap(f_, o.f_);
    std::swap(x_, o.x_);
    return *this;
  }
  Minimizer(const Minimizer &) = delete;
  Minimizer &operator=(const Minimizer &) = delete;
  ~Minimizer() {
    if (!m_) return;
    try {
      sync();
    } catch (...) {
    }
    deft_cg_destroy(m_);
  }

  // MinimizerInterface::improve_energy: false once the outermost wrapper stops
  bool improve_energy(bool verbose = false) {
    int more = 0;
    check(deft_cg_improve_energy(m_, verbose ? 1 : 0, &more));
    if (!more) sync();
    return more != 0;
  }
  double energy() {
    double e;
    check(deft_cg_energy(m_, &e));
    sync();
    return e;
  }
  std::vector<double> grad() const { return vec(&deft_cg_grad); }
  std::vector<double> pgrad() const { return vec(&deft_cg_pgrad); }
  // minimize(f, gd, newx): start over (iteration count, step, direction), from *newx if given
  void minimize(std::vector<double> *newx = nullptr) {
    if (newx) {
      x_ = newx;
      f_->set_Veff(*newx);
    }
    check(deft_cg_minimize(m_));
    check(deft_cg_invalidate(m_));
  }
  void print_info(const char *prefix = "") const {
    const deft_minimize_stats s = stats();
    std::printf("%s==== Iteration %d ====\n%sEnergy = %.16g\n%sstep = %g\n", prefix, s.iter, prefix, s.energy, prefix,
                s.step);
  }
  deft_minimize_stats stats() const {
    deft_minimize_stats s;
    check(deft_cg_stats_get(m_, &s));
    return s;
  }
  // the current Veff into *x
  void sync() {
    if (x_) *x_ = f_->Veff();
  }
  void set_precision(double p) { check(deft_cg_set_precision(m_, p)); }
  void set_maxiter(int n) { check(deft_cg_set_maxiter(m_, n)); }

 private:
  std::vector<double> vec(int (*fn)(deft_cg, double *)) const {
    std::vector<double> v((size_t)f_->description().NxNyNz);
    check(fn(m_, v.data()));
    return v;
  }
  deft_cg m_ = nullptr;
  WhiteBearFluid *f_ = nullptr;
  std::vector<double> *x_ = nullptr;
};

// ConjugateGradient(f, gd, kT, &x, QuadraticLineMinimizer, stepsize) (src/ConjugateGradient.cpp:164-166)
inline Minimizer ConjugateGradient(WhiteBearFluid &f, const GridDescription &, double kT, std::vector<double> *x,
                                   LineMinimizerKind = QuadraticLineMinimizer, double stepsize = 10.0) {
  return Minimizer(f, kT, x, false, stepsize);
}
// PreconditionedConjugateGradient(f, gd, kT, &x, QuadraticLineMinimizer, stepsize) (:168-171)
inline Minimizer PreconditionedConjugateGradient(WhiteBearFluid &f, const GridDescription &, double kT,
                                                 std::vector<double> *x, LineMinimizerKind = QuadraticLineMinimizer,
                                                 double stepsize = 10.0) {
  return Minimizer(f, kT, x, true, stepsize);
}
// Precision(precision, m) (src/Precision.cpp:154-156) and MaxIter(maxiter, m) (src/Minimizer.cpp)
inline Minimizer Precision(double precision, Minimizer m) {
  m.set_precision(precision);
  return m;
}
inline Minimizer MaxIter(int maxiter, Minimizer m) {
  m.set_maxiter(maxiter);
  return m;
}

// ---------------------------------------------------------------------------
// driver output formats (papers/contact/figs/inner-sphere.cpp, walls.cpp)
// ---------------------------------------------------------------------------
namespace detail {
inline double cartesian(const GridDescription &gd, int x, int y, int z, int axis) {
  return x * gd.Lat.a[axis] / gd.Nx + y * gd.Lat.a[3 + axis] / gd.Ny + z * gd.Lat.a[6 + axis] / gd.Nz;
}
inline double at(const GridDescription &gd, const std::vector<double> &f, int x, int y, int z) {
  x %= gd.Nx;
  y %= gd.Ny;
  z %= gd.Nz;
  if (x < 0) x += gd.Nx;
  if (y < 0) y += gd.Ny;
  if (z < 0) z += gd.Nz;
  return f[(size_t)gd.index(x, y, z)];
}
// closest_value (inner-sphere.cpp:78-101) for an orthorhombic cell: the value at the
// grid point nearest to t = delta among those with (p - t).delta > 0; 0 when there is
// none (delta = 0).  Such a point lies within 2.9 spacings of t, so the window of
// grid points from floor(t/h) - 2 to floor(t/h) + 3 holds it; where that window
// reaches past the reference's scan range [-N/2, N/2) the whole range is scanned
// instead.  The order is the reference's (x, y, z ascending), so ties resolve alike.
inline double closest_value(const GridDescription &gd, const std::vector<double> &f, const double delta[3]) {
  const double h[3] = {gd.Lat.a[0] / gd.Nx, gd.Lat.a[4] / gd.Ny, gd.Lat.a[8] / gd.Nz};
  const int N[3] = {gd.Nx, gd.Ny, gd.Nz};
  int lo[3], hi[3];
  bool clamped = false;
  for (int i = 0; i < 3; ++i) {
    const int c = (int)std::floor(delta[i] / h[i]);
    lo[i] = c - 2;
    hi[i] = c + 3;
    if (lo[i] < -N[i] / 2 || hi[i] > N[i] / 2 - 1) clamped = true;
  }
  if (clamped)
    for (int i = 0; i < 3; ++i) {
      lo[i] = -N[i] / 2;
      hi[i] = N[i] / 2 - 1;
    }
  double closest = 1e300, best = 0.0;
  for (int x = lo[0]; x <= hi[0]; ++x)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int z = lo[2]; z <= hi[2]; ++z) {
        const double p[3] = {x * h[0], y * h[1], z * h[2]};
        double dot = 0.0, d2 = 0.0;
        for (int i = 0; i < 3; ++i) {
          const double r = p[i] - delta[i];
          dot += r * delta[i];
          d2 += r * r;
        }
        const double dist = std::sqrt(d2);
        if (dot > 0 && dist < closest) {
          best = at(gd, f, x, y, z);
          closest = dist;
        }
      }
  return best;
}
// radial_average (inner-sphere.cpp:104-111): mean of closest_value over the
// twenty icosahedron directions at this radius
inline double radial_average(const GridDescription &gd, const std::vector<double> &f, double radius) {
  const double g = (1 + std::sqrt(5.0)) / 2;
  static const double ico[20][3] = {{1, 1, 1},      {-1, 1, 1},     {1, -1, 1},     {1, 1, -1},     {1, -1, -1},
                                    {-1, 1, -1},    {-1, -1, 1},    {-1, -1, -1},   {0, 1 / g, g},  {0, 1 / g, -g},
                                    {0, -1 / g, g}, {0, -1 / g, -g}, {g, 0, 1 / g}, {-g, 0, 1 / g}, {g, 0, -1 / g},
                                    {-g, 0, -1 / g}, {1 / g, g, 0}, {1 / g, -g, 0}, {-1 / g, g, 0}, {-1 / g, -g, 0}};
  double total = 0.0;
  for (int i = 0; i < 20; ++i) {
    const double n = std::sqrt(ico[i][0] * ico[i][0] + ico[i][1] * ico[i][1] + ico[i][2] * ico[i][2]);
    const double d[3] = {radius * ico[i][0] / n, radius * ico[i][1] / n, radius * ico[i][2] / n};
    total += closest_value(gd, f, d);
  }
  return total / 20;
}
}  // namespace detail

// radial_plot (inner-sphere.cpp:149-172): y of grid point (0, y, 0) and the seven
// fields there, for y < Ny/2, "%g" tab-separated.  Returns false if the file
// cannot be created (the reference prints a message and returns).
inline bool radial_plot(const char *fname, const GridDescription &gd, const std::vector<std::vector<double>> &f) {
  if (f.size() != 7) throw Error(DEFT_ERR_ARG, "deft_amd: radial_plot takes seven fields");
  FILE *out = std::fopen(fname, "w");
  if (!out) {
    std::fprintf(stderr, "Unable to create file %s!\n", fname);
    return false;
  }
  for (int y = 0; y < gd.Ny / 2; y++) {
    std::fprintf(out, "%g", detail::cartesian(gd, 0, y, 0, 1));
    for (const auto &v : f) std::fprintf(out, "\t%g", v[(size_t)gd.index(0, y, 0)]);
    std::fprintf(out, "\n");
  }
  std::fclose(out);
  return true;
}

// radial_plot2 (inner-sphere.cpp:114-147): the seven fields averaged over twenty
// directions at radii r = 0, dr, ... below Nx dr / 2, dr = diameter /
// floor(diameter / |a2/Ny|) -- the driver's "-mean.dat" file.  Orthorhombic cells.
inline bool radial_plot2(const char *fname, const GridDescription &gd, double diameter,
                         const std::vector<std::vector<double>> &f) {
  if (f.size() != 7) throw Error(DEFT_ERR_ARG, "deft_amd: radial_plot2 takes seven fields");
  FILE *out = std::fopen(fname, "w");
  if (!out) {
    std::fprintf(stderr, "Unable to create file %s!\n", fname);
    return false;
  }
  const double dr = diameter / std::floor(diameter / (gd.Lat.length(1) / gd.Ny));
  for (double r = 0; r < gd.Nx * dr / 2; r += dr) {
    std::fprintf(out, "%g", r);
    for (const auto &v : f) std::fprintf(out, "\t%g", detail::radial_average(gd, v, r));
    std::fprintf(out, "\n");
  }
  std::fclose(out);
  return true;
}

// z_plot (walls.cpp:55-80): z of grid point (0, 0, z) and eight fields there, z < Nz/2.
inline bool z_plot(const char *fname, const GridDescription &gd, const std::vector<std::vector<double>> &f) {
  if (f.size() != 8) throw Error(DEFT_ERR_ARG, "deft_amd: z_plot takes eight fields");
  FILE *out = std::fopen(fname, "w");
  if (!out) {
    std::fprintf(stderr, "Unable to create file %s!\n", fname);
    return false;
  }
  for (int z = 0; z < gd.Nz / 2; z++) {
    std::fprintf(out, "%g", detail::cartesian(gd, 0, 0, z, 2));
    for (const auto &v : f) std::fprintf(out, "\t%g", v[(size_t)gd.index(0, 0, z)]);
    std::fprintf(out, "\n");
  }
  std::fclose(out);
  return true;
}

}  // namespace deft_amd
