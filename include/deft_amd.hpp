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
// The functional is OfEffectivePotential(HardSpheresWBnotensor(R) + IdealGas +
// ChemicalPotential(mu)) with an external potential (src/HardSpheres.cpp:137-159,
// src/EffectivePotentialToDensity.cpp:28-90); x is Veff, stored on the GPU.
// Vectors are std::vector<double> (the reference uses Eigen VectorXd).  Unlike
// the reference, failures throw deft_amd::Error (the reference has no error
// channel besides NaN values and asserts).
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
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
// the reference's rule N = 1 + int(e/100 + |a|/spacing), rounded up to the next
// edge the FFT kernels support (the reference's FFTW takes any N).
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

  static int edge(double a, double spacing) {
    int n = 1 + int(std::exp(1.0) / 100 + a / spacing);
    while (!deft_supported_length(n) && n < 4096) ++n;
    return n;
  }
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
  // Functional::integralgrad(kT, gd, x, g, pg): g += dF/dVeff, pg += preconditioned
  void integralgrad(double, const GridDescription &, const std::vector<double> &x, std::vector<double> *g,
                    std::vector<double> *pg = nullptr) {
    size_check(x);
    if (!g) throw Error(DEFT_ERR_ARG, "deft_amd: null gradient");
    size_check(*g);
    if (pg) size_check(*pg);
    check(deft_functional_grad(plan_, x.data(), g->data(), pg ? pg->data() : nullptr, nullptr));
  }
  // per-term energies of the last evaluation (print_summary, src/Functional.cpp:281-316)
  const deft_energy_terms &last_terms() const { return terms_; }

 private:
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

}  // namespace deft_amd
