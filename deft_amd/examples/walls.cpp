// papers/contact/figs/walls.cpp:83-132 (run_walls) through the C++ host layer and
// the old-API minimizer seam, for one packing fraction and functional:
//   mu from the bulk of that functional (find_chemical_potential), a dw x dw x zmax
//   cell (dw = 0.001, zmax = width + 2 spacing = 36) on GridDescription(lat, 0.01)
//   -- 1 x 1 x 3601 points --, the wall as constrain(notinwall, f) (|z| <= spacing:
//   Vext = NaN), Veff started at -ln(eta constraint/(4 pi/3) + 1e-4 eta/(4 pi/3)),
//   Precision(|f_bulk dw^2 width| 1e-4, PreconditionedConjugateGradient(f, gd, 1,
//   &potential, QuadraticLineMinimizer)) for at most 100 calls, then the density and
//   the contact-density estimators written with z_plot into <outdir>/walls<name>-<eta>.dat
//   (density, density, gSigmaS, YuWu_S, gSigmaA, n0, Gross, nA; for WBm2 the Mark II
//   forms of gSigmaS and gSigmaA).
//
// usage: walls eta WB|WBT|WBm2 outdir
// Prints one JSON line: energy, improve_energy calls, particles per area.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "deft_amd.hpp"

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s eta WB|WBT|WBm2 outdir\n", argv[0]);
    return 2;
  }
  const double eta = std::atof(argv[1]);
  const std::string name = argv[2], outdir = argv[3];
  int functional;
  if (name == "WB")
    functional = DEFT_WHITE_BEAR;
  else if (name == "WBT")
    functional = DEFT_WHITE_BEAR_TENSOR;
  else if (name == "WBm2")
    functional = DEFT_WHITE_BEAR_MARK_II;
  else {
    std::fprintf(stderr, "functional must be WB, WBT or WBm2\n");
    return 2;
  }
  try {
    using namespace deft_amd;
    const double width = 30, dw = 0.001, spacing = 3;
    const double nb = eta / (4 * M_PI / 3);
    double fbulk, mu, p;
    check(deft_bulk_state(functional, nb, 1.0, 1.0, &fbulk, &mu, &p));
    const double zmax = width + 2 * spacing;
    const GridDescription gd(Lattice::orthorhombic(dw, dw, zmax), 0.01);
    WhiteBearFluid f(gd, 1.0, 1.0, mu, DEFT_FFT_FP64);
    f.set_functional(functional);
    std::vector<double> vext((size_t)gd.NxNyNz), potential((size_t)gd.NxNyNz);
    for (int x = 0; x < gd.Nx; ++x)
      for (int y = 0; y < gd.Ny; ++y)
        for (int z = 0; z < gd.Nz; ++z) {
          double r[3];
          gd.position(x, y, z, r);
          const double zz = r[2] - zmax * std::floor(r[2] / zmax + 0.5);  // Wigner-Seitz cell
          const double constraint = std::fabs(zz) > spacing ? 1.0 : 0.0;  // notinwall
          const size_t i = (size_t)gd.index(x, y, z);
          vext[i] = constraint == 0 ? NAN : 0.0;  // constrain(constraint, f)
          potential[i] = -std::log((eta * constraint + 1e-4 * eta) / (4 * M_PI / 3));
        }
    f.set_Vext(vext);
    // (fhs + IdealGas() + ChemicalPotential(mu))(1, n_b) dw^2 width = -p dw^2 width
    const double approx_energy = (fbulk - mu * nb) * dw * dw * width;
    const double precision = std::fabs(approx_energy * 1e-4);
    int calls = 0;
    double energy;
    {
      Minimizer min = Precision(precision, PreconditionedConjugateGradient(f, gd, 1, &potential, QuadraticLineMinimizer));
      for (int i = 0; min.improve_energy(false) && i < 100; i++) calls = i + 1;
      energy = min.energy();
    }
    // the estimators read the density only; the tensor functional's density goes
    // through a White-Bear plan, as walls.cpp applies gSigmaS to it
    std::vector<std::vector<double>> c;
    if (functional == DEFT_WHITE_BEAR_TENSOR) {
      WhiteBearFluid wb(gd, 1.0, 1.0, mu, DEFT_FFT_FP64);
      wb.set_Vext(vext);
      wb.set_Veff(potential);
      c = wb.contact_fields();
    } else {
      c = f.contact_fields();
    }
    // c: n, n0, gSigmaS, YuWu_S, nA, gSigmaA, Gross
    char fname[1024];
    std::snprintf(fname, sizeof fname, "%s/walls%s-%04.2f.dat", outdir.c_str(), name.c_str(), eta);
    z_plot(fname, gd, {c[0], c[0], c[2], c[3], c[5], c[1], c[6], c[4]});
    // particles per area over the first half of the cell (walls.cpp:135-147)
    const double sep = zmax / gd.Nz;
    double per_area = 0;
    for (int j = 0; j < gd.Nz / 2; ++j) per_area += c[0][(size_t)j] * sep;
    std::printf("{\"energy\": %.17g, \"calls\": %d, \"per_area\": %.17g, \"grid\": [%d, %d, %d]}\n", energy, calls,
                per_area, gd.Nx, gd.Ny, gd.Nz);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
