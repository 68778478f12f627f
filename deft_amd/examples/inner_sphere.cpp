// papers/contact/figs/inner-sphere.cpp:177-252 (run_spherical_solute) through the
// C++ host layer and the old-API minimizer seam, White-Bear functional:
//   mu from the bulk (find_chemical_potential), a periodic (diameter + 16)^3 cell
//   on GridDescription(lat, 0.1), the solute as constrain(mask, f) (mask = 0
//   within diameter/2 of the origin: Vext = NaN there), Veff started at
//   -ln(n_b mask + 1e-4 n_b), Precision(|p V| 1e-10, PreconditionedConjugateGradient(
//   f, gd, 1, &potential, QuadraticLineMinimizer)) for at most 200 calls, then the
//   density and the contact-density estimators written with radial_plot and
//   radial_plot2 into <outdir>/inner-sphereWB-<diameter>-<eta>{,-mean}.dat.
//
// usage: inner_sphere diameter eta outdir
// Prints one JSON line: energy, improve_energy calls, particle number.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "deft_amd.hpp"

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s diameter eta outdir\n", argv[0]);
    return 2;
  }
  const double diameter = std::atof(argv[1]), eta = std::atof(argv[2]);
  const std::string outdir = argv[3];
  try {
    using namespace deft_amd;
    const double padding = 16;
    const double meandensity = eta / (4 * M_PI / 3);
    const double mu = find_chemical_potential(meandensity, 1.0, 1.0);
    const double xmax = diameter + padding;
    const GridDescription gd(Lattice::orthorhombic(xmax, xmax, xmax), 0.1);
    WhiteBearFluid f(gd, 1.0, 1.0, mu, DEFT_FFT_FP64);
    // constraint.Set(notinwall) at Wigner-Seitz positions (src/Grid.cpp:43-52)
    std::vector<double> vext((size_t)gd.NxNyNz), potential((size_t)gd.NxNyNz);
    for (int x = 0; x < gd.Nx; ++x)
      for (int y = 0; y < gd.Ny; ++y)
        for (int z = 0; z < gd.Nz; ++z) {
          double r[3];
          gd.position(x, y, z, r);
          double d2 = 0;
          for (int i = 0; i < 3; ++i) {
            r[i] -= xmax * std::floor(r[i] / xmax + 0.5);
            d2 += r[i] * r[i];
          }
          const double constraint = std::sqrt(d2) > diameter / 2 ? 1.0 : 0.0;
          const size_t i = (size_t)gd.index(x, y, z);
          vext[i] = constraint == 0 ? NAN : 0.0;  // constrain(constraint, f)
          potential[i] = -std::log(meandensity * constraint + 1e-4 * meandensity);
        }
    f.set_Vext(vext);
    // approx_energy = (fhs + IdealGas() + ChemicalPotential(mu))(1, n_b) xmax^3 = -p V
    const double approx_energy = -pressure(meandensity, 1.0, 1.0) * xmax * xmax * xmax;
    const double precision = std::fabs(approx_energy * 1e-10);
    int calls = 0;
    double energy;
    {
      Minimizer min = Precision(precision, PreconditionedConjugateGradient(f, gd, 1, &potential, QuadraticLineMinimizer));
      for (int i = 0; min.improve_energy(false) && i < 200; i++) calls = i + 1;
      energy = min.energy();
    }
    const std::vector<std::vector<double>> fields = f.contact_fields();
    double Ntot = 0;
    for (double v : fields[0]) Ntot += v;
    Ntot *= gd.dvolume;
    char name[1024];
    std::snprintf(name, sizeof name, "%s/inner-sphereWB-%04.1f-%04.2f.dat", outdir.c_str(), diameter, eta);
    radial_plot(name, gd, fields);
    std::snprintf(name, sizeof name, "%s/inner-sphereWB-%04.1f-%04.2f-mean.dat", outdir.c_str(), diameter, eta);
    radial_plot2(name, gd, diameter, fields);
    std::printf("{\"energy\": %.17g, \"calls\": %d, \"N\": %.17g, \"grid\": [%d, %d, %d]}\n", energy, calls, Ntot,
                gd.Nx, gd.Ny, gd.Nz);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
