// Hard-sphere fluid around a spherical hard solute, driven through the C++
// host layer the way deft's papers/contact/figs/inner-sphere.cpp:177-211
// drives the FFTW path: bulk mu from the equation of state, a hard
// constraint for the solute, Veff started at the bulk value, then
// improve_energy until the minimizer stops.
//
// usage: solute N L radius eta [mixed]
//   N      grid edge (cubic cell of edge L), radius: excluded-centre radius,
//   eta    bulk packing fraction, mixed: 1 = FP32 -> FP64 precision ladder
// Prints one JSON line: energy, iterations, particle number, density along x.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "deft_amd.hpp"

int main(int argc, char **argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s N L radius eta [mixed]\n", argv[0]);
    return 2;
  }
  const int N = std::atoi(argv[1]);
  const double L = std::atof(argv[2]), radius = std::atof(argv[3]), eta = std::atof(argv[4]);
  const bool mixed = argc > 5 && std::atoi(argv[5]) != 0;
  try {
    using namespace deft_amd;
    const double R = 1.0, kT = 1.0;
    const double nb = eta / (4 * M_PI / 3 * R * R * R);
    const double mu = find_chemical_potential(nb, R, kT);
    GridDescription gd(Lattice::orthorhombic(L, L, L), N, N, N);
    WhiteBearFluid f(gd, R, kT, mu, mixed ? DEFT_FFT_FP32 : DEFT_FFT_FP64);
    std::vector<double> vext((size_t)gd.NxNyNz), veff((size_t)gd.NxNyNz, -kT * std::log(nb));
    for (int x = 0; x < N; ++x)
      for (int y = 0; y < N; ++y)
        for (int z = 0; z < N; ++z) {
          double r[3];
          gd.position(x, y, z, r);
          double d2 = 0;
          for (int i = 0; i < 3; ++i) {
            if (r[i] >= L / 2) r[i] -= L;  // nearest image of the solute at the origin
            d2 += r[i] * r[i];
          }
          vext[((size_t)x * N + y) * N + z] = std::sqrt(d2) < radius ? INFINITY : 0.0;
        }
    f.set_Vext(vext);
    f.set_Veff(veff);
    Minimize min(&f);
    min.precondition(true);
    min.set_relative_precision(1e-11);
    if (mixed) min.set_mixed_precision(1e-6);
    while (min.improve_energy(false)) {
    }
    const double energy = min.energy();
    const std::vector<double> n = f.density();
    double Ntot = 0;
    for (double v : n) Ntot += v;
    Ntot *= gd.dvolume;
    std::printf("{\"energy\": %.17g, \"iterations\": %d, \"N\": %.17g, \"density_x\": [", energy,
                min.get_iteration_count(), Ntot);
    for (int x = 0; x < N; ++x) std::printf("%s%.17g", x ? ", " : "", n[(size_t)x * N * N]);
    std::printf("]}\n");
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
