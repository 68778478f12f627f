// Writes radial_plot2 (include/deft_amd.hpp) of seven fields read from stdin:
//   Nx Ny Nz Lx Ly Lz diameter, then 7 fields of Nx*Ny*Nz doubles
// into the file named on the command line.
#include <cstdio>
#include <vector>

#include "deft_amd.hpp"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  int n[3];
  double L[3], diameter;
  if (std::scanf("%d %d %d %lf %lf %lf %lf", &n[0], &n[1], &n[2], &L[0], &L[1], &L[2], &diameter) != 7) return 3;
  const deft_amd::GridDescription gd(deft_amd::Lattice::orthorhombic(L[0], L[1], L[2]), n[0], n[1], n[2]);
  std::vector<std::vector<double>> f(7, std::vector<double>((size_t)gd.NxNyNz));
  for (auto &v : f)
    for (double &x : v)
      if (std::scanf("%lf", &x) != 1) return 3;
  return deft_amd::radial_plot2(argv[1], gd, diameter, f) ? 0 : 1;
}
