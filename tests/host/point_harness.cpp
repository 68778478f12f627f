// Host build of the HIP path's pointwise White-Bear code (deft_amd/csrc/wb_point.hpp):
// reads lines "functional v..." on stdin (5 weighted densities n3 n2 vx vy vz, or 11
// for the tensor functional 3: ... Txx Tyy Tzz Txy Tyz Tzx) and prints p1 p2 p3, the
// partials and the energy-only p1 p2 p3 with 17 significant digits.  A line "-1 x"
// prints log1m(x) = log(1 - x) instead.
#include <cstdio>

#include "../../deft_amd/csrc/wb_point.hpp"

int main() {
  int fn;
  while (std::scanf("%d", &fn) == 1) {
    if (fn == -1) {
      double x;
      if (std::scanf("%lf", &x) != 1) return 1;
      std::printf("%.17g\n", deft::log1m(x));
      continue;
    }
    const int nw = fn == 3 ? 11 : 5;
    double v[11];
    for (int i = 0; i < nw; ++i)
      if (std::scanf("%lf", &v[i]) != 1) return 1;
    const deft::WBConst c = deft::make_wb_const(1.0, fn);
    double p1, p2, p3, d[11], e1, e2, e3;
    if (nw == 11) {
      deft::wbt_phi_partials(v, c, p1, p2, p3, d);
      deft::wbt_phi(v, c, e1, e2, e3);
    } else {
      deft::wb_phi_partials(v[0], v[1], v[2], v[3], v[4], c, p1, p2, p3, d);
      deft::wb_phi(v[0], v[1], v[2], v[3], v[4], c, e1, e2, e3);
    }
    std::printf("%.17g %.17g %.17g", p1, p2, p3);
    for (int i = 0; i < nw; ++i) std::printf(" %.17g", d[i]);
    std::printf(" %.17g %.17g %.17g\n", e1, e2, e3);
  }
  return 0;
}
