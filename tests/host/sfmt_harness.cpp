// Host build of the soft-FMT point code (deft_amd/csrc/sfmt.hpp).  Reads
// commands on stdin and prints values with 17 significant digits:
//   c sigma epsilon kT xi   set the constants (prints alpha)
//   r k                     the seven radial weight functions at |k| = k
//   g u v                   exp(-v^2) Re erf(u + iv)
//   w x y                   the Faddeeva function w(x + iy): re im
//   p v0 .. v15             Phi1 Phi2 Phi3 and the sixteen partials
#include <cstdio>
#include <cstring>

#include "../../deft_amd/csrc/sfmt.hpp"

int main() {
  char cmd[8];
  deft::SFMTConst c = deft::make_sfmt_const(1.0, 1.0, 1.0, 0.1);
  while (std::scanf("%7s", cmd) == 1) {
    if (!std::strcmp(cmd, "c")) {
      double s, e, kT, xi;
      if (std::scanf("%lf %lf %lf %lf", &s, &e, &kT, &xi) != 4) return 1;
      c = deft::make_sfmt_const(s, e, kT, xi);
      std::printf("%.17g\n", c.alpha);
    } else if (!std::strcmp(cmd, "r")) {
      double k, r[7];
      if (std::scanf("%lf", &k) != 1) return 1;
      deft::sfmt_radial(k, c, r);
      for (int i = 0; i < 7; ++i) std::printf(i ? " %.17g" : "%.17g", r[i]);
      std::printf("\n");
    } else if (!std::strcmp(cmd, "g")) {
      double u, v;
      if (std::scanf("%lf %lf", &u, &v) != 2) return 1;
      std::printf("%.17g\n", deft::sfmt_erf_term(u, v, c));
    } else if (!std::strcmp(cmd, "w")) {
      double x, y, re, im;
      if (std::scanf("%lf %lf", &x, &y) != 2) return 1;
      deft::faddeeva_w(x, y, c, re, im);
      std::printf("%.17g %.17g\n", re, im);
    } else if (!std::strcmp(cmd, "p")) {
      double v[16], d[16], p1, p2, p3;
      for (int i = 0; i < 16; ++i)
        if (std::scanf("%lf", &v[i]) != 1) return 1;
      deft::sfmt_phi(v, p1, p2, p3);
      deft::sfmt_phi_partials(v, d);
      std::printf("%.17g %.17g %.17g", p1, p2, p3);
      for (int i = 0; i < 16; ++i) std::printf(" %.17g", d[i]);
      std::printf("\n");
    } else {
      return 1;
    }
  }
  return 0;
}
