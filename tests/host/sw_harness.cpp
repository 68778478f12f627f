// Host build of the square-well point code (deft_amd/csrc/sw.hpp).  Commands
// on stdin, values printed with 17 significant digits:
//   x R eps lambda k            the five kernels Xi_0..Xi_4 at |k| = k
//   a R n3 n2 vx vy vz          gSigmaA's eight numerator fields
//   j R n3 n2 vx vy vz          their Jacobian d a_j/d w_m by dual numbers, m-major (5 x 8)
//   h R eps lambda kT n lognq   the uniform liquid's f and df/dn
#include <cstdio>
#include <cstring>

#include "../../deft_amd/csrc/sw.hpp"

int main() {
  char cmd[8];
  while (std::scanf("%7s", cmd) == 1) {
    if (!std::strcmp(cmd, "x")) {
      double R, e, l, k, x[5];
      if (std::scanf("%lf %lf %lf %lf", &R, &e, &l, &k) != 4) return 1;
      deft::sw_xi(k, deft::make_sw_const(R, e, l), x);
      for (int i = 0; i < 5; ++i) std::printf(i ? " %.17g" : "%.17g", x[i]);
      std::printf("\n");
    } else if (!std::strcmp(cmd, "a") || !std::strcmp(cmd, "j")) {
      double R, w[5];
      if (std::scanf("%lf %lf %lf %lf %lf %lf", &R, &w[0], &w[1], &w[2], &w[3], &w[4]) != 6) return 1;
      if (cmd[0] == 'a') {
        double a[8];
        deft::sw_afields(w[0], w[1], w[2], w[3], w[4], R, a);
        for (int i = 0; i < 8; ++i) std::printf(i ? " %.17g" : "%.17g", a[i]);
      } else {
        for (int m = 0; m < 5; ++m) {
          deft::Dual in[5], a[8];
          for (int t = 0; t < 5; ++t) in[t] = {w[t], t == m ? 1.0 : 0.0};
          deft::sw_afields(in[0], in[1], in[2], in[3], in[4], R, a);
          for (int i = 0; i < 8; ++i) std::printf((m || i) ? " %.17g" : "%.17g", a[i].d);
        }
      }
      std::printf("\n");
    } else if (!std::strcmp(cmd, "h")) {
      double R, e, l, kT, n, lq, f, d;
      if (std::scanf("%lf %lf %lf %lf %lf %lf", &R, &e, &l, &kT, &n, &lq) != 6) return 1;
      deft::sw_bulk(deft::make_sw_const(R, e, l), n, kT, lq, &f, &d);
      std::printf("%.17g %.17g\n", f, d);
    } else {
      return 1;
    }
  }
  return 0;
}
