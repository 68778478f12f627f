// Host build of the HIP path's pointwise White-Bear code (deft_amd/csrc/wb_point.hpp):
// reads "functional n3 n2 vx vy vz" lines on stdin, prints p1 p2 p3 d0..d4 and the
// energy-only p1 p2 p3 with 17 significant digits.
#include <cstdio>

#include "../../deft_amd/csrc/wb_point.hpp"

int main() {
  int fn;
  double n3, n2, vx, vy, vz;
  while (std::scanf("%d %lf %lf %lf %lf %lf", &fn, &n3, &n2, &vx, &vy, &vz) == 6) {
    const deft::WBConst c = deft::make_wb_const(1.0, fn);
    double p1, p2, p3, d[5], e1, e2, e3;
    deft::wb_phi_partials(n3, n2, vx, vy, vz, c, p1, p2, p3, d);
    deft::wb_phi(n3, n2, vx, vy, vz, c, e1, e2, e3);
    std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", p1, p2, p3, d[0], d[1],
                d[2], d[3], d[4], e1, e2, e3);
  }
  return 0;
}
