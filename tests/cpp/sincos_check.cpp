// Accuracy of dfm::sincosRed (defmm_amd/csrc/sincos_red.hpp) against long
// double sinl / cosl: prints the largest absolute errors over the ranges the
// kernels use (phases of up to a million radians), multiples of pi/2 and
// tiny arguments.
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <random>

#include "../../defmm_amd/csrc/sincos_red.hpp"

int main() {
  std::mt19937_64 g(12345);
  double worst_s = 0, worst_c = 0, worst_x = 0;
  auto check = [&](double x) {
    double s, c;
    dfm::sincosRed(x, s, c);
    long double xs = sinl((long double)x), xc = cosl((long double)x);
    double es = (double)fabsl((long double)s - xs), ec = (double)fabsl((long double)c - xc);
    if (es > worst_s) { worst_s = es; worst_x = x; }
    if (ec > worst_c) { worst_c = ec; worst_x = x; }
  };
  for (double R : {1.0, 10.0, 1000.0, 1.0e5, 1.0e6}) {
    std::uniform_real_distribution<double> u(-R, R);
    for (int i = 0; i < 2000000; i++) check(u(g));
  }
  for (int k = -200000; k <= 200000; k++) {
    double x = k * 1.5707963267948966;
    check(x);
    check(std::nextafter(x, 1e300));
    check(std::nextafter(x, -1e300));
  }
  for (double x = 1e-300; x < 1.0; x *= 1.7) { check(x); check(-x); }
  check(0.0);
  check(2.0e6);  // library branch
  printf("max_abs_err_sin %.3e max_abs_err_cos %.3e at %.17g\n", worst_s, worst_c, worst_x);
  // float variant (arguments up to 2^20)
  double wf = 0;
  std::uniform_real_distribution<double> uf(-1.0e6, 1.0e6);
  for (int i = 0; i < 4000000; i++) {
    double x = i < 2000000 ? uf(g) : (i - 3000000) * 1e-3;
    float s, c;
    dfm::sincosRedF(x, s, c);
    long double xs = sinl((long double)x), xc = cosl((long double)x);
    wf = std::max(wf, (double)std::max(fabsl((long double)s - xs), fabsl((long double)c - xc)));
  }
  printf("max_abs_err_float %.3e\n", wf);
  return 0;
}
