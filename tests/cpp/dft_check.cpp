// The 15-point DFT of M2F and F2L at order 8 (defmm_amd/csrc/dft.hpp),
// compiled for the CPU: forward and inverse, all 15 inputs and the pruned
// 8-input form, in double and in float, against the direct sum in long
// double on random inputs and on unit impulses.  Prints the largest error
// relative to the largest output magnitude per case.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../defmm_amd/csrc/dft.hpp"

struct D2 { double x, y; };
struct F2 { float x, y; };

template <typename T2, int S, int NIN>
static double check(std::mt19937_64& g, int trials) {
  typedef decltype(T2::x) R;
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  const long double pi = 3.141592653589793238462643383279502884L;
  double worst = 0;
  for (int t = 0; t < trials + NIN; t++) {
    T2 x[15];
    long double xr[15], xi[15];
    for (int n = 0; n < 15; n++) {
      double a = 0, b = 0;
      if (n < NIN) {
        if (t < NIN) { a = (n == t); }       // impulses first
        else { a = u(g); b = u(g); }
      }
      x[n].x = (R)a; x[n].y = (R)b;
      xr[n] = (long double)x[n].x; xi[n] = (long double)x[n].y;
      if (n >= NIN) { x[n].x = (R)NAN; x[n].y = (R)NAN; }  // must not be read
    }
    dfm::dft15<S, NIN>(x);
    long double big = 0, err = 0;
    for (int k = 0; k < 15; k++) {
      long double sr = 0, si = 0;
      for (int n = 0; n < NIN; n++) {
        const long double ang = S * 2 * pi * (long double)((n * k) % 15) / 15;
        sr += xr[n] * cosl(ang) - xi[n] * sinl(ang);
        si += xr[n] * sinl(ang) + xi[n] * cosl(ang);
      }
      big = fmaxl(big, sqrtl(sr * sr + si * si));
      const long double dr = (long double)x[k].x - sr, di = (long double)x[k].y - si;
      err = fmaxl(err, sqrtl(dr * dr + di * di));
    }
    if (!(err == err)) return 1e300;  // NaN: a zero input was read
    worst = std::fmax(worst, (double)(err / big));
  }
  return worst;
}

int main() {
  std::mt19937_64 g(2024);
  std::printf("f64_fwd15 %.3e\n", check<D2, -1, 15>(g, 20000));
  std::printf("f64_inv15 %.3e\n", check<D2, +1, 15>(g, 20000));
  std::printf("f64_fwd8 %.3e\n", check<D2, -1, 8>(g, 20000));
  std::printf("f64_inv8 %.3e\n", check<D2, +1, 8>(g, 20000));
  std::printf("f32_fwd15 %.3e\n", check<F2, -1, 15>(g, 20000));
  std::printf("f32_inv15 %.3e\n", check<F2, +1, 15>(g, 20000));
  std::printf("f32_fwd8 %.3e\n", check<F2, -1, 8>(g, 20000));
  std::printf("f32_inv8 %.3e\n", check<F2, +1, 8>(g, 20000));
  return 0;
}
