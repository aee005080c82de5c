// sincos_red.hpp -- sin and cos in double precision for the phases of the
// near field and of P2M / L2P, written from the mathematics:
//   x = n * (pi/2) + y,  n = rint(x * 2/pi),
//   y = (x - n * P1) - n * P2 with pi/2 = P1 + P2 split into two doubles and
//   both steps fused multiply-adds (Cody and Waite's additive reduction; the
//   first step is exact for |n| < 2^20: its result is a multiple of 2^-52
//   below 1 in magnitude, or a Sterbenz difference when |n| = 1),
//   sin y, cos y: Taylor series on |y| <= pi/4 truncated after y^15/15! and
//   y^16/16! (remainders below 5e-17 and 2e-18), evaluated by Horner in y^2,
//   sin x, cos x: sin y and cos y exchanged and signed by n mod 4.
// Arguments beyond 2^20 in magnitude (a million radians) use the library.
// The coefficients are 1/k! rounded to double.
#ifndef DEFMM_SINCOS_RED_HPP
#define DEFMM_SINCOS_RED_HPP

#include <cmath>

#if defined(__HIPCC__)
#define DFM_HD __host__ __device__
#else
#define DFM_HD
#endif

namespace dfm {

DFM_HD inline void sincosPoly(double y, double& sy, double& cy) {
  const double z = y * y;
  double ps = -7.647163731819816e-13;    // -1/15!
  ps = fma(ps, z, 1.6059043836821613e-10);  // 1/13!
  ps = fma(ps, z, -2.505210838544172e-08);  // -1/11!
  ps = fma(ps, z, 2.7557319223985893e-06);  // 1/9!
  ps = fma(ps, z, -0.0001984126984126984);  // -1/7!
  ps = fma(ps, z, 0.008333333333333333);    // 1/5!
  ps = fma(ps, z, -0.16666666666666666);    // -1/3!
  sy = fma(ps * z, y, y);
  double pc = 4.779477332387385e-14;     // 1/16!
  pc = fma(pc, z, -1.1470745597729725e-11);  // -1/14!
  pc = fma(pc, z, 2.08767569878681e-09);     // 1/12!
  pc = fma(pc, z, -2.755731922398589e-07);   // -1/10!
  pc = fma(pc, z, 2.48015873015873e-05);     // 1/8!
  pc = fma(pc, z, -0.001388888888888889);    // -1/6!
  pc = fma(pc, z, 0.041666666666666664);     // 1/4!
  pc = fma(pc, z, -0.5);                     // -1/2!
  cy = fma(pc, z, 1.0);
}

// s = sin(x), c = cos(x) for |x| <= 2^20 (no range check)
DFM_HD inline void sincosRedSmall(double x, double& s, double& c) {
  const double n = rint(x * 0.6366197723675814);   // 2/pi
  double y = fma(-n, 1.5707963267948966, x);        // P1 = pi/2 rounded
  y = fma(-n, 6.123233995736766e-17, y);            // P2 = pi/2 - P1 rounded
  double sy, cy;
  sincosPoly(y, sy, cy);
  const int q = (int)n & 3;
  const double a = (q & 1) ? cy : sy, b = (q & 1) ? sy : cy;
  s = (q & 2) ? -a : a;
  c = ((q + 1) & 2) ? -b : b;
}

// s = sin(x), c = cos(x)
DFM_HD inline void sincosRed(double x, double& s, double& c) {
  if (!(fabs(x) <= 1048576.0)) {  // also NaN
#if defined(__HIP_DEVICE_COMPILE__)
    sincos(x, &s, &c);
#else
    s = std::sin(x);
    c = std::cos(x);
#endif
    return;
  }
  sincosRedSmall(x, s, c);
}

// Float sine and cosine of a double argument (the mixed-precision phases):
// the same reduction by pi/2 in double, then float Taylor polynomials on
// |y| <= pi/4 truncated after y^9/9! and y^8/8! (remainders below 2e-9 and
// 3e-8, under the float rounding of 6e-8).  |x| <= 2^20.
DFM_HD inline void sincosRedF(double x, float& s, float& c) {
  const double n = rint(x * 0.6366197723675814);
  double yd = fma(-n, 1.5707963267948966, x);
  yd = fma(-n, 6.123233995736766e-17, yd);
  const float y = (float)yd, z = y * y;
  float ps = 2.7557319e-06f;               // 1/9!
  ps = fmaf(ps, z, -0.0001984127f);        // -1/7!
  ps = fmaf(ps, z, 0.008333334f);          // 1/5!
  ps = fmaf(ps, z, -0.16666667f);          // -1/3!
  const float sy = fmaf(ps * z, y, y);
  float pc = 2.4801588e-05f;               // 1/8!
  pc = fmaf(pc, z, -0.0013888889f);        // -1/6!
  pc = fmaf(pc, z, 0.041666668f);          // 1/4!
  pc = fmaf(pc, z, -0.5f);                 // -1/2!
  const float cy = fmaf(pc, z, 1.0f);
  const int q = (int)n & 3;
  const float a = (q & 1) ? cy : sy, b = (q & 1) ? sy : cy;
  s = (q & 2) ? -a : a;
  c = ((q + 1) & 2) ? -b : b;
}

}  // namespace dfm

#endif
