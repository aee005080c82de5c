// dft.hpp -- short DFTs of the Fourier-domain M2L in registers: the lines of
// the 3D transforms of M2F (forward, circulant.hpp:95-106 of the reference)
// and F2L (inverse, circulant.hpp:108-118) for the transform length
// P = 2L - 1 = 15 of order 8, written from the mathematics:
//
//   y_k = sum_{n<N} x_n w^{nk},   w = exp(S * 2 pi i / N),
//   S = -1 for the forward transform, +1 for the inverse (unscaled).
//
// N = 15 = 3 * 5 by the prime-factor (Good-Thomas) algorithm, which needs no
// twiddle factors: with n = (5 n1 + 3 n2) mod 15 and k = (10 k1 + 6 k2) mod
// 15 (Chinese remaindering: k = k1 mod 3, k = k2 mod 5), n k = 5 n1 k1 +
// 3 n2 k2 (mod 15), so the 15-point DFT is five 3-point DFTs over n1
// followed by three 5-point DFTs over n2.  The 3- and 5-point DFTs use the
// symmetric sums and differences of their inputs (x1 +- x2; x1 +- x4,
// x2 +- x3) and the constants cos and sin of 2 pi / 3, 2 pi / 5 and 4 pi / 5.
// 168 real operations per 15-point line (fused multiply-adds counted once)
// against 480 for the direct sum over 8 nonzero inputs.
//
// Inputs n >= NIN are structurally zero (the M2F lines embed L = 8 values in
// 15): the 3-point stage skips them at compile time.  Outputs that a caller
// does not read are removed by the compiler.
#ifndef DEFMM_DFT_HPP
#define DEFMM_DFT_HPP

#if defined(__HIPCC__)
#define DFM_HD_INLINE __host__ __device__ __forceinline__
#define DFM_UNROLL _Pragma("unroll")
#else
#define DFM_HD_INLINE inline
#define DFM_UNROLL
#endif

namespace dfm {

// 3-point DFT in place; Zi: input i is zero (its register is not read).
template <int S, bool Z0, bool Z1, bool Z2, typename T2>
DFM_HD_INLINE void dft3(T2& x0, T2& x1, T2& x2) {
  typedef decltype(x0.x) R;
  const R h = (R)0.86602540378443864676;  // sin(2 pi / 3)
  if constexpr (Z0 && Z1 && Z2) {
    x0.x = x0.y = (R)0;
    x1 = x2 = x0;
  } else if constexpr (Z1 && Z2) {
    x1 = x0;
    x2 = x0;
    return;
  } else {
    T2 t, d;  // x1 + x2, x1 - x2
    if constexpr (Z1) {
      t = x2;
      d.x = -x2.x;
      d.y = -x2.y;
    } else if constexpr (Z2) {
      t = x1;
      d = x1;
    } else {
      t.x = x1.x + x2.x;
      t.y = x1.y + x2.y;
      d.x = x1.x - x2.x;
      d.y = x1.y - x2.y;
    }
    T2 m;  // x0 - t / 2
    if constexpr (Z0) {
      x0 = t;
      m.x = (R)-0.5 * t.x;
      m.y = (R)-0.5 * t.y;
    } else {
      m.x = x0.x - (R)0.5 * t.x;
      m.y = x0.y - (R)0.5 * t.y;
      x0.x += t.x;
      x0.y += t.y;
    }
    // y1 = m + S i h d, y2 = m - S i h d;  i d = (-d.y, d.x)
    const R sh = S > 0 ? h : -h;
    x1.x = m.x - sh * d.y;
    x1.y = m.y + sh * d.x;
    x2.x = m.x + sh * d.y;
    x2.y = m.y - sh * d.x;
  }
}

// 5-point DFT in place.
template <int S, typename T2>
DFM_HD_INLINE void dft5(T2& x0, T2& x1, T2& x2, T2& x3, T2& x4) {
  typedef decltype(x0.x) R;
  const R c1 = (R)0.30901699437494742410;   // cos(2 pi / 5)
  const R c2 = (R)-0.80901699437494742410;  // cos(4 pi / 5)
  const R s1 = (R)(S > 0 ? 0.95105651629515357212 : -0.95105651629515357212);  // S sin(2 pi / 5)
  const R s2 = (R)(S > 0 ? 0.58778525229247312917 : -0.58778525229247312917);  // S sin(4 pi / 5)
  T2 t1, t2, t3, t4;
  t1.x = x1.x + x4.x; t1.y = x1.y + x4.y;
  t2.x = x2.x + x3.x; t2.y = x2.y + x3.y;
  t3.x = x1.x - x4.x; t3.y = x1.y - x4.y;
  t4.x = x2.x - x3.x; t4.y = x2.y - x3.y;
  T2 a1, a2, b1, b2;
  a1.x = x0.x + c1 * t1.x + c2 * t2.x;
  a1.y = x0.y + c1 * t1.y + c2 * t2.y;
  a2.x = x0.x + c2 * t1.x + c1 * t2.x;
  a2.y = x0.y + c2 * t1.y + c1 * t2.y;
  // the imaginary parts before the factor i: S (s1 t3 + s2 t4), S (s2 t3 - s1 t4)
  b1.x = s1 * t3.x + s2 * t4.x;
  b1.y = s1 * t3.y + s2 * t4.y;
  b2.x = s2 * t3.x - s1 * t4.x;
  b2.y = s2 * t3.y - s1 * t4.y;
  x0.x += t1.x + t2.x;
  x0.y += t1.y + t2.y;
  // y1 = a1 + i b1, y4 = a1 - i b1, y2 = a2 + i b2, y3 = a2 - i b2
  x1.x = a1.x - b1.y; x1.y = a1.y + b1.x;
  x4.x = a1.x + b1.y; x4.y = a1.y - b1.x;
  x2.x = a2.x - b2.y; x2.y = a2.y + b2.x;
  x3.x = a2.x + b2.y; x3.y = a2.y - b2.x;
}

constexpr bool pfaZero15(int n1, int n2, int nin) { return (5 * n1 + 3 * n2) % 15 >= nin; }

// 15-point DFT in place (x[n], n >= NIN, are taken as zero and not read).
template <int S, int NIN, typename T2>
DFM_HD_INLINE void dft15(T2 (&x)[15]) {
  T2 a[3][5];
DFM_UNROLL
  for (int n1 = 0; n1 < 3; n1++)
DFM_UNROLL
    for (int n2 = 0; n2 < 5; n2++)
      if ((5 * n1 + 3 * n2) % 15 < NIN) a[n1][n2] = x[(5 * n1 + 3 * n2) % 15];
#define DFM_PFA_COL(n2)                                                                     \
  dft3<S, pfaZero15(0, n2, NIN), pfaZero15(1, n2, NIN), pfaZero15(2, n2, NIN)>(a[0][n2],    \
                                                                               a[1][n2],    \
                                                                               a[2][n2]);
  DFM_PFA_COL(0) DFM_PFA_COL(1) DFM_PFA_COL(2) DFM_PFA_COL(3) DFM_PFA_COL(4)
#undef DFM_PFA_COL
DFM_UNROLL
  for (int k1 = 0; k1 < 3; k1++) dft5<S>(a[k1][0], a[k1][1], a[k1][2], a[k1][3], a[k1][4]);
DFM_UNROLL
  for (int k1 = 0; k1 < 3; k1++)
DFM_UNROLL
    for (int k2 = 0; k2 < 5; k2++) x[(10 * k1 + 6 * k2) % 15] = a[k1][k2];
}

}  // namespace dfm

#endif
