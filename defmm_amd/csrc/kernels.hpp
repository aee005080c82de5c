// Device kernels of the FMM evaluation pass (gfx950).  Every kernel works on
// flattened lists built by plan.cpp; the real type R is double (DEFMM_FP64)
// or float (DEFMM_MIXED: storage and arithmetic in float, phase arguments
// reduced in double).  Complex values are stored interleaved (R2 = double2 or
// float2).  Operator formulas follow the reference operators cited per kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sincos_red.hpp"

namespace dfm {

constexpr int LMAX = 10;  // largest interpolation order supported on device

template <typename R> struct C2T;
template <> struct C2T<double> { typedef double2 T; };
template <> struct C2T<float> { typedef float2 T; };

template <typename R>
__device__ __forceinline__ typename C2T<R>::T c2(R a, R b) {
  typename C2T<R>::T r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename A, typename B>
__device__ __forceinline__ A cmul(A a, B b) {
  A r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
// a * conj(b)
template <typename A, typename B>
__device__ __forceinline__ A cmulc(A a, B b) {
  A r;
  r.x = a.x * b.x + a.y * b.y;
  r.y = a.y * b.x - a.x * b.y;
  return r;
}

// acc += u * f with four explicit fused multiply-adds in a fixed order, so
// that every kernel forming the same sums rounds them the same way
// (whatever contraction the compiler would choose in each context).
template <typename A>
__device__ __forceinline__ void cmacFma(A& acc, const A u, const A f) {
  acc.x = fma(u.x, f.x, acc.x);
  acc.x = fma(-u.y, f.y, acc.x);
  acc.y = fma(u.x, f.y, acc.y);
  acc.y = fma(u.y, f.x, acc.y);
}

// exp(i*arg).  double: sincosRed (sincos_red.hpp).  float: the argument is
// reduced in double first (sincosRedF; beyond 2^20 rad modulo 2*pi), so that
// a float evaluation keeps an absolute phase error near 1e-7 for arguments of
// hundreds of radians.
__device__ __forceinline__ void expi(double arg, double& c, double& s) { sincosRed(arg, s, c); }
__device__ __forceinline__ void expi(double arg, float& c, float& s) {
  if (fabs(arg) <= 1048576.0) {
    sincosRedF(arg, s, c);      // reduction by pi/2 in double, float polynomials
    return;
  }
  const double inv2pi = 0.15915494309189533577;
  double t = arg * inv2pi;
  t -= rint(t);                 // t in [-1/2, 1/2] turns
  sincospif((float)(2.0 * t), &s, &c);
}

// Constants of an interpolation order, in the plan's real type:
//   xi[L]          grid points -1 + 2 i/(L-1)
//   iden[L]        1 / prod_{j != i} (xi_i - xi_j)
//   A0, A1[L*L]    M2M 1D matrices [parent*L + child], offset -, +
//   B0, B1[L*L]    L2L 1D matrices [child*L + parent], offset -, +
//   tw[2*P]        cos, sin of 2*pi*m/P (m = 0..P-1)
template <typename R>
struct OpConsts {
  const R* xi;
  const R* iden;
  const R* A[2];
  const R* B[2];
  const R* twc;
  const R* tws;
};

// Accumulator of V_MFMA_F32_16X16X4_F32 (4 floats per lane).
typedef float mfma_v4f __attribute__((ext_vector_type(4)));

// Slots of a source leaf accumulated per P2M pass (they share the basis
// products and the leaf's staged points; fewer from L = 9 for registers).
constexpr int p2mSlotsPerPass(int L) { return L >= 9 ? 2 : 4; }

// Lagrange weights S_i(t), t in [-1, 1], i < L.
template <typename R>
__device__ __forceinline__ void lagrangeAll(R t, int L, const R* xi, const R* iden, R* out) {
  for (int i = 0; i < LMAX; i++) {
    if (i >= L) break;
    R p = iden[i];
    for (int j = 0; j < LMAX; j++) {
      if (j >= L) break;
      if (j != i) p *= (t - xi[j]);
    }
    out[i] = p;
  }
}

// Cell tables of one tree (device copies of plan.hpp Tree).
struct DevCells {
  const int* lvl;
  const int* pbeg;
  const int* nprt;
  const double* ctr;  // 3 per cell
  const double* rad;
  const int* slot0;
  const int* nslot;
  const int* slotDir;
};

// Level tables.
struct DevLevels {
  const int* isHF;
  const int* dirOff;   // first direction of the level's stage in dirs
  const int* sigRow0;  // first phase-table row of the level (-1: none)
};

}  // namespace dfm
