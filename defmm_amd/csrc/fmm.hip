// MI355X (gfx950) executor of the directional FMM evaluation pass and its C
// ABI (include/defmm_amd.h).  One evaluation = gather charges into tree order,
// P2M, M2M per level, M2F, fused M2L+F2L, L2L per level, L2P, P2P near field,
// scatter -- the sequence of gemv (reference include/interface.hpp:153-170)
// with every tree walk replaced by a flat list built at plan time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/defmm_amd.h"
#include "errors.hpp"
#include "grid.hpp"
#include "dft.hpp"
#include "kernels.hpp"
#include "plan.hpp"
#include "xfer.hpp"

namespace dfm {

thread_local std::string g_err;  // declared in errors.hpp
static std::atomic<int> g_setupMode{DEFMM_SETUP_DEVICE};  // defmm_set_setup_mode

// Work-item grids (grid.hpp): lists of millions of items use a 2D grid; a
// block's item index is blockIdx.y * gridDim.x + blockIdx.x.
__device__ __forceinline__ int itemIndex() { return (int)(blockIdx.y * gridDim.x + blockIdx.x); }
static inline dim3 itemGrid(int64_t n) {
  const Grid2 g = itemGrid2(n);
  return dim3(g.x, g.y);
}

// ----------------------------------------------------------------------------
// kernels
// ----------------------------------------------------------------------------

// fill a complex array with zeros (used instead of hipMemsetAsync so that the
// sequence can be captured in a graph)
// Grid-stride over n elements with a capped grid (zeroGrid), so that no
// dispatch approaches 2^32 work-items whatever n is.
template <typename R2>
__global__ void k_zero(R2* a, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    R2 z;
    z.x = 0;
    z.y = 0;
    a[i] = z;
  }
}

// Tree::loadRHS (tree.hpp:149-154): q_tree[i] = q_in[perm[i]]
template <typename R>
__global__ void k_gather(const double2* __restrict__ qin, const int* __restrict__ perm,
                         typename C2T<R>::T* __restrict__ q, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double2 v = qin[perm[i]];
    q[i] = c2<R>((R)v.x, (R)v.y);
  }
}

// Tree::writeLHS (tree.hpp:156-160): p_out[perm[i]] = p_tree[i]
template <typename R>
__global__ void k_scatter(const typename C2T<R>::T* __restrict__ p, const int* __restrict__ perm,
                          double2* __restrict__ pout, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    typename C2T<R>::T v = p[i];
    double2 o;
    o.x = (double)v.x;
    o.y = (double)v.y;
    pout[perm[i]] = o;
  }
}

// P2M (P2M.hpp:113-136 with the leaf matrices of mapOnTrees.hpp:36-83):
//   HF leaf, slot of direction d:
//     M[I] = conj(sig2_d[I]) * sum_j S_I((y_j-c)/r) exp(-i kappa d.(y_j-c)) q_j
//   LF leaf: M[I] = sum_j S_I((y_j-c)/r) q_j
// One block per (leaf, pass of p2mSlotsPerPass(L) slots): particles are
// staged 64 at a time in LDS, and the pass's slots (directions) are
// accumulated together, sharing the basis products.  (One block per leaf,
// looping over its passes, left most of the GPU idle: 1,089 leaves with up
// to 384 directions at 20,000 points / 32 wavelengths.)
template <typename R, int L>
__global__ __launch_bounds__(128) void k_p2m(const int2* __restrict__ items, DevCells S,
                                             DevLevels lv, const double* __restrict__ px,
                                             const double* __restrict__ py,
                                             const double* __restrict__ pz,
                                             const typename C2T<R>::T* __restrict__ q,
                                             const double* __restrict__ dirs,
                                             const typename C2T<R>::T* __restrict__ sig2,
                                             OpConsts<R> oc, double kappa,
                                             typename C2T<R>::T* __restrict__ M, int nItems) {
  typedef typename C2T<R>::T R2;
  constexpr int LL = L * L * L;
  constexpr int NI = (LL + 127) / 128;
  constexpr int C = p2mSlotsPerPass(L);  // slots per pass: they share the basis products
  __shared__ R Sw[3][L][64];
  __shared__ R2 w[C][64];
  __shared__ double rel[3][64];
  const int item = itemIndex();
  if (item >= nItems) return;
  const int2 it = items[item];
  const int c = it.x, k0 = it.y;
  const int lvl = S.lvl[c], pb = S.pbeg[c], n = S.nprt[c], s0 = S.slot0[c], ns = S.nslot[c];
  const double cx = S.ctr[3 * c], cy = S.ctr[3 * c + 1], cz = S.ctr[3 * c + 2], rad = S.rad[c];
  const int hf = lv.isHF[lvl];
  const int doff = lv.dirOff[lvl], srow = lv.sigRow0[lvl];
  const int tid = threadIdx.x;
  {
    const int kc = min(C, ns - k0);
    R2 acc[NI][C];
#pragma unroll
    for (int m = 0; m < NI; m++)
#pragma unroll
      for (int u = 0; u < C; u++) acc[m][u] = c2<R>(0, 0);
    for (int j0 = 0; j0 < n; j0 += 64) {
      const int cn = min(64, n - j0);
      __syncthreads();
      if (tid < cn) {
        const int j = tid;
        double r0 = px[pb + j0 + j] - cx, r1 = py[pb + j0 + j] - cy, r2 = pz[pb + j0 + j] - cz;
        rel[0][j] = r0;
        rel[1][j] = r1;
        rel[2][j] = r2;
        R t[LMAX];
        lagrangeAll<R>((R)(r0 / rad), L, oc.xi, oc.iden, t);
#pragma unroll
        for (int i = 0; i < L; i++) Sw[0][i][j] = t[i];
        lagrangeAll<R>((R)(r1 / rad), L, oc.xi, oc.iden, t);
#pragma unroll
        for (int i = 0; i < L; i++) Sw[1][i][j] = t[i];
        lagrangeAll<R>((R)(r2 / rad), L, oc.xi, oc.iden, t);
#pragma unroll
        for (int i = 0; i < L; i++) Sw[2][i][j] = t[i];
      }
      __syncthreads();
      // w[u][j] = q_j exp(-i kappa d_u.(y_j - c)) for the pass's slots u
      for (int x = tid; x < C * 64; x += 128) {
        const int u = x >> 6, j = x & 63;
        R2 val = c2<R>(0, 0);
        if (j < cn && u < kc) {
          const R2 qq = q[pb + j0 + j];
          if (hf) {
            const int d = S.slotDir[s0 + k0 + u];
            const double dx = dirs[3 * (doff + d)], dy = dirs[3 * (doff + d) + 1],
                         dz = dirs[3 * (doff + d) + 2];
            R cs, sn;
            expi(-kappa * (dx * rel[0][j] + dy * rel[1][j] + dz * rel[2][j]), cs, sn);
            val = cmul(qq, c2<R>(cs, sn));
          } else {
            val = qq;
          }
        }
        w[u][j] = val;
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < NI; m++) {
        const int I = tid + 128 * m;
        if (I < LL) {
          const int i0 = I / (L * L), i1 = (I / L) % L, i2 = I % L;
          for (int j = 0; j < cn; j++) {
            const R s = Sw[0][i0][j] * Sw[1][i1][j] * Sw[2][i2][j];
#pragma unroll
            for (int u = 0; u < C; u++) {
              if (u < kc) {
                acc[m][u].x += s * w[u][j].x;
                acc[m][u].y += s * w[u][j].y;
              }
            }
          }
        }
      }
    }
#pragma unroll
    for (int m = 0; m < NI; m++) {
      const int I = tid + 128 * m;
      if (I < LL) {
#pragma unroll
        for (int u = 0; u < C; u++) {
          if (u < kc) {
            R2 v = acc[m][u];
            if (hf) v = cmulc(v, sig2[(size_t)(srow + S.slotDir[s0 + k0 + u]) * LL + I]);
            M[(size_t)(s0 + k0 + u) * LL + I] = v;
          }
        }
      }
    }
  }
}

// P2M for 16 right-hand sides on the matrix cores (float path), leaves of at
// most 64 points:
//   M_s,r = conj(sig2_s) .* (S W_s,r),  W_s,r[j] = q_r[j] exp(-i kappa d_s.(y_j - c))
// with S (L^3 x points) the Lagrange weights (P2M.hpp:113-136; LF leaves
// without the phases).  One block of 4 waves per (leaf, range of up to 32
// slots); the leaf's points, 1D weights and charges are staged in LDS, and
// wave w takes the range's slots w mod 4.  For a slot the wave writes W (64
// points x 16 right-hand sides) to LDS, then for every 16-row block of L^3
// runs ceil(n / 4) steps of V_MFMA_F32_16X16X4_F32 for the real and the
// imaginary part (A: lane l row I = 16 m + l % 16, point 4 k + l / 16, the
// product of the point's three 1D weights; B: point 4 k + l / 16,
// right-hand side l % 16) and stores the 16 x 16 result times conj(sig2_s).
constexpr int kP2MMfmaNR = 16, kP2MMfmaSlots = 32, kP2MMfmaPoints = 64;
template <int L>
__global__ __launch_bounds__(256) void k_p2m_mfma(
    const int2* __restrict__ items, DevCells S, DevLevels lv, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ pz,
    const float2* __restrict__ q, size_t qStride, const double* __restrict__ dirs,
    const float2* __restrict__ sig2, OpConsts<float> oc, double kappa,
    float2* __restrict__ M, size_t mStride, int nItems) {
  constexpr int LL = L * L * L, MT = (LL + 15) / 16, NP = kP2MMfmaPoints;
  __shared__ float Wt[3][NP][L];
  __shared__ double rel[3][NP];
  __shared__ float2 qs[NP][kP2MMfmaNR];
  __shared__ float2 Bw[4][NP][kP2MMfmaNR];
  const int item = itemIndex();
  if (item >= nItems) return;
  const int2 it = items[item];
  const int c = it.x, k0 = it.y;
  const int lvl = S.lvl[c], pb = S.pbeg[c], n = min(S.nprt[c], NP), s0 = S.slot0[c],
            ns = S.nslot[c];
  const double cx = S.ctr[3 * c], cy = S.ctr[3 * c + 1], cz = S.ctr[3 * c + 2], rad = S.rad[c];
  const int hf = lv.isHF[lvl];
  const int doff = lv.dirOff[lvl], srow = lv.sigRow0[lvl];
  const int kc = min(kP2MMfmaSlots, ns - k0);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  if (tid < NP) {
    const int i = tid;
    double r0 = 0., r1 = 0., r2 = 0.;
    if (i < n) {
      r0 = px[pb + i] - cx;
      r1 = py[pb + i] - cy;
      r2 = pz[pb + i] - cz;
    }
    rel[0][i] = r0;
    rel[1][i] = r1;
    rel[2][i] = r2;
    float t[LMAX];
    const float on = i < n ? 1.f : 0.f;
    lagrangeAll<float>((float)(r0 / rad), L, oc.xi, oc.iden, t);
#pragma unroll
    for (int a = 0; a < L; a++) Wt[0][i][a] = t[a] * on;
    lagrangeAll<float>((float)(r1 / rad), L, oc.xi, oc.iden, t);
#pragma unroll
    for (int a = 0; a < L; a++) Wt[1][i][a] = t[a];
    lagrangeAll<float>((float)(r2 / rad), L, oc.xi, oc.iden, t);
#pragma unroll
    for (int a = 0; a < L; a++) Wt[2][i][a] = t[a];
  }
  for (int x = tid; x < NP * kP2MMfmaNR; x += 256) {
    const int i = x / kP2MMfmaNR, r = x % kP2MMfmaNR;
    qs[i][r] = i < n ? q[(size_t)r * qStride + pb + i] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const int lr = lane & 15, j = lane >> 4;
  const int ksn = (n + 3) / 4;
  for (int u = wave; u < kc; u += 4) {
    const int sl = s0 + k0 + u;
    const int d = hf ? S.slotDir[sl] : 0;
    // W of this slot: lane = point
    {
      float2 ph = make_float2(1.f, 0.f);
      if (hf) {
        const double dx = dirs[3 * (doff + d)], dy = dirs[3 * (doff + d) + 1],
                     dz = dirs[3 * (doff + d) + 2];
        float cs, sn;
        expi(-kappa * (dx * rel[0][lane] + dy * rel[1][lane] + dz * rel[2][lane]), cs, sn);
        ph = make_float2(cs, sn);
      }
#pragma unroll
      for (int r = 0; r < kP2MMfmaNR; r++) Bw[wave][lane][r] = cmul(qs[lane][r], ph);
    }
    // the wave's LDS writes are visible to its own lanes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const float2* sgr = sig2 + (size_t)(srow + d) * LL;
    for (int mt = 0; mt < MT; mt++) {
      const int I = 16 * mt + lr;
      const bool rowOn = I < LL;
      const int a0 = rowOn ? I / (L * L) : 0, a1 = rowOn ? (I / L) % L : 0,
                a2 = rowOn ? I % L : 0;
      mfma_v4f dre = {0.f, 0.f, 0.f, 0.f}, dim = {0.f, 0.f, 0.f, 0.f};
      for (int ks = 0; ks < ksn; ks++) {
        const int k = 4 * ks + j;
        const float a = rowOn ? Wt[0][k][a0] * Wt[1][k][a1] * Wt[2][k][a2] : 0.f;
        const float2 b = Bw[wave][k][lr];
        dre = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b.x, dre, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b.y, dim, 0, 0, 0);
      }
      // rows 16 mt + 4 j + v, right-hand side lr
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int Io = 16 * mt + 4 * j + v;
        if (Io < LL) {
          float2 val = make_float2(dre[v], dim[v]);
          if (hf) val = cmulc(val, sgr[Io]);
          M[(size_t)lr * mStride + (size_t)sl * LL + Io] = val;
        }
      }
    }
    // this slot's W is read before the next slot's is written
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Separable translation of one level (M2M: M2M.hpp:34-124 with the sweeps of
// interpolation.hpp:244-361; L2L: L2L.hpp:34-127 with interpolation.hpp:
// 395-513).  For every output slot o:
//   out[o] (+)= sum_e edv_e * conj(post_e) .* Sweep_{sgn_e}(pre_e .* in[e])
// where Sweep applies the L x L matrix mat[(sgn >> k) & 1] along axis k
// (axis 2 first, as the reference's three GEMM + rotation passes do), and
// pre/post are rows of the phase tables (absent for LF levels).
template <typename R, int L>
__global__ __launch_bounds__(128) void k_trans(
    const int* __restrict__ out, const int* __restrict__ beg, const int* __restrict__ in,
    const int* __restrict__ sgn, const int* __restrict__ sigIdx, const double2* __restrict__ edv,
    const typename C2T<R>::T* __restrict__ src, typename C2T<R>::T* __restrict__ dst,
    const typename C2T<R>::T* __restrict__ preTab, const typename C2T<R>::T* __restrict__ postTab,
    const R* __restrict__ mat0, const R* __restrict__ mat1, int accumulate, int nItems,
    size_t rStride, int nr) {
  typedef typename C2T<R>::T R2;
  constexpr int LL = L * L * L;
  constexpr int NI = (LL + 127) / 128;
  __shared__ R2 a[LL], b[LL];
  __shared__ R mt[2][L * L];
  const int tid = threadIdx.x;
  for (int x = tid; x < L * L; x += 128) {
    mt[0][x] = mat0[x];
    mt[1][x] = mat1[x];
  }
  // consecutive outputs (the directions of one cell) read the same input
  // slots: give each XCD a contiguous part of the output order so that they
  // meet in one L2 (bijection of [0, n8), identity on the tail)
  const int blk = itemIndex();
  const int nAll = nItems * nr;  // (output, right-hand side) pairs
  if (blk >= nAll) return;
  const int n8 = nAll & ~7;
  const int oa = blk < n8 ? (blk & 7) * (n8 >> 3) + (blk >> 3) : blk;
  const int rr = oa / nItems, o = oa - rr * nItems;
  src += rr * rStride;
  dst += rr * rStride;
  const int e0 = beg[o], e1 = beg[o + 1];
  R2 acc[NI];
#pragma unroll
  for (int m = 0; m < NI; m++) acc[m] = c2<R>(0, 0);
  // the input row (and its phase row) of entry e + 1 is loaded into
  // registers while entry e is swept, hiding the load latency
  R2 nv[NI], np[NI];
  auto fetch = [&](int e) {
    const int s = in[e], row = sigIdx[e];
#pragma unroll
    for (int m = 0; m < NI; m++) {
      const int I = min(tid + 128 * m, LL - 1);
      nv[m] = src[(size_t)s * LL + I];
      np[m] = row >= 0 ? preTab[(size_t)row * LL + I] : c2<R>(1, 0);
    }
  };
  if (e0 < e1) fetch(e0);
  for (int e = e0; e < e1; e++) {
    const int row = sigIdx[e], sg = sgn[e];
    const double2 ed = edv[e];
    const R2 edr = c2<R>((R)ed.x, (R)ed.y);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < NI; m++) {
      const int I = tid + 128 * m;
      if (I < LL) a[I] = cmul(nv[m], np[m]);
    }
    if (e + 1 < e1) fetch(e + 1);
    __syncthreads();
    {
      const R* M2 = mt[(sg >> 2) & 1];
#pragma unroll
      for (int m = 0; m < NI; m++) {
        const int I = tid + 128 * m;
        if (I < LL) {
          const int i2 = I % L, base = I - i2;
          R2 t = c2<R>(0, 0);
#pragma unroll
          for (int j = 0; j < L; j++) {
            t.x += M2[i2 * L + j] * a[base + j].x;
            t.y += M2[i2 * L + j] * a[base + j].y;
          }
          b[I] = t;
        }
      }
    }
    __syncthreads();
    {
      const R* M1 = mt[(sg >> 1) & 1];
#pragma unroll
      for (int m = 0; m < NI; m++) {
        const int I = tid + 128 * m;
        if (I < LL) {
          const int i1 = (I / L) % L, base = I - i1 * L;
          R2 t = c2<R>(0, 0);
#pragma unroll
          for (int j = 0; j < L; j++) {
            t.x += M1[i1 * L + j] * b[base + j * L].x;
            t.y += M1[i1 * L + j] * b[base + j * L].y;
          }
          a[I] = t;
        }
      }
    }
    __syncthreads();
    {
      const R* M0 = mt[sg & 1];
#pragma unroll
      for (int m = 0; m < NI; m++) {
        const int I = tid + 128 * m;
        if (I < LL) {
          const int i0 = I / (L * L), base = I - i0 * L * L;
          R2 t = c2<R>(0, 0);
#pragma unroll
          for (int j = 0; j < L; j++) {
            t.x += M0[i0 * L + j] * a[base + j * L * L].x;
            t.y += M0[i0 * L + j] * a[base + j * L * L].y;
          }
          if (row >= 0) t = cmulc(t, postTab[(size_t)row * LL + I]);
          t = cmul(t, edr);
          acc[m].x += t.x;
          acc[m].y += t.y;
        }
      }
    }
  }
  const int os = out[o];
#pragma unroll
  for (int m = 0; m < NI; m++) {
    const int I = tid + 128 * m;
    if (I < LL) {
      R2 v = acc[m];
      if (accumulate) {
        R2 old = dst[(size_t)os * LL + I];
        v.x += old.x;
        v.y += old.y;
      }
      dst[(size_t)os * LL + I] = v;
    }
  }
}

// The same translation with one lane per line (used where it is the faster
// of the two: float, and L >= 6 in double).  One wave per output slot; an
// entry's input row times its pre-phase is staged in LDS; in each sweep lane
// l < L^2 owns one whole line along the swept axis, reads its L values once
// and forms all L outputs (k_trans reads L values per output).  The entry's
// 1D matrix is the same for every lane, so it comes through scalar loads
// rather than LDS.
// Threads per block: the L^2 lines rounded up to whole waves (one wave up
// to L = 8, two at L = 9, 10).
template <int L>
constexpr int transLinesTB() {
  return (L * L + 63) / 64 * 64;
}
template <typename R, int L>
__global__ __launch_bounds__(transLinesTB<L>()) void k_trans_lines(
    const int* __restrict__ out, const int* __restrict__ beg, const int* __restrict__ in,
    const int* __restrict__ sgn, const int* __restrict__ sigIdx, const double2* __restrict__ edv,
    const typename C2T<R>::T* __restrict__ src, typename C2T<R>::T* __restrict__ dst,
    const typename C2T<R>::T* __restrict__ preTab, const typename C2T<R>::T* __restrict__ postTab,
    const R* __restrict__ mat0, const R* __restrict__ mat1, int accumulate, int nItems,
    size_t rStride, int nr) {
  typedef typename C2T<R>::T R2;
  constexpr int LL = L * L * L, NL = L * L, TB = transLinesTB<L>();
  constexpr int NI = (LL + TB - 1) / TB;
  static_assert(NL <= TB, "one lane per line");
  __shared__ R2 a[LL], b[LL];
  const int lane = threadIdx.x;
  // XCD-contiguous output order, as in k_trans
  const int blk = itemIndex();
  const int nAll = nItems * nr;  // (output, right-hand side) pairs
  if (blk >= nAll) return;
  const int n8 = nAll & ~7;
  const int oa = blk < n8 ? (blk & 7) * (n8 >> 3) + (blk >> 3) : blk;
  const int rr = oa / nItems, o = oa - rr * nItems;
  src += rr * rStride;
  dst += rr * rStride;
  const int e0 = beg[o], e1 = beg[o + 1];
  // the line of this lane; lanes >= L^2 sweep line 0 and store nothing
  const bool own = lane < NL;
  const int ln = own ? lane : 0;
  const int hi = ln / L, lo = ln % L;
  R2 acc[L];  // elements I = i * L^2 + ln
#pragma unroll
  for (int i = 0; i < L; i++) acc[i] = c2<R>(0, 0);
  // the input row (and its phase row) of entry e + 1 is loaded into
  // registers while entry e is swept, hiding the load latency
  R2 nv[NI], np[NI];
  auto fetch = [&](int e) {
    const int s = in[e], row = sigIdx[e];
#pragma unroll
    for (int m = 0; m < NI; m++) {
      const int I = min(lane + TB * m, LL - 1);
      nv[m] = src[(size_t)s * LL + I];
      np[m] = row >= 0 ? preTab[(size_t)row * LL + I] : c2<R>(1, 0);
    }
  };
  // w[i] = sum_k M[i * L + k] * v[k] along one line
  auto line = [&](const R* __restrict__ M, const R2* v, R2* w) {
#pragma unroll
    for (int i = 0; i < L; i++) {
      R2 t = c2<R>(0, 0);
#pragma unroll
      for (int k = 0; k < L; k++) {
        const R m = M[i * L + k];
        t.x += m * v[k].x;
        t.y += m * v[k].y;
      }
      w[i] = t;
    }
  };
  if (e0 < e1) fetch(e0);
  for (int e = e0; e < e1; e++) {
    const int row = sigIdx[e], sg = sgn[e];
    const double2 ed = edv[e];
    const R2 edr = c2<R>((R)ed.x, (R)ed.y);
    // post-phase of this lane's outputs, loaded ahead of the sweeps
    R2 pv[L];
#pragma unroll
    for (int i = 0; i < L; i++)
      pv[i] = row >= 0 ? postTab[(size_t)row * LL + i * NL + ln] : c2<R>(1, 0);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < NI; m++) {
      const int I = lane + TB * m;
      if (I < LL) a[I] = cmul(nv[m], np[m]);
    }
    if (e + 1 < e1) fetch(e + 1);
    __syncthreads();
    R2 v[L], w[L];
    // axis 2: line (i0, i1) = (hi, lo), elements ln * L + k
#pragma unroll
    for (int k = 0; k < L; k++) v[k] = a[ln * L + k];
    line((sg & 4) ? mat1 : mat0, v, w);
    if (own)
#pragma unroll
      for (int i = 0; i < L; i++) b[ln * L + i] = w[i];
    __syncthreads();
    // axis 1: line (i0, i2) = (hi, lo), elements hi * L^2 + k * L + lo
#pragma unroll
    for (int k = 0; k < L; k++) v[k] = b[hi * NL + k * L + lo];
    line((sg & 2) ? mat1 : mat0, v, w);
    if (own)
#pragma unroll
      for (int i = 0; i < L; i++) a[hi * NL + i * L + lo] = w[i];
    __syncthreads();
    // axis 0: line (i1, i2) = ln, elements k * L^2 + ln
#pragma unroll
    for (int k = 0; k < L; k++) v[k] = a[k * NL + ln];
    line((sg & 1) ? mat1 : mat0, v, w);
#pragma unroll
    for (int i = 0; i < L; i++) {
      const R2 t = cmul(cmulc(w[i], pv[i]), edr);
      acc[i].x += t.x;
      acc[i].y += t.y;
    }
  }
  if (!own) return;
  const int os = out[o];
#pragma unroll
  for (int i = 0; i < L; i++) {
    const int I = i * NL + ln;
    R2 v = acc[i];
    if (accumulate) {
      R2 old = dst[(size_t)os * LL + I];
      v.x += old.x;
      v.y += old.y;
    }
    dst[(size_t)os * LL + I] = v;
  }
}


// M2F (M2F.hpp:35-46, circulant.hpp:95-106): F = DFT_P^3 of M embedded in
// the (2L-1)^3 grid at I2i(I, P), unnormalised, exp(-2 pi i jk/P).  The
// embedding is zero outside the L^3 corner, so the three axis passes sum over
// L inputs only.
// Rows per block: as many as fit the two intermediate arrays in 48 KB of LDS
// (at most 4).
template <typename R, int L>
constexpr int m2fRows() {
  constexpr int P = 2 * L - 1;
  constexpr int bytes = (L * L * P + L * P * P) * 2 * (int)sizeof(R);
  constexpr int r = 49152 / bytes;
  return r < 1 ? 1 : r > 4 ? 4 : r;
}
// Each pass gives one lane a whole line along its axis: the lane reads the
// line's L inputs once and forms all P outputs, with twiddles whose indices
// are compile-time constants (held in registers, not read from LDS); the last
// pass stores F straight from registers.  For P = 15 (order 8) a line is the
// prime-factor 15-point DFT of dft.hpp (168 instead of 480 operations).
template <typename R, int L>
// Blocks for all nrhs right-hand sides of a launch: block b transforms rows
// of right-hand side b / (blocks per right-hand side) (M + r * mStride into
// F + r * fStride).
__global__ __launch_bounds__(256) void k_m2f(const int* __restrict__ m2fSlot,
                                             const typename C2T<R>::T* __restrict__ M,
                                             typename C2T<R>::T* __restrict__ F,
                                             const R* __restrict__ twc,
                                             const R* __restrict__ tws, int nRows,
                                             size_t mStride, size_t fStride, int nrhs) {
  typedef typename C2T<R>::T R2;
  constexpr int P = 2 * L - 1, LL = L * L * L, PP = P * P * P, NR = m2fRows<R, L>();
  __shared__ R2 T1[NR][L * L * P], T2[NR][L * P * P];
  const int tid = threadIdx.x;
  const int perRhs = (nRows + NR - 1) / NR;
  const int bi = itemIndex();
  if (bi >= perRhs * nrhs) return;
  const int rhs = bi / perRhs;
  M += rhs * mStride;
  F += rhs * fStride;
  const int row0 = (bi - rhs * perRhs) * NR;
  const int nr = min(NR, nRows - row0);
  R cw[P], sw[P];
#pragma unroll
  for (int m = 0; m < P; m++) {
    cw[m] = twc[m];
    sw[m] = tws[m];
  }
  // t += v * exp(-2 pi i m / P)
  auto mac = [&](R2& t, const R2 v, const int m) {
    t.x += v.x * cw[m] + v.y * sw[m];
    t.y += v.y * cw[m] - v.x * sw[m];
  };
  // axis 2: line (k, i0, i1) of M -> T1[k][(i0 * L + i1) * P + k2]
  for (int x = tid; x < nr * L * L; x += 256) {
    const int k = x / (L * L), r = x - k * (L * L);
    const R2* Mr = M + (size_t)m2fSlot[row0 + k] * LL + r * L;
    R2 v[L];
#pragma unroll
    for (int j = 0; j < L; j++) v[j] = Mr[j];
    if constexpr (P == 15) {
      R2 y[P];
#pragma unroll
      for (int j = 0; j < L; j++) y[j] = v[j];
      dft15<-1, L>(y);
#pragma unroll
      for (int k2 = 0; k2 < P; k2++) T1[k][r * P + k2] = y[k2];
    } else {
#pragma unroll
      for (int k2 = 0; k2 < P; k2++) {
        R2 t = c2<R>(0, 0);
#pragma unroll
        for (int j = 0; j < L; j++) mac(t, v[j], (j * k2) % P);
        T1[k][r * P + k2] = t;
      }
    }
  }
  __syncthreads();
  // axis 1: line (k, i0, k2) -> T2[k][(i0 * P + k1) * P + k2]
  for (int x = tid; x < nr * L * P; x += 256) {
    const int k = x / (L * P), r = x - k * (L * P);
    const int i0 = r / P, k2 = r - i0 * P;
    R2 v[L];
#pragma unroll
    for (int j = 0; j < L; j++) v[j] = T1[k][(i0 * L + j) * P + k2];
    if constexpr (P == 15) {
      R2 y[P];
#pragma unroll
      for (int j = 0; j < L; j++) y[j] = v[j];
      dft15<-1, L>(y);
#pragma unroll
      for (int k1 = 0; k1 < P; k1++) T2[k][(i0 * P + k1) * P + k2] = y[k1];
    } else {
#pragma unroll
      for (int k1 = 0; k1 < P; k1++) {
        R2 t = c2<R>(0, 0);
#pragma unroll
        for (int j = 0; j < L; j++) mac(t, v[j], (j * k1) % P);
        T2[k][(i0 * P + k1) * P + k2] = t;
      }
    }
  }
  __syncthreads();
  // axis 0: line (k, k1, k2) -> F[row0 + k][k0 * P^2 + k1 * P + k2]
  for (int x = tid; x < nr * P * P; x += 256) {
    const int k = x / (P * P), r = x - k * (P * P);
    R2 v[L];
#pragma unroll
    for (int j = 0; j < L; j++) v[j] = T2[k][j * P * P + r];
    R2* Fr = F + (size_t)(row0 + k) * PP + r;
    if constexpr (P == 15) {
      R2 y[P];
#pragma unroll
      for (int j = 0; j < L; j++) y[j] = v[j];
      dft15<-1, L>(y);
#pragma unroll
      for (int k0 = 0; k0 < P; k0++) Fr[k0 * P * P] = y[k0];
    } else {
#pragma unroll
      for (int k0 = 0; k0 < P; k0++) {
        R2 t = c2<R>(0, 0);
#pragma unroll
        for (int j = 0; j < L; j++) mac(t, v[j], (j * k0) % P);
        Fr[k0 * P * P] = t;
      }
    }
  }
}

// Fused M2L + F2L (M2L.hpp:114-157, F2L.hpp:35-46, circulant.hpp:108-118) for
// a group of up to G target slots of one level and direction:
//   J_g[w] = sum over the group's entries e read by target g of
//            U_{tab}[perm_h(w)] * F_{src_e}[w]
//   L_g[I] = (1/PP) * IDFT(J_g)[I2i(I, P)]          (written, not accumulated)
// Each entry's F row is loaded once for all targets of the group that read it.
// J never leaves LDS; the inverse transform only produces the L^3 corner.
// perm_h (permutations.hpp:127-150): with h = z + 8 s, component k of the
// permuted index is w[sigma_s[k]], replaced by P - w[sigma_s[k]] when bit
// sigma_s[k] of z is set and it is nonzero.  Rather than permuting per
// frequency, the table store holds one row per (table t, axis order s, flips
// z) that the plan uses, row[w] = U_t[perm_h(w)], so every load reads
// frequency w of a row at a fixed per-thread offset (at 1M points / 128
// wavelengths / L = 5: 779,700 rows, 9.1 GB in FP64).  The device code of an
// entry's target is its row << 3.  kVariantOfSigma maps the 3 bits of sigma
// in h to the axis-order index s (permutations.hpp:87-120).
static const int kVariantOfSigma[8] = {0, 1, 0, 2, 3, 0, 4, 5};

// M2L block shape for order L: NW frequencies per thread (at most 3 where a
// block of <= 1024 threads covers PP, fewer when 256 threads suffice), TB
// threads, G target slots per group.  A thread keeps NW F values, G x NW
// table values and G x NW accumulators in registers.
// From L = 8 (k_m2l_col) a thread owns one column of the frequency cube: NW
// = P frequencies, TB = P^2 threads rounded up to whole waves, one target.
template <int L>
struct M2LShape {
  static constexpr int P = 2 * L - 1;
  static constexpr int PP = P * P * P;
  static constexpr bool kColumns = L >= 8;
  static constexpr int NWmin = (PP + 1023) / 1024;
  static constexpr int NWpref = (PP + 255) / 256 < 3 ? (PP + 255) / 256 : 3;
  static constexpr int NW = kColumns ? P : NWmin > NWpref ? NWmin : NWpref;
  static constexpr int TB =
      kColumns ? (P * P + 63) / 64 * 64 : ((PP + NW - 1) / NW + 63) / 64 * 64;
  static constexpr int G = m2lGroupSize(L);
};
// With NR right-hand sides per launch a thread holds G x NW x NR
// accumulators in the same shape (measured faster than one frequency per
// thread in 768-thread blocks: 144 vs 187 ms per right-hand side at 1M
// points / 128 wavelengths / L = 5, float, two right-hand sides).
template <int L, int NR>
struct M2LShapeN {
  static constexpr int PP = M2LShape<L>::PP;
  static constexpr int NW = M2LShape<L>::NW;
  static constexpr int TB = M2LShape<L>::TB;
  static constexpr int G = M2LShape<L>::G;
};
// Right-hand sides per M2L launch: at most 16 (double) or 32 (float)
// complex accumulators per thread.
// (double from L = 7: one, float at L = 7: two; wider kernels spill at the
// occupancy asked for below, and at L = 8 in double without it)
template <typename R, int L>
constexpr int m2lMaxNR() {
  if (M2LShape<L>::kColumns) return sizeof(R) == 8 ? 1 : 2;
  constexpr int NW = M2LShape<L>::NW;
  constexpr int cap = (sizeof(R) == 8 ? 16 : 32) / (M2LShape<L>::G * NW);
  constexpr int w = cap >= 8 ? 8 : cap >= 4 ? 4 : cap >= 2 ? 2 : 1;
  return (sizeof(R) == 8 && L >= 7) ? 1 : (L == 7 && w > 2) ? 2 : w;
}
static int m2lMaxNRAt(int prec, int L) {
  switch (L) {
#define DFM_NR_CASE(l) \
    case l: return prec == DEFMM_FP64 ? m2lMaxNR<double, l>() : m2lMaxNR<float, l>();
    DFM_NR_CASE(2) DFM_NR_CASE(3) DFM_NR_CASE(4) DFM_NR_CASE(5) DFM_NR_CASE(6)
    DFM_NR_CASE(7) DFM_NR_CASE(8) DFM_NR_CASE(9) DFM_NR_CASE(10)
#undef DFM_NR_CASE
    default: return 1;
  }
}

// Waves per SIMD the compiler is asked to fit: at L = 7 (768-thread blocks)
// float with two right-hand sides took 105 VGPRs, one block per CU; 6 waves
// per SIMD (two blocks) fit without spilling in groups of 3 targets (in
// double they spill, so double keeps one block per CU).
template <typename R, int L, int NR>
constexpr int m2lWavesPerEU() {
  return (L == 7 && sizeof(R) == 4) ? 6 : 1;
}

template <int PP>
__device__ __forceinline__ size_t tableRow(int code) {
  return (size_t)(code >> 3) * PP;
}


// NR right-hand sides per launch: right-hand side r reads F + r * fStride
// and writes Lout + r * lStride; each table value is loaded once for all.
// SL (sliced): the entries are swept once per frequency slice m (a thread's
// m-th frequency) instead of once for all NW slices, so at any time a work
// group reads 1/NW of each of its rows: the rows that neighbouring groups
// share stay in L2 longer.  Two entries' loads are issued together.  Each
// accumulator sums the same terms in the same order with the same fused
// multiply-adds (cmacFma): results are identical.
template <typename R, int L, int NR, bool SL>
__global__ __launch_bounds__((M2LShapeN<L, NR>::TB))
__attribute__((amdgpu_waves_per_eu(m2lWavesPerEU<R, L, NR>(), 8))) void k_m2l(
    const int* __restrict__ groupBeg, const int* __restrict__ groupTarget,
    const int* __restrict__ gSrcF, const int* __restrict__ gMask, const int* __restrict__ gCode,
    const typename C2T<R>::T* __restrict__ F, size_t fStride,
    const typename C2T<R>::T* __restrict__ U, typename C2T<R>::T* __restrict__ Lout,
    size_t lStride, const R* __restrict__ twc, const R* __restrict__ tws, int nItems) {
  typedef typename C2T<R>::T R2;
  constexpr int P = 2 * L - 1, LL = L * L * L, PP = P * P * P;
  constexpr int NW = M2LShapeN<L, NR>::NW, TB = M2LShapeN<L, NR>::TB, G = M2LShapeN<L, NR>::G;
  __shared__ R2 J[PP], T1[L * P * P], T2[L * L * P];
  __shared__ R wc[P], ws[P];
  const int tid = threadIdx.x;
  // Workgroups are dealt to the 8 XCDs round-robin; give each XCD a
  // contiguous part of the (direction-major) group order so that its L2 sees
  // the groups that share source rows.  A bijection of [0, n8), n8 = n
  // rounded down to a multiple of 8, and the identity on the tail.
  const int b = itemIndex();
  if (b >= nItems) return;
  const int n8 = nItems & ~7;
  const int grp = b < n8 ? (b & 7) * (n8 >> 3) + (b >> 3) : b;
  if (tid < P) {
    wc[tid] = twc[tid];
    ws[tid] = tws[tid];
  }
  // frequencies of this thread, clamped to PP - 1: every load is in bounds
  // and unconditional, and the accumulators of w >= PP are never stored
  unsigned wl[NW];
#pragma unroll
  for (int m = 0; m < NW; m++) wl[m] = (unsigned)min(tid + TB * m, PP - 1);
  R2 acc[G][NW][NR];
#pragma unroll
  for (int g = 0; g < G; g++)
#pragma unroll
    for (int m = 0; m < NW; m++)
#pragma unroll
      for (int r = 0; r < NR; r++) acc[g][m][r] = c2<R>(0, 0);
  const int e0 = groupBeg[grp], e1 = groupBeg[grp + 1];
  if constexpr (SL) {
    // one slice m at a time; entries e, e + 1 loaded together
    auto ld = [&](int e, int m, R2 (&f)[NR], R2 (&u)[G], int& mask) {
      const R2* Fr = F + (size_t)gSrcF[e] * PP + wl[m];
      mask = gMask[e];
#pragma unroll
      for (int r = 0; r < NR; r++) f[r] = Fr[r * fStride];
#pragma unroll
      for (int g = 0; g < G; g++)
        if ((mask >> g) & 1) u[g] = U[tableRow<PP>(gCode[(size_t)e * G + g]) + wl[m]];
    };
    auto mac = [&](int m, const R2 (&f)[NR], const R2 (&u)[G], int mask) {
#pragma unroll
      for (int g = 0; g < G; g++) {
        if (!((mask >> g) & 1)) continue;
#pragma unroll
        for (int r = 0; r < NR; r++) cmacFma(acc[g][m][r], u[g], f[r]);
      }
    };
#pragma unroll
    for (int m = 0; m < NW; m++) {
      int e = e0;
      for (; e + 1 < e1; e += 2) {
        R2 fa[NR], ua[G], fb[NR], ub[G];
        int ma, mb;
        ld(e, m, fa, ua, ma);
        ld(e + 1, m, fb, ub, mb);
        mac(m, fa, ua, ma);
        mac(m, fb, ub, mb);
      }
      if (e < e1) {
        R2 fa[NR], ua[G];
        int ma;
        ld(e, m, fa, ua, ma);
        mac(m, fa, ua, ma);
      }
    }
  } else
  for (int e = e0; e < e1; e++) {
    const R2* Fr = F + (size_t)gSrcF[e] * PP;
    const int mask = gMask[e];
    int code[G];
#pragma unroll
    for (int g = 0; g < G; g++) code[g] = gCode[(size_t)e * G + g];
    // issue all of the entry's loads before the first use
    R2 f[NW][NR], u[G][NW];
#pragma unroll
    for (int m = 0; m < NW; m++)
#pragma unroll
      for (int r = 0; r < NR; r++) f[m][r] = Fr[r * fStride + wl[m]];
#pragma unroll
    for (int g = 0; g < G; g++) {
      // (u of a target absent from the entry is left unset and not read)
      if ((mask >> g) & 1) {
        const R2* Ur = U + tableRow<PP>(code[g]);
#pragma unroll
        for (int m = 0; m < NW; m++) u[g][m] = Ur[wl[m]];
      }
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (!((mask >> g) & 1)) continue;
#pragma unroll
      for (int m = 0; m < NW; m++)
#pragma unroll
        for (int r = 0; r < NR; r++) cmacFma(acc[g][m][r], u[g][m], f[m][r]);
    }
  }
  const R inv = (R)(1.0 / (double)PP);
#pragma unroll
  for (int g = 0; g < G; g++) {
    const int t = groupTarget[(size_t)grp * G + g];
    if (t < 0) continue;
#pragma unroll
    for (int rr = 0; rr < NR; rr++) {
    R2* Lr = Lout + rr * lStride + (size_t)t * LL;
    __syncthreads();
#pragma unroll
    for (int m = 0; m < NW; m++) {
      const int w = tid + TB * m;
      if (w < PP) J[w] = acc[g][m][rr];
    }
    __syncthreads();
    // axis 0: T1[(k0*P + w1)*P + w2], k0 < L
    for (int x = tid; x < L * P * P; x += TB) {
      const int r = x % (P * P), k0 = x / (P * P);
      R2 s = c2<R>(0, 0);
#pragma unroll
      for (int v0 = 0; v0 < P; v0++) {
        const int m = (v0 * k0) % P;
        const R2 v = J[v0 * P * P + r];
        s.x += v.x * wc[m] - v.y * ws[m];
        s.y += v.y * wc[m] + v.x * ws[m];
      }
      T1[x] = s;
    }
    __syncthreads();
    // axis 1: T2[(k0*L + k1)*P + w2], k1 < L
    for (int x = tid; x < L * L * P; x += TB) {
      const int v2 = x % P, k1 = (x / P) % L, k0 = x / (L * P);
      R2 s = c2<R>(0, 0);
#pragma unroll
      for (int v1 = 0; v1 < P; v1++) {
        const int m = (v1 * k1) % P;
        const R2 v = T1[(k0 * P + v1) * P + v2];
        s.x += v.x * wc[m] - v.y * ws[m];
        s.y += v.y * wc[m] + v.x * ws[m];
      }
      T2[x] = s;
    }
    __syncthreads();
    // axis 2 and scaling: L[(k0*L + k1)*L + k2]
    for (int x = tid; x < LL; x += TB) {
      const int k2 = x % L, r = x / L;
      R2 s = c2<R>(0, 0);
#pragma unroll
      for (int v2 = 0; v2 < P; v2++) {
        const int m = (v2 * k2) % P;
        const R2 v = T2[r * P + v2];
        s.x += v.x * wc[m] - v.y * ws[m];
        s.y += v.y * wc[m] + v.x * ws[m];
      }
      Lr[x] = c2<R>(s.x * inv, s.y * inv);
    }
    }
  }
}

// Fused M2L + F2L for orders L >= 8 (same lists and formulas as k_m2l, one
// target per group): thread t < P^2 owns the column (w1, w2) = t of the
// frequency cube, all P values of w0, so the first axis of the inverse DFT
// is summed in registers and the Fourier-domain local expansion J never
// goes to LDS (J, T1 and T2 in LDS would need 198 KB in double at L = 10;
// T1 and T2 take 88 KB; at L = 8 the grouped kernel's 98 KB allowed one
// block per CU, whose inverse DFT then overlapped no loads).  Loads of a
// column are P^2 elements apart and coalesced across the block's threads.
// Each target sums its entries in the order of the grouped kernel, so both
// give bit-identical results.
template <typename R, int L, int NR>
__global__ __launch_bounds__(M2LShape<L>::TB) void k_m2l_col(
    const int* __restrict__ groupBeg, const int* __restrict__ groupTarget,
    const int* __restrict__ gSrcF, const int* __restrict__ gMask, const int* __restrict__ gCode,
    const typename C2T<R>::T* __restrict__ F, size_t fStride,
    const typename C2T<R>::T* __restrict__ U, typename C2T<R>::T* __restrict__ Lout,
    size_t lStride, const R* __restrict__ twc, const R* __restrict__ tws, int nItems) {
  typedef typename C2T<R>::T R2;
  constexpr int P = 2 * L - 1, P2 = P * P, LL = L * L * L, PP = P * P * P;
  constexpr int TB = M2LShape<L>::TB;
  static_assert(M2LShape<L>::G == 1, "one target per group");
  // order 8 runs each F2L step for all NR right-hand sides at once
  constexpr int NT = P == 15 ? NR : 1;
  __shared__ R2 T1[NT][L * P2], T2[NT][L * L * P];
  __shared__ R wc[P], ws[P];
  const int tid = threadIdx.x;
  const int b = itemIndex();
  if (b >= nItems) return;
  // XCD-contiguous group order, as in k_m2l
  const int n8 = nItems & ~7;
  const int grp = b < n8 ? (b & 7) * (n8 >> 3) + (b >> 3) : b;
  if (tid < P) {
    wc[tid] = twc[tid];
    ws[tid] = tws[tid];
  }
  // threads beyond the cube's columns load a valid column and store nothing
  const int col = min(tid, P2 - 1);
  R2 acc[P][NR];
#pragma unroll
  for (int m = 0; m < P; m++)
#pragma unroll
    for (int r = 0; r < NR; r++) acc[m][r] = c2<R>(0, 0);
  const int e0 = groupBeg[grp], e1 = groupBeg[grp + 1];
  for (int e = e0; e < e1; e++) {
    const R2* Fr = F + (size_t)gSrcF[e] * PP + col;
    const R2* Ur = U + tableRow<PP>(gCode[e]) + col;
    // all of the entry's loads are issued before the first use
    R2 f[P][NR], u[P];
#pragma unroll
    for (int m = 0; m < P; m++)
#pragma unroll
      for (int r = 0; r < NR; r++) f[m][r] = Fr[r * fStride + m * P2];
#pragma unroll
    for (int m = 0; m < P; m++) u[m] = Ur[m * P2];
    // keep every load above this point: the scheduler would otherwise sink
    // them next to their first use and keep only a few in flight per wave
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int m = 0; m < P; m++)
#pragma unroll
      for (int r = 0; r < NR; r++) cmacFma(acc[m][r], u[m], f[m][r]);
  }
  const int t = groupTarget[grp];
  if (t < 0) return;  // (uniform over the block)
  const R inv = (R)(1.0 / (double)PP);
  if constexpr (P == 15) {
    // prime-factor 15-point lines; each step for all right-hand sides, one
    // barrier per step: axis 0 in registers, then a lane per line of axis 1
    // (NR * L * P lines) and of axis 2 (NR * L^2 lines)
    if (tid < P2)
#pragma unroll
      for (int rr = 0; rr < NR; rr++) {
        R2 y[P];
#pragma unroll
        for (int m = 0; m < P; m++) y[m] = acc[m][rr];
        dft15<1, P>(y);
#pragma unroll
        for (int k0 = 0; k0 < L; k0++) T1[rr][k0 * P2 + tid] = y[k0];
      }
    __syncthreads();
    for (int x = tid; x < NR * L * P; x += TB) {
      const int rr = x / (L * P), z = x - rr * (L * P);
      const int k0 = z / P, v2 = z - k0 * P;
      R2 y[P];
#pragma unroll
      for (int v1 = 0; v1 < P; v1++) y[v1] = T1[rr][(k0 * P + v1) * P + v2];
      dft15<1, P>(y);
#pragma unroll
      for (int k1 = 0; k1 < L; k1++) T2[rr][(k0 * L + k1) * P + v2] = y[k1];
    }
    __syncthreads();
    for (int x = tid; x < NR * L * L; x += TB) {
      const int rr = x / (L * L), z = x - rr * (L * L);
      R2 y[P];
#pragma unroll
      for (int v2 = 0; v2 < P; v2++) y[v2] = T2[rr][z * P + v2];
      dft15<1, P>(y);
      R2* Lr = Lout + rr * lStride + (size_t)t * LL + z * L;
#pragma unroll
      for (int k2 = 0; k2 < L; k2++) Lr[k2] = c2<R>(y[k2].x * inv, y[k2].y * inv);
    }
    return;
  }
  __syncthreads();  // twiddles in LDS
#pragma unroll
  for (int rr = 0; rr < NR; rr++) {
    R2* Lr = Lout + rr * lStride + (size_t)t * LL;
    if (rr > 0) __syncthreads();  // T1 / T2 of the previous right-hand side consumed
    // axis 0 in registers: T1[k0 * P^2 + col] = sum_w0 J[w0, col] e^{+2 pi i w0 k0 / P}
    if (tid < P2) {
#pragma unroll
      for (int k0 = 0; k0 < L; k0++) {
        R2 s = c2<R>(0, 0);
#pragma unroll
        for (int m = 0; m < P; m++) {
          const int q = (m * k0) % P;
          const R2 v = acc[m][rr];
          s.x += v.x * wc[q] - v.y * ws[q];
          s.y += v.y * wc[q] + v.x * ws[q];
        }
        T1[0][k0 * P2 + tid] = s;
      }
    }
    __syncthreads();
    // axis 1: T2[(k0*L + k1)*P + w2], k1 < L
    for (int x = tid; x < L * L * P; x += TB) {
      const int v2 = x % P, k1 = (x / P) % L, k0 = x / (L * P);
      R2 s = c2<R>(0, 0);
#pragma unroll
      for (int v1 = 0; v1 < P; v1++) {
        const int q = (v1 * k1) % P;
        const R2 v = T1[0][(k0 * P + v1) * P + v2];
        s.x += v.x * wc[q] - v.y * ws[q];
        s.y += v.y * wc[q] + v.x * ws[q];
      }
      T2[0][x] = s;
    }
    __syncthreads();
    // axis 2 and scaling: L[(k0*L + k1)*L + k2]
    for (int x = tid; x < LL; x += TB) {
      const int k2 = x % L, r = x / L;
      R2 s = c2<R>(0, 0);
#pragma unroll
      for (int v2 = 0; v2 < P; v2++) {
        const int q = (v2 * k2) % P;
        const R2 v = T2[0][r * P + v2];
        s.x += v.x * wc[q] - v.y * ws[q];
        s.y += v.y * wc[q] + v.x * ws[q];
      }
      Lr[x] = c2<R>(s.x * inv, s.y * inv);
    }
  }
}

// The same grouped M2L + F2L without a stored F: each entry's source row is
// transformed on the fly from its M expansion (the M2F of M2F.hpp:35-46,
// circulant.hpp:95-106: pruned 3-pass DFT of the embedded L^3 corner, in
// LDS), so only M (L^3) is read per entry and no F store exists.
template <typename R, int L>
__global__ __launch_bounds__(M2LShape<L>::TB) void k_m2l_fly(const int* __restrict__ groupBeg,
                                                 const int* __restrict__ groupTarget,
                                                 const int* __restrict__ gSrcSlot,
                                                 const int* __restrict__ gMask,
                                                 const int* __restrict__ gCode,
                                                 const typename C2T<R>::T* __restrict__ M,
                                                 const typename C2T<R>::T* __restrict__ U,
                                                 typename C2T<R>::T* __restrict__ Lout,
                                                 const R* __restrict__ twc,
                                                 const R* __restrict__ tws, int nItems) {
  typedef typename C2T<R>::T R2;
  constexpr int P = 2 * L - 1, LL = L * L * L, PP = P * P * P;
  constexpr int NW = M2LShape<L>::NW, TB = M2LShape<L>::TB, G = M2LShape<L>::G;
  constexpr int NB = L * P * P;  // >= L*L*P
  __shared__ R2 sM[LL], B1[NB], B2[NB], J[PP];
  __shared__ R wc[P], ws[P];
  const int tid = threadIdx.x;
  const int b = itemIndex();
  if (b >= nItems) return;
  const int n8 = nItems & ~7;
  const int grp = b < n8 ? (b & 7) * (n8 >> 3) + (b >> 3) : b;
  if (tid < P) {
    wc[tid] = twc[tid];
    ws[tid] = tws[tid];
  }
  unsigned wl[NW];
  int w0[NW];
#pragma unroll
  for (int m = 0; m < NW; m++) {
    wl[m] = (unsigned)min(tid + TB * m, PP - 1);
    w0[m] = (int)wl[m] / (P * P);
  }
  R2 acc[G][NW];
#pragma unroll
  for (int g = 0; g < G; g++)
#pragma unroll
    for (int m = 0; m < NW; m++) acc[g][m] = c2<R>(0, 0);
  const int e0 = groupBeg[grp], e1 = groupBeg[grp + 1];
  for (int e = e0; e < e1; e++) {
    const R2* Mr = M + (size_t)gSrcSlot[e] * LL;
    const int mask = gMask[e];
    // the table values are loaded first: they arrive while the F row is
    // computed in LDS
    R2 u[G][NW];
#pragma unroll
    for (int g = 0; g < G; g++) {
      const R2* Ur = U + tableRow<PP>(gCode[(size_t)e * G + g]);
      if ((mask >> g) & 1) {
#pragma unroll
        for (int m = 0; m < NW; m++) u[g][m] = Ur[wl[m]];
      } else {
#pragma unroll
        for (int m = 0; m < NW; m++) u[g][m] = c2<R>(0, 0);
      }
    }
    __syncthreads();
    for (int I = tid; I < LL; I += TB) sM[I] = Mr[I];
    __syncthreads();
    // axis 2: B1[(i0*L + i1)*P + k2]
    for (int x = tid; x < L * L * P; x += TB) {
      const int k2 = x % P, r = x / P;
      R2 t = c2<R>(0, 0);
#pragma unroll
      for (int j = 0; j < L; j++) {
        const int m = (j * k2) % P;
        const R2 v = sM[r * L + j];
        t.x += v.x * wc[m] + v.y * ws[m];
        t.y += v.y * wc[m] - v.x * ws[m];
      }
      B1[x] = t;
    }
    __syncthreads();
    // axis 1: B2[(i0*P + k1)*P + k2]
    for (int x = tid; x < L * P * P; x += TB) {
      const int k2 = x % P, k1 = (x / P) % P, i0 = x / (P * P);
      R2 t = c2<R>(0, 0);
#pragma unroll
      for (int j = 0; j < L; j++) {
        const int m = (j * k1) % P;
        const R2 v = B1[(i0 * L + j) * P + k2];
        t.x += v.x * wc[m] + v.y * ws[m];
        t.y += v.y * wc[m] - v.x * ws[m];
      }
      B2[x] = t;
    }
    __syncthreads();
    // axis 0 for this thread's frequencies, then the Hadamard products
    R2 f[NW];
#pragma unroll
    for (int m = 0; m < NW; m++) {
      const int w = tid + TB * m;
      R2 t = c2<R>(0, 0);
      if (w < PP) {
        const int r = w % (P * P);
#pragma unroll
        for (int j = 0; j < L; j++) {
          const int mm = (j * w0[m]) % P;
          const R2 v = B2[j * P * P + r];
          t.x += v.x * wc[mm] + v.y * ws[mm];
          t.y += v.y * wc[mm] - v.x * ws[mm];
        }
      }
      f[m] = t;
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (!((mask >> g) & 1)) continue;
#pragma unroll
      for (int m = 0; m < NW; m++) {
        acc[g][m].x += u[g][m].x * f[m].x - u[g][m].y * f[m].y;
        acc[g][m].y += u[g][m].x * f[m].y + u[g][m].y * f[m].x;
      }
    }
  }
  const R inv = (R)(1.0 / (double)PP);
#pragma unroll
  for (int g = 0; g < G; g++) {
    const int t = groupTarget[(size_t)grp * G + g];
    if (t < 0) continue;
    R2* Lr = Lout + (size_t)t * LL;
    __syncthreads();
#pragma unroll
    for (int m = 0; m < NW; m++) {
      const int w = tid + TB * m;
      if (w < PP) J[w] = acc[g][m];
    }
    __syncthreads();
    for (int x = tid; x < L * P * P; x += TB) {
      const int r = x % (P * P), k0 = x / (P * P);
      R2 s = c2<R>(0, 0);
#pragma unroll
      for (int v0 = 0; v0 < P; v0++) {
        const int m = (v0 * k0) % P;
        const R2 v = J[v0 * P * P + r];
        s.x += v.x * wc[m] - v.y * ws[m];
        s.y += v.y * wc[m] + v.x * ws[m];
      }
      B1[x] = s;
    }
    __syncthreads();
    for (int x = tid; x < L * L * P; x += TB) {
      const int v2 = x % P, k1 = (x / P) % L, k0 = x / (L * P);
      R2 s = c2<R>(0, 0);
#pragma unroll
      for (int v1 = 0; v1 < P; v1++) {
        const int m = (v1 * k1) % P;
        const R2 v = B1[(k0 * P + v1) * P + v2];
        s.x += v.x * wc[m] - v.y * ws[m];
        s.y += v.y * wc[m] + v.x * ws[m];
      }
      B2[x] = s;
    }
    __syncthreads();
    for (int x = tid; x < LL; x += TB) {
      const int k2 = x % L, r = x / L;
      R2 s = c2<R>(0, 0);
#pragma unroll
      for (int v2 = 0; v2 < P; v2++) {
        const int m = (v2 * k2) % P;
        const R2 v = B2[r * P + v2];
        s.x += v.x * wc[m] - v.y * ws[m];
        s.y += v.y * wc[m] + v.x * ws[m];
      }
      Lr[x] = c2<R>(s.x * inv, s.y * inv);
    }
  }
}

// L = 0 for the target slots that receive no M2L (their L gets only L2L).
// (grid-stride with a capped grid, as k_zero)
template <typename R, int L>
__global__ __launch_bounds__(256) void k_zero_slots(const int* __restrict__ slots, int n,
                                                    typename C2T<R>::T* __restrict__ Lout) {
  constexpr int LL = L * L * L;
  const int64_t total = (int64_t)n * LL, stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += stride) {
    const int s = slots[i / LL];
    Lout[(size_t)s * LL + (i % LL)] = c2<R>(0, 0);
  }
}

// L2P (L2P.hpp:121-173 with mapOnTrees.hpp:85-134):
//   HF: p_i += sum_d exp(i kappa d.(x_i-c)) sum_I S_I((x_i-c)/r) sig2_d[I] L_d[I]
//   LF: p_i += sum_I S_I((x_i-c)/r) L[I]
// One block of 4 waves per leaf.  The leaf's slots are dealt to NG slot
// groups: one per wave, or two per wave (half-waves of 32 lanes) when the
// leaf holds at most 32 points; NG slots at a time are staged in LDS (all of
// the block's threads load them) and each lane evaluates one point against
// its group's slot.  The groups' sums are added in group order at the end,
// so the result does not depend on timing.  (One wave per leaf, looping over
// its slots one at a time, kept 1,089 waves busy at 20,000 points / 32
// wavelengths, a third of their lanes idle.)
template <typename R, int L>
constexpr int l2pGroupsMax() {
  return L * L * L * 2 * (int)sizeof(R) <= 8192 ? 8 : 4;
}
template <typename R, int L>
__global__ __launch_bounds__(256) void k_l2p(const int* __restrict__ leaves, DevCells T,
                                             DevLevels lv, const double* __restrict__ px,
                                             const double* __restrict__ py,
                                             const double* __restrict__ pz,
                                             const double* __restrict__ dirs,
                                             const typename C2T<R>::T* __restrict__ sig2,
                                             const typename C2T<R>::T* __restrict__ Lst,
                                             OpConsts<R> oc, double kappa,
                                             typename C2T<R>::T* __restrict__ p, int nItems) {
  typedef typename C2T<R>::T R2;
  constexpr int LL = L * L * L, NWV = 4, NGMAX = l2pGroupsMax<R, L>();
  __shared__ R2 buf[NGMAX][LL];
  __shared__ R2 red[NGMAX][64];
  const int item = itemIndex();
  if (item >= nItems) return;
  const int c = leaves[item];
  const int lvl = T.lvl[c], pb = T.pbeg[c], n = T.nprt[c], s0 = T.slot0[c], ns = T.nslot[c];
  const double cx = T.ctr[3 * c], cy = T.ctr[3 * c + 1], cz = T.ctr[3 * c + 2], rad = T.rad[c];
  const int hf = lv.isHF[lvl];
  const int doff = lv.dirOff[lvl], srow = lv.sigRow0[lvl];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // slot groups per wave (uniform over the block) and this lane's group and point
  const int spw = (n <= 32 && 2 * NWV <= NGMAX) ? 2 : 1;
  const int NG = NWV * spw < NGMAX ? NWV * spw : NGMAX;
  const int pts = 64 / spw;
  const int g = wave * spw + (spw == 2 ? lane >> 5 : 0);
  const int pl = spw == 2 ? (lane & 31) : lane;
  for (int i0 = 0; i0 < n; i0 += pts) {
    const int i = i0 + pl;
    const bool act = i < n && g < NG;
    double r0 = 0., r1 = 0., r2 = 0.;
    if (act) {
      r0 = px[pb + i] - cx;
      r1 = py[pb + i] - cy;
      r2 = pz[pb + i] - cz;
    }
    R sx[LMAX], sy[LMAX], sz[LMAX];
    lagrangeAll<R>((R)(r0 / rad), L, oc.xi, oc.iden, sx);
    lagrangeAll<R>((R)(r1 / rad), L, oc.xi, oc.iden, sy);
    lagrangeAll<R>((R)(r2 / rad), L, oc.xi, oc.iden, sz);
    R2 pacc = c2<R>(0, 0);
    for (int k0 = 0; k0 < ns; k0 += NG) {
      const int kc = min(NG, ns - k0);
      __syncthreads();
      // stage the kc slots' L .* sig2 rows
      for (int x = tid; x < kc * LL; x += 256) {
        const int u = x / LL, I = x - u * LL, k = s0 + k0 + u;
        const R2 v = Lst[(size_t)k * LL + I];
        buf[u][I] = hf ? cmul(v, sig2[(size_t)(srow + T.slotDir[k]) * LL + I]) : v;
      }
      __syncthreads();
      if (g < kc) {
        const R2* b = buf[g];
        R2 s = c2<R>(0, 0);
#pragma unroll
        for (int a = 0; a < L; a++)
#pragma unroll
          for (int bb = 0; bb < L; bb++) {
            R2 s1 = c2<R>(0, 0);
#pragma unroll
            for (int e = 0; e < L; e++) {
              const R2 v = b[(a * L + bb) * L + e];
              s1.x += sz[e] * v.x;
              s1.y += sz[e] * v.y;
            }
            const R t = sx[a] * sy[bb];
            s.x += t * s1.x;
            s.y += t * s1.y;
          }
        if (hf) {
          const int d = T.slotDir[s0 + k0 + g];
          const double dx = dirs[3 * (doff + d)], dy = dirs[3 * (doff + d) + 1],
                       dz = dirs[3 * (doff + d) + 2];
          R cs, sn;
          expi(kappa * (dx * r0 + dy * r1 + dz * r2), cs, sn);
          s = cmul(s, c2<R>(cs, sn));
        }
        pacc.x += s.x;
        pacc.y += s.y;
      }
    }
    // the groups' sums, in group order
    __syncthreads();
    if (g < NG) red[g][pl] = pacc;
    __syncthreads();
    if (tid < pts && i0 + tid < n) {
      R2 t = red[0][tid];
      for (int gg = 1; gg < NG; gg++) {
        t.x += red[gg][tid].x;
        t.y += red[gg][tid].y;
      }
      R2 o = p[pb + i0 + tid];
      o.x += t.x;
      o.y += t.y;
      p[pb + i0 + tid] = o;
    }
  }
}

// L2P for 16 right-hand sides on the matrix cores (float path).  For a
// target leaf and each of its slots s, C_s = S B_s with S (points x L^3,
// real) the Lagrange weights S_I((x_i - c)/r) and B_s (L^3 x 16, complex) the
// slot's local expansions times its phase row, B_s[I][r] = sig2_s[I]
// L_s,r[I]; then p_r[i] += exp(i kappa d_s.(x_i - c)) C_s[i][r] (the sums of
// L2P.hpp:121-173).  One block of 4 waves per (leaf, tile of up to TP
// points): the tile's S is staged in LDS once for all slots; wave w takes
// the slots k = w mod 4, each as ceil(L^3 / 4) steps of V_MFMA_F32_16X16X4_F32
// (real and imaginary part of B, TP / 16 row blocks sharing each B operand)
// with A = S (lane l: point l % 16 of a row block, weight index 4 k + l / 16)
// and B (lane l: weight index 4 k + l / 16, right-hand side l % 16).  The
// waves' sums are added in wave order at the end (deterministic).
constexpr int kL2PMfmaNR = 16;
template <int L>
constexpr int l2pMfmaTile() {
  return L * L * L <= 512 ? 32 : 16;
}
template <int L>
__global__ __launch_bounds__(256) void k_l2p_mfma(
    const int2* __restrict__ items, DevCells T, DevLevels lv, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ pz,
    const double* __restrict__ dirs, const float2* __restrict__ sig2,
    const float2* __restrict__ Lst, size_t lStride, OpConsts<float> oc, double kappa,
    float2* __restrict__ p, size_t pStride, int nItems) {
  constexpr int LL = L * L * L, KS = (LL + 3) / 4, TP = l2pMfmaTile<L>(), NB = TP / 16;
  // S (TP x 4 KS floats) and, once every wave is past the slot loop, the
  // waves' sums (4 x TP x 16 complex) share one array
  constexpr int NS = TP * KS * 4, NRED = 4 * TP * kL2PMfmaNR * 2;
  __shared__ float smem[NS > NRED ? NS : NRED];
  float(*S)[KS * 4] = reinterpret_cast<float(*)[KS * 4]>(smem);
  float2(*red)[TP][kL2PMfmaNR] = reinterpret_cast<float2(*)[TP][kL2PMfmaNR]>(smem);
  __shared__ float W[3][TP][L];
  __shared__ double rel[3][TP];
  const int item = itemIndex();
  if (item >= nItems) return;
  const int2 it = items[item];
  const int c = it.x, i0 = it.y;
  const int lvl = T.lvl[c], pb = T.pbeg[c], n = T.nprt[c], s0 = T.slot0[c], ns = T.nslot[c];
  const double cx = T.ctr[3 * c], cy = T.ctr[3 * c + 1], cz = T.ctr[3 * c + 2], rad = T.rad[c];
  const int hf = lv.isHF[lvl];
  const int doff = lv.dirOff[lvl], srow = lv.sigRow0[lvl];
  const int nt = min(TP, n - i0);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // Lagrange weights of the tile's points (zero rows beyond the leaf)
  if (tid < TP) {
    const int i = tid;
    double r0 = 0., r1 = 0., r2 = 0.;
    if (i < nt) {
      r0 = px[pb + i0 + i] - cx;
      r1 = py[pb + i0 + i] - cy;
      r2 = pz[pb + i0 + i] - cz;
    }
    rel[0][i] = r0;
    rel[1][i] = r1;
    rel[2][i] = r2;
    float t[LMAX];
    const float on = i < nt ? 1.f : 0.f;
    lagrangeAll<float>((float)(r0 / rad), L, oc.xi, oc.iden, t);
#pragma unroll
    for (int a = 0; a < L; a++) W[0][i][a] = t[a] * on;
    lagrangeAll<float>((float)(r1 / rad), L, oc.xi, oc.iden, t);
#pragma unroll
    for (int a = 0; a < L; a++) W[1][i][a] = t[a];
    lagrangeAll<float>((float)(r2 / rad), L, oc.xi, oc.iden, t);
#pragma unroll
    for (int a = 0; a < L; a++) W[2][i][a] = t[a];
  }
  __syncthreads();
  for (int x = tid; x < TP * KS * 4; x += 256) {
    const int i = x / (KS * 4), I = x - i * (KS * 4);
    float v = 0.f;
    if (I < LL) v = W[0][i][I / (L * L)] * W[1][i][(I / L) % L] * W[2][i][I % L];
    S[i][I] = v;
  }
  __syncthreads();
  const int lr = lane & 15, j = lane >> 4;
  float2 pacc[NB][4];
#pragma unroll
  for (int b = 0; b < NB; b++)
#pragma unroll
    for (int v = 0; v < 4; v++) pacc[b][v] = make_float2(0.f, 0.f);
  for (int k = wave; k < ns; k += 4) {
    const int sl = s0 + k;
    const int d = hf ? T.slotDir[sl] : 0;
    const float2* sg = sig2 + (size_t)(srow + d) * LL;
    const float2* Ls = Lst + (size_t)lr * lStride + (size_t)sl * LL;  // right-hand side lr
    mfma_v4f cre[NB], cim[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
      cre[b] = mfma_v4f{0.f, 0.f, 0.f, 0.f};
      cim[b] = mfma_v4f{0.f, 0.f, 0.f, 0.f};
    }
    // KU steps' loads of L and of the phase row are issued together, ahead
    // of their matrix-core steps (one load in flight per step would leave
    // the loop bound by the load latency)
    constexpr int KU = 8;
    for (int ks0 = 0; ks0 < KS; ks0 += KU) {
      float2 lv2[KU], sv[KU];
#pragma unroll
      for (int u = 0; u < KU; u++) {
        const int kk = 4 * (ks0 + u) + j;
        const bool on = ks0 + u < KS && kk < LL;
        lv2[u] = on ? Ls[kk] : make_float2(0.f, 0.f);
        sv[u] = (on && hf) ? sg[kk] : make_float2(1.f, 0.f);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < KU; u++) {
        if (ks0 + u >= KS) break;
        const int kk = 4 * (ks0 + u) + j;
        const float2 bv = hf ? cmul(lv2[u], sv[u]) : lv2[u];
#pragma unroll
        for (int b = 0; b < NB; b++) {
          const float a = S[16 * b + lr][kk];
          cre[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv.x, cre[b], 0, 0, 0);
          cim[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv.y, cim[b], 0, 0, 0);
        }
      }
    }
    // this lane holds C[point 16 b + 4 j + v][right-hand side lr]
    double dx = 0., dy = 0., dz = 0.;
    if (hf) {
      dx = dirs[3 * (doff + d)];
      dy = dirs[3 * (doff + d) + 1];
      dz = dirs[3 * (doff + d) + 2];
    }
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = 16 * b + 4 * j + v;
        float2 cv = make_float2(cre[b][v], cim[b][v]);
        if (hf) {
          float cs, sn;
          expi(kappa * (dx * rel[0][i] + dy * rel[1][i] + dz * rel[2][i]), cs, sn);
          cv = cmul(cv, make_float2(cs, sn));
        }
        pacc[b][v].x += cv.x;
        pacc[b][v].y += cv.y;
      }
  }
  __syncthreads();  // every wave is done with S
#pragma unroll
  for (int b = 0; b < NB; b++)
#pragma unroll
    for (int v = 0; v < 4; v++) red[wave][16 * b + 4 * j + v][lr] = pacc[b][v];
  __syncthreads();
  for (int x = tid; x < TP * kL2PMfmaNR; x += 256) {
    const int i = x / kL2PMfmaNR, r = x % kL2PMfmaNR;
    if (i < nt) {
      float2 t = red[0][i][r];
#pragma unroll
      for (int w = 1; w < 4; w++) {
        t.x += red[w][i][r].x;
        t.y += red[w][i][r].y;
      }
      float2* o = p + (size_t)r * pStride + pb + i0 + i;
      float2 ov = *o;
      ov.x += t.x;
      ov.y += t.y;
      *o = ov;
    }
  }
}

// P2P near field (P2P.hpp:42-79): p_i += sum_j exp(i kappa r)/(4 pi r) q_j,
// r = |x_i - y_j|, term 0 when r^2 < 1e-16.  A work item is up to 64 targets
// of one leaf against a run of its source list; with nt targets, the wave's
// lanes form floor(64/nt) groups that split the sources.  Sources are staged
// in LDS 64 at a time.  Each item writes per-target partial sums (double).
struct P2PItem {
  int leaf;  // index into the leaf list
  int t0;    // first target of the chunk (leaf-relative)
  int r;     // first source range
  int off;   // offset into that range
  int cnt;   // number of sources
};

// Kernel value G = exp(i kappa r) / (4 pi r) between a target and a source
// (P2P.hpp:61-75), 0 when r^2 < 1e-16.
// Small: kappa * r <= 2^20 for every pair (checked on the host), so the
// double sine and cosine need no range check.
template <typename R, bool Small>
__device__ __forceinline__ void p2pKernel(double xt, double yt, double zt, double xs, double ys,
                                          double zs, double kappa, R& co, R& si);
template <>
__device__ __forceinline__ void p2pKernel<double, true>(double xt, double yt, double zt,
                                                        double xs, double ys, double zs,
                                                        double kappa, double& co, double& si) {
  const double x = xt - xs, y = yt - ys, z = zt - zs;
  const double R2v = x * x + y * y + z * z;
  double K = rsqrt(R2v);
  if (R2v < 1.e-16) K = 0.;
  const double r = K * R2v * kappa;
  K *= 0.079577471545947667884;  // 1/(4 pi)
  sincosRedSmall(r, si, co);
  co *= K;
  si *= K;
}
template <>
__device__ __forceinline__ void p2pKernel<double, false>(double xt, double yt, double zt,
                                                         double xs, double ys, double zs,
                                                         double kappa, double& co, double& si) {
  const double x = xt - xs, y = yt - ys, z = zt - zs;
  const double R2v = x * x + y * y + z * z;
  double K = rsqrt(R2v);
  if (R2v < 1.e-16) K = 0.;
  const double r = K * R2v * kappa;
  K *= 0.079577471545947667884;  // 1/(4 pi)
  sincosRed(r, si, co);
  co *= K;
  si *= K;
}
template <>
__device__ __forceinline__ void p2pKernel<float, true>(double xt, double yt, double zt,
                                                       double xs, double ys, double zs,
                                                       double kappa, float& co, float& si);
template <>
__device__ __forceinline__ void p2pKernel<float, false>(double xt, double yt, double zt,
                                                        double xs, double ys, double zs,
                                                        double kappa, float& co, float& si) {
  const double x = xt - xs, y = yt - ys, z = zt - zs;
  const double R2v = x * x + y * y + z * z;
  // r to ~1e-14 relative: float reciprocal square root, one Newton step in double
  const float y0 = (R2v < 1.e-16) ? 0.f : rsqrtf((float)R2v);
  const double r0 = R2v * (double)y0;
  const double r = fma(0.5 * (double)y0, fma(-r0, r0, R2v), r0);
  const float K = y0 * 0.079577471545947667884f;
  expi(kappa * r, co, si);
  co *= K;
  si *= K;
}
template <>
__device__ __forceinline__ void p2pKernel<float, true>(double xt, double yt, double zt,
                                                       double xs, double ys, double zs,
                                                       double kappa, float& co, float& si) {
  const double x = xt - xs, y = yt - ys, z = zt - zs;
  const double R2v = x * x + y * y + z * z;
  const float y0 = (R2v < 1.e-16) ? 0.f : rsqrtf((float)R2v);
  const double r0 = R2v * (double)y0;
  const double r = fma(0.5 * (double)y0, fma(-r0, r0, R2v), r0);
  const float K = y0 * 0.079577471545947667884f;
  sincosRedF(kappa * r, si, co);
  co *= K;
  si *= K;
}

// NR right-hand sides: charges q + r * qStride; the partial sums of item k,
// right-hand side r, lane t at partial[(k * NR + r) * 64 + t].
template <typename R, int NR>
__global__ __launch_bounds__(64) void k_p2p(const P2PItem* __restrict__ items,
                                            const int* __restrict__ leafCell,
                                            const int* __restrict__ tpbeg,
                                            const int* __restrict__ tnprt,
                                            const double* __restrict__ tx,
                                            const double* __restrict__ ty,
                                            const double* __restrict__ tz,
                                            const double* __restrict__ sx,
                                            const double* __restrict__ sy,
                                            const double* __restrict__ sz,
                                            const typename C2T<R>::T* __restrict__ q,
                                            size_t qStride,
                                            const int* __restrict__ rb, const int* __restrict__ re,
                                            double kappa, int smallArg,
                                            double2* __restrict__ partial, int nItems) {
  typedef typename C2T<R>::T R2;
  __shared__ double lx[64], ly[64], lz[64];
  __shared__ R2 lq[NR][64];
  __shared__ double2 red[NR][64];
  const int item = itemIndex();
  if (item >= nItems) return;
  const P2PItem it = items[item];
  const int c = leafCell[it.leaf];
  const int tb = tpbeg[c] + it.t0;
  const int nt = min(64, tnprt[c] - it.t0);
  const int G = 64 / nt;
  const int lane = threadIdx.x;
  const int ti = lane % nt, g = lane / nt;
  const bool active = g < G;
  const double xt = tx[tb + ti], yt = ty[tb + ti], zt = tz[tb + ti];
  double accr[NR], acci[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) accr[k] = acci[k] = 0.;
  int r = it.r, off = it.off;
  for (int pos = 0; pos < it.cnt; pos += 64) {
    const int tileN = min(64, it.cnt - pos);
    __syncthreads();
    if (lane < tileN) {
      int rr = r, oo = off + lane;
      int len = re[rr] - rb[rr];
      while (oo >= len) {
        oo -= len;
        rr++;
        len = re[rr] - rb[rr];
      }
      const int j = rb[rr] + oo;
      lx[lane] = sx[j];
      ly[lane] = sy[j];
      lz[lane] = sz[j];
#pragma unroll
      for (int k = 0; k < NR; k++) lq[k][lane] = q[k * qStride + j];
    }
    {
      int adv = tileN;
      while (adv > 0) {
        const int rem = re[r] - rb[r] - off;
        if (adv < rem) {
          off += adv;
          adv = 0;
        } else {
          adv -= rem;
          r++;
          off = 0;
        }
      }
    }
    __syncthreads();
    if (active) {
      R pr[NR], pi[NR];
#pragma unroll
      for (int k = 0; k < NR; k++) pr[k] = pi[k] = 0;
      auto term = [&](int j, auto small) {
        R co, si;
        p2pKernel<R, decltype(small)::value>(xt, yt, zt, lx[j], ly[j], lz[j], kappa, co, si);
#pragma unroll
        for (int k = 0; k < NR; k++) {
          const R2 qs = lq[k][j];
          pr[k] += co * qs.x - si * qs.y;
          pi[k] += co * qs.y + si * qs.x;
        }
      };
      if (smallArg) {
#pragma unroll 2
        for (int j = g; j < tileN; j += G) term(j, std::true_type());
      } else {
        for (int j = g; j < tileN; j += G) term(j, std::false_type());
      }
#pragma unroll
      for (int k = 0; k < NR; k++) {
        accr[k] += (double)pr[k];
        acci[k] += (double)pi[k];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NR; k++) red[k][lane] = make_double2(accr[k], acci[k]);
  __syncthreads();
  if (lane < nt) {
#pragma unroll
    for (int k = 0; k < NR; k++) {
      double2 s = red[k][lane];
      for (int gg = 1; gg < G; gg++) {
        s.x += red[k][gg * nt + lane].x;
        s.y += red[k][gg * nt + lane].y;
      }
      partial[((size_t)item * NR + k) * 64 + lane] = s;
    }
  }
}

// P2P for 16 right-hand sides per launch on the matrix cores (float, the
// mixed-precision path).  An item's near-field sum is the product of the
// kernel matrix G (targets x sources, complex, evaluated by the VALU one
// element per lane, double phase reduction as in k_p2p) and the charge matrix
// Q (sources x 16 right-hand sides):
//   P_re += G_re Q_re - G_im Q_im,   P_im += G_re Q_im + G_im Q_re,
// four V_MFMA_F32_16X16X4_F32 per 4 sources.  Operand layouts of that
// instruction: A (16 x 4) lane l holds row l % 16, column l / 16; B (4 x 16)
// lane l holds row l / 16, column l % 16; D (16 x 16) lane l holds column
// l % 16, rows 4 (l / 16) + 0..3.  So lane l evaluates G for target
// 16 w + l % 16 (wave w of the block's 4) and source 4 k + l / 16 of the
// current K-step k, and ends holding right-hand side l % 16 of targets
// 16 w + 4 (l / 16) + 0..3.  Sources are staged 64 at a time in LDS (the
// charges as [source][right-hand side]); each 64-source tile's sums are added
// to double accumulators (as k_p2p does) and the item's partial sums go to
// the layout k_p2p_reduce<float, 16> reads.
constexpr int kP2PMfmaNR = 16;

template <bool Small>
__global__ __launch_bounds__(256) void k_p2p_mfma(
    const P2PItem* __restrict__ items, const int* __restrict__ leafCell,
    const int* __restrict__ tpbeg, const int* __restrict__ tnprt, const double* __restrict__ tx,
    const double* __restrict__ ty, const double* __restrict__ tz, const double* __restrict__ sx,
    const double* __restrict__ sy, const double* __restrict__ sz, const float2* __restrict__ q,
    size_t qStride, const int* __restrict__ rb, const int* __restrict__ re, double kappa,
    double2* __restrict__ partial, int nItems) {
  __shared__ double lx[64], ly[64], lz[64];
  __shared__ int lj[64];
  __shared__ float2 lq[64][kP2PMfmaNR];
  const int item = itemIndex();
  if (item >= nItems) return;
  const P2PItem it = items[item];
  const int c = leafCell[it.leaf];
  const int tb = tpbeg[c] + it.t0;
  const int nt = min(64, tnprt[c] - it.t0);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int ti = wave * 16 + (lane & 15);  // target of this lane's A row
  const bool tok = ti < nt;
  const bool waveOn = wave * 16 < nt;      // uniform over the wave
  double xt = 0., yt = 0., zt = 0.;
  if (tok) {
    xt = tx[tb + ti];
    yt = ty[tb + ti];
    zt = tz[tb + ti];
  }
  const int kq = lane >> 4, rq = lane & 15;
  mfma_v4f dre = {0.f, 0.f, 0.f, 0.f}, dim = {0.f, 0.f, 0.f, 0.f};
  double accr[4] = {0., 0., 0., 0.}, acci[4] = {0., 0., 0., 0.};
  int r = it.r, off = it.off;
  for (int pos = 0; pos < it.cnt; pos += 64) {
    const int tileN = min(64, it.cnt - pos);
    __syncthreads();
    if (tid < 64) {
      int j = -1;
      if (tid < tileN) {
        int rr = r, oo = off + tid;
        int len = re[rr] - rb[rr];
        while (oo >= len) {
          oo -= len;
          rr++;
          len = re[rr] - rb[rr];
        }
        j = rb[rr] + oo;
        lx[tid] = sx[j];
        ly[tid] = sy[j];
        lz[tid] = sz[j];
      }
      lj[tid] = j;
    }
    {
      int adv = tileN;
      while (adv > 0) {
        const int rem = re[r] - rb[r] - off;
        if (adv < rem) {
          off += adv;
          adv = 0;
        } else {
          adv -= rem;
          r++;
          off = 0;
        }
      }
    }
    __syncthreads();
    for (int x = tid; x < 64 * kP2PMfmaNR; x += 256) {
      const int sj = x / kP2PMfmaNR, rr = x % kP2PMfmaNR;
      const int j = lj[sj];
      lq[sj][rr] = j >= 0 ? q[rr * qStride + j] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    if (waveOn) {
      for (int k = 0; k < 16; k++) {
        const int j = 4 * k + kq;
        float gr = 0.f, gi = 0.f;
        if (tok && j < tileN) p2pKernel<float, Small>(xt, yt, zt, lx[j], ly[j], lz[j], kappa, gr, gi);
        const float2 b = lq[j][rq];
        dre = __builtin_amdgcn_mfma_f32_16x16x4f32(gr, b.x, dre, 0, 0, 0);
        dre = __builtin_amdgcn_mfma_f32_16x16x4f32(-gi, b.y, dre, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f32_16x16x4f32(gr, b.y, dim, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f32_16x16x4f32(gi, b.x, dim, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; v++) {
        accr[v] += (double)dre[v];
        acci[v] += (double)dim[v];
        dre[v] = 0.f;
        dim[v] = 0.f;
      }
    }
  }
  if (waveOn)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int t = wave * 16 + 4 * kq + v;
      if (t < nt)
        partial[((size_t)item * kP2PMfmaNR + rq) * 64 + t] = make_double2(accr[v], acci[v]);
    }
}

// Sum the partial results of a leaf's items into p (items of a leaf are
// contiguous, ordered by target chunk).
// NR right-hand sides (partial layout of k_p2p<R, NR>); right-hand side r
// adds into p + r * pStride.
template <typename R, int NR>
__global__ __launch_bounds__(64) void k_p2p_reduce(const int* __restrict__ itemBeg,
                                                   const P2PItem* __restrict__ items,
                                                   const int* __restrict__ leafCell,
                                                   const int* __restrict__ tpbeg,
                                                   const int* __restrict__ tnprt,
                                                   const double2* __restrict__ partial,
                                                   typename C2T<R>::T* __restrict__ p,
                                                   size_t pStride, int nItems) {
  typedef typename C2T<R>::T R2;
  const int leaf = itemIndex();
  if (leaf >= nItems) return;
  const int c = leafCell[leaf];
  const int pb = tpbeg[c], n = tnprt[c];
  const int ib = itemBeg[leaf], ie = itemBeg[leaf + 1];
  for (int t0 = 0; t0 < n; t0 += 64) {
    const int i = t0 + threadIdx.x;
    double sr[NR], si[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) sr[r] = si[r] = 0.;
    for (int k = ib; k < ie; k++) {
      if (items[k].t0 != t0) continue;
#pragma unroll
      for (int r = 0; r < NR; r++) {
        const double2 v = partial[((size_t)k * NR + r) * 64 + threadIdx.x];
        sr[r] += v.x;
        si[r] += v.y;
      }
    }
    if (i < n) {
#pragma unroll
      for (int r = 0; r < NR; r++) {
        R2 o = p[r * pStride + pb + i];
        o.x = (R)((double)o.x + sr[r]);
        o.y = (R)((double)o.y + si[r]);
        p[r * pStride + pb + i] = o;
      }
    }
  }
}

// ----------------------------------------------------------------------------
// executor
// ----------------------------------------------------------------------------

// A failed call is reported here and its error consumed: HIP also keeps it
// as the thread's last error, which the next launch check
// (hipGetLastError) would otherwise report again, e.g. after an allocation
// that did not fit and whose previous settings were restored.
#define HIPCK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      (void)hipGetLastError();                                                \
      g_err = std::string(#x) + ": " + hipGetErrorString(e_);                 \
      return e_ == hipErrorOutOfMemory ? ERR_MEMORY : ERR_HIP;                \
    }                                                                         \
  } while (0)

struct DevTrans {
  int n = 0;  // output slots
  int level = 0, hf = 0;
  int *out = nullptr, *beg = nullptr, *in = nullptr, *sgn = nullptr, *sigIdx = nullptr;
  double2* edv = nullptr;
};

// k_trans_lines where it is the faster one (measured: float at every order,
// double from L = 6; k_trans for double at L <= 5)
template <typename R, int L>
// One level's translations for nr right-hand sides in one launch (right-hand
// side r of src and dst at r * rStride).
static void launchTrans(const DevTrans& t, const typename C2T<R>::T* src,
                        typename C2T<R>::T* dst, const typename C2T<R>::T* pre,
                        const typename C2T<R>::T* post, const R* m0, const R* m1, int accumulate,
                        size_t rStride, int nr, hipStream_t s) {
  // (output, right-hand side) pairs are counted in int: chunks of right-hand sides
  const int per = (int)std::max<int64_t>(1, std::min<int64_t>(nr, INT32_MAX / std::max(1, t.n)));
  for (int r0 = 0; r0 < nr; r0 += per) {
    const int c = std::min(per, nr - r0);
    const int64_t n = (int64_t)t.n * c;
    const typename C2T<R>::T* sr = src + r0 * rStride;
    typename C2T<R>::T* dr = dst + r0 * rStride;
    if (sizeof(R) == 4 || L >= 6)
      hipLaunchKernelGGL((k_trans_lines<R, L>), itemGrid(n), dim3(transLinesTB<L>()), 0, s,
                         t.out, t.beg, t.in, t.sgn, t.sigIdx, t.edv, sr, dr, pre, post, m0, m1,
                         accumulate, t.n, rStride, c);
    else
      hipLaunchKernelGGL((k_trans<R, L>), itemGrid(n), dim3(128), 0, s, t.out, t.beg, t.in,
                         t.sgn, t.sigIdx, t.edv, sr, dr, pre, post, m0, m1, accumulate, t.n,
                         rStride, c);
  }
}

struct Exec {
  int prec = 0;
  int L = 0, P = 0, LL = 0, PP = 0;
  double kappa = 0;
  int64_t ns = 0, nt = 0;
  Stats st;
  int device = 0;
  std::vector<void*> mem;
  int64_t bytes = 0;
  hipStream_t own = nullptr;
  hipEvent_t ev[DEFMM_NSTAGES + 1] = {};
  bool evRecorded = false;

  double *sx = nullptr, *sy = nullptr, *sz = nullptr, *tx = nullptr, *ty = nullptr, *tz = nullptr;
  int *sperm = nullptr, *tperm = nullptr;
  DevCells S{}, T{};
  DevLevels lv{};
  double* dirs = nullptr;
  void *sig = nullptr, *sig2 = nullptr, *U = nullptr;
  int* perm = nullptr;
  void* consts = nullptr;  // xi, iden, A0, A1, B0, B1, twc, tws (real type)
  int *p2m = nullptr, *l2p = nullptr;
  int np2m = 0, nl2p = 0;
  int2* p2mItems = nullptr;  // (source leaf, first slot of a pass)
  int np2mItems = 0;
  // matrix-core P2M: (leaf, first slot of a range of 32) for leaves of <= 64
  // points, and the vector-unit items of the larger leaves
  int2 *p2mMfmaItems = nullptr, *p2mBigItems = nullptr;
  int np2mMfma = 0, np2mBig = 0;
  bool p2mMfma = true;       // float P2M in chunks of 16 right-hand sides on the matrix cores
  bool m2lSliced = false;    // grouped M2L sweeps its entries once per frequency slice
  // (target leaf, first point) of the matrix-core L2P's point tiles, per
  // tile size (16 or 32 points: the order's l2pMfmaTile)
  int2* l2pTiles = nullptr;
  int nl2pTiles = 0;
  bool l2pMfma = true;       // float L2P in chunks of 16 right-hand sides on the matrix cores
  std::vector<DevTrans> m2m, l2l;
  int *m2fSlot = nullptr;
  int nm2f = 0;
  int *m2lZero = nullptr;
  int nM2LTargets = 0, nM2LZero = 0;
  int *gBeg = nullptr, *gTarget = nullptr, *gMask = nullptr, *gCode = nullptr;
  // per group entry, the row it reads: a global F row (stored), a source
  // slot (on the fly) or a workspace row of its batch (batched)
  int* gSrc = nullptr;
  int nGroups = 0;
  int m2lG = 0;
  int nTslots = 0, nSslots = 0;
  P2PItem* items = nullptr;
  int nitems = 0;
  int *leafCell = nullptr, *itemBeg = nullptr, *rb = nullptr, *re = nullptr;
  int nleaves = 0;
  double2* partial = nullptr;
  void *q = nullptr, *p = nullptr, *ML = nullptr, *F = nullptr;
  void* Lb = nullptr;        // local expansions (aliases ML when F is stored)
  size_t mStride = 0, lStride = 0, fStride = 0;  // elements between right-hand sides
  int m2lMode = DEFMM_M2L_STORED;
  int smallArg = 0;          // kappa * (largest point distance) <= 2^20
  bool p2pMfma = true;       // float P2P in chunks of 16 right-hand sides on the matrix cores
  bool m2lAuto = true;       // mode chosen by free memory (until set explicitly)
  int nr = 1;                // right-hand sides the expansion stores hold
  int batchRows = 16384;     // requested workspace rows of the batched M2L
  int wsRows = 0;            // rows of the workspace in use
  int* bRows = nullptr;      // batched M2L: source slots of the batches' rows
  std::vector<int> hBatchGroupBeg, hBatchRowBeg;
  // host copies of the grouped entries' global F rows, the F-row -> source
  // slot map and the group bounds (the mode-dependent lists are built from them)
  ListVec hGSrcF;
  std::vector<int> hM2fSlot, hGroupBeg;
  double2 *qio = nullptr, *pio = nullptr;
  double stageMs[DEFMM_NSTAGES] = {};
  // host copies needed to re-upload the typed buffers in another precision
  std::vector<cplx> hSig, hSig2, hTab;
  std::vector<int> hRowKey;  // per table-store row: (table * 6 + axis order) * 8 + flips
  std::vector<double> hA[2], hB[2];
  size_t nMLslots = 0;
  std::vector<void*> typedMem;
  int64_t typedBytes = 0;
  bool typedPhase = false;  // allocations go to typedMem
  // The typed stores (expansions, tables, charges, potentials) are complete:
  // cleared before they are released, set only after a successful upload.
  // Executions refuse to run while it is false.
  bool valid = false;
  // Completion of the last execution and the stream it ran on: an execution
  // on another stream waits for it (every execution shares the stores), and
  // the setters and destroy wait for it before releasing memory.
  hipEvent_t done = nullptr;
  hipStream_t lastStream = nullptr;
  bool doneRecorded = false;
  // set while the pass is captured into a graph (defmm_batch_create): no
  // events are recorded or waited on inside the capture
  bool capturing = false;
  // incremented by every successful upload of the typed stores: a captured
  // graph holds their addresses and is stale once it changes
  uint64_t generation = 0;

  void freeTyped() {
    valid = false;
    for (void* m : typedMem) (void)hipFree(m);
    typedMem.clear();
    bytes -= typedBytes;
    typedBytes = 0;
    sig = sig2 = U = consts = nullptr;
    ML = F = Lb = q = p = nullptr;
    gSrc = bRows = nullptr;
    partial = qio = pio = nullptr;
  }
  // wait until the last execution, on whatever stream, has finished
  int quiesce() {
    if (hipSetDevice(device) != hipSuccess) return ERR_HIP;
    if (doneRecorded && hipEventSynchronize(done) != hipSuccess) return ERR_HIP;
    if (own && hipStreamSynchronize(own) != hipSuccess) return ERR_HIP;
    return OK;
  }
  ~Exec() {
    (void)quiesce();
    for (void* m : typedMem) (void)hipFree(m);
    for (void* m : mem) (void)hipFree(m);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (done) (void)hipEventDestroy(done);
    if (own) (void)hipStreamDestroy(own);
  }

  // device memory budget (bytes, 0: none), checked by every allocation
  int64_t memLimit = 0;
  int overBudget(size_t nbytes) {
    if (memLimit > 0 && bytes + (int64_t)nbytes > memLimit) {
      g_err = "the plan's device memory limit (" + std::to_string(memLimit) +
              " bytes) would be exceeded";
      return ERR_MEMORY;
    }
    return OK;
  }
  template <class X>
  int up(X*& dst, const X* src, size_t n) {
    dst = nullptr;
    if (n == 0) n = 1;
    void* d = nullptr;
    if (overBudget(n * sizeof(X))) return ERR_MEMORY;
    HIPCK(hipMalloc(&d, n * sizeof(X)));
    (typedPhase ? typedMem : mem).push_back(d);
    bytes += (int64_t)(n * sizeof(X));
    if (typedPhase) typedBytes += (int64_t)(n * sizeof(X));
    dst = (X*)d;
    if (src) HIPCK(copyLarge(d, src, n * sizeof(X), true));
    return OK;
  }
  template <class X, class A>
  int upv(X*& dst, const std::vector<X, A>& v) {
    return up(dst, v.empty() ? nullptr : v.data(), v.size());
  }
  int alloc(void*& dst, size_t nbytes) {
    void* d = nullptr;
    if (nbytes == 0) nbytes = 16;
    if (overBudget(nbytes)) return ERR_MEMORY;
    HIPCK(hipMalloc(&d, nbytes));
    (typedPhase ? typedMem : mem).push_back(d);
    bytes += (int64_t)nbytes;
    if (typedPhase) typedBytes += (int64_t)nbytes;
    dst = d;
    return OK;
  }
};

template <typename R>
static typename C2T<R>::T c2h(double a, double b) {
  typename C2T<R>::T r;
  r.x = (R)a;
  r.y = (R)b;
  return r;
}

template <typename R>
static std::vector<typename C2T<R>::T> toR2(const std::vector<cplx>& v) {
  std::vector<typename C2T<R>::T> o(v.size());
  for (size_t i = 0; i < v.size(); i++) {
    o[i].x = (R)v[i].real();
    o[i].y = (R)v[i].imag();
  }
  return o;
}

#define CKR(x)               \
  do {                       \
    int r_ = (x);            \
    if (r_ != OK) return r_; \
  } while (0)

// Batched M2L: cut the groups, in processing order, into batches whose
// distinct source F rows number at most wsRows; per batch, the rows (source
// slots to transform) in first-use order and, per group entry, its row in
// the batch.  A group is never split.
static int buildBatches(Exec& X, std::vector<int>& local, std::vector<int>& rows) {
  const std::vector<int>& gb = X.hGroupBeg;
  const int ng = (int)gb.size() - 1;
  int maxRows = 1;
  for (int g = 0; g < ng; g++) maxRows = std::max(maxRows, gb[g + 1] - gb[g]);
  X.wsRows = std::max(X.batchRows, maxRows);
  const int W = X.wsRows;
  local.assign(X.hGSrcF.size(), -1);
  rows.clear();
  X.hBatchGroupBeg.assign(1, 0);
  X.hBatchRowBeg.assign(1, 0);
  // inBatch[r]: last batch that holds F row r, rowOf[r]: its row there;
  // stamp[r]: last group that counted r
  std::vector<int> inBatch(X.hM2fSlot.size(), -1), rowOf(X.hM2fSlot.size(), -1),
      stamp(X.hM2fSlot.size(), -1);
  int b = 0, nIn = 0;
  for (int g = 0; g < ng; g++) {
    int fresh = 0;
    for (int e = gb[g]; e < gb[g + 1]; e++) {
      int r = X.hGSrcF[e];
      if (inBatch[r] != b && stamp[r] != g) { stamp[r] = g; fresh++; }
    }
    if (nIn > 0 && nIn + fresh > W) {
      X.hBatchGroupBeg.push_back(g);
      X.hBatchRowBeg.push_back((int)rows.size());
      b++;
      nIn = 0;
    }
    for (int e = gb[g]; e < gb[g + 1]; e++) {
      int r = X.hGSrcF[e];
      if (inBatch[r] != b) {
        inBatch[r] = b;
        rowOf[r] = nIn++;
        rows.push_back(X.hM2fSlot[r]);
      }
      local[e] = rowOf[r];
    }
  }
  if (ng > 0) {
    X.hBatchGroupBeg.push_back(ng);
    X.hBatchRowBeg.push_back((int)rows.size());
  }
  // checks: every entry has a row inside its batch, every batch fits
  for (size_t k = 0; k + 1 < X.hBatchGroupBeg.size(); k++) {
    int r0 = X.hBatchRowBeg[k], r1 = X.hBatchRowBeg[k + 1];
    if (r1 - r0 > W || r1 < r0) { g_err = "internal: M2L batch exceeds its workspace"; return ERR_STATE; }
    for (int e = gb[X.hBatchGroupBeg[k]]; e < gb[X.hBatchGroupBeg[k + 1]]; e++)
      if (local[e] < 0 || local[e] >= r1 - r0 || rows[r0 + local[e]] != X.hM2fSlot[X.hGSrcF[e]]) {
        g_err = "internal: M2L batch row out of range";
        return ERR_STATE;
      }
  }
  X.st.m2l_batches = (int64_t)X.hBatchGroupBeg.size() - 1;
  X.st.m2l_batch_rows = (int64_t)rows.size();
  return OK;
}

// Rows of the M2L table store from the kernel tables (set-up, on the GPU):
// row k has key (t * 6 + s) * 8 + z: table t, axis order s, flips z;
//   row[i] = U_t[(a_{s0} * P + a_{s1}) * P + a_{s2}],  i = (i0 * P + i1) * P + i2,
// with a_j = P - i_j where bit j of z is set and i_j != 0, else i_j
// (the index permutation of permutations.hpp:127-150 applied once per row).
__constant__ int kTabAxisOrder[6][3] = {{0, 1, 2}, {1, 0, 2}, {1, 2, 0},
                                        {0, 2, 1}, {2, 0, 1}, {2, 1, 0}};
template <typename R2>
__global__ void k_table_rows(const int* __restrict__ keys, int64_t nrows,
                             const R2* __restrict__ tab, int P, R2* __restrict__ rows) {
  const int64_t PP = (int64_t)P * P * P, n = nrows * PP;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride) {
    const int64_t k = x / PP;
    const int i = (int)(x - k * PP);
    const int key = keys[k];
    const int t = key / 48, s = (key / 8) % 6, z = key % 8;
    int a[3] = {i / (P * P), (i / P) % P, i % P};
#pragma unroll
    for (int j = 0; j < 3; j++)
      if (((z >> j) & 1) && a[j]) a[j] = P - a[j];
    const int src = (a[kTabAxisOrder[s][0]] * P + a[kTabAxisOrder[s][1]]) * P +
                    a[kTabAxisOrder[s][2]];
    rows[x] = tab[(int64_t)t * PP + src];
  }
}

template <typename R>
static int uploadTyped(Exec& X) {
  typedef typename C2T<R>::T R2;
  const int L = X.L, P = X.P;
  X.typedPhase = true;
  {
    auto a = toR2<R>(X.hSig);
    R2* d;
    CKR(X.upv(d, a));
    X.sig = d;
    a = toR2<R>(X.hSig2);
    CKR(X.upv(d, a));
    X.sig2 = d;
    // one row per (table, axis order, flips) in use, expanded on the GPU
    // from the kernel tables (k_table_rows); a plan without M2L keeps one
    // zero row that is never read
    const int PP = X.PP;
    const size_t ntab = X.hTab.size() / (size_t)PP;
    const size_t nrows = X.hRowKey.size();
    for (int key : X.hRowKey)
      if (ntab > 0 && (key < 0 || (size_t)(key / 48) >= ntab)) {
        g_err = "internal: M2L table id out of range";
        return ERR_STATE;
      }
    void* rows = nullptr;
    CKR(X.alloc(rows, nrows * (size_t)PP * sizeof(R2)));
    X.U = rows;
    if (ntab == 0) {
      HIPCK(hipMemset(rows, 0, nrows * (size_t)PP * sizeof(R2)));
    } else {
      // the kernel tables and the keys are only needed here
      std::vector<R2> tab = toR2<R>(X.hTab);
      R2* dTab = nullptr;
      int* dKeys = nullptr;
      if (X.overBudget(tab.size() * sizeof(R2) + nrows * sizeof(int))) return ERR_MEMORY;
      HIPCK(hipMalloc(&dTab, tab.size() * sizeof(R2)));
      int rc = OK;
      if (hipMalloc(&dKeys, nrows * sizeof(int)) != hipSuccess ||
          hipMemcpy(dTab, tab.data(), tab.size() * sizeof(R2), hipMemcpyHostToDevice) !=
              hipSuccess ||
          hipMemcpy(dKeys, X.hRowKey.data(), nrows * sizeof(int), hipMemcpyHostToDevice) !=
              hipSuccess) {
        (void)hipGetLastError();
        g_err = "M2L table upload failed";
        rc = ERR_HIP;
      } else {
        const int64_t n = (int64_t)nrows * PP;
        hipLaunchKernelGGL(k_table_rows<R2>, dim3(zeroGrid(n, 256)), dim3(256), 0, 0, dKeys,
                           (int64_t)nrows, dTab, X.P, (R2*)rows);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
          g_err = "M2L table rows kernel failed";
          rc = ERR_HIP;
        }
      }
      (void)hipFree(dKeys);
      (void)hipFree(dTab);
      if (rc != OK) return rc;
    }
  }
  {
    std::vector<R> c;
    std::vector<double> xi(L), iden(L);
    for (int i = 0; i < L; i++) xi[i] = -1. + 2. * ((double)i / (double)(L - 1));
    for (int i = 0; i < L; i++) {
      double pr = 1.;
      for (int j = 0; j < L; j++)
        if (j != i) pr *= (xi[i] - xi[j]);
      iden[i] = 1. / pr;
    }
    for (double v : xi) c.push_back((R)v);
    for (double v : iden) c.push_back((R)v);
    for (int s = 0; s < 2; s++)
      for (double v : X.hA[s]) c.push_back((R)v);
    for (int s = 0; s < 2; s++)
      for (double v : X.hB[s]) c.push_back((R)v);
    for (int m = 0; m < P; m++) c.push_back((R)std::cos(2. * M_PI * (double)m / (double)P));
    for (int m = 0; m < P; m++) c.push_back((R)std::sin(2. * M_PI * (double)m / (double)P));
    R* d;
    CKR(X.upv(d, c));
    X.consts = d;
  }
  // right-hand side r of every store at base + r * stride
  const size_t nr = (size_t)X.nr;
  if (X.m2lMode == DEFMM_M2L_STORED) {
    // M is dead after M2F and every target slot's L is written by M2L or
    // zeroed: one buffer holds M, then L
    X.mStride = X.lStride = X.nMLslots * X.LL;
    X.fStride = (size_t)std::max(1, X.nm2f) * X.PP;
    CKR(X.alloc(X.ML, nr * X.mStride * sizeof(R2)));
    CKR(X.alloc(X.F, nr * X.fStride * sizeof(R2)));
    X.Lb = X.ML;
    CKR(X.upv(X.gSrc, X.hGSrcF));
  } else {
    // M2L reads M while writing L: two buffers
    X.mStride = (size_t)std::max(1, X.nSslots) * X.LL;
    X.lStride = (size_t)std::max(1, X.nTslots) * X.LL;
    CKR(X.alloc(X.ML, nr * X.mStride * sizeof(R2)));
    CKR(X.alloc(X.Lb, nr * X.lStride * sizeof(R2)));
    X.F = nullptr;
    if (X.m2lMode == DEFMM_M2L_ON_THE_FLY) {
      std::vector<int> ss(X.hGSrcF.size());
      for (size_t e = 0; e < ss.size(); e++) ss[e] = X.hM2fSlot[X.hGSrcF[e]];
      CKR(X.upv(X.gSrc, ss));
    } else {
      std::vector<int> local, rows;
      CKR(buildBatches(X, local, rows));
      CKR(X.upv(X.gSrc, local));
      CKR(X.upv(X.bRows, rows));
      // the workspace holds the right-hand sides of one M2L launch
      X.fStride = (size_t)X.wsRows * X.PP;
      const size_t nw = std::min(nr, (size_t)m2lMaxNRAt(X.prec, X.L));
      CKR(X.alloc(X.F, nw * X.fStride * sizeof(R2)));
    }
  }
  CKR(X.alloc(X.q, nr * (size_t)X.ns * sizeof(R2)));
  CKR(X.alloc(X.p, nr * (size_t)X.nt * sizeof(R2)));
  {
    void* d;
    CKR(X.alloc(d, nr * (size_t)std::max(1, X.nitems) * 64 * sizeof(double2)));
    X.partial = (double2*)d;
    CKR(X.alloc(d, nr * (size_t)X.ns * sizeof(double2)));
    X.qio = (double2*)d;
    CKR(X.alloc(d, nr * (size_t)X.nt * sizeof(double2)));
    X.pio = (double2*)d;
  }
  X.typedPhase = false;
  return OK;
}

// Default M2L variant: stored F rows (one M2F and one M2L launch, the
// faster variant) when they fit in 90% of the free device memory with the
// shared M/L store for all right-hand sides, else cache-sized batches
// (separate M and L stores).  Call with the typed stores released.
static int chooseM2LMode(Exec& X) {
  if (!X.m2lAuto) return OK;
  size_t freeB = 0, totalB = 0;
  if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) {
    g_err = "hipMemGetInfo failed";
    return ERR_HIP;
  }
  if (X.memLimit > 0)
    freeB = std::min(freeB, (size_t)std::max<int64_t>(0, X.memLimit - X.bytes));
  const size_t r2 = X.prec == DEFMM_FP64 ? 16 : 8;
  const size_t storedB =
      (size_t)X.nr * (((size_t)X.nm2f * X.PP + X.nMLslots * X.LL) * r2 + ((size_t)X.ns + X.nt) * (r2 + 16) +
                      (size_t)X.nitems * 64 * 16) +
      (size_t)X.hGSrcF.size() * sizeof(int);
  X.m2lMode = (double)storedB < 0.9 * (double)freeB ? DEFMM_M2L_STORED : DEFMM_M2L_BATCHED;
  return OK;
}

static int uploadPrecision(Exec& X) {
  X.freeTyped();
  int rc = X.prec == DEFMM_FP64 ? uploadTyped<double>(X) : uploadTyped<float>(X);
  X.typedPhase = false;
  if (rc != OK) {
    X.freeTyped();  // a partial upload leaves nothing behind
    return rc;
  }
  X.valid = true;
  X.generation++;
  return OK;
}

// Re-upload the typed stores after a change of width, precision or M2L
// variant.  On failure the previous settings are restored and uploaded again,
// so that the plan stays usable; if that fails too, the plan stays invalid
// and executions return DEFMM_ESTATE.
static int reupload(Exec& X, int oldNr, int oldPrec, int oldMode, bool oldAuto, int oldRows) {
  X.evRecorded = false;
  X.freeTyped();  // the variant is chosen by the memory free without them
  int rc = chooseM2LMode(X);
  if (rc == OK) rc = uploadPrecision(X);
  if (rc == OK) return OK;
  const std::string why = g_err;
  X.nr = oldNr;
  X.prec = oldPrec;
  X.m2lMode = oldMode;
  X.m2lAuto = oldAuto;
  X.batchRows = oldRows;
  if (uploadPrecision(X) == OK)
    g_err = why + " (the plan keeps its previous settings)";
  else
    g_err = why + " (the previous settings could not be restored: the plan is unusable)";
  return rc;
}

static int upload(Exec& X, Plan& pl) {
  X.L = pl.L; X.P = pl.P; X.LL = pl.LL; X.PP = pl.PP; X.kappa = pl.kappa;
  {
    // every near-field phase kappa * r is below kappa times the diagonal of
    // the box holding all points
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (const Tree* t : {&pl.S, &pl.T})
      for (size_t i = 0; i < t->x.size(); i++) {
        const double v[3] = {t->x[i], t->y[i], t->z[i]};
        for (int d = 0; d < 3; d++) {
          lo[d] = std::min(lo[d], v[d]);
          hi[d] = std::max(hi[d], v[d]);
        }
      }
    double diag2 = 0.;
    for (int d = 0; d < 3; d++) diag2 += (hi[d] - lo[d]) * (hi[d] - lo[d]);
    X.smallArg = std::fabs(pl.kappa) * std::sqrt(diag2) * (1. + 1e-12) <= 1048576.0 ? 1 : 0;
  }
  X.st = pl.st;
  X.ns = (int64_t)pl.S.x.size();
  X.nt = (int64_t)pl.T.x.size();
  CKR(X.upv(X.sx, pl.S.x)); CKR(X.upv(X.sy, pl.S.y)); CKR(X.upv(X.sz, pl.S.z));
  CKR(X.upv(X.tx, pl.T.x)); CKR(X.upv(X.ty, pl.T.y)); CKR(X.upv(X.tz, pl.T.z));
  CKR(X.upv(X.sperm, pl.S.perm)); CKR(X.upv(X.tperm, pl.T.perm));
  for (int w = 0; w < 2; w++) {
    Tree& t = w ? pl.T : pl.S;
    DevCells& d = w ? X.T : X.S;
    int* ip;
    double* dp;
    CKR(X.upv(ip, t.lvl)); d.lvl = ip;
    CKR(X.upv(ip, t.pbeg)); d.pbeg = ip;
    CKR(X.upv(ip, t.nprt)); d.nprt = ip;
    CKR(X.upv(dp, t.ctr)); d.ctr = dp;
    CKR(X.upv(dp, t.rad)); d.rad = dp;
    CKR(X.upv(ip, t.slot0)); d.slot0 = ip;
    CKR(X.upv(ip, t.nslot)); d.nslot = ip;
    CKR(X.upv(ip, t.slotDir)); d.slotDir = ip;
  }
  X.nSslots = pl.S.nslots;
  X.nTslots = pl.T.nslots;
  {
    std::vector<int> dirOff(pl.nlvl, 0);
    std::vector<double> dd;
    std::vector<int> stageOff;
    for (size_t e = 0; e < pl.dirs.size(); e++) {
      stageOff.push_back((int)(dd.size() / 3));
      for (auto& v : pl.dirs[e]) { dd.push_back(v[0]); dd.push_back(v[1]); dd.push_back(v[2]); }
    }
    for (int l = 0; l < pl.nlvl; l++)
      if (pl.isHF[l] && pl.HFStage[l] < (int)stageOff.size()) dirOff[l] = stageOff[pl.HFStage[l]];
    int* ip;
    CKR(X.upv(X.dirs, dd));
    CKR(X.upv(ip, pl.isHF)); X.lv.isHF = ip;
    CKR(X.upv(ip, dirOff)); X.lv.dirOff = ip;
    CKR(X.upv(ip, pl.sigRow0)); X.lv.sigRow0 = ip;
  }
  CKR(X.upv(X.perm, pl.perm));
  CKR(X.upv(X.p2m, pl.p2mCell)); X.np2m = (int)pl.p2mCell.size();
  {
    // P2M work items: every pass of p2mSlotsPerPass(L) slots of a leaf
    const int C = p2mSlotsPerPass(pl.L);
    std::vector<int2> it;
    for (int c : pl.p2mCell)
      for (int k0 = 0; k0 < std::max(1, pl.S.nslot[c]); k0 += C) it.push_back(make_int2(c, k0));
    CKR(X.upv(X.p2mItems, it));
    X.np2mItems = (int)it.size();
    std::vector<int2> mf, big;
    for (int c : pl.p2mCell) {
      const int nsl = std::max(1, pl.S.nslot[c]);
      if (pl.S.nprt[c] <= kP2MMfmaPoints)
        for (int k0 = 0; k0 < nsl; k0 += kP2MMfmaSlots) mf.push_back(make_int2(c, k0));
      else
        for (int k0 = 0; k0 < nsl; k0 += C) big.push_back(make_int2(c, k0));
    }
    CKR(X.upv(X.p2mMfmaItems, mf));
    X.np2mMfma = (int)mf.size();
    CKR(X.upv(X.p2mBigItems, big));
    X.np2mBig = (int)big.size();
  }
  CKR(X.upv(X.l2p, pl.l2pCell)); X.nl2p = (int)pl.l2pCell.size();
  {
    // tiles of the matrix-core L2P: 32 points up to L^3 = 512, else 16
    const int tp = pl.LL <= 512 ? 32 : 16;
    std::vector<int2> it;
    for (int c : pl.l2pCell)
      for (int i0 = 0; i0 < std::max(1, pl.T.nprt[c]); i0 += tp) it.push_back(make_int2(c, i0));
    CKR(X.upv(X.l2pTiles, it));
    X.nl2pTiles = (int)it.size();
  }
  for (int w = 0; w < 2; w++) {
    auto& src = w ? pl.l2l : pl.m2m;
    auto& dst = w ? X.l2l : X.m2m;
    for (auto& tl : src) {
      DevTrans d;
      d.n = (int)tl.out.size();
      d.level = tl.level;
      d.hf = tl.hf;
      CKR(X.upv(d.out, tl.out)); CKR(X.upv(d.beg, tl.beg)); CKR(X.upv(d.in, tl.in));
      CKR(X.upv(d.sgn, tl.sgn)); CKR(X.upv(d.sigIdx, tl.sigIdx));
      // std::complex<double> is laid out as two doubles (real, imaginary)
      static_assert(sizeof(cplx) == sizeof(double2), "complex layout");
      CKR(X.up(d.edv, reinterpret_cast<const double2*>(tl.edv.data()), tl.edv.size()));
      dst.push_back(d);
    }
  }
  CKR(X.upv(X.m2fSlot, pl.m2fSlot)); X.nm2f = (int)pl.m2fSlot.size();
  // the M2L kernels read the grouped lists; the per-target CSR and the
  // processing order stay on the host
  X.nM2LTargets = (int)pl.m2lOrder.size();
  CKR(X.upv(X.m2lZero, pl.m2lZero)); X.nM2LZero = (int)pl.m2lZero.size();
  CKR(X.upv(X.gBeg, pl.m2lGroupBeg)); X.nGroups = (int)pl.m2lGroupBeg.size() - 1;
  X.m2lG = pl.m2lG;
  CKR(X.upv(X.gTarget, pl.m2lGroupTarget));
  CKR(X.upv(X.gMask, pl.m2lGMask));
  {
    // (table id << 6 | h) -> row of the materialised table store << 3: one
    // row per (table, axis order, flips) that some entry reads, in key order
    // (codes of absent group targets point at row 0, never read).  Marking
    // the keys in use and mapping the codes run in parallel over ranges of
    // entries; the result does not depend on the thread count.
    const size_t nkeys = (size_t)std::max(1, pl.ntab) * 48;
    const int G = pl.m2lG;
    const size_t ne = pl.m2lGMask.size();
    auto keyOf = [&](int c) {
      return ((size_t)(c >> 6) * 6 + kVariantOfSigma[(c >> 3) & 7]) * 8 + (c & 7);
    };
    const unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    auto parallel = [&](auto&& fn) {
      std::vector<std::thread> th;
      for (unsigned t = 0; t < nth; t++) th.emplace_back(fn, ne * t / nth, ne * (t + 1) / nth, t);
      for (auto& x : th) x.join();
    };
    // per-thread marks of the keys in use (no shared writes), then merged
    std::vector<std::vector<uint8_t>> used(nth);
    std::vector<int> bad(nth, 0);
    parallel([&](size_t e0, size_t e1, unsigned t) {
      used[t].assign(nkeys, 0);
      for (size_t e = e0; e < e1; e++)
        for (int g = 0; g < G; g++)
          if ((pl.m2lGMask[e] >> g) & 1) {
            const size_t key = keyOf(pl.m2lGCode[e * G + g]);
            if (key >= nkeys) bad[t] = 1;
            else used[t][key] = 1;
          }
    });
    for (int b : bad)
      if (b) {
        g_err = "internal: M2L table code out of range";
        return ERR_STATE;
      }
    std::vector<int> rowOf(nkeys, -1);
    X.hRowKey.clear();
    for (size_t k = 0; k < nkeys; k++) {
      bool u = false;
      for (unsigned t = 0; t < nth && !u; t++) u = used[t][k] != 0;
      if (u) {
        rowOf[k] = (int)X.hRowKey.size();
        X.hRowKey.push_back((int)k);
      }
    }
    std::vector<std::vector<uint8_t>>().swap(used);
    ListVec dc(pl.m2lGCode.size());  // every element written below
    parallel([&](size_t e0, size_t e1, unsigned) {
      for (size_t e = e0; e < e1; e++)
        for (int g = 0; g < G; g++)
          dc[e * G + g] = ((pl.m2lGMask[e] >> g) & 1) ? rowOf[keyOf(pl.m2lGCode[e * G + g])] << 3 : 0;
    });
    if (X.hRowKey.empty()) X.hRowKey.push_back(0);
    X.st.m2l_table_rows = (int64_t)X.hRowKey.size();
    CKR(X.upv(X.gCode, dc));
  }
  X.hGSrcF.swap(pl.m2lGSrcF);
  X.hM2fSlot = pl.m2fSlot;
  X.hGroupBeg = pl.m2lGroupBeg;
  // P2P work items: 64-target chunks x runs of <= budget sources.  The
  // budget is 4,096 sources unless that leaves fewer than 16,384 items (too
  // few waves to fill 1,024 SIMDs: the 20,000-point cube has 3,249); then it
  // is halved, down to 512.
  {
    std::vector<P2PItem> items;
    std::vector<int> itemBeg;
    const Tree& T = pl.T;
    auto build = [&](int kBudget) {
      items.clear();
      itemBeg.clear();
      for (size_t lf = 0; lf < pl.p2pLeaf.size(); lf++) {
        itemBeg.push_back((int)items.size());
        int c = pl.p2pLeaf[lf];
        for (int t0 = 0; t0 < T.nprt[c]; t0 += 64) {
          int r = pl.p2pBeg[lf], off = 0, cnt = 0, r0 = r, off0 = 0;
          for (; r < pl.p2pBeg[lf + 1]; r++) {
            int len = pl.p2pSrcE[r] - pl.p2pSrcB[r];
            off = 0;
            while (off < len) {
              int take = std::min(len - off, kBudget - cnt);
              if (cnt == 0) { r0 = r; off0 = off; }
              cnt += take;
              off += take;
              if (cnt == kBudget) {
                items.push_back({(int)lf, t0, r0, off0, cnt});
                cnt = 0;
              }
            }
          }
          if (cnt) items.push_back({(int)lf, t0, r0, off0, cnt});
        }
      }
      itemBeg.push_back((int)items.size());
    };
    for (int budget = 4096;; budget /= 2) {
      build(budget);
      if (items.size() >= 16384 || budget <= 512) break;
    }
    // the kernel walks ranges from (r, off) for cnt sources: they must stay
    // inside the item's leaf list
    for (const P2PItem& it : items) {
      int rEnd = pl.p2pBeg[it.leaf + 1];
      if (it.r < pl.p2pBeg[it.leaf] || it.r >= rEnd || it.cnt < 1 || it.t0 < 0 ||
          it.t0 >= T.nprt[pl.p2pLeaf[it.leaf]]) {
        g_err = "internal: P2P item out of range";
        return ERR_STATE;
      }
      int64_t avail = -(int64_t)it.off;
      for (int r = it.r; r < rEnd; r++) avail += pl.p2pSrcE[r] - pl.p2pSrcB[r];
      if (it.off < 0 || it.off >= pl.p2pSrcE[it.r] - pl.p2pSrcB[it.r] || it.cnt > avail) {
        g_err = "internal: P2P item exceeds its source list";
        return ERR_STATE;
      }
    }
    X.nitems = (int)items.size();
    X.nleaves = (int)pl.p2pLeaf.size();
    X.st.p2p_items = X.nitems;
    CKR(X.upv(X.items, items));
    CKR(X.upv(X.itemBeg, itemBeg));
    CKR(X.upv(X.leafCell, pl.p2pLeaf));
    CKR(X.upv(X.rb, pl.p2pSrcB));
    CKR(X.upv(X.re, pl.p2pSrcE));
  }
  X.hSig.swap(pl.sig);
  X.hSig2.swap(pl.sig2);
  X.hTab.swap(pl.m2lTab);
  for (int s = 0; s < 2; s++) { X.hA[s] = pl.A[s]; X.hB[s] = pl.B[s]; }
  X.nMLslots = (size_t)std::max(pl.S.nslots, pl.T.nslots);
  CKR(chooseM2LMode(X));
  return uploadPrecision(X);
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Launchers of the NR-wide kernels: the widest NR in {8, 4, 2, 1} that is
// <= n and allowed, for right-hand sides r0 .. r0 + NR - 1.
template <typename R, int L, int NR>
static void launchM2L(Exec& X, hipStream_t s, int g0, int ng, const typename C2T<R>::T* F,
                      typename C2T<R>::T* Lb, const OpConsts<R>& oc) {
  typedef typename C2T<R>::T R2;
  constexpr int G = M2LShapeN<L, NR>::G;
  if constexpr (M2LShape<L>::kColumns)
    hipLaunchKernelGGL((k_m2l_col<R, L, NR>), itemGrid(ng), dim3(M2LShape<L>::TB), 0, s,
                       X.gBeg + g0, X.gTarget + (size_t)g0 * G, X.gSrc, X.gMask, X.gCode, F,
                       X.fStride, (const R2*)X.U, Lb, X.lStride, oc.twc, oc.tws, ng);
  else if (X.m2lSliced)
    hipLaunchKernelGGL((k_m2l<R, L, NR, true>), itemGrid(ng), dim3(M2LShapeN<L, NR>::TB), 0, s,
                       X.gBeg + g0, X.gTarget + (size_t)g0 * G, X.gSrc, X.gMask, X.gCode, F,
                       X.fStride, (const R2*)X.U, Lb, X.lStride, oc.twc, oc.tws, ng);
  else
    hipLaunchKernelGGL((k_m2l<R, L, NR, false>), itemGrid(ng), dim3(M2LShapeN<L, NR>::TB), 0, s,
                       X.gBeg + g0, X.gTarget + (size_t)g0 * G, X.gSrc, X.gMask, X.gCode, F,
                       X.fStride, (const R2*)X.U, Lb, X.lStride, oc.twc, oc.tws, ng);
}

// Right-hand sides of the next M2L launch when n are left.
template <typename R, int L>
constexpr int m2lChunkWidth(int n) {
  constexpr int mx = m2lMaxNR<R, L>();
  return (mx >= 8 && n >= 8) ? 8 : (mx >= 4 && n >= 4) ? 4 : (mx >= 2 && n >= 2) ? 2 : 1;
}

// M2L of groups g0 .. g0 + ng - 1 for right-hand sides r0 .. r0 + n' - 1
// (n' = m2lChunkWidth(n), returned); their F rows start at F (right-hand
// side r0 + i at F + i * fStride).
template <typename R, int L>
static int m2lChunk(Exec& X, hipStream_t s, int g0, int ng, int r0, int n, const OpConsts<R>& oc,
                    const typename C2T<R>::T* F) {
  typedef typename C2T<R>::T R2;
  R2* Lb = (R2*)X.Lb + (size_t)r0 * X.lStride;
  constexpr int mx = m2lMaxNR<R, L>();
  if (mx >= 8 && n >= 8) { launchM2L<R, L, (mx >= 8 ? 8 : 1)>(X, s, g0, ng, F, Lb, oc); return 8; }
  if (mx >= 4 && n >= 4) { launchM2L<R, L, (mx >= 4 ? 4 : 1)>(X, s, g0, ng, F, Lb, oc); return 4; }
  if (mx >= 2 && n >= 2) { launchM2L<R, L, (mx >= 2 ? 2 : 1)>(X, s, g0, ng, F, Lb, oc); return 2; }
  launchM2L<R, L, 1>(X, s, g0, ng, F, Lb, oc);
  return 1;
}

template <typename R, int NR>
static void launchP2P(Exec& X, hipStream_t s, int r0) {
  typedef typename C2T<R>::T R2;
  const R2* q = (const R2*)X.q + (size_t)r0 * X.ns;
  R2* p = (R2*)X.p + (size_t)r0 * X.nt;
  hipLaunchKernelGGL((k_p2p<R, NR>), itemGrid(X.nitems), dim3(64), 0, s, X.items, X.leafCell,
                     X.T.pbeg, X.T.nprt, X.tx, X.ty, X.tz, X.sx, X.sy, X.sz, q, (size_t)X.ns, X.rb,
                     X.re, X.kappa, X.smallArg, X.partial, X.nitems);
  hipLaunchKernelGGL((k_p2p_reduce<R, NR>), itemGrid(X.nleaves), dim3(64), 0, s, X.itemBeg,
                     X.items, X.leafCell, X.T.pbeg, X.T.nprt, (const double2*)X.partial, p,
                     (size_t)X.nt, X.nleaves);
}

// 16 right-hand sides r0 .. r0 + 15 of the float path on the matrix cores
static void launchP2PMfma(Exec& X, hipStream_t s, int r0) {
  const float2* q = (const float2*)X.q + (size_t)r0 * X.ns;
  float2* p = (float2*)X.p + (size_t)r0 * X.nt;
  if (X.smallArg)
    hipLaunchKernelGGL((k_p2p_mfma<true>), itemGrid(X.nitems), dim3(256), 0, s, X.items,
                       X.leafCell, X.T.pbeg, X.T.nprt, X.tx, X.ty, X.tz, X.sx, X.sy, X.sz, q,
                       (size_t)X.ns, X.rb, X.re, X.kappa, X.partial, X.nitems);
  else
    hipLaunchKernelGGL((k_p2p_mfma<false>), itemGrid(X.nitems), dim3(256), 0, s, X.items,
                       X.leafCell, X.T.pbeg, X.T.nprt, X.tx, X.ty, X.tz, X.sx, X.sy, X.sz, q,
                       (size_t)X.ns, X.rb, X.re, X.kappa, X.partial, X.nitems);
  hipLaunchKernelGGL((k_p2p_reduce<float, kP2PMfmaNR>), itemGrid(X.nleaves), dim3(64), 0, s,
                     X.itemBeg, X.items, X.leafCell, X.T.pbeg, X.T.nprt,
                     (const double2*)X.partial, p, (size_t)X.nt, X.nleaves);
}

// One evaluation pass for nrhs <= X.nr right-hand sides: qin / pout hold
// nrhs x n complex128 values, right-hand side after right-hand side, in
// input order.  Stages as in the reference's gemv; M2L and P2P process up to
// 8 right-hand sides per launch, the other stages one per launch.
template <typename R, int L>
static int runTyped(Exec& X, const double2* qin, double2* pout, int parts, int nrhs,
                    hipStream_t s) {
  typedef typename C2T<R>::T R2;
  constexpr int P = 2 * L - 1;
  const R* cst = (const R*)X.consts;
  OpConsts<R> oc;
  oc.xi = cst;
  oc.iden = cst + L;
  oc.A[0] = cst + 2 * L;
  oc.A[1] = oc.A[0] + L * L;
  oc.B[0] = oc.A[1] + L * L;
  oc.B[1] = oc.B[0] + L * L;
  oc.twc = oc.B[1] + L * L;
  oc.tws = oc.twc + P;
  if (nrhs < 1 || nrhs > X.nr) {
    g_err = "right-hand sides exceed the plan's width";
    return ERR_ARG;
  }
  if (X.m2lG != M2LShape<L>::G) {
    g_err = "internal: M2L group size mismatch";
    return ERR_STATE;
  }
  const size_t ns = (size_t)X.ns, nt = (size_t)X.nt;
  auto qr = [&](int r) { return (R2*)X.q + r * ns; };
  auto pr = [&](int r) { return (R2*)X.p + r * nt; };
  auto Mr = [&](int r) { return (R2*)X.ML + r * X.mStride; };
  auto Lr = [&](int r) { return (R2*)X.Lb + r * X.lStride; };
  auto Fr = [&](int r) { return (R2*)X.F + r * X.fStride; };
  const R2* sig = (const R2*)X.sig;
  const R2* sig2 = (const R2*)X.sig2;
  // (a failed record surfaces at the launch check below or in stage_times)
  auto mark = [&](int k) {
    if (!X.capturing) (void)hipEventRecord(X.ev[k], s);
  };
  mark(0);
  hipLaunchKernelGGL(k_zero<R2>, dim3(zeroGrid((int64_t)nrhs * (int64_t)nt, 256)), dim3(256), 0,
                     s, (R2*)X.p, (int64_t)nrhs * (int64_t)nt);
  for (int r = 0; r < nrhs; r++)
    hipLaunchKernelGGL(k_gather<R>, dim3(nblk(X.ns, 256)), dim3(256), 0, s, qin + r * ns,
                       X.sperm, qr(r), (int)X.ns);
  mark(1);
  const bool far = (parts & 2) != 0, near = (parts & 1) != 0;

  if (far && X.np2mItems)
    for (int r = 0; r < nrhs;) {
      if constexpr (sizeof(R) == 4) {
        if (nrhs - r >= kP2MMfmaNR && X.p2mMfma) {
          // leaves of <= 64 points on the matrix cores, the others per
          // right-hand side on the vector units
          if (X.np2mMfma)
            hipLaunchKernelGGL((k_p2m_mfma<L>), itemGrid(X.np2mMfma), dim3(256), 0, s,
                               X.p2mMfmaItems, X.S, X.lv, X.sx, X.sy, X.sz,
                               (const float2*)qr(r), (size_t)ns, X.dirs, (const float2*)sig2, oc,
                               X.kappa, (float2*)Mr(r), X.mStride, X.np2mMfma);
          if (X.np2mBig)
            for (int rr = r; rr < r + kP2MMfmaNR; rr++)
              hipLaunchKernelGGL((k_p2m<R, L>), itemGrid(X.np2mBig), dim3(128), 0, s,
                                 X.p2mBigItems, X.S, X.lv, X.sx, X.sy, X.sz, (const R2*)qr(rr),
                                 X.dirs, sig2, oc, X.kappa, Mr(rr), X.np2mBig);
          r += kP2MMfmaNR;
          continue;
        }
      }
      hipLaunchKernelGGL((k_p2m<R, L>), itemGrid(X.np2mItems), dim3(128), 0, s, X.p2mItems, X.S,
                         X.lv, X.sx, X.sy, X.sz, (const R2*)qr(r), X.dirs, sig2, oc, X.kappa,
                         Mr(r), X.np2mItems);
      r++;
    }
  mark(2);
  if (far)
    for (auto& t : X.m2m)
      launchTrans<R, L>(t, (const R2*)Mr(0), Mr(0), sig, sig2, oc.A[0], oc.A[1], 0, X.mStride,
                        nrhs, s);
  mark(3);
  const bool stored = X.m2lMode == DEFMM_M2L_STORED;
  // M2F blocks of nr right-hand sides in one launch (their count in int)
  auto launchM2F = [&](const int* slots, int nRows, int r0, int nr, R2* F0) {
    const int64_t perRhs = (nRows + m2fRows<R, L>() - 1) / m2fRows<R, L>();
    const int per = (int)std::max<int64_t>(1, std::min<int64_t>(nr, INT32_MAX / std::max<int64_t>(1, perRhs)));
    for (int a = 0; a < nr; a += per) {
      const int c = std::min(per, nr - a);
      hipLaunchKernelGGL((k_m2f<R, L>), itemGrid(perRhs * c), dim3(256), 0, s, slots,
                         (const R2*)Mr(r0 + a), F0 + (size_t)a * X.fStride, oc.twc, oc.tws, nRows,
                         X.mStride, X.fStride, c);
    }
  };
  if (far && X.nm2f && stored) launchM2F(X.m2fSlot, X.nm2f, 0, nrhs, Fr(0));
  mark(4);
  if (far && X.nGroups) {
    if (stored) {
      // L (aliasing M) is written for every target slot from here on
      for (int r0 = 0; r0 < nrhs;)
        r0 += m2lChunk<R, L>(X, s, 0, X.nGroups, r0, nrhs - r0, oc, Fr(r0));
    } else if (X.m2lMode == DEFMM_M2L_ON_THE_FLY) {
      if constexpr (!M2LShape<L>::kColumns) {
        for (int r = 0; r < nrhs; r++)
          hipLaunchKernelGGL((k_m2l_fly<R, L>), itemGrid(X.nGroups), dim3(M2LShape<L>::TB), 0,
                             s, X.gBeg, X.gTarget, X.gSrc, X.gMask, X.gCode, (const R2*)Mr(r),
                             (const R2*)X.U, Lr(r), oc.twc, oc.tws, X.nGroups);
      } else {
        g_err = "the on-the-fly M2L variant supports orders up to 7";
        return ERR_ARG;
      }
    } else {
      // per batch of groups and per chunk of right-hand sides of one M2L
      // launch: the batch's F rows of the chunk into the workspace, then the
      // batch's M2L over them.  The workspace holds one chunk, so its rows
      // are read back while still in the caches, and the next chunk
      // overwrites them there.  (Stream order separates an M2L launch from
      // the next M2F into the same rows.)
      const size_t nb = X.hBatchGroupBeg.size() - 1;
      for (size_t b = 0; b < nb; b++) {
        const int g0 = X.hBatchGroupBeg[b], g1 = X.hBatchGroupBeg[b + 1];
        const int w0 = X.hBatchRowBeg[b], w1 = X.hBatchRowBeg[b + 1];
        for (int r0 = 0; r0 < nrhs;) {
          const int c = m2lChunkWidth<R, L>(nrhs - r0);
          launchM2F(X.bRows + w0, w1 - w0, r0, c, Fr(0));
          r0 += m2lChunk<R, L>(X, s, g0, g1 - g0, r0, c, oc, Fr(0));
        }
      }
    }
  }
  mark(5);
  if (far && X.nM2LZero)
    for (int r = 0; r < nrhs; r++)
      hipLaunchKernelGGL((k_zero_slots<R, L>), dim3(zeroGrid((int64_t)X.nM2LZero * X.LL, 256)),
                         dim3(256), 0, s, X.m2lZero, X.nM2LZero, Lr(r));
  if (far)
    for (auto& t : X.l2l)
      launchTrans<R, L>(t, (const R2*)Lr(0), Lr(0), sig2, sig, oc.B[0], oc.B[1], 1, X.lStride,
                        nrhs, s);
  mark(6);
  if (far && X.nl2p)
    for (int r = 0; r < nrhs;) {
      if constexpr (sizeof(R) == 4) {
        if (nrhs - r >= kL2PMfmaNR && X.l2pMfma) {
          hipLaunchKernelGGL((k_l2p_mfma<L>), itemGrid(X.nl2pTiles), dim3(256), 0, s,
                             X.l2pTiles, X.T, X.lv, X.tx, X.ty, X.tz, X.dirs, (const float2*)sig2,
                             (const float2*)Lr(r), X.lStride, oc, X.kappa, (float2*)pr(r),
                             (size_t)nt, X.nl2pTiles);
          r += kL2PMfmaNR;
          continue;
        }
      }
      hipLaunchKernelGGL((k_l2p<R, L>), itemGrid(X.nl2p), dim3(256), 0, s, X.l2p, X.T, X.lv,
                         X.tx, X.ty, X.tz, X.dirs, sig2, (const R2*)Lr(r), oc, X.kappa, pr(r),
                         X.nl2p);
      r++;
    }
  mark(7);
  if (near && X.nitems) {
    // near field: per-item partial sums, then folded into p per target leaf,
    // up to 8 right-hand sides per kernel value on the VALU, 16 on the
    // matrix cores in float.  (Run on a second stream beside M2L it measured
    // no faster: both contend for the same CUs.)
    for (int r0 = 0; r0 < nrhs;) {
      const int n = nrhs - r0;
      if (sizeof(R) == 4 && n >= kP2PMfmaNR && X.p2pMfma) {
        launchP2PMfma(X, s, r0);
        r0 += kP2PMfmaNR;
      } else if (n >= 8) { launchP2P<R, 8>(X, s, r0); r0 += 8; }
      else if (n >= 4) { launchP2P<R, 4>(X, s, r0); r0 += 4; }
      else if (n >= 2) { launchP2P<R, 2>(X, s, r0); r0 += 2; }
      else { launchP2P<R, 1>(X, s, r0); r0 += 1; }
    }
  }
  mark(8);
  for (int r = 0; r < nrhs; r++)
    hipLaunchKernelGGL(k_scatter<R>, dim3(nblk(X.nt, 256)), dim3(256), 0, s, (const R2*)pr(r),
                       X.tperm, pout + r * nt, (int)X.nt);
  mark(9);
  if (!X.capturing) X.evRecorded = true;
  HIPCK(hipGetLastError());
  return OK;
}

template <typename R>
static int runR(Exec& X, const double2* qin, double2* pout, int parts, int nrhs, hipStream_t s) {
  switch (X.L) {
    case 2: return runTyped<R, 2>(X, qin, pout, parts, nrhs, s);
    case 3: return runTyped<R, 3>(X, qin, pout, parts, nrhs, s);
    case 4: return runTyped<R, 4>(X, qin, pout, parts, nrhs, s);
    case 5: return runTyped<R, 5>(X, qin, pout, parts, nrhs, s);
    case 6: return runTyped<R, 6>(X, qin, pout, parts, nrhs, s);
    case 7: return runTyped<R, 7>(X, qin, pout, parts, nrhs, s);
    case 8: return runTyped<R, 8>(X, qin, pout, parts, nrhs, s);
    case 9: return runTyped<R, 9>(X, qin, pout, parts, nrhs, s);
    case 10: return runTyped<R, 10>(X, qin, pout, parts, nrhs, s);
  }
  g_err = "interpolation order outside 2..10";
  return ERR_ARG;
}

// nrhs right-hand sides in chunks of the plan's width
static int run(Exec& X, const double2* qin, double2* pout, int parts, int nrhs, hipStream_t s) {
  if (!X.valid) {
    g_err = "the plan's device stores are not allocated (a failed reallocation left it unusable)";
    return ERR_STATE;
  }
  HIPCK(hipSetDevice(X.device));
  // every execution uses the plan's stores: one on another stream than the
  // last starts after it (a captured pass is ordered by its batch instead)
  if (!X.capturing && X.doneRecorded && s != X.lastStream) HIPCK(hipStreamWaitEvent(s, X.done, 0));
  for (int r0 = 0; r0 < nrhs; r0 += X.nr) {
    const int n = std::min(X.nr, nrhs - r0);
    const double2* q = qin + (size_t)r0 * X.ns;
    double2* p = pout + (size_t)r0 * X.nt;
    int rc = X.prec == DEFMM_FP64 ? runR<double>(X, q, p, parts, n, s)
                                  : runR<float>(X, q, p, parts, n, s);
    if (rc) return rc;
  }
  if (X.capturing) return OK;
  HIPCK(hipEventRecord(X.done, s));
  X.lastStream = s;
  X.doneRecorded = true;
  return OK;
}

// Unit charges for the autotuner's passes.
__global__ void k_fill_one(double2* a, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    a[i] = make_double2(1.0, 0.0);
}

// Median time (ms) of `reps` passes of nrhs right-hand sides on the plan's
// staging buffers (charges in qio, potentials to pio), on its own stream.
static int timePasses(Exec& X, int nrhs, int reps, double& ms) {
  hipEvent_t a = nullptr, b = nullptr;
  HIPCK(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    g_err = "hipEventCreate failed";
    return ERR_HIP;
  }
  std::vector<double> t;
  int rc = OK;
  // one untimed pass first (first-touch of the stores)
  for (int k = -1; k < reps && rc == OK; k++) {
    if (hipEventRecord(a, X.own) != hipSuccess) { rc = ERR_HIP; break; }
    rc = run(X, X.qio, X.pio, 3, nrhs, X.own);
    if (rc) break;
    float f = 0.f;
    if (hipEventRecord(b, X.own) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
        hipEventElapsedTime(&f, a, b) != hipSuccess) {
      g_err = "autotune: event timing failed";
      rc = ERR_HIP;
      break;
    }
    if (k >= 0) t.push_back(f);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return rc;
  std::sort(t.begin(), t.end());
  ms = t[t.size() / 2];
  return OK;
}

// Work of one plan's pass in arbitrary units (near-field pairs plus M2L
// pair-frequencies, per right-hand side), for balancing graph branches.
static double passWork(const Exec& X, int nrhs) {
  return (double)nrhs * ((double)X.st.p2p_pairs + (double)X.st.m2l_entries * X.PP +
                         (double)(X.st.src_slots + X.st.trg_slots) * X.LL * X.L);
}

}  // namespace dfm

// ----------------------------------------------------------------------------
// C ABI
// ----------------------------------------------------------------------------
using dfm::Exec;
using dfm::g_err;

struct defmm_plan_s {
  Exec x;
};

// A batch of independent plans evaluated by one HIP graph (defmm_batch_*).
struct defmm_batch_s {
  std::vector<defmm_plan> plans;
  std::vector<uint64_t> gens;     // plan generations at capture
  std::vector<int> branchOf;      // branch of each plan
  std::vector<hipStream_t> branch;
  std::vector<hipEvent_t> evs;    // fork + one join per branch
  hipStream_t cs = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int device = 0;
  ~defmm_batch_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto e : evs) (void)hipEventDestroy(e);
    for (auto b : branch) (void)hipStreamDestroy(b);
    if (cs) (void)hipStreamDestroy(cs);
  }
};

extern "C" {

const char* defmm_strerror(int code) {
  switch (code) {
    case DEFMM_OK: return "success";
    case DEFMM_EARG: return "invalid argument";
    case DEFMM_ENOPRECOMP: return "no M2L level: tree too shallow for this wavenumber";
    case DEFMM_EDEPTH: return "octree deeper than 30 levels";
    case DEFMM_EHIP: return "HIP runtime error";
    case DEFMM_EMEMORY: return "out of memory";
    case DEFMM_ETABLE: return "M2L offset outside the precomputed tables";
    case DEFMM_ENODEVICE: return "no GPU visible";
    case DEFMM_ESTATE: return "call out of sequence";
  }
  return "unknown error";
}

const char* defmm_last_error(void) { return g_err.c_str(); }

int defmm_set_setup_mode(int mode) {
  if (mode != DEFMM_SETUP_DEVICE && mode != DEFMM_SETUP_HOST) {
    g_err = "unknown set-up mode";
    return -DEFMM_EARG;
  }
  return dfm::g_setupMode.exchange(mode);
}

int defmm_plan_create(const double* src_xyz, int64_t n_src, const double* trg_xyz,
                      int64_t n_trg, int order, int ncrit, double kappa, int lvmax,
                      int precision, defmm_plan* out) {
  if (!out) { g_err = "null output pointer"; return DEFMM_EARG; }
  *out = nullptr;
  if (precision != DEFMM_FP64 && precision != DEFMM_MIXED) { g_err = "bad precision"; return DEFMM_EARG; }
  if (order < 2 || order > dfm::LMAX) { g_err = "order must be in 2..10"; return DEFMM_EARG; }
  if (lvmax < -1 || (lvmax < 0 && ncrit < 1)) { g_err = "bad ncrit/lvmax"; return DEFMM_EARG; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { g_err = "no HIP device"; return DEFMM_ENODEVICE; }
  dfm::Plan* pl = new dfm::Plan();
  std::string msg;
  int rc;
  {
    std::unique_ptr<dfm::SetupDevice> sd(
        dfm::g_setupMode.load() == DEFMM_SETUP_DEVICE ? dfm::newSetupDevice() : nullptr);
    rc = dfm::buildPlan(*pl, src_xyz, n_src, trg_xyz, n_trg, order, lvmax >= 0 ? -lvmax : ncrit,
                        kappa, msg, sd.get());
  }
  if (rc == dfm::kSetupFallback) {
    msg.clear();
    rc = dfm::buildPlan(*pl, src_xyz, n_src, trg_xyz, n_trg, order, lvmax >= 0 ? -lvmax : ncrit,
                        kappa, msg);
  }
  if (rc != dfm::OK) {
    g_err = msg;
    delete pl;
    return rc;
  }
  defmm_plan_s* h = new defmm_plan_s();
  h->x.prec = precision;
  if (hipGetDevice(&h->x.device) != hipSuccess) {
    g_err = "hipGetDevice failed";
    delete h;
    delete pl;
    return DEFMM_EHIP;
  }
  const auto tu0 = std::chrono::steady_clock::now();
  rc = dfm::upload(h->x, *pl);
  h->x.st.ms_upload =
      (int64_t)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tu0).count();
  delete pl;
  if (rc == dfm::OK) {
    if (hipStreamCreate(&h->x.own) != hipSuccess) rc = DEFMM_EHIP;
    if (rc == dfm::OK && hipEventCreateWithFlags(&h->x.done, hipEventDisableTiming) != hipSuccess)
      rc = DEFMM_EHIP;
    for (auto& e : h->x.ev)
      if (rc == dfm::OK && hipEventCreate(&e) != hipSuccess) rc = DEFMM_EHIP;
  }
  if (rc != dfm::OK) {
    delete h;
    return rc;
  }
  *out = h;
  return DEFMM_OK;
}

int defmm_execute_device(defmm_plan plan, const void* q_dev, void* p_dev, void* stream) {
  if (!plan || !q_dev || !p_dev) { g_err = "null argument"; return DEFMM_EARG; }
  return dfm::run(plan->x, (const double2*)q_dev, (double2*)p_dev, 3, 1, (hipStream_t)stream);
}

int defmm_execute_batch_device(defmm_plan plan, const void* q_dev, void* p_dev, int nrhs,
                               void* stream) {
  if (!plan || !q_dev || !p_dev) { g_err = "null argument"; return DEFMM_EARG; }
  if (nrhs < 1) { g_err = "nrhs must be positive"; return DEFMM_EARG; }
  return dfm::run(plan->x, (const double2*)q_dev, (double2*)p_dev, 3, nrhs, (hipStream_t)stream);
}

static int execHost(defmm_plan plan, const double* q, double* p, int parts, int nrhs) {
  if (!plan || !q || !p) { g_err = "null argument"; return DEFMM_EARG; }
  if (parts < 1 || parts > 3) { g_err = "parts must be 1, 2 or 3"; return DEFMM_EARG; }
  if (nrhs < 1) { g_err = "nrhs must be positive"; return DEFMM_EARG; }
  Exec& X = plan->x;
  if (!X.valid) {
    g_err = "the plan's device stores are not allocated (a failed reallocation left it unusable)";
    return DEFMM_ESTATE;
  }
  if (hipSetDevice(X.device) != hipSuccess) { g_err = "hipSetDevice failed"; return DEFMM_EHIP; }
  // the staging buffers are written on the plan's own stream: after the last
  // execution, whatever stream it ran on
  if (X.doneRecorded && X.lastStream != X.own &&
      hipStreamWaitEvent(X.own, X.done, 0) != hipSuccess) {
    g_err = "hipStreamWaitEvent failed";
    return DEFMM_EHIP;
  }
  for (int r0 = 0; r0 < nrhs; r0 += X.nr) {
    const int n = std::min(X.nr, nrhs - r0);
    if (hipMemcpyAsync(X.qio, q + 2 * (size_t)r0 * X.ns, n * X.ns * sizeof(double2),
                       hipMemcpyHostToDevice, X.own) != hipSuccess) {
      g_err = "copy of charges failed";
      return DEFMM_EHIP;
    }
    int rc = dfm::run(X, X.qio, X.pio, parts, n, X.own);
    if (rc) return rc;
    if (hipMemcpyAsync(p + 2 * (size_t)r0 * X.nt, X.pio, n * X.nt * sizeof(double2),
                       hipMemcpyDeviceToHost, X.own) != hipSuccess ||
        hipStreamSynchronize(X.own) != hipSuccess) {
      g_err = std::string("execution failed: ") + hipGetErrorString(hipGetLastError());
      return DEFMM_EHIP;
    }
  }
  return DEFMM_OK;
}

int defmm_execute(defmm_plan plan, const double* q, double* p) {
  return execHost(plan, q, p, 3, 1);
}

int defmm_execute_batch(defmm_plan plan, const double* q, double* p, int nrhs) {
  return execHost(plan, q, p, 3, nrhs);
}

int defmm_execute_parts(defmm_plan plan, const double* q, double* p, int parts) {
  return execHost(plan, q, p, parts, 1);
}

int defmm_plan_set_rhs(defmm_plan plan, int nrhs) {
  if (!plan) { g_err = "null plan"; return DEFMM_EARG; }
  if (nrhs < 1 || nrhs > 64) { g_err = "nrhs must be in 1..64"; return DEFMM_EARG; }
  Exec& X = plan->x;
  if (nrhs == X.nr && X.valid) return DEFMM_OK;
  if (X.quiesce() != dfm::OK) {
    g_err = "device synchronisation failed";
    return DEFMM_EHIP;
  }
  const int oldNr = X.nr;
  X.nr = nrhs;
  return dfm::reupload(X, oldNr, X.prec, X.m2lMode, X.m2lAuto, X.batchRows);
}

int defmm_plan_stats(defmm_plan plan, int64_t* out, int n) {
  if (!plan || !out) { g_err = "null argument"; return DEFMM_EARG; }
  const dfm::Stats& s = plan->x.st;
  int64_t v[DEFMM_NSTATS] = {s.p2p_calls, s.p2p_pairs, s.hf_m2l, s.lf_m2l, s.src_slots,
                             s.trg_slots, s.hf_tables, s.lf_tables, s.hf_leaves_p2m,
                             s.lf_leaves_p2m, s.hf_m2m, s.lf2hf_m2m, s.lf_m2m, s.hf_l2p,
                             s.lf_l2p, s.hf_l2l, s.hf2lf_l2l, s.lf_l2l, s.nlvl, s.cpt,
                             s.max_stage, s.ncells_t, s.ncells_s, s.p2p_items, s.m2f_slots,
                             s.m2l_entries, s.m2l_targets, s.m2l_groups,
                             s.m2l_group_entries, s.m2l_batches, s.m2l_batch_rows,
                             s.ms_tree, s.ms_walk, s.ms_tables, s.ms_trans, s.ms_m2l,
                             s.ms_p2p, s.ms_upload, s.m2l_table_rows, s.setup_device};
  for (int i = 0; i < n && i < DEFMM_NSTATS; i++) out[i] = v[i];
  return DEFMM_OK;
}

int defmm_stage_times(defmm_plan plan, double* ms, int n) {
  if (!plan || !ms) { g_err = "null argument"; return DEFMM_EARG; }
  Exec& X = plan->x;
  if (!X.evRecorded) { g_err = "no execution recorded"; return DEFMM_ESTATE; }
  if (hipEventSynchronize(X.ev[DEFMM_NSTAGES]) != hipSuccess) { g_err = "event sync failed"; return DEFMM_EHIP; }
  for (int k = 0; k < DEFMM_NSTAGES && k < n; k++) {
    float t = 0.f;
    hipEvent_t a = X.ev[k], b = X.ev[k + 1];
    if (hipEventElapsedTime(&t, a, b) != hipSuccess) { g_err = "event timing failed"; return DEFMM_EHIP; }
    ms[k] = t;
  }
  return DEFMM_OK;
}

int defmm_plan_set_precision(defmm_plan plan, int precision) {
  if (!plan) { g_err = "null plan"; return DEFMM_EARG; }
  if (precision != DEFMM_FP64 && precision != DEFMM_MIXED) { g_err = "bad precision"; return DEFMM_EARG; }
  Exec& X = plan->x;
  if (precision == X.prec && X.valid) return DEFMM_OK;
  if (X.quiesce() != dfm::OK) {
    g_err = "device synchronisation failed";
    return DEFMM_EHIP;
  }
  const int oldPrec = X.prec;
  X.prec = precision;
  return dfm::reupload(X, X.nr, oldPrec, X.m2lMode, X.m2lAuto, X.batchRows);
}

int defmm_plan_set_m2l(defmm_plan plan, int mode) {
  if (!plan) { g_err = "null plan"; return DEFMM_EARG; }
  if (mode != DEFMM_M2L_ON_THE_FLY && mode != DEFMM_M2L_STORED && mode != DEFMM_M2L_BATCHED) {
    g_err = "bad M2L mode";
    return DEFMM_EARG;
  }
  Exec& X = plan->x;
  if (mode == DEFMM_M2L_ON_THE_FLY && X.L > 7) {
    g_err = "the on-the-fly M2L variant supports orders up to 7";
    return DEFMM_EARG;
  }
  if (mode == X.m2lMode && X.valid) {
    X.m2lAuto = false;
    return DEFMM_OK;
  }
  if (X.quiesce() != dfm::OK) {
    g_err = "device synchronisation failed";
    return DEFMM_EHIP;
  }
  const int oldMode = X.m2lMode;
  const bool oldAuto = X.m2lAuto;
  X.m2lMode = mode;
  X.m2lAuto = false;
  return dfm::reupload(X, X.nr, X.prec, oldMode, oldAuto, X.batchRows);
}

int defmm_plan_get_m2l(defmm_plan plan, int* mode) {
  if (!plan || !mode) { g_err = "null argument"; return DEFMM_EARG; }
  *mode = plan->x.m2lMode;
  return DEFMM_OK;
}

int defmm_plan_set_m2l_batch_rows(defmm_plan plan, int rows) {
  if (!plan) { g_err = "null plan"; return DEFMM_EARG; }
  if (rows < 1) { g_err = "batch rows must be positive"; return DEFMM_EARG; }
  Exec& X = plan->x;
  const int oldRows = X.batchRows;
  X.batchRows = rows;
  if (X.m2lMode != DEFMM_M2L_BATCHED && X.valid) return DEFMM_OK;
  if (X.quiesce() != dfm::OK) {
    g_err = "device synchronisation failed";
    return DEFMM_EHIP;
  }
  return dfm::reupload(X, X.nr, X.prec, X.m2lMode, X.m2lAuto, oldRows);
}

int defmm_plan_set_option(defmm_plan plan, int option, int64_t value) {
  if (!plan) { g_err = "null plan"; return DEFMM_EARG; }
  Exec& X = plan->x;
  switch (option) {
    case DEFMM_OPT_P2P_MFMA:
    case DEFMM_OPT_L2P_MFMA:
    case DEFMM_OPT_P2M_MFMA:
      if (value != 0 && value != 1) { g_err = "the option takes 0 or 1"; return DEFMM_EARG; }
      if (X.quiesce() != dfm::OK) { g_err = "device synchronisation failed"; return DEFMM_EHIP; }
      (option == DEFMM_OPT_P2P_MFMA ? X.p2pMfma
                                    : option == DEFMM_OPT_L2P_MFMA ? X.l2pMfma : X.p2mMfma) =
          value != 0;
      return DEFMM_OK;
    case DEFMM_OPT_M2L_SLICED:
      if (value != 0 && value != 1) { g_err = "the option takes 0 or 1"; return DEFMM_EARG; }
      if (X.quiesce() != dfm::OK) { g_err = "device synchronisation failed"; return DEFMM_EHIP; }
      X.m2lSliced = value != 0;
      return DEFMM_OK;
  }
  g_err = "unknown option";
  return DEFMM_EARG;
}

int defmm_plan_autotune(defmm_plan plan, int nrhs, int reps, char* report, int report_len) {
  if (!plan) { g_err = "null plan"; return DEFMM_EARG; }
  Exec& X = plan->x;
  if (nrhs < 1 || nrhs > X.nr) { g_err = "nrhs must be in 1 .. the plan's width"; return DEFMM_EARG; }
  if (reps < 1 || reps > 100) { g_err = "reps must be in 1..100"; return DEFMM_EARG; }
  if (!X.valid) { g_err = "the plan is unusable"; return DEFMM_ESTATE; }
  if (X.quiesce() != dfm::OK) { g_err = "device synchronisation failed"; return DEFMM_EHIP; }
  std::string rep;
  char buf[256];
  auto fill = [&]() -> int {
    hipLaunchKernelGGL(dfm::k_fill_one, dim3(dfm::zeroGrid((int64_t)nrhs * X.ns, 256)), dim3(256),
                       0, X.own, X.qio, (int64_t)nrhs * X.ns);
    return hipGetLastError() == hipSuccess ? dfm::OK : dfm::ERR_HIP;
  };
  // 1. M2L storage variant
  {
    struct Cand { int mode, rows; };
    std::vector<Cand> cands = {{DEFMM_M2L_STORED, X.batchRows}, {DEFMM_M2L_BATCHED, 2048},
                               {DEFMM_M2L_BATCHED, 4096},  {DEFMM_M2L_BATCHED, 8192},
                               {DEFMM_M2L_BATCHED, 16384}, {DEFMM_M2L_BATCHED, 32768},
                               {DEFMM_M2L_BATCHED, 65536}};
    const int mode0 = X.m2lMode, rows0 = X.batchRows;
    const bool auto0 = X.m2lAuto;
    double best = 1e300;
    Cand keep = {mode0, rows0};
    for (const Cand& c : cands) {
      X.m2lMode = c.mode;
      X.batchRows = c.rows;
      X.m2lAuto = false;
      int rc = dfm::reupload(X, X.nr, X.prec, mode0, auto0, rows0);
      if (rc == DEFMM_EMEMORY) {
        std::snprintf(buf, sizeof buf, "m2l %s/%d: does not fit; ", c.mode == DEFMM_M2L_STORED ? "stored" : "batched", c.rows);
        rep += buf;
        continue;
      }
      if (rc) return rc;
      double ms = 0.;
      if ((rc = fill()) || (rc = dfm::timePasses(X, nrhs, reps, ms))) return rc;
      std::snprintf(buf, sizeof buf, "m2l %s/%d: %.3f ms; ", c.mode == DEFMM_M2L_STORED ? "stored" : "batched", c.rows, ms);
      rep += buf;
      if (ms < best) { best = ms; keep = c; }
    }
    X.m2lMode = keep.mode;
    X.batchRows = keep.rows;
    X.m2lAuto = false;
    int rc = dfm::reupload(X, X.nr, X.prec, mode0, auto0, rows0);
    if (rc) return rc;
    std::snprintf(buf, sizeof buf, "kept m2l %s/%d\n", keep.mode == DEFMM_M2L_STORED ? "stored" : "batched", keep.rows);
    rep += buf;
  }
  // 2. sliced sweep of the grouped M2L (orders <= 7)
  if (X.L <= 7 && X.m2lMode != DEFMM_M2L_ON_THE_FLY) {
    double t[2] = {0., 0.};
    for (int on = 0; on < 2; on++) {
      X.m2lSliced = on != 0;
      int rc = fill();
      if (rc || (rc = dfm::timePasses(X, nrhs, reps, t[on]))) return rc;
    }
    X.m2lSliced = t[1] < t[0];
    std::snprintf(buf, sizeof buf, "m2l_sliced off %.3f ms, on %.3f ms: kept %s\n", t[0], t[1],
                  X.m2lSliced ? "on" : "off");
    rep += buf;
  }
  // 3. matrix cores per stage (mixed precision, chunks of 16 right-hand sides)
  if (X.prec == DEFMM_MIXED && nrhs >= dfm::kP2PMfmaNR) {
    struct Knob { const char* name; bool* v; };
    Knob knobs[3] = {{"p2p_mfma", &X.p2pMfma}, {"l2p_mfma", &X.l2pMfma}, {"p2m_mfma", &X.p2mMfma}};
    for (Knob& k : knobs) {
      double t[2] = {0., 0.};
      for (int on = 0; on < 2; on++) {
        *k.v = on != 0;
        int rc = fill();
        if (rc || (rc = dfm::timePasses(X, nrhs, reps, t[on]))) return rc;
      }
      *k.v = t[1] <= t[0];
      std::snprintf(buf, sizeof buf, "%s off %.3f ms, on %.3f ms: kept %s\n", k.name, t[0], t[1],
                    *k.v ? "on" : "off");
      rep += buf;
    }
  }
  X.evRecorded = false;
  if (report && report_len > 0) {
    std::strncpy(report, rep.c_str(), (size_t)report_len - 1);
    report[report_len - 1] = 0;
  }
  return DEFMM_OK;
}

int defmm_plan_set_memory_limit(defmm_plan plan, int64_t bytes) {
  if (!plan) { g_err = "null plan"; return DEFMM_EARG; }
  if (bytes < 0) { g_err = "memory limit must be >= 0"; return DEFMM_EARG; }
  plan->x.memLimit = bytes;
  return DEFMM_OK;
}

int defmm_plan_device_bytes(defmm_plan plan, int64_t* bytes) {
  if (!plan || !bytes) { g_err = "null argument"; return DEFMM_EARG; }
  *bytes = plan->x.bytes;
  return DEFMM_OK;
}

int defmm_plan_build_stats(const double* src_xyz, int64_t n_src, const double* trg_xyz,
                           int64_t n_trg, int order, int ncrit, double kappa, int lvmax,
                           int64_t* out, int n) {
  if (!out) { g_err = "null output pointer"; return DEFMM_EARG; }
  if (order < 2 || order > dfm::LMAX) { g_err = "order must be in 2..10"; return DEFMM_EARG; }
  if (lvmax < -1 || (lvmax < 0 && ncrit < 1)) { g_err = "bad ncrit/lvmax"; return DEFMM_EARG; }
  dfm::Plan* pl = new dfm::Plan();
  std::string msg;
  int rc = dfm::buildPlan(*pl, src_xyz, n_src, trg_xyz, n_trg, order, lvmax >= 0 ? -lvmax : ncrit,
                          kappa, msg);
  if (rc != dfm::OK) {
    g_err = msg;
    delete pl;
    return rc;
  }
  defmm_plan_s tmp;
  tmp.x.st = pl->st;
  delete pl;
  return defmm_plan_stats(&tmp, out, n);
}

int defmm_batch_create(const defmm_plan* plans, int n, const void* const* q_dev,
                       void* const* p_dev, const int* nrhs, int nbranch, defmm_batch* out) {
  if (!out) { g_err = "null output pointer"; return DEFMM_EARG; }
  *out = nullptr;
  if (!plans || !q_dev || !p_dev || !nrhs || n < 1) { g_err = "null argument or empty batch"; return DEFMM_EARG; }
  if (nbranch < 0) { g_err = "nbranch must be >= 0"; return DEFMM_EARG; }
  for (int i = 0; i < n; i++) {
    if (!plans[i] || !q_dev[i] || !p_dev[i] || nrhs[i] < 1) {
      g_err = "instance " + std::to_string(i) + ": null plan or array, or nrhs < 1";
      return DEFMM_EARG;
    }
    if (!plans[i]->x.valid) {
      g_err = "instance " + std::to_string(i) + ": the plan is unusable";
      return DEFMM_ESTATE;
    }
    if (plans[i]->x.device != plans[0]->x.device) {
      g_err = "all plans of a batch must live on one device";
      return DEFMM_EARG;
    }
    for (int j = 0; j < i; j++)
      if (plans[j] == plans[i]) {
        g_err = "a plan appears twice in the batch (its stores cannot serve two instances)";
        return DEFMM_EARG;
      }
  }
  defmm_batch_s* b = new defmm_batch_s();
  b->device = plans[0]->x.device;
  auto fail = [&](int code, const std::string& msg) {
    g_err = msg;
    delete b;
    return code;
  };
  if (hipSetDevice(b->device) != hipSuccess) return fail(DEFMM_EHIP, "hipSetDevice failed");
  const int nb = std::max(1, std::min(n, nbranch > 0 ? nbranch : 8));
  // longest processing time first: instances by decreasing work, each to
  // the least loaded branch
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::vector<double> work(n);
  for (int i = 0; i < n; i++) work[i] = dfm::passWork(plans[i]->x, nrhs[i]);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return work[a] > work[c]; });
  std::vector<double> load(nb, 0.);
  b->branchOf.assign(n, 0);
  for (int i : order) {
    const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    b->branchOf[i] = k;
    load[k] += work[i];
  }
  if (hipStreamCreate(&b->cs) != hipSuccess) return fail(DEFMM_EHIP, "hipStreamCreate failed");
  b->branch.assign(nb, nullptr);
  for (auto& st : b->branch)
    if (hipStreamCreate(&st) != hipSuccess) return fail(DEFMM_EHIP, "hipStreamCreate failed");
  b->evs.assign(nb + 1, nullptr);
  for (auto& e : b->evs)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return fail(DEFMM_EHIP, "hipEventCreate failed");
  // every plan's last execution has finished before the capture reads its state
  for (int i = 0; i < n; i++)
    if (plans[i]->x.quiesce() != dfm::OK) return fail(DEFMM_EHIP, "device synchronisation failed");
  if (hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(DEFMM_EHIP, "hipStreamBeginCapture failed");
  int rc = DEFMM_OK;
  std::string why;
  if (hipEventRecord(b->evs[0], b->cs) != hipSuccess) rc = DEFMM_EHIP;
  for (int k = 0; k < nb && !rc; k++)
    if (hipStreamWaitEvent(b->branch[k], b->evs[0], 0) != hipSuccess) rc = DEFMM_EHIP;
  for (int k = 0; k < nb && !rc; k++)
    for (int i : order) {
      if (b->branchOf[i] != k || rc) continue;
      Exec& X = plans[i]->x;
      X.capturing = true;
      rc = dfm::run(X, (const double2*)q_dev[i], (double2*)p_dev[i], 3, nrhs[i], b->branch[k]);
      X.capturing = false;
      if (rc) why = "instance " + std::to_string(i) + ": " + g_err;
    }
  for (int k = 0; k < nb && !rc; k++)
    if (hipEventRecord(b->evs[k + 1], b->branch[k]) != hipSuccess ||
        hipStreamWaitEvent(b->cs, b->evs[k + 1], 0) != hipSuccess)
      rc = DEFMM_EHIP;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->cs, &g);
  b->graph = g;
  if (rc) return fail(rc, why.empty() ? "capture of the batch failed" : why);
  if (ec != hipSuccess || !g)
    return fail(DEFMM_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  const hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (ei != hipSuccess)
    return fail(ei == hipErrorOutOfMemory ? DEFMM_EMEMORY : DEFMM_EHIP,
                std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  b->plans.assign(plans, plans + n);
  for (int i = 0; i < n; i++) b->gens.push_back(plans[i]->x.generation);
  *out = b;
  return DEFMM_OK;
}

int defmm_batch_execute(defmm_batch b, void* stream) {
  if (!b) { g_err = "null batch"; return DEFMM_EARG; }
  if (hipSetDevice(b->device) != hipSuccess) { g_err = "hipSetDevice failed"; return DEFMM_EHIP; }
  hipStream_t s = (hipStream_t)stream;
  for (size_t i = 0; i < b->plans.size(); i++) {
    Exec& X = b->plans[i]->x;
    if (!X.valid || X.generation != b->gens[i]) {
      g_err = "instance " + std::to_string(i) +
              ": its plan was reallocated since the batch was created; create the batch again";
      return DEFMM_ESTATE;
    }
    if (X.doneRecorded && X.lastStream != s && hipStreamWaitEvent(s, X.done, 0) != hipSuccess) {
      g_err = "hipStreamWaitEvent failed";
      return DEFMM_EHIP;
    }
  }
  const hipError_t e = hipGraphLaunch(b->exec, s);
  if (e != hipSuccess) {
    g_err = std::string("hipGraphLaunch: ") + hipGetErrorString(e);
    return DEFMM_EHIP;
  }
  for (defmm_plan p : b->plans) {
    Exec& X = p->x;
    if (hipEventRecord(X.done, s) != hipSuccess) {
      g_err = "hipEventRecord failed";
      return DEFMM_EHIP;
    }
    X.lastStream = s;
    X.doneRecorded = true;
    X.evRecorded = false;  // stage times are not recorded inside the graph
  }
  return DEFMM_OK;
}

int defmm_batch_destroy(defmm_batch b) {
  if (!b) return DEFMM_OK;
  (void)hipSetDevice(b->device);
  for (defmm_plan p : b->plans) (void)p->x.quiesce();
  delete b;
  return DEFMM_OK;
}

int defmm_plan_destroy(defmm_plan plan) {
  if (!plan) return DEFMM_OK;
  (void)plan->x.quiesce();
  delete plan;
  return DEFMM_OK;
}

}  // extern "C"
