// setup.hip -- the plan builder's octrees, dual walk, M2L and P2P lists on
// the GPU.
//
// The octrees are built level by level with the recursive build's rules
// (sort.hpp:98-162 of the reference): each splitting cell's points are
// partitioned stably by octant on the device (k_tree_count, k_tree_part),
// the children and their centres are made on the host (splitCell), and the
// cells are finally numbered in pre-order (treeFromLevels), which gives the
// recursive build's tree exactly.
//
// The host builder (plan.cpp) walks the source and target octrees from the
// roots (hrzpass.hpp:61-91 and the set-up walk hrzpass.hpp:111-138 of the
// reference): a pair of same-level cells is an M2L pair when it passes the
// level's admissibility test (hrzpass.hpp:39-56), a P2P pair when it does not
// and either cell is a leaf, and otherwise the pairs of their children are
// visited.  Here the walk runs level by level: the sources that target cell
// i meets at level l are the children of the sources its parent expanded at
// level l - 1.  One wavefront handles one target cell; it takes its
// parent's expanded sources in order, eight per step with one lane per
// child, classifies the 64 candidate pairs and appends each class in lane
// order (ballot + prefix count).  A count pass sizes every target's segment
// first, so each target's M2L, P2P and expansion lists come out in
// ascending source cell order without a sort: within one target the host's
// depth-first walk meets its sources in that same order (sources that first
// differ at some level are visited in the order of their child index there,
// which is pre-order).
//
// The M2L pairs stay on the device.  Each offset iZ = idx(target) -
// idx(source) a level meets gets a representative pair (the smallest
// (target, source), by a 64-bit atomic minimum over a dense window of
// offsets); the device searches each offset's cone from it (k_cones), the
// host turns the representatives into table ids and permutation hashes
// (offsetCodes, plan.cpp), and the device then marks the
// effective directions of both cells of every HF pair.  After the host has
// numbered the slots, the device builds the M2L CSR by target slot: a count
// per slot (atomics), the host's prefix sum, then a stable placement in
// which the wavefront of a target cell places its pairs in segment order,
// one distinct target slot at a time (a target slot belongs to one cell,
// so its cursor has one writer).  The grouped M2L lists and the P2P lists
// are merges of lists that are already in order (k_groups, k_p2p_leaf).
//
// Every decision is computed as on the host: the HF test in double
// precision with the same operations in the same order and no contraction
// (this file is compiled with -ffp-contract=off), the LF test and the
// offsets in integers.  The resulting plan is the host builder's, entry for
// entry.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "plan.hpp"
#include "xfer.hpp"

namespace dfm {
namespace {

constexpr int kWaves = 4;  // wavefronts (target cells) per 256-thread block

struct DTree {
  const int* idx;      // 3 per cell
  const double* ctr;   // 3 per cell
  const double* rad;
  const int* son0;
  const int* nson;
  const int* sons;
};

struct WalkArgs {
  DTree T, S;
  int nt;              // target cells at this level
  const int* tcell;    // their cell ids
  const int* tpar;     // their parents' positions in the previous level's list
  const int* eBeg;     // previous level: sources expanded per target (CSR)
  const int* eSrc;
  int kind;            // 0: no M2L at this level, 1: HF test, 2: LF test
  double kappa;
  int* cM;             // count pass: per target M2L, P2P and expansion counts
  int* cP;
  int* cE;
  const int* oM;       // write pass: per target offsets (exclusive sums)
  const int* oP;
  const int* oE;
  int* mSrc;
  int* pSrc;
  int* eOut;
  unsigned long long* rep;  // representative pair per offset of the window
  int W;
  int* overflow;       // set when an M2L offset falls outside the window
};

// hrzpass.hpp:51-56, the operations of the host's macHF in the same order
__device__ __forceinline__ bool macHF(double t0, double t1, double t2, const double* sc,
                                      double sr, double kappa) {
#pragma clang fp contract(off)
  const double d0 = t0 - sc[0], d1 = t1 - sc[1], d2 = t2 - sc[2];
  double s = 0.;
  s += d0 * d0;
  s += d1 * d1;
  s += d2 * d2;
  const double tmp = (1. - 1.e-8) * (sqrt(s) - 2. * sr);
  const double tt = 2. * sr;
  const double a = tt * tt * kappa;
  return tmp >= (a < tt ? tt : a);
}

// hrzpass.hpp:39-49
__device__ __forceinline__ bool macLF(int z0, int z1, int z2) {
  const int test = abs(z0) + abs(z1) + abs(z2);
  const int ncmp = (z0 != 0) + (z1 != 0) + (z2 != 0);
  return test > ncmp;
}

__device__ __forceinline__ int winIndex(int z0, int z1, int z2, int W) {
  const int n = 2 * W + 1;
  return ((z0 + W) * n + (z1 + W)) * n + (z2 + W);
}

template <bool WRITE>
__global__ __launch_bounds__(256) void k_walk(WalkArgs a) {
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (t >= a.nt) return;  // uniform per wavefront
  const int i = a.tcell[t], p = a.tpar[t];
  const int eb = a.eBeg[p], ee = a.eBeg[p + 1];
  const int ti0 = a.T.idx[3 * i], ti1 = a.T.idx[3 * i + 1], ti2 = a.T.idx[3 * i + 2];
  const double tc0 = a.T.ctr[3 * i], tc1 = a.T.ctr[3 * i + 1], tc2 = a.T.ctr[3 * i + 2];
  const bool tleaf = a.T.nson[i] == 0;
  const unsigned long long lt = (1ull << lane) - 1ull;
  int nM = 0, nP = 0, nE = 0;
  int bM = 0, bP = 0, bE = 0;
  if (WRITE) {
    bM = a.oM[t];
    bP = a.oP[t];
    bE = a.oE[t];
  }
  for (int e0 = eb; e0 < ee; e0 += 8) {
    const int e = e0 + (lane >> 3), b = lane & 7;
    int j = -1;
    if (e < ee) {
      const int q = a.eSrc[e];
      if (b < a.S.nson[q]) j = a.S.sons[a.S.son0[q] + b];
    }
    int cls = 0, z0 = 0, z1 = 0, z2 = 0;
    if (j >= 0) {
      z0 = ti0 - a.S.idx[3 * j];
      z1 = ti1 - a.S.idx[3 * j + 1];
      z2 = ti2 - a.S.idx[3 * j + 2];
      bool far = false;
      if (a.kind == 1) far = macHF(tc0, tc1, tc2, a.S.ctr + 3 * j, a.S.rad[j], a.kappa);
      else if (a.kind == 2) far = macLF(z0, z1, z2);
      cls = far ? 1 : (tleaf || a.S.nson[j] == 0) ? 2 : 3;
    }
    const unsigned long long mM = __ballot(cls == 1), mP = __ballot(cls == 2),
                             mE = __ballot(cls == 3);
    if (WRITE) {
      if (cls == 1) {
        a.mSrc[bM + nM + __popcll(mM & lt)] = j;
        const int W = a.W;
        if (abs(z0) <= W && abs(z1) <= W && abs(z2) <= W) {
          unsigned long long* r = a.rep + winIndex(z0, z1, z2, W);
          const unsigned long long v = ((unsigned long long)(unsigned)i << 32) | (unsigned)j;
          if (v < __atomic_load_n(r, __ATOMIC_RELAXED)) atomicMin(r, v);
        } else {
          atomicOr(a.overflow, 1);
        }
      } else if (cls == 2) {
        a.pSrc[bP + nP + __popcll(mP & lt)] = j;
      } else if (cls == 3) {
        a.eOut[bE + nE + __popcll(mE & lt)] = j;
      }
    }
    nM += __popcll(mM);
    nP += __popcll(mP);
    nE += __popcll(mE);
  }
  if (!WRITE && lane == 0) {
    a.cM[t] = nM;
    a.cP[t] = nP;
    a.cE[t] = nE;
  }
}

// Octree build, one level: the points of each splitting cell, in chunks of
// kTreeChunk consecutive positions, a block per chunk and kTreePer points per
// thread in position order.  k_tree_count writes each point's octant (the
// host builder's test x_k > c_k) and the chunk's octant counts;
// k_tree_part moves each point to its octant's range at the offset the host
// derived from the counts plus its rank among the chunk's points of that
// octant that precede it (a stable partition).
constexpr int kTreePer = 16;
constexpr int kTreeChunk = 256 * kTreePer;

struct TreeChunk {
  int cell;   // index among the level's splitting cells
  int start;  // first position of the chunk
  int len;
};

__global__ __launch_bounds__(256) void k_tree_count(const TreeChunk* __restrict__ ch,
                                                    const double* __restrict__ cctr,
                                                    const int* __restrict__ perm,
                                                    const double* __restrict__ xyz,
                                                    uint8_t* __restrict__ code, int* __restrict__ cnt) {
  __shared__ int s[8];
  const TreeChunk c = ch[blockIdx.x];
  if (threadIdx.x < 8) s[threadIdx.x] = 0;
  __syncthreads();
  const double c0 = cctr[3 * c.cell], c1 = cctr[3 * c.cell + 1], c2 = cctr[3 * c.cell + 2];
  int local[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < kTreePer; k++) {
    const int i = threadIdx.x * kTreePer + k;
    if (i >= c.len) break;
    const size_t p = (size_t)perm[c.start + i];
    const int o = (xyz[3 * p] > c0 ? 1 : 0) | (xyz[3 * p + 1] > c1 ? 2 : 0) | (xyz[3 * p + 2] > c2 ? 4 : 0);
    code[c.start + i] = (uint8_t)o;
    local[o]++;
  }
#pragma unroll
  for (int o = 0; o < 8; o++)
    if (local[o]) atomicAdd(&s[o], local[o]);
  __syncthreads();
  if (threadIdx.x < 8) cnt[(size_t)blockIdx.x * 8 + threadIdx.x] = s[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_tree_part(const TreeChunk* __restrict__ ch,
                                                   const int* __restrict__ off,
                                                   const uint8_t* __restrict__ code,
                                                   const int* __restrict__ perm, int* __restrict__ out) {
  __shared__ int sc[8][256];
  const TreeChunk c = ch[blockIdx.x];
  const int t = threadIdx.x;
  int local[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < kTreePer; k++) {
    const int i = t * kTreePer + k;
    if (i >= c.len) break;
    local[code[c.start + i]]++;
  }
#pragma unroll
  for (int o = 0; o < 8; o++) sc[o][t] = local[o];
  __syncthreads();
  if (t < 8) {
    // exclusive prefix over the threads, one octant per thread
    int run = 0;
    for (int u = 0; u < 256; u++) {
      const int v = sc[t][u];
      sc[t][u] = run;
      run += v;
    }
  }
  __syncthreads();
  int pos[8];
#pragma unroll
  for (int o = 0; o < 8; o++) pos[o] = off[(size_t)blockIdx.x * 8 + o] + sc[o][t];
  for (int k = 0; k < kTreePer; k++) {
    const int i = t * kTreePer + k;
    if (i >= c.len) break;
    const int o = code[c.start + i];
    out[pos[o]++] = perm[c.start + i];
  }
}

// Cone of each offset met (M2L.hpp:189-197, the host's bestCone): the
// direction nearest to the representative pair's normalised centre
// difference, the first within 1e-12 of the best winning, with the host's
// operations in the host's order (no contraction in this file; double
// square root and division correctly rounded).
__global__ __launch_bounds__(256) void k_cones(const unsigned long long* __restrict__ rep, int n,
                                               const double* __restrict__ Tctr,
                                               const double* __restrict__ Sctr,
                                               const double* __restrict__ dirs, int ndir,
                                               int* __restrict__ cone) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const unsigned long long r = rep[k];
  if (r == ~0ull) {
    cone[k] = -1;
    return;
  }
  const int i = (int)(r >> 32), j = (int)(r & 0xFFFFFFFFull);
  double f0 = Tctr[3 * i] - Sctr[3 * j], f1 = Tctr[3 * i + 1] - Sctr[3 * j + 1],
         f2 = Tctr[3 * i + 2] - Sctr[3 * j + 2];
  double t = 0.;
  t += f0 * f0;
  t += f1 * f1;
  t += f2 * f2;
  const double nn = sqrt(t);
  f0 /= nn;
  f1 /= nn;
  f2 /= nn;
  double best = 1.e30;
  int ind = -1;
  for (int d = 0; d < ndir; d++) {
    const double e0 = dirs[3 * d] - f0, e1 = dirs[3 * d + 1] - f1, e2 = dirs[3 * d + 2] - f2;
    double u = 0.;
    u += e0 * e0;
    u += e1 * e1;
    u += e2 * e2;
    const double m = sqrt(u);
    if (m < best - 1.e-12) {
      best = m;
      ind = d;
    }
  }
  cone[k] = ind;
}

// The M2L pairs of one level, CSR by target cell (the level's target list).
struct PairArgs {
  int nt;
  const int* tcell;
  const int* mBeg;
  const int* mSrc;
  const int* Tidx;
  const int* Sidx;
  const unsigned long long* code;  // per window offset: id | hash << 32 | (cone + 1) << 40
  int W;
};

// Effective directions (M2L.hpp:189-197): both cells of an HF pair need the
// pair's cone.
__global__ __launch_bounds__(256) void k_marks(PairArgs a, const int64_t* Tmark0,
                                               const int64_t* Smark0, uint8_t* Tm, uint8_t* Sm) {
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (t >= a.nt) return;
  const int i = a.tcell[t];
  const int ti0 = a.Tidx[3 * i], ti1 = a.Tidx[3 * i + 1], ti2 = a.Tidx[3 * i + 2];
  const int64_t tm = Tmark0[i];
  for (int e = a.mBeg[t] + lane; e < a.mBeg[t + 1]; e += 64) {
    const int j = a.mSrc[e];
    const unsigned long long c =
        a.code[winIndex(ti0 - a.Sidx[3 * j], ti1 - a.Sidx[3 * j + 1], ti2 - a.Sidx[3 * j + 2], a.W)];
    const int cone = (int)(c >> 40) - 1;
    Tm[tm + cone] = 1;
    Sm[Smark0[j] + cone] = 1;
  }
}

struct DSlots {
  const int* slot0;
  const int* nslot;
  const uint8_t* lf;     // the cell's single slot is an LF slot
  const int64_t* mark0;
  const int* slotIdx;    // per (cell, direction) of the marks
};

// Tree::slotOf
__device__ __forceinline__ int slotOf(const DSlots& s, int c, int d) {
  if (s.nslot[c] == 0) return -1;
  if (s.lf[c]) return s.slot0[c];
  const int64_t m = s.mark0[c];
  return (d >= 0 && m >= 0) ? s.slotIdx[m + d] : -1;
}

struct ListArgs {
  PairArgs pa;
  int hf;
  DSlots TS, SS;
  int* cnt;          // per target slot
  uint8_t* used;     // per source slot
  int* missing;
  int* cur;          // per target slot: next free entry
  const int* m2fIndex;
  int* outSrcF;
  int* outTab;
  int* outHash;
};

__device__ __forceinline__ void pairSlots(const ListArgs& a, int i, int ti0, int ti1, int ti2,
                                          int j, int& ts, int& ss, int& id, int& h) {
  const PairArgs& p = a.pa;
  const unsigned long long c =
      p.code[winIndex(ti0 - p.Sidx[3 * j], ti1 - p.Sidx[3 * j + 1], ti2 - p.Sidx[3 * j + 2], p.W)];
  const int cone = a.hf ? (int)(c >> 40) - 1 : -1;
  id = (int)(c & 0xFFFFFFFFu);
  h = (int)((c >> 32) & 63);
  ts = slotOf(a.TS, i, cone);
  ss = slotOf(a.SS, j, cone);
}

__global__ __launch_bounds__(256) void k_list_count(ListArgs a) {
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (t >= a.pa.nt) return;
  const int i = a.pa.tcell[t];
  const int ti0 = a.pa.Tidx[3 * i], ti1 = a.pa.Tidx[3 * i + 1], ti2 = a.pa.Tidx[3 * i + 2];
  for (int e = a.pa.mBeg[t] + lane; e < a.pa.mBeg[t + 1]; e += 64) {
    int ts, ss, id, h;
    pairSlots(a, i, ti0, ti1, ti2, a.pa.mSrc[e], ts, ss, id, h);
    if (ts < 0 || ss < 0) {
      atomicAdd(a.missing, 1);
    } else {
      atomicAdd(a.cnt + ts, 1);
      a.used[ss] = 1;
    }
  }
}

// Stable placement: the pairs of a target cell in segment order, each at its
// target slot's cursor.  Per 64-pair chunk, the lanes that share the first
// unplaced lane's slot take consecutive entries (one atomic on the cursor,
// whose returned value orders successive chunks), until every lane is placed.
__global__ __launch_bounds__(256) void k_list_place(ListArgs a) {
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (t >= a.pa.nt) return;
  const int i = a.pa.tcell[t];
  const int ti0 = a.pa.Tidx[3 * i], ti1 = a.pa.Tidx[3 * i + 1], ti2 = a.pa.Tidx[3 * i + 2];
  const unsigned long long lt = (1ull << lane) - 1ull;
  const int mb = a.pa.mBeg[t], me = a.pa.mBeg[t + 1];
  for (int e0 = mb; e0 < me; e0 += 64) {
    const int e = e0 + lane;
    const bool v = e < me;
    int ts = -1, ss = -1, id = 0, h = 0;
    if (v) pairSlots(a, i, ti0, ti1, ti2, a.pa.mSrc[e], ts, ss, id, h);
    unsigned long long todo = __ballot(v);
    int pos = 0;
    while (todo) {
      const int lead = __ffsll((long long)todo) - 1;
      const int s0 = __shfl(ts, lead);
      const unsigned long long m = __ballot(v && ts == s0) & todo;
      int base = 0;
      if (lane == lead) base = atomicAdd(a.cur + s0, __popcll(m));
      base = __shfl(base, lead);
      if ((m >> lane) & 1ull) pos = base + __popcll(m & lt);
      todo &= ~m;
    }
    if (v) {
      a.outSrcF[pos] = a.m2fIndex[ss];
      a.outTab[pos] = id;
      a.outHash[pos] = h;
    }
  }
}

// Grouped M2L lists: per group of up to G target slots, the union of their
// source rows in ascending order, each with the mask of the targets reading
// it and their (table << 6 | hash) codes (0 for targets not reading it).  A
// thread per group merges its targets' lists, which ascend strictly (a
// source row repeated within one target would make the host's grouping
// differ: it is flagged and the plan goes to the host builder).
struct GroupArgs {
  int ngroups;
  const int* target;  // ngroups x G target slots, -1 padded
  const int* beg;     // M2L CSR by target slot
  const int* srcF;
  const int* tab;
  const int* hash;
  int* count;         // count pass: entries per group
  const int* gBeg;    // write pass: first entry of each group
  int* outSrc;
  int* outMask;
  int* outCode;       // G per entry
  int* bad;
};

template <int G, bool WRITE>
__global__ __launch_bounds__(256) void k_groups(GroupArgs a) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= a.ngroups) return;
  constexpr int kEnd = 0x7FFFFFFF;
  int h[G], e[G], v[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    const int t = a.target[(size_t)gi * G + g];
    h[g] = t >= 0 ? a.beg[t] : 0;
    e[g] = t >= 0 ? a.beg[t + 1] : 0;
    v[g] = h[g] < e[g] ? a.srcF[h[g]] : kEnd;
  }
  int n = 0;
  const int o = WRITE ? a.gBeg[gi] : 0;
  for (;;) {
    int m = kEnd;
#pragma unroll
    for (int g = 0; g < G; g++) m = min(m, v[g]);
    if (m == kEnd) break;
    int mask = 0;
#pragma unroll
    for (int g = 0; g < G; g++) {
      int code = 0;
      if (v[g] == m) {
        mask |= 1 << g;
        code = (a.tab[h[g]] << 6) | a.hash[h[g]];
        h[g]++;
        const int nv = h[g] < e[g] ? a.srcF[h[g]] : kEnd;
        if (nv <= m) atomicOr(a.bad, 1);
        v[g] = nv;
      }
      if (WRITE) a.outCode[(size_t)(o + n) * G + g] = code;
    }
    if (WRITE) {
      a.outSrc[o + n] = m;
      a.outMask[o + n] = mask;
    }
    n++;
  }
  if (!WRITE) a.count[gi] = n;
}

// P2P lists (P2P.hpp:42-79 over whole cell ranges): a target leaf's near
// sources are those of the P2P pairs of the leaf and of each of its
// ancestors (a pair with a non-leaf target covers all its leaves).  Each such
// list is in ascending source cell order; a thread per leaf merges them by
// source range (pbeg, pbeg + nprt), as the host sorts them, and joins
// touching ranges, keeping others (overlapping ones included) separate.
struct P2PArgs {
  int nleaf;
  const int* leaf;          // target leaves in pre-order
  const int* parent;        // per target cell, -1 at the root
  const int* lvl;
  const int2* seg;          // per target cell: its P2P pairs [x, y) in its level's list
  const int* const* src;    // per level: the P2P pairs' source cells
  const int* Spbeg;
  const int* Snprt;
  const int* Tnprt;
  int* count;               // count pass: ranges per leaf
  unsigned long long* pairs;  // count pass: sum of n_target x n_source over the pairs
  const int* beg;           // write pass: first range of each leaf
  int* outB;
  int* outE;
};

template <bool WRITE>
__global__ __launch_bounds__(256) void k_p2p_leaf(P2PArgs a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = k < a.nleaf;
  unsigned long long pts = 0;
  if (valid) {
    const int leaf = a.leaf[k];
    int hl[32], hp[32], he[32];
    int nh = 0;
    for (int c = leaf; c >= 0 && nh < 32; c = a.parent[c]) {
      const int2 sg = a.seg[c];
      if (sg.y > sg.x) {
        hl[nh] = a.lvl[c];
        hp[nh] = sg.x;
        he[nh] = sg.y;
        nh++;
      }
    }
    int n = 0;
    const int o = WRITE ? a.beg[k] : 0;
    int cb = 0, ce = 0;
    bool open = false;
    for (;;) {
      int best = -1, bb = 0x7FFFFFFF, be = 0x7FFFFFFF;
      for (int h = 0; h < nh; h++) {
        if (hp[h] >= he[h]) continue;
        const int j = a.src[hl[h]][hp[h]];
        const int rb = a.Spbeg[j], re = rb + a.Snprt[j];
        if (rb < bb || (rb == bb && re < be)) {
          best = h;
          bb = rb;
          be = re;
        }
      }
      if (best < 0) break;
      hp[best]++;
      pts += (unsigned long long)(be - bb);
      if (open && bb == ce) {
        ce = be;
      } else {
        if (open) {
          if (WRITE) {
            a.outB[o + n] = cb;
            a.outE[o + n] = ce;
          }
          n++;
        }
        cb = bb;
        ce = be;
        open = true;
      }
    }
    if (open) {
      if (WRITE) {
        a.outB[o + n] = cb;
        a.outE[o + n] = ce;
      }
      n++;
    }
    if (!WRITE) a.count[k] = n;
    pts *= (unsigned long long)a.Tnprt[leaf];
  }
  if (!WRITE) {
    // one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) pts += __shfl_down(pts, off);
    if ((threadIdx.x & 63) == 0 && pts) atomicAdd(a.pairs, pts);
  }
}

class GpuSetup final : public SetupDevice {
 public:
  ~GpuSetup() override {
    for (void* p : owned_) (void)hipFree(p);
  }
  int tree(Tree& t, const double* xyz, int64_t n, const double ctr[3], double rad, int ncrit,
           int lvmax, std::string& msg) override;
  int walk(Plan& pl, std::vector<std::array<int, 4>>& tabKeys, std::vector<int>& pt,
           std::vector<int>& ps, std::string& msg) override;
  int m2lLists(Plan& pl, std::string& msg) override;
  int m2lGroups(Plan& pl, std::string& msg) override;
  int p2pLists(Plan& pl, std::string& msg) override;

 private:
  // P2P pairs the walk found: per level the source cells, per target cell
  // its range in its level's list
  std::vector<int*> pLv_;
  std::vector<int2> pSeg_;
  // the M2L CSR entries left on the device by m2lLists
  int* dSrcF_ = nullptr;
  int* dTab_ = nullptr;
  int* dHash_ = nullptr;
  struct Level {
    int l = 0, nt = 0, kind = 0, W = -1;
    int* tcell = nullptr;
    int* mBeg = nullptr;   // nt + 1
    int* mSrc = nullptr;
    unsigned long long* code = nullptr;
    int64_t nm = 0;
  };
  std::vector<Level> lv_;
  std::vector<void*> owned_;
  int* Tidx_ = nullptr;
  int* Sidx_ = nullptr;

  template <class T>
  hipError_t alloc(T*& p, size_t n) {
    p = nullptr;
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
    if (e == hipSuccess) owned_.push_back((void*)p);
    else p = nullptr;
    return e;
  }
  template <class T>
  hipError_t put(T*& d, const T* h, size_t n) {
    hipError_t e = alloc(d, n);
    if (e != hipSuccess || n == 0) return e;
    return hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice);
  }
  template <class T>
  hipError_t put(T*& d, const std::vector<T>& h) {
    return put(d, h.data(), h.size());
  }
  void release(void* p) {
    if (!p) return;
    for (size_t k = 0; k < owned_.size(); k++)
      if (owned_[k] == p) {
        (void)hipFree(p);
        owned_[k] = owned_.back();
        owned_.pop_back();
        return;
      }
  }
};

// A failed HIP call hands the plan back to the host builder (the error is
// consumed, so that later calls do not report it).
#define SCK(call)                                                                   \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      (void)hipGetLastError();                                                      \
      msg = std::string("device set-up: ") + hipGetErrorString(e_) + " at " #call; \
      return kSetupFallback;                                                        \
    }                                                                               \
  } while (0)

inline unsigned blocksFor(int nt) { return (unsigned)((nt + kWaves - 1) / kWaves); }

int GpuSetup::tree(Tree& t, const double* xyz, int64_t n, const double ctr[3], double rad,
                   int ncrit, int lvmax, std::string& msg) {
  double* dxyz = nullptr;
  int *perm = nullptr, *perm2 = nullptr;
  uint8_t* code = nullptr;
  SCK(put(dxyz, xyz, 3 * (size_t)n));
  {
    std::vector<int> id(n);
    for (int64_t i = 0; i < n; i++) id[i] = (int)i;
    SCK(put(perm, id));
  }
  SCK(alloc(perm2, (size_t)n));
  SCK(alloc(code, (size_t)n));
  std::vector<std::vector<LevelCell>> lev(1);
  LevelCell r;
  r.b = 0;
  r.n = n;
  for (int q = 0; q < 3; q++) r.ctr[q] = ctr[q];
  r.rad = rad;
  lev[0].push_back(r);
  for (int l = 0;; l++) {
    std::vector<int> split;
    for (int c = 0; c < (int)lev[l].size(); c++)
      if (!cellStops(lev[l][c].n, l, ncrit, lvmax)) split.push_back(c);
    if (split.empty()) break;
    if (l >= 30) return ERR_DEPTH;  // the children would be deeper than 30 levels
    std::vector<TreeChunk> chunks;
    std::vector<double> cc(3 * split.size());
    for (size_t s = 0; s < split.size(); s++) {
      const LevelCell& c = lev[l][split[s]];
      for (int q = 0; q < 3; q++) cc[3 * s + q] = c.ctr[q];
      for (int64_t o = 0; o < c.n; o += kTreeChunk)
        chunks.push_back({(int)s, (int)(c.b + o), (int)std::min<int64_t>(kTreeChunk, c.n - o)});
    }
    TreeChunk* dch = nullptr;
    double* dcc = nullptr;
    int *cnt = nullptr, *off = nullptr;
    SCK(put(dch, chunks));
    SCK(put(dcc, cc));
    SCK(alloc(cnt, chunks.size() * 8));
    k_tree_count<<<(unsigned)chunks.size(), 256, 0, 0>>>(dch, dcc, perm, dxyz, code, cnt);
    SCK(hipGetLastError());
    std::vector<int> hc(chunks.size() * 8);
    SCK(hipMemcpy(hc.data(), cnt, hc.size() * sizeof(int), hipMemcpyDeviceToHost));
    // octant totals per cell, then each chunk's first position per octant
    std::vector<int64_t> tot(split.size() * 8, 0);
    for (size_t k = 0; k < chunks.size(); k++)
      for (int o = 0; o < 8; o++) tot[chunks[k].cell * 8 + o] += hc[k * 8 + o];
    std::vector<int64_t> run(split.size() * 8);
    for (size_t s = 0; s < split.size(); s++) {
      int64_t b = lev[l][split[s]].b;
      for (int o = 0; o < 8; o++) {
        run[s * 8 + o] = b;
        b += tot[s * 8 + o];
      }
    }
    std::vector<int> ho(chunks.size() * 8);
    for (size_t k = 0; k < chunks.size(); k++)
      for (int o = 0; o < 8; o++) {
        ho[k * 8 + o] = (int)run[chunks[k].cell * 8 + o];
        run[chunks[k].cell * 8 + o] += hc[k * 8 + o];
      }
    SCK(put(off, ho));
    // positions outside the splitting cells keep their points
    SCK(hipMemcpy(perm2, perm, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice));
    k_tree_part<<<(unsigned)chunks.size(), 256, 0, 0>>>(dch, off, code, perm, perm2);
    SCK(hipGetLastError());
    std::swap(perm, perm2);
    release(dch);
    release(dcc);
    release(cnt);
    release(off);
    std::vector<LevelCell> next;
    for (size_t s = 0; s < split.size(); s++) splitCell(lev[l][split[s]], &tot[s * 8], next);
    lev.push_back(std::move(next));
  }
  std::vector<int> hp(n);
  SCK(copyLarge(hp.data(), perm, (size_t)n * sizeof(int), false));
  release(dxyz);
  release(perm);
  release(perm2);
  release(code);
  treeFromLevels(t, lev, hp, xyz);
  return OK;
}

int GpuSetup::walk(Plan& pl, std::vector<std::array<int, 4>>& tabKeys, std::vector<int>& pt,
                   std::vector<int>& ps, std::string& msg) {
  const Tree &T = pl.T, &S = pl.S;
  std::vector<int> W(pl.nlvl, -1);
  for (int l = 0; l < pl.nlvl; l++) {
    W[l] = m2lWindow(pl, l);
    if (pl.precomp[l] && W[l] < 0) {
      msg = "device set-up: offset window too large";
      return kSetupFallback;
    }
  }
  pt.clear();
  ps.clear();
  int64_t nHF = 0, nLF = 0, nP2P = 0;
  // the root pair (level 0 has no M2L)
  pLv_.assign(pl.nlvl, nullptr);
  pSeg_.assign(T.ncells(), make_int2(0, 0));
  if (T.leaf(0) || S.leaf(0)) {
    const int root = 0;
    SCK(put(pLv_[0], &root, 1));
    pSeg_[0] = make_int2(0, 1);
    pl.st.hf_m2l = pl.st.lf_m2l = 0;
    pl.st.p2p_calls = 1;
    tabKeys.clear();
    return OK;
  }
  // target cells per level (cell order) and their parents' positions
  const int nT = T.ncells();
  std::vector<int> pos(nT, 0), parent(nT, -1);
  std::vector<std::vector<int>> cells(pl.nlvl);
  for (int c = 0; c < nT; c++) {
    pos[c] = (int)cells[T.lvl[c]].size();
    cells[T.lvl[c]].push_back(c);
    for (int a = 0; a < T.nson[c]; a++) parent[T.sons[T.son0[c] + a]] = c;
  }
  double *Tctr = nullptr, *Sctr = nullptr, *Srad = nullptr;
  int *Tnson = nullptr, *Snson = nullptr, *Sson0 = nullptr, *Ssons = nullptr;
  SCK(put(Tidx_, T.idx));
  SCK(put(Tctr, T.ctr));
  SCK(put(Tnson, T.nson));
  SCK(put(Sidx_, S.idx));
  SCK(put(Sctr, S.ctr));
  SCK(put(Srad, S.rad));
  SCK(put(Snson, S.nson));
  SCK(put(Sson0, S.son0));
  SCK(put(Ssons, S.sons));
  int* overflow = nullptr;
  SCK(alloc(overflow, 1));
  SCK(hipMemset(overflow, 0, sizeof(int)));
  std::vector<unsigned long long*> rep(pl.nlvl, nullptr);
  for (int l = 0; l < pl.nlvl; l++) {
    if (W[l] < 0) continue;
    const size_t n = (size_t)(2 * W[l] + 1) * (2 * W[l] + 1) * (2 * W[l] + 1);
    SCK(alloc(rep[l], n));
    SCK(hipMemset(rep[l], 0xFF, n * sizeof(unsigned long long)));
  }
  // level 0: the root target expands the root source
  int *eBeg = nullptr, *eSrc = nullptr;
  {
    const int b[2] = {0, 1}, s0 = 0;
    SCK(put(eBeg, b, 2));
    SCK(put(eSrc, &s0, 1));
  }
  int64_t nE = 1, nmTotal = 0;
  const DTree dT = {Tidx_, Tctr, nullptr, nullptr, Tnson, nullptr};
  const DTree dS = {Sidx_, Sctr, Srad, Sson0, Snson, Ssons};
  for (int l = 1; l < pl.nlvl && nE > 0; l++) {
    const int nt = (int)cells[l].size();
    if (nt == 0) break;
    std::vector<int> tpar(nt);
    for (int t = 0; t < nt; t++) tpar[t] = pos[parent[cells[l][t]]];
    Level L;
    L.l = l;
    L.nt = nt;
    L.kind = pl.precomp[l] ? ((l && pl.isHF[l]) ? 1 : 2) : 0;
    L.W = W[l];
    int *tparD = nullptr, *cnt = nullptr;
    SCK(put(L.tcell, cells[l]));
    SCK(put(tparD, tpar));
    SCK(alloc(cnt, (size_t)3 * nt));
    WalkArgs a = {};
    a.T = dT;
    a.S = dS;
    a.nt = nt;
    a.tcell = L.tcell;
    a.tpar = tparD;
    a.eBeg = eBeg;
    a.eSrc = eSrc;
    a.kind = L.kind;
    a.kappa = pl.kappa;
    a.cM = cnt;
    a.cP = cnt + nt;
    a.cE = cnt + 2 * (size_t)nt;
    a.rep = rep[l];
    a.W = L.W;
    a.overflow = overflow;
    k_walk<false><<<blocksFor(nt), 256, 0, 0>>>(a);
    SCK(hipGetLastError());
    std::vector<int> hc((size_t)3 * nt);
    SCK(hipMemcpy(hc.data(), cnt, hc.size() * sizeof(int), hipMemcpyDeviceToHost));
    release(cnt);
    // exclusive sums: oM | oP | oE, nt + 1 each
    std::vector<int> off((size_t)3 * (nt + 1));
    int64_t tot[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) {
      for (int t = 0; t < nt; t++) {
        off[(size_t)k * (nt + 1) + t] = (int)tot[k];
        tot[k] += hc[(size_t)k * nt + t];
        if (tot[k] >= ((int64_t)1 << 31)) {
          msg = "device set-up: more than 2^31 pairs at one level";
          return kSetupFallback;
        }
      }
      off[(size_t)k * (nt + 1) + nt] = (int)tot[k];
    }
    int* offD = nullptr;
    SCK(put(offD, off));
    SCK(alloc(L.mSrc, (size_t)tot[0]));
    int *pSrc = nullptr, *eOut = nullptr;
    SCK(alloc(pSrc, (size_t)tot[1]));
    SCK(alloc(eOut, (size_t)tot[2]));
    a.oM = offD;
    a.oP = offD + (nt + 1);
    a.oE = offD + 2 * (size_t)(nt + 1);
    a.mSrc = L.mSrc;
    a.pSrc = pSrc;
    a.eOut = eOut;
    k_walk<true><<<blocksFor(nt), 256, 0, 0>>>(a);
    SCK(hipGetLastError());
    // P2P pairs stay on the device for the P2P lists
    pLv_[l] = pSrc;
    for (int t = 0; t < nt; t++)
      pSeg_[cells[l][t]] = make_int2(off[(size_t)(nt + 1) + t], off[(size_t)(nt + 1) + t + 1]);
    release(tparD);
    // M2L segments: keep the offsets on both sides
    SCK(alloc(L.mBeg, (size_t)nt + 1));
    SCK(hipMemcpy(L.mBeg, offD, ((size_t)nt + 1) * sizeof(int), hipMemcpyDeviceToDevice));
    release(offD);
    L.nm = tot[0];
    if (L.kind == 1) nHF += tot[0];
    if (L.kind == 2) nLF += tot[0];
    nP2P += tot[1];
    nmTotal += tot[0];
    // the next level expands this level's expansions
    release(eBeg);
    release(eSrc);
    std::vector<int> eb(off.begin() + 2 * (size_t)(nt + 1), off.end());
    SCK(put(eBeg, eb));
    eSrc = eOut;
    nE = tot[2];
    if (L.nm) lv_.push_back(L);
    else {
      release(L.tcell);
      release(L.mSrc);
      release(L.mBeg);
    }
  }
  release(eBeg);
  release(eSrc);
  int ovf = 0;
  SCK(hipMemcpy(&ovf, overflow, sizeof(int), hipMemcpyDeviceToHost));
  if (ovf) {
    msg = "device set-up: an M2L offset outside the level's window";
    return kSetupFallback;
  }
  if (nmTotal >= ((int64_t)1 << 31)) {
    msg = "M2L lists exceed 2^31 entries";
    return ERR_MEMORY;
  }
  // table ids, hashes and cones of the offsets met
  std::vector<std::vector<uint64_t>> hrep(pl.nlvl), code;
  for (int l = 0; l < pl.nlvl; l++) {
    if (!rep[l]) continue;
    const size_t n = (size_t)(2 * W[l] + 1) * (2 * W[l] + 1) * (2 * W[l] + 1);
    hrep[l].resize(n);
    SCK(hipMemcpy(hrep[l].data(), rep[l], n * sizeof(uint64_t), hipMemcpyDeviceToHost));
  }
  // cones of the HF levels' offsets on the device
  std::vector<std::vector<int>> cones(pl.nlvl);
  for (int l = 0; l < pl.nlvl; l++) {
    if (!rep[l] || !(l && pl.isHF[l])) continue;
    const std::vector<std::array<double, 3>>& D = pl.dirs[pl.HFStage[l]];
    std::vector<double> dd(3 * D.size());
    for (size_t d = 0; d < D.size(); d++)
      for (int q = 0; q < 3; q++) dd[3 * d + q] = D[d][q];
    const int n = (int)hrep[l].size();
    double* dD = nullptr;
    int* dc = nullptr;
    SCK(put(dD, dd));
    SCK(alloc(dc, (size_t)n));
    k_cones<<<(unsigned)((n + 255) / 256), 256, 0, 0>>>(rep[l], n, Tctr, Sctr, dD, (int)D.size(), dc);
    SCK(hipGetLastError());
    cones[l].resize(n);
    SCK(hipMemcpy(cones[l].data(), dc, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    release(dD);
    release(dc);
  }
  int r = offsetCodes(pl, W, hrep, code, tabKeys, msg, &cones);
  if (r != OK) return r;
  for (int l = 0; l < pl.nlvl; l++) {
    if (!rep[l]) continue;
    SCK(hipMemcpy(rep[l], code[l].data(), code[l].size() * sizeof(uint64_t),
                  hipMemcpyHostToDevice));
  }
  for (Level& L : lv_) L.code = rep[L.l];
  // effective-direction marks of both trees
  {
    uint8_t *Tm = nullptr, *Sm = nullptr;
    int64_t *Tmark0 = nullptr, *Smark0 = nullptr;
    SCK(alloc(Tm, T.marks.size()));
    SCK(alloc(Sm, S.marks.size()));
    if (Tm) SCK(hipMemset(Tm, 0, T.marks.size()));
    if (Sm) SCK(hipMemset(Sm, 0, S.marks.size()));
    SCK(put(Tmark0, T.mark0));
    SCK(put(Smark0, S.mark0));
    for (const Level& L : lv_) {
      if (L.kind != 1) continue;
      PairArgs pa = {L.nt, L.tcell, L.mBeg, L.mSrc, Tidx_, Sidx_, L.code, L.W};
      k_marks<<<blocksFor(L.nt), 256, 0, 0>>>(pa, Tmark0, Smark0, Tm, Sm);
      SCK(hipGetLastError());
    }
    if (Tm) SCK(copyLarge(pl.T.marks.data(), Tm, T.marks.size(), false));
    if (Sm) SCK(copyLarge(pl.S.marks.data(), Sm, S.marks.size(), false));
    release(Tm);
    release(Sm);
    release(Tmark0);
    release(Smark0);
  }
  release(Tctr);
  release(Sctr);
  release(Srad);
  release(Tnson);
  release(Snson);
  release(Sson0);
  release(Ssons);
  pl.st.hf_m2l = nHF;
  pl.st.lf_m2l = nLF;
  pl.st.p2p_calls = nP2P;
  return OK;
}

int GpuSetup::m2lLists(Plan& pl, std::string& msg) {
  const Tree &T = pl.T, &S = pl.S;
  const int nts = T.nslots, nss = S.nslots;
  int64_t nm = 0;
  for (const Level& L : lv_) nm += L.nm;
  pl.m2lBeg.assign((size_t)nts + 1, 0);
  pl.m2lEntries = 0;
  pl.m2fIndex.assign(nss, -1);
  pl.m2fSlot.clear();
  pl.m2lSrcF.clear();
  pl.m2lTabId.clear();
  pl.m2lHash.clear();
  if (nm == 0) return OK;
  auto upSlots = [&](const Tree& t, DSlots& d) -> hipError_t {
    std::vector<uint8_t> lf(t.ncells(), 0);
    for (int c = 0; c < t.ncells(); c++) lf[c] = t.nslot[c] > 0 && t.slotDir[t.slot0[c]] < 0;
    int *s0 = nullptr, *ns = nullptr, *si = nullptr;
    uint8_t* lfD = nullptr;
    int64_t* m0 = nullptr;
    hipError_t e;
    if ((e = put(s0, t.slot0)) != hipSuccess) return e;
    if ((e = put(ns, t.nslot)) != hipSuccess) return e;
    if ((e = put(lfD, lf)) != hipSuccess) return e;
    if ((e = put(m0, t.mark0)) != hipSuccess) return e;
    if ((e = put(si, t.slotIdx)) != hipSuccess) return e;
    d = {s0, ns, lfD, m0, si};
    return hipSuccess;
  };
  ListArgs a = {};
  SCK(upSlots(T, a.TS));
  SCK(upSlots(S, a.SS));
  int* missing = nullptr;
  SCK(alloc(a.cnt, (size_t)nts));
  SCK(alloc(a.used, (size_t)nss));
  SCK(alloc(missing, 1));
  SCK(hipMemset(a.cnt, 0, (size_t)nts * sizeof(int)));
  SCK(hipMemset(a.used, 0, (size_t)nss));
  SCK(hipMemset(missing, 0, sizeof(int)));
  a.missing = missing;
  for (const Level& L : lv_) {
    a.pa = {L.nt, L.tcell, L.mBeg, L.mSrc, Tidx_, Sidx_, L.code, L.W};
    a.hf = L.kind == 1;
    k_list_count<<<blocksFor(L.nt), 256, 0, 0>>>(a);
    SCK(hipGetLastError());
  }
  int nmiss = 0;
  SCK(hipMemcpy(&nmiss, missing, sizeof(int), hipMemcpyDeviceToHost));
  if (nmiss) {
    msg = "internal: missing M2L slot";
    return ERR_STATE;
  }
  {
    std::vector<int> c(nts);
    SCK(hipMemcpy(c.data(), a.cnt, (size_t)nts * sizeof(int), hipMemcpyDeviceToHost));
    int64_t run = 0;
    for (int s = 0; s < nts; s++) {
      run += c[s];
      pl.m2lBeg[s + 1] = (int)run;
    }
    if (run != nm) {
      msg = "internal: M2L entry count mismatch";
      return ERR_STATE;
    }
    std::vector<uint8_t> used(nss);
    SCK(hipMemcpy(used.data(), a.used, (size_t)nss, hipMemcpyDeviceToHost));
    for (int s = 0; s < nss; s++)
      if (used[s]) {
        pl.m2fIndex[s] = (int)pl.m2fSlot.size();
        pl.m2fSlot.push_back(s);
      }
  }
  release(a.used);
  a.used = nullptr;
  // cursors start at the slots' first entries (the counts are no longer needed)
  SCK(hipMemcpy(a.cnt, pl.m2lBeg.data(), (size_t)nts * sizeof(int), hipMemcpyHostToDevice));
  a.cur = a.cnt;
  int* f = nullptr;
  SCK(put(f, pl.m2fIndex));
  a.m2fIndex = f;
  SCK(alloc(a.outSrcF, (size_t)nm));
  SCK(alloc(a.outTab, (size_t)nm));
  SCK(alloc(a.outHash, (size_t)nm));
  for (const Level& L : lv_) {
    a.pa = {L.nt, L.tcell, L.mBeg, L.mSrc, Tidx_, Sidx_, L.code, L.W};
    a.hf = L.kind == 1;
    k_list_place<<<blocksFor(L.nt), 256, 0, 0>>>(a);
    SCK(hipGetLastError());
  }
  // the entries stay on the device: the grouped lists are built from them
  pl.m2lEntries = nm;
  dSrcF_ = a.outSrcF;
  dTab_ = a.outTab;
  dHash_ = a.outHash;
  return OK;
}

template <int G>
static void launchGroups(const GroupArgs& a, bool write) {
  const unsigned nb = (unsigned)((a.ngroups + 255) / 256);
  if (write) k_groups<G, true><<<nb, 256, 0, 0>>>(a);
  else k_groups<G, false><<<nb, 256, 0, 0>>>(a);
}

int GpuSetup::m2lGroups(Plan& pl, std::string& msg) {
  const int G = pl.m2lG;
  const int64_t ng = (int64_t)pl.m2lGroupTarget.size() / G;
  pl.m2lGroupBeg.assign(ng + 1, 0);
  pl.m2lGSrcF.clear();
  pl.m2lGMask.clear();
  pl.m2lGCode.clear();
  if (ng == 0) return OK;
  if (G != 1 && G != 3 && G != 4) {
    msg = "device set-up: unsupported M2L group size";
    return kSetupFallback;
  }
  if (pl.m2lEntries > 0 && !dSrcF_) {
    msg = "device set-up: M2L lists not on the device";
    return kSetupFallback;
  }
  GroupArgs a = {};
  a.ngroups = (int)ng;
  int *tgt = nullptr, *beg = nullptr, *cnt = nullptr, *bad = nullptr;
  SCK(put(tgt, pl.m2lGroupTarget));
  SCK(put(beg, pl.m2lBeg));
  SCK(alloc(cnt, (size_t)ng + 1));
  SCK(alloc(bad, 1));
  SCK(hipMemset(bad, 0, sizeof(int)));
  a.target = tgt;
  a.beg = beg;
  a.srcF = dSrcF_;
  a.tab = dTab_;
  a.hash = dHash_;
  a.count = cnt;
  a.bad = bad;
  auto launch = [&](bool write) {
    if (G == 1) launchGroups<1>(a, write);
    else if (G == 3) launchGroups<3>(a, write);
    else launchGroups<4>(a, write);
    return hipGetLastError();
  };
  SCK(launch(false));
  std::vector<int> c(ng);
  SCK(hipMemcpy(c.data(), cnt, (size_t)ng * sizeof(int), hipMemcpyDeviceToHost));
  int flag = 0;
  SCK(hipMemcpy(&flag, bad, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) {
    msg = "device set-up: a target's M2L sources do not strictly ascend";
    return kSetupFallback;
  }
  int64_t run = 0;
  for (int64_t g = 0; g < ng; g++) {
    run += c[g];
    if (run >= ((int64_t)1 << 31)) {
      msg = "M2L lists exceed 2^31 entries";
      return ERR_MEMORY;
    }
    pl.m2lGroupBeg[g + 1] = (int)run;
  }
  SCK(hipMemcpy(cnt, pl.m2lGroupBeg.data(), ((size_t)ng + 1) * sizeof(int),
                hipMemcpyHostToDevice));
  a.gBeg = cnt;
  SCK(alloc(a.outSrc, (size_t)run));
  SCK(alloc(a.outMask, (size_t)run));
  SCK(alloc(a.outCode, (size_t)run * G));
  SCK(launch(true));
  pl.m2lGSrcF.resize(run);
  pl.m2lGMask.resize(run);
  pl.m2lGCode.resize((size_t)run * G);
  SCK(copyLarge(pl.m2lGSrcF.data(), a.outSrc, (size_t)run * sizeof(int), false));
  SCK(copyLarge(pl.m2lGMask.data(), a.outMask, (size_t)run * sizeof(int), false));
  SCK(copyLarge(pl.m2lGCode.data(), a.outCode, (size_t)run * G * sizeof(int), false));
  return OK;
}

int GpuSetup::p2pLists(Plan& pl, std::string& msg) {
  const Tree &T = pl.T, &S = pl.S;
  pl.p2pLeaf.clear();
  pl.p2pBeg.assign(1, 0);
  pl.p2pSrcB.clear();
  pl.p2pSrcE.clear();
  std::vector<int> leaves, parent(T.ncells(), -1);
  for (int c = 0; c < T.ncells(); c++) {
    if (T.leaf(c)) leaves.push_back(c);
    for (int k = 0; k < T.nson[c]; k++) parent[T.sons[T.son0[c] + k]] = c;
  }
  const int nl = (int)leaves.size();
  if ((int)pLv_.size() != pl.nlvl || (int)pSeg_.size() != T.ncells()) {
    msg = "device set-up: P2P pairs not on the device";
    return kSetupFallback;
  }
  P2PArgs a = {};
  a.nleaf = nl;
  int *dl = nullptr, *dp = nullptr, *dlv = nullptr, *dsb = nullptr, *dsn = nullptr, *dtn = nullptr;
  int2* dseg = nullptr;
  int** dsrc = nullptr;
  int* cnt = nullptr;
  unsigned long long* pairs = nullptr;
  SCK(put(dl, leaves));
  SCK(put(dp, parent));
  SCK(put(dlv, T.lvl));
  SCK(put(dseg, pSeg_));
  SCK(put(dsrc, pLv_));
  SCK(put(dsb, S.pbeg));
  SCK(put(dsn, S.nprt));
  SCK(put(dtn, T.nprt));
  SCK(alloc(cnt, (size_t)nl + 1));
  SCK(alloc(pairs, 1));
  SCK(hipMemset(pairs, 0, sizeof(unsigned long long)));
  a.leaf = dl;
  a.parent = dp;
  a.lvl = dlv;
  a.seg = dseg;
  a.src = dsrc;
  a.Spbeg = dsb;
  a.Snprt = dsn;
  a.Tnprt = dtn;
  a.count = cnt;
  a.pairs = pairs;
  const unsigned nb = (unsigned)((nl + 255) / 256);
  k_p2p_leaf<false><<<nb, 256, 0, 0>>>(a);
  SCK(hipGetLastError());
  std::vector<int> c(nl);
  SCK(hipMemcpy(c.data(), cnt, (size_t)nl * sizeof(int), hipMemcpyDeviceToHost));
  unsigned long long np = 0;
  SCK(hipMemcpy(&np, pairs, sizeof(np), hipMemcpyDeviceToHost));
  pl.st.p2p_pairs = (int64_t)np;
  std::vector<int> beg((size_t)nl + 1, 0);
  int64_t run = 0;
  for (int k = 0; k < nl; k++) {
    beg[k] = (int)run;
    run += c[k];
    if (run >= ((int64_t)1 << 31)) {
      msg = "P2P lists exceed 2^31 ranges";
      return ERR_MEMORY;
    }
    if (c[k]) {
      pl.p2pLeaf.push_back(leaves[k]);
      pl.p2pBeg.push_back((int)run);
    }
  }
  beg[nl] = (int)run;
  SCK(hipMemcpy(cnt, beg.data(), ((size_t)nl + 1) * sizeof(int), hipMemcpyHostToDevice));
  a.beg = cnt;
  SCK(alloc(a.outB, (size_t)run));
  SCK(alloc(a.outE, (size_t)run));
  k_p2p_leaf<true><<<nb, 256, 0, 0>>>(a);
  SCK(hipGetLastError());
  pl.p2pSrcB.resize(run);
  pl.p2pSrcE.resize(run);
  SCK(copyLarge(pl.p2pSrcB.data(), a.outB, (size_t)run * sizeof(int), false));
  SCK(copyLarge(pl.p2pSrcE.data(), a.outE, (size_t)run * sizeof(int), false));
  return OK;
}

#undef SCK

}  // namespace

SetupDevice* newSetupDevice() { return new GpuSetup(); }

namespace {
// memcpy split over host threads
void parallelCopy(char* dst, const char* src, size_t n) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t per = std::max<size_t>((size_t)4 << 20, (n + hw - 1) / hw);
  std::vector<std::thread> th;
  for (size_t o = per; o < n; o += per)
    th.emplace_back([=] { std::memcpy(dst + o, src + o, std::min(per, n - o)); });
  std::memcpy(dst, src, std::min(per, n));
  for (auto& t : th) t.join();
}
}  // namespace

hipError_t copyLarge(void* dst, const void* src, size_t bytes, bool toDevice) {
  const hipMemcpyKind kind = toDevice ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  const size_t chunk = (size_t)64 << 20;
  if (bytes < chunk) return hipMemcpy(dst, src, bytes, kind);
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return e;
  void* pin[2] = {nullptr, nullptr};
  hipStream_t s = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool staged = hipHostMalloc(&pin[0], chunk, hipHostMallocDefault) == hipSuccess &&
                hipHostMalloc(&pin[1], chunk, hipHostMallocDefault) == hipSuccess &&
                hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
                hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) == hipSuccess &&
                hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) == hipSuccess;
  if (!staged) {
    (void)hipGetLastError();
    e = hipMemcpy(dst, src, bytes, kind);
  } else {
    const size_t nch = (bytes + chunk - 1) / chunk;
    auto len = [&](size_t k) { return std::min(chunk, bytes - k * chunk); };
    char* d = (char*)dst;
    const char* h = (const char*)src;
    if (toDevice) {
      // chunk k: host -> pin[k % 2] (after chunk k - 2's device copy), pin -> device
      for (size_t k = 0; k < nch && e == hipSuccess; k++) {
        const int b = (int)(k & 1);
        if (k >= 2) e = hipEventSynchronize(ev[b]);
        if (e != hipSuccess) break;
        parallelCopy((char*)pin[b], h + k * chunk, len(k));
        e = hipMemcpyAsync(d + k * chunk, pin[b], len(k), kind, s);
        if (e == hipSuccess) e = hipEventRecord(ev[b], s);
      }
    } else {
      // chunk k: device -> pin[k % 2]; chunk k + 1 is in flight while chunk k
      // moves from pin to host
      e = hipMemcpyAsync(pin[0], h, len(0), kind, s);
      if (e == hipSuccess) e = hipEventRecord(ev[0], s);
      for (size_t k = 0; k < nch && e == hipSuccess; k++) {
        const int b = (int)(k & 1);
        e = hipEventSynchronize(ev[b]);
        if (e == hipSuccess && k + 1 < nch) {
          e = hipMemcpyAsync(pin[b ^ 1], h + (k + 1) * chunk, len(k + 1), kind, s);
          if (e == hipSuccess) e = hipEventRecord(ev[b ^ 1], s);
        }
        if (e == hipSuccess) parallelCopy(d + k * chunk, (const char*)pin[b], len(k));
      }
    }
    hipError_t e2 = hipStreamSynchronize(s);
    if (e == hipSuccess) e = e2;
  }
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  if (s) (void)hipStreamDestroy(s);
  for (void* p : pin)
    if (p) (void)hipHostFree(p);
  return e;
}

}  // namespace dfm
