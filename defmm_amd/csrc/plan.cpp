// Host plan builder.  Restates the reference's setup (prcmpt,
// include/interface.hpp:100-151) as list construction: instead of marking
// cells and later re-walking the trees for every evaluation, one dual walk
// records every M2L and P2P pair, and the per-stage lists are flattened into
// CSR arrays that the device executes level by level.
//
// Compile without floating-point contraction (-ffp-contract=off): the octree
// split tests, the admissibility test and the best-cone search are discrete
// decisions taken on rounded doubles.
#include "plan.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>
#include <tuple>
#include <unordered_map>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace dfm {

static const cplx I1(0., 1.);
static const double PI4 = 4. * M_PI;

static inline double nrm3(const double v[3]) {
  double t = 0.;
  for (int k = 0; k < 3; k++) t += v[k] * v[k];
  return std::sqrt(t);
}

// Equispaced Lagrange basis on [-alpha, alpha] (thetis_polynomials.hpp:81-94).
static double lagrange(int i, double x, int n, double alpha) {
  double res = 1., nx = x / alpha;
  double pi = -1. + 2. * ((double)i / (double)(n - 1));
  for (int j = 0; j < n; j++) {
    if (j == i) continue;
    double pj = -1. + 2. * ((double)j / (double)(n - 1));
    res *= (nx - pj) / (pi - pj);
  }
  return res;
}
static inline double gridPoint(int i, int n, double alpha) {
  return alpha * (-1. + 2. * ((double)i / (double)(n - 1)));
}

int Tree::slotOf(int c, int d) const {
  if (nslot[c] == 0) return -1;
  if (slotDir[slot0[c]] < 0) return slot0[c];
  if (!slotIdx.empty()) return (d >= 0 && mark0[c] >= 0) ? slotIdx[mark0[c] + d] : -1;
  // directions of a cell's slots are ascending
  const int* b = slotDir.data() + slot0[c];
  const int* e = b + nslot[c];
  const int* it = std::lower_bound(b, e, d);
  if (it == e || *it != d) return -1;
  return slot0[c] + (int)(it - b);
}

// ------------------------------------------------------------- octree ----
// Stable octant partition with the split test x_k > c_k, children in octant
// order 0..7 (bit k = axis k), a cell is a leaf when it holds fewer than
// Ncrit points (sort.hpp:49,98-162,187-220) or, with a fixed depth, when it
// reaches that level (sort.hpp:51).  Cells are numbered in pre-order.
namespace {
struct Builder {
  Tree& t;
  int ncrit, lvmax;
  std::vector<int> tmp;
  std::vector<double> px, py, pz;
  int err = OK;
  Builder(Tree& tr, int nc, int lm) : t(tr), ncrit(nc), lvmax(lm) {}

  void cell(int b, int n, const double ctr[3], double rad, const int idx[3], int lvl) {
    if (err) return;
    if (lvl > 30) { err = ERR_DEPTH; return; }
    int me = (int)t.lvl.size();
    t.lvl.push_back(lvl);
    t.nprt.push_back(n);
    t.pbeg.push_back(b);
    t.son0.push_back(0);
    t.nson.push_back(0);
    for (int k = 0; k < 3; k++) { t.idx.push_back(idx[k]); t.ctr.push_back(ctr[k]); }
    t.rad.push_back(rad);
    if ((int)t.nPerLevel.size() <= lvl) t.nPerLevel.push_back(0);
    t.nPerLevel[lvl]++;
    bool stop = lvmax < 0 ? (n < ncrit) : (lvl >= lvmax);
    if (stop) return;
    int histo[8] = {0}, off[8], add[8] = {0};
    std::vector<unsigned char> code(n);
    for (int i = 0; i < n; i++) {
      int p = t.perm[b + i];
      int bc = (px[p] > ctr[0] ? 1 : 0) | (py[p] > ctr[1] ? 2 : 0) | (pz[p] > ctr[2] ? 4 : 0);
      code[i] = (unsigned char)bc;
      histo[bc]++;
    }
    off[0] = 0;
    for (int i = 1; i < 8; i++) off[i] = off[i - 1] + histo[i - 1];
    for (int i = 0; i < n; i++) tmp[b + off[code[i]] + (add[code[i]]++)] = t.perm[b + i];
    std::copy(tmp.begin() + b, tmp.begin() + b + n, t.perm.begin() + b);
    double nr = rad / 2.;
    std::vector<int> kids;
    for (int i = 0; i < 8; i++) {
      if (!histo[i]) continue;
      double nc[3];
      int ni[3];
      for (int k = 0; k < 3; k++) {
        nc[k] = ctr[k] + nr * (((i >> k) & 1) * 2 - 1);
        ni[k] = 2 * idx[k] + ((i >> k) & 1);
      }
      kids.push_back((int)t.lvl.size());
      cell(b + off[i], histo[i], nc, nr, ni, lvl + 1);
      if (err) return;
    }
    // sons are stored after recursion: pre-order ids are known only then
    t.son0[me] = (int)t.sons.size();
    t.nson[me] = (int)kids.size();
    for (int k : kids) t.sons.push_back(k);
  }
};
}  // namespace

static int buildTree(Tree& t, const double* xyz, int64_t n, const double ctr[3], double rad,
                     int ncrit, int lvmax) {
  Builder bd(t, ncrit, lvmax);
  bd.px.resize(n); bd.py.resize(n); bd.pz.resize(n);
  for (int64_t i = 0; i < n; i++) {
    bd.px[i] = xyz[3 * i]; bd.py[i] = xyz[3 * i + 1]; bd.pz[i] = xyz[3 * i + 2];
  }
  t.perm.resize(n);
  std::iota(t.perm.begin(), t.perm.end(), 0);
  bd.tmp.resize(n);
  int idx0[3] = {0, 0, 0};
  bd.cell(0, (int)n, ctr, rad, idx0, 0);
  if (bd.err) return bd.err;
  t.x.resize(n); t.y.resize(n); t.z.resize(n);
  for (int64_t i = 0; i < n; i++) {
    int p = t.perm[i];
    t.x[i] = bd.px[p]; t.y[i] = bd.py[p]; t.z[i] = bd.pz[p];
  }
  return OK;
}

// Level-synchronous octree (the device builder's): per level the cells in
// parent order and, within a parent, octant order; a splitting cell's points
// are partitioned stably by octant, so every cell's points are a contiguous
// range of the final point order, as in the recursive build above.
bool cellStops(int64_t n, int lvl, int ncrit, int lvmax) {
  return lvmax < 0 ? (n < ncrit) : (lvl >= lvmax);
}

// Children of a splitting cell from its octant counts, with the recursive
// build's centre arithmetic (sort.hpp:98-162).
void splitCell(LevelCell& c, const int64_t cnt[8], std::vector<LevelCell>& next) {
  const double nr = c.rad / 2.;
  int64_t off = c.b;
  c.kid0 = (int)next.size();
  c.nkid = 0;
  for (int i = 0; i < 8; i++) {
    if (!cnt[i]) continue;
    LevelCell k;
    k.b = off;
    k.n = cnt[i];
    for (int q = 0; q < 3; q++) {
      k.ctr[q] = c.ctr[q] + nr * (((i >> q) & 1) * 2 - 1);
      k.idx[q] = 2 * c.idx[q] + ((i >> q) & 1);
    }
    k.rad = nr;
    next.push_back(k);
    c.nkid++;
    off += cnt[i];
  }
}

// The recursive build's layout from a level-synchronous one: cells numbered
// in pre-order, a cell's son list appended after its children's (as the
// recursion appends it), points in the final order.
void treeFromLevels(Tree& t, const std::vector<std::vector<LevelCell>>& lv, std::vector<int>& perm,
                    const double* xyz) {
  const int nl = (int)lv.size();
  std::vector<std::vector<int>> size(nl), pre(nl);
  for (int l = nl - 1; l >= 0; l--) {
    size[l].assign(lv[l].size(), 1);
    for (size_t c = 0; c < lv[l].size(); c++)
      for (int k = 0; k < lv[l][c].nkid; k++) size[l][c] += size[l + 1][lv[l][c].kid0 + k];
  }
  for (int l = 0; l < nl; l++) pre[l].assign(lv[l].size(), 0);
  for (int l = 0; l + 1 < nl; l++)
    for (size_t c = 0; c < lv[l].size(); c++) {
      int p = pre[l][c] + 1;
      for (int k = 0; k < lv[l][c].nkid; k++) {
        pre[l + 1][lv[l][c].kid0 + k] = p;
        p += size[l + 1][lv[l][c].kid0 + k];
      }
    }
  const int nc = size[0][0];
  t.lvl.assign(nc, 0); t.nprt.assign(nc, 0); t.pbeg.assign(nc, 0);
  t.son0.assign(nc, 0); t.nson.assign(nc, 0);
  t.idx.assign(3 * (size_t)nc, 0); t.ctr.assign(3 * (size_t)nc, 0.); t.rad.assign(nc, 0.);
  t.nPerLevel.assign(nl, 0);
  for (int l = 0; l < nl; l++)
    for (size_t c = 0; c < lv[l].size(); c++) {
      const LevelCell& x = lv[l][c];
      const int id = pre[l][c];
      t.lvl[id] = l;
      t.nprt[id] = (int)x.n;
      t.pbeg[id] = (int)x.b;
      t.nson[id] = x.nkid;
      for (int q = 0; q < 3; q++) { t.idx[3 * id + q] = x.idx[q]; t.ctr[3 * id + q] = x.ctr[q]; }
      t.rad[id] = x.rad;
      t.nPerLevel[l]++;
    }
  t.sons.clear();
  std::function<void(int, int)> post = [&](int l, int c) {
    const LevelCell& x = lv[l][c];
    for (int k = 0; k < x.nkid; k++) post(l + 1, x.kid0 + k);
    if (x.nkid) {
      t.son0[pre[l][c]] = (int)t.sons.size();
      for (int k = 0; k < x.nkid; k++) t.sons.push_back(pre[l + 1][x.kid0 + k]);
    }
  };
  post(0, 0);
  const int64_t n = (int64_t)perm.size();
  t.x.resize(n); t.y.resize(n); t.z.resize(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    const int64_t p = perm[i];
    t.x[i] = xyz[3 * p]; t.y[i] = xyz[3 * p + 1]; t.z[i] = xyz[3 * p + 2];
  }
  t.perm.swap(perm);
}

// --------------------------------------------------------- directions ----
// newdirections.hpp:75-171: stage 0 = +x,-x,+y,-y,+z,-z; stage e >= 1 =
// [x-face(+), x-face(-), z-face(+), z-face(-), y-face(+), y-face(-)], each
// face refined from its parent cone by (0, +-r, +-r), r = 2^-(e+1).
static void makeDirections(std::vector<std::vector<std::array<double, 3>>>& D, int n) {
  D.assign(n, {});
  for (int e = 0; e < n; e++) D[e].resize((size_t)6 << (2 * e));
  for (int k = 0; k < 3; k++) {
    D[0][2 * k] = {0., 0., 0.};
    D[0][2 * k][k] = 1.;
    D[0][2 * k + 1] = {0., 0., 0.};
    D[0][2 * k + 1][k] = -1.;
  }
  for (int e = 1; e < n; e++) {
    double r = 1. / double(1 << (e + 1));
    int nprev = 1 << (2 * (e - 1));
    for (int i = 0; i < nprev; i++)
      for (int q = 0; q < 4; q++) {
        std::array<double, 3> d = D[e - 1][i];
        d[1] += ((q & 1) ? -1. : 1.) * r;
        d[2] += ((q & 2) ? -1. : 1.) * r;
        double nn = nrm3(d.data());
        for (int k = 0; k < 3; k++) d[k] /= nn;
        D[e][i * 4 + q] = d;
      }
    int face = 1 << (2 * e);
    for (int i = 0; i < face; i++) {
      D[e][face + i] = D[e][i];
      D[e][face + i][0] = -D[e][i][0];
    }
    for (int j = 1; j < 3; j++)
      for (int i = 0; i < 2 * face; i++)
        for (int k = 0; k < 3; k++) D[e][j * 2 * face + i][k] = D[e][i][(k + j) % 3];
  }
}

// First direction at minimal distance; mirror-image cones of a symmetric
// offset are equidistant, so distances within 1e-12 of the best count as ties.
static int bestCone(const std::vector<std::array<double, 3>>& D, const double dir[3]) {
  double best = 1.e30;
  int ind = -1;
  for (size_t i = 0; i < D.size(); i++) {
    double d[3] = {D[i][0] - dir[0], D[i][1] - dir[1], D[i][2] - dir[2]};
    double nn = nrm3(d);
    if (nn < best - 1.e-12) { best = nn; ind = (int)i; }
  }
  return ind;
}

// --------------------------------------------------------- symmetries ----
// hypercube.hpp:108-136 and permutations.hpp:87-150,208-232.
static const int kSigma[8][3] = {{0, 1, 2}, {1, 0, 2}, {0, 1, 2}, {1, 2, 0},
                                 {0, 2, 1}, {0, 1, 2}, {2, 0, 1}, {2, 1, 0}};
static inline void fundamental(const int X[3], int R[3], int& hsh) {
  int a0 = std::abs(X[0]), a1 = std::abs(X[1]), a2 = std::abs(X[2]);
  int h1 = (a0 < a1 ? 1 : 0) | (a0 < a2 ? 2 : 0) | (a1 < a2 ? 4 : 0);
  int h0 = (X[0] < 0 ? 1 : 0) | (X[1] < 0 ? 2 : 0) | (X[2] < 0 ? 4 : 0);
  hsh = h0 + (h1 << 3);
  for (int k = 0; k < 3; k++) R[k] = std::abs(X[kSigma[h1][k]]);
}

// ---------------------------------------------------------------- DFT ----
void dft3(int P, const cplx* in, cplx* out, int sign) {
  int PP = P * P * P;
  std::vector<cplx> tw(P * P), a(in, in + PP), b(PP);
  for (int j = 0; j < P; j++)
    for (int k = 0; k < P; k++) {
      double ang = 2. * M_PI * (double)((j * k) % P) / (double)P;
      tw[j * P + k] = cplx(std::cos(ang), sign * std::sin(ang));
    }
  int strides[3] = {P * P, P, 1};
  for (int ax = 2; ax >= 0; ax--) {
    int s = strides[ax];
    for (int i = 0; i < PP; i++) {
      int k = (i / s) % P;
      int base = i - k * s;
      cplx acc = 0.;
      for (int j = 0; j < P; j++) acc += a[base + j * s] * tw[j * P + k];
      b[i] = acc;
    }
    a.swap(b);
  }
  std::copy(a.begin(), a.end(), out);
}

// ---------------------------------------------------------- the build ----
// The offsets iZ = idx(target) - idx(source) of M2L pairs at a level are
// bounded by the parents' non-admissible range: below the first
// precomputation level by the parents' HF (or LF) admissibility reach, at
// that level by the level's width.  Returns the half-width W of a window
// |iZ_k| <= W that holds them, or -1 when the level has no M2L or the
// window would exceed 2^26 offsets.
int m2lWindow(const Plan& pl, int l) {
  if (l < 0 || l >= pl.nlvl || !pl.precomp[l]) return -1;
  int W = (1 << std::min(l, 20));
  if (l > pl.maxStage) {
    int pv = l - 1;
    double tt = 2. * (pl.boxRad / double(1 << pv));
    double reach = (pl.isHF[pv] && pv) ? std::max(tt * tt * pl.kappa, tt) / tt + 1. : 1.;
    int bp = (int)std::ceil(reach * 1.01) + 1;
    W = std::min(W, 2 * bp + 2);
  }
  size_t n = (size_t)(2 * W + 1) * (2 * W + 1) * (2 * W + 1);
  return n > ((size_t)1 << 26) ? -1 : W;
}

// Table ids, permutation hashes and cones of the M2L offsets a device walk
// met: rep[l] holds per window offset (index ((iZ0 + W) n + iZ1 + W) n +
// iZ2 + W, n = 2W + 1) the smallest pair (target << 32 | source) with that
// offset, or ~0.  The table keys (level, fundamental offset) are sorted and
// numbered as the host walk numbers them; code[l] gets, per offset met,
// table id | hash << 32 | (cone + 1) << 40 (cone -1 at LF levels), the cone
// chosen from the representative pair as the host walk chooses it from the
// first pair it meets with that offset (M2L.hpp:189-197), or taken from
// cones (per level and offset) when the caller searched them elsewhere.
int offsetCodes(const Plan& pl, const std::vector<int>& W,
                const std::vector<std::vector<uint64_t>>& rep,
                std::vector<std::vector<uint64_t>>& code,
                std::vector<std::array<int, 4>>& tabKeys, std::string& msg,
                const std::vector<std::vector<int>>* cones) {
  tabKeys.clear();
  code.assign(pl.nlvl, {});
  auto offsetOf = [](int64_t k, int Wl, int iZ[3]) {
    const int64_t n = 2 * Wl + 1;
    iZ[2] = (int)(k % n) - Wl;
    iZ[1] = (int)((k / n) % n) - Wl;
    iZ[0] = (int)(k / (n * n)) - Wl;
  };
  for (int l = 0; l < pl.nlvl; l++) {
    if (W[l] < 0 || rep[l].empty()) continue;
    for (int64_t k = 0; k < (int64_t)rep[l].size(); k++) {
      if (rep[l][k] == ~(uint64_t)0) continue;
      int iZ[3], nI[3], h;
      offsetOf(k, W[l], iZ);
      fundamental(iZ, nI, h);
      if (!(l && pl.isHF[l]))
        for (int q = 0; q < 3; q++)
          if (nI[q] >= pl.nCellPerDirMax[l]) {
            msg = "M2L offset outside the table range";
            return ERR_TABLE;
          }
      tabKeys.push_back({l, nI[0], nI[1], nI[2]});
    }
  }
  std::sort(tabKeys.begin(), tabKeys.end());
  tabKeys.erase(std::unique(tabKeys.begin(), tabKeys.end()), tabKeys.end());
  for (int l = 0; l < pl.nlvl; l++) {
    if (W[l] < 0 || rep[l].empty()) continue;
    const bool hf = l && pl.isHF[l];
    code[l].assign(rep[l].size(), 0);
    // a cone search over all directions of the level per offset met
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < (int64_t)rep[l].size(); k++) {
      const uint64_t r = rep[l][k];
      if (r == ~(uint64_t)0) continue;
      int iZ[3], nI[3], h;
      offsetOf(k, W[l], iZ);
      fundamental(iZ, nI, h);
      const std::array<int, 4> key = {l, nI[0], nI[1], nI[2]};
      const int id = (int)(std::lower_bound(tabKeys.begin(), tabKeys.end(), key) - tabKeys.begin());
      int cn = -1;
      if (hf && cones) {
        cn = (*cones)[l][k];
      } else if (hf) {
        const int i = (int)(r >> 32), j = (int)(r & 0xFFFFFFFFu);
        double fZ[3];
        for (int q = 0; q < 3; q++) fZ[q] = pl.T.ctr[3 * i + q] - pl.S.ctr[3 * j + q];
        double nn = nrm3(fZ);
        for (int q = 0; q < 3; q++) fZ[q] /= nn;
        cn = bestCone(pl.dirs[pl.HFStage[l]], fZ);
      }
      code[l][k] = (uint64_t)(uint32_t)id | ((uint64_t)h << 32) | ((uint64_t)(cn + 1) << 40);
    }
  }
  return OK;
}

namespace {
struct PairLists {
  // M2L pairs in walk order: target cell, source cell, table id, hash, cone
  std::vector<int> mt, ms, mtab, mh, mcone;
  // P2P pairs in walk order
  std::vector<int> pt, ps;
};

// A walker owns its table-id map and caches; several walkers run on
// disjoint cell pairs in parallel (their table ids are local and are renumbered
// afterwards).  Marks are shared and set with atomic stores.
struct Walker {
  Plan& pl;
  PairLists* out;
  int64_t nHF = 0, nLF = 0, nP2P = 0, pairsP2P = 0;
  // task mode: pairs at level taskLevel are not walked but recorded as tasks
  int taskLevel = -1;
  std::vector<std::array<int, 2>>* tasks = nullptr;
  // table ids keyed by (level, fundamental offset)
  std::unordered_map<uint64_t, int> tabId;
  std::vector<std::array<int, 4>> tabKey;  // level, a, b, c
  std::unordered_map<uint64_t, int> cone;  // (table id, hash) -> cone
  // per-level dense cache of offsets met so far: packed table id, hash, cone
  std::vector<std::vector<uint64_t>> win;
  std::vector<int> winW;
  int err = OK;
  Walker(Plan& p, PairLists* o) : pl(p), out(o) {
    win.resize(pl.nlvl);
    winW.assign(pl.nlvl, -1);
    for (int l = 0; l < pl.nlvl; l++) {
      winW[l] = m2lWindow(pl, l);
      if (winW[l] >= 0) win[l].assign((size_t)(2 * winW[l] + 1) * (2 * winW[l] + 1) * (2 * winW[l] + 1), 0);
    }
  }

  static void mark(Tree& t, int c, int cn) {
    __atomic_store_n(&t.marks[t.mark0[c] + cn], (uint8_t)1, __ATOMIC_RELAXED);
  }
  static uint64_t key4(int l, int a, int b, int c) {
    return ((uint64_t)l << 48) | ((uint64_t)(uint32_t)a << 32) | ((uint64_t)(uint32_t)b << 16) |
           (uint64_t)(uint32_t)c;
  }
  int table(int l, const int nI[3]) {
    uint64_t k = key4(l, nI[0], nI[1], nI[2]);
    auto it = tabId.find(k);
    if (it != tabId.end()) return it->second;
    int id = (int)tabKey.size();
    tabId[k] = id;
    tabKey.push_back({l, nI[0], nI[1], nI[2]});
    return id;
  }
  // hrzpass.hpp:51-56
  bool macHF(int i, int j) const {
    const Tree &T = pl.T, &S = pl.S;
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = T.ctr[3 * i + k] - S.ctr[3 * j + k];
    double tmp = (1. - 1.e-8) * (nrm3(d) - 2. * S.rad[j]);
    double tt = 2. * S.rad[j];
    return tmp >= std::max(tt * tt * pl.kappa, tt);
  }
  // hrzpass.hpp:39-49
  bool macLF(int i, int j) const {
    int test = 0, ncmp = 3;
    for (int k = 0; k < 3; k++) {
      int d = pl.T.idx[3 * i + k] - pl.S.idx[3 * j + k];
      ncmp -= (d ? 0 : 1);
      test += std::abs(d);
    }
    return test > ncmp;
  }
  void m2l(int i, int j, bool hf) {
    int lvl = pl.T.lvl[i];
    int iZ[3], nI[3], h;
    for (int k = 0; k < 3; k++) iZ[k] = pl.T.idx[3 * i + k] - pl.S.idx[3 * j + k];
    uint64_t* slotw = nullptr;
    int W = winW[lvl];
    if (W >= 0 && std::abs(iZ[0]) <= W && std::abs(iZ[1]) <= W && std::abs(iZ[2]) <= W) {
      size_t n1 = 2 * W + 1;
      slotw = &win[lvl][((size_t)(iZ[0] + W) * n1 + (size_t)(iZ[1] + W)) * n1 + (size_t)(iZ[2] + W)];
      if (*slotw) {
        uint64_t v = *slotw;
        int id = (int)(v & 0xFFFFFFFFu), hh = (int)((v >> 32) & 63), cn = (int)(v >> 40) - 1;
        if (hf) {
          mark(pl.T, i, cn);
          mark(pl.S, j, cn);
          nHF++;
        } else {
          nLF++;
        }
        out->mt.push_back(i); out->ms.push_back(j); out->mtab.push_back(id);
        out->mh.push_back(hh); out->mcone.push_back(cn);
        return;
      }
    }
    fundamental(iZ, nI, h);
    if (!hf)
      for (int k = 0; k < 3; k++)
        if (nI[k] >= pl.nCellPerDirMax[lvl]) err = ERR_TABLE;
    int id = table(lvl, nI);
    int cn = -1;
    if (hf) {
      // M2L.hpp:189-197: cone of the first pair met with this offset
      uint64_t ck = ((uint64_t)id << 8) | (uint64_t)h;
      auto it = cone.find(ck);
      if (it == cone.end()) {
        double fZ[3];
        for (int k = 0; k < 3; k++) fZ[k] = pl.T.ctr[3 * i + k] - pl.S.ctr[3 * j + k];
        double nn = nrm3(fZ);
        for (int k = 0; k < 3; k++) fZ[k] /= nn;
        cn = bestCone(pl.dirs[pl.HFStage[lvl]], fZ);
        cone[ck] = cn;
      } else cn = it->second;
      mark(pl.T, i, cn);
      mark(pl.S, j, cn);
      nHF++;
    } else {
      nLF++;
    }
    out->mt.push_back(i); out->ms.push_back(j); out->mtab.push_back(id);
    out->mh.push_back(h); out->mcone.push_back(cn);
    if (slotw) *slotw = (uint64_t)(uint32_t)id | ((uint64_t)h << 32) | ((uint64_t)(cn + 1) << 40);
  }
  // hrzpass.hpp:61-91 (and the setup walk hrzpass.hpp:111-138, which visits
  // the HF pairs in the same order)
  void walk(int i, int j) {
    if (err) return;
    const Tree &T = pl.T, &S = pl.S;
    int l = T.lvl[i];
    if (tasks && l == taskLevel) {
      tasks->push_back({i, j});
      return;
    }
    if (pl.precomp[l] && l == S.lvl[j]) {
      if (l && pl.isHF[l]) {
        if (macHF(i, j)) { m2l(i, j, true); return; }
      } else if (macLF(i, j)) {
        m2l(i, j, false);
        return;
      }
    }
    if (T.leaf(i) || S.leaf(j)) {
      out->pt.push_back(i);
      out->ps.push_back(j);
      nP2P++;
      pairsP2P += (int64_t)T.nprt[i] * S.nprt[j];
      return;
    }
    for (int a = 0; a < T.nson[i]; a++)
      for (int b = 0; b < S.nson[j]; b++) walk(T.sons[T.son0[i] + a], S.sons[S.son0[j] + b]);
  }
};
}  // namespace

// dwnpass.hpp:80-104: a child needs direction d when its parent has any of
// the refined directions 4d..4d+3.  Level by level from the top, the cells of
// a level in parallel (a child's marks are written by its parent only).
static void propagate(Plan& pl, Tree& t) {
  const int n = t.ncells();
  std::vector<std::vector<int>> byLevel(pl.nlvl);
  for (int c = 0; c < n; c++) byLevel[t.lvl[c]].push_back(c);
  for (int l = 0; l + 1 < pl.nlvl; l++) {
    if (!(pl.isHF[l] && pl.precomp[l]) || !pl.isHF[l + 1]) continue;
    const std::vector<int>& cs = byLevel[l];
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < (int64_t)cs.size(); k++) {
      const int c = cs[k];
      for (int a = 0; a < t.nson[c]; a++) {
        int s = t.sons[t.son0[c] + a];
        for (int dt = 0; dt < pl.nDir[l + 1]; dt++)
          for (int q = 0; q < 4; q++)
            if (t.marks[t.mark0[c] + 4 * dt + q]) t.marks[t.mark0[s] + dt] = 1;
      }
    }
  }
}

// initdevs.hpp:187-245: HF cells get one slot per marked direction, in
// ascending direction order; LF cells at precomputation levels one slot.
// Slots are numbered in cell order: counts per cell, a prefix sum, then the
// cells filled in parallel.
static void numberSlots(Plan& pl, Tree& t) {
  const int n = t.ncells();
  t.slot0.assign(n, 0);
  t.nslot.assign(n, 0);
#pragma omp parallel for schedule(dynamic, 256)
  for (int c = 0; c < n; c++) {
    const int l = t.lvl[c];
    if (!pl.precomp[l]) continue;
    if (pl.isHF[l]) {
      int k = 0;
      for (int d = 0; d < pl.nDir[l]; d++) k += t.marks[t.mark0[c] + d] ? 1 : 0;
      t.nslot[c] = k;
    } else {
      t.nslot[c] = 1;
    }
  }
  int64_t run = 0;
  for (int c = 0; c < n; c++) {
    t.slot0[c] = (int)run;
    run += t.nslot[c];
  }
  t.nslots = (int)run;
  t.slotCell.assign(run, 0);
  t.slotDir.assign(run, -1);
  t.slotIdx.assign(t.marks.size(), -1);
#pragma omp parallel for schedule(dynamic, 256)
  for (int c = 0; c < n; c++) {
    const int l = t.lvl[c];
    if (!pl.precomp[l]) continue;
    int s = t.slot0[c];
    if (pl.isHF[l]) {
      for (int d = 0; d < pl.nDir[l]; d++)
        if (t.marks[t.mark0[c] + d]) {
          t.slotCell[s] = c;
          t.slotDir[s] = d;
          t.slotIdx[t.mark0[c] + d] = s;
          s++;
        }
    } else {
      t.slotCell[s] = c;
    }
  }
}

int validatePlan(const Plan& pl, std::string& msg);

int buildPlan(Plan& pl, const double* src, int64_t ns, const double* trg, int64_t nt, int L,
              int ncrit, double kappa, std::string& msg, SetupDevice* dev) {
  const int lvmax = ncrit < 0 ? -ncrit : -1;  // negative Ncrit: fixed depth
  // kappa = 0 is the Laplace kernel 1/(4 pi r) (the reference's
  // thetis::kernels::Laplace, include/thetis/thetis_kernels.hpp:35): every
  // level is low-frequency and M2L starts at level 2, as in the reference's
  // initFrequency (include/tables/frequency.hpp:43-129)
  if (!src || !trg) {
    msg = "null point array";
    return ERR_ARG;
  }
  if (ns <= 0 || nt <= 0 || ns > (int64_t)2000000000 || nt > (int64_t)2000000000) {
    msg = "n_src and n_trg must be in 1 .. 2e9";
    return ERR_ARG;
  }
  if (L < 2 || L > 12 || ncrit == 0 || !(kappa >= 0) || !std::isfinite(kappa)) {
    msg = "invalid argument (order, Ncrit or kappa)";
    return ERR_ARG;
  }
  for (const auto& [xyz, n] : {std::make_pair(src, ns), std::make_pair(trg, nt)})
    for (int64_t i = 0; i < 3 * n; i++)
      if (!std::isfinite(xyz[i])) {
        msg = "non-finite point coordinate";
        return ERR_ARG;
      }
  pl = Plan();
  pl.L = L; pl.P = 2 * L - 1; pl.LL = L * L * L; pl.PP = pl.P * pl.P * pl.P;
  pl.ncrit = ncrit; pl.kappa = kappa;
  const int LL = pl.LL, PP = pl.PP, P = pl.P;

  // particles.hpp:79-100: one cubic box around both clouds
  double mn[3], mx[3];
  for (int k = 0; k < 3; k++) mn[k] = mx[k] = trg[k];
  auto upd = [&](const double* p, int64_t n) {
    for (int64_t i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) {
        double v = p[3 * i + k];
        mn[k] = v < mn[k] ? v : mn[k];
        mx[k] = v > mx[k] ? v : mx[k];
      }
  };
  upd(trg, nt);
  upd(src, ns);
  double r1 = 0., r2 = 0.;
  for (int k = 0; k < 3; k++) pl.boxCtr[k] = (mn[k] + mx[k]) / 2;
  r1 = pl.boxCtr[0] - mn[0]; r2 = mx[0] - pl.boxCtr[0];
  for (int k = 1; k < 3; k++) {
    r1 = std::max(r1, pl.boxCtr[k] - mn[k]);
    r2 = std::max(r2, mx[k] - pl.boxCtr[k]);
  }
  pl.boxRad = std::max(r1, r2);
  if (!(pl.boxRad > 0.)) { msg = "all points coincide"; return ERR_ARG; }

  auto tick = [t0 = std::chrono::steady_clock::now()]() mutable {
    auto t1 = std::chrono::steady_clock::now();
    double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    t0 = t1;
    return (int64_t)ms;
  };
  int e;
  if (dev) {
    e = dev->tree(pl.T, trg, nt, pl.boxCtr, pl.boxRad, ncrit, lvmax, msg);
    if (!e) e = dev->tree(pl.S, src, ns, pl.boxCtr, pl.boxRad, ncrit, lvmax, msg);
    if (e == kSetupFallback) return e;
  } else {
    e = buildTree(pl.T, trg, nt, pl.boxCtr, pl.boxRad, ncrit, lvmax);
    if (!e) e = buildTree(pl.S, src, ns, pl.boxCtr, pl.boxRad, ncrit, lvmax);
  }
  if (e) { msg = "octree deeper than 30 levels (too many coincident points for Ncrit)"; return e; }

  // frequency.hpp:43-129, 158-184 (eta = 1)
  int nlvl = std::max((int)pl.T.nPerLevel.size(), (int)pl.S.nPerLevel.size());
  pl.nlvl = nlvl;
  pl.isHF.assign(nlvl, 0); pl.HFStage.assign(nlvl, 0); pl.nDir.assign(nlvl, 1);
  pl.nCellPerDirMax.assign(nlvl, -1); pl.precomp.assign(nlvl, 0);
  double sd = 2. * pl.boxRad;
  int cpt = 0;
  while (sd * kappa > 1.) { sd /= 2.; cpt++; }
  pl.cpt = cpt;
  pl.noHF = !(cpt > 1);
  if (cpt) makeDirections(pl.dirs, cpt);
  for (int l = 0; l < nlvl && l <= cpt; l++) { pl.isHF[l] = cpt - l + 1; pl.HFStage[l] = cpt - l; }
  for (int l = 0; l < nlvl; l++)
    pl.nDir[l] = (pl.isHF[l] && l) ? (int)pl.dirs[pl.HFStage[l]].size() : 1;
  int ms = 0;
  while (1. < kappa * (2. * pl.boxRad) / double(1 << ms) / double(1 << ms)) {
    ms++;
    if (ms >= nlvl) { msg = "no precomputation level: the tree is too shallow for kappa"; return ERR_NO_PRECOMP; }
  }
  ms = std::max(ms, 2);
  if (ms >= nlvl) { msg = "no precomputation level: the tree has fewer than 3 levels"; return ERR_NO_PRECOMP; }
  pl.maxStage = ms;
  for (int l = 0; l < nlvl; l++) pl.precomp[l] = (l >= ms);
  for (int l = 0; l < nlvl; l++)
    if (pl.precomp[l])
      pl.nCellPerDirMax[l] = pl.isHF[l] ? (1 << l) : (pl.isHF[l - 1] ? 8 : 4);

  // effective-direction marks
  for (Tree* t : {&pl.S, &pl.T}) {
    t->mark0.assign(t->ncells(), -1);
    int64_t tot = 0;
    for (int c = 0; c < t->ncells(); c++) {
      int l = t->lvl[c];
      if (pl.precomp[l] && pl.isHF[l]) { t->mark0[c] = tot; tot += pl.nDir[l]; }
    }
    t->marks.assign(tot, 0);
  }

  pl.st.ms_tree = tick();
  // dual walk
  PairLists pr;
  std::vector<std::array<int, 4>> tabKeys;  // (level, fundamental offset) by table id
  if (dev) {
    // on the device: marks, table keys, statistics and P2P pairs here; the
    // M2L pairs stay on the device for its M2L lists
    int r = dev->walk(pl, tabKeys, pr.pt, pr.ps, msg);
    if (r != OK) return r;
    pl.st.setup_device = 1;
  } else {
    // The walk from the roots is split at a level with enough cell pairs:
    // a serial walk records the pairs reached at that level as tasks (and
    // lists the pairs decided above it); the tasks are walked in parallel,
    // each into its own lists, and all lists are joined in serial-walk
    // order, so the result does not depend on the thread count.
    int taskLevel = std::min(3, nlvl - 1);
    std::vector<std::array<int, 2>> tasks;
    std::vector<PairLists> head(1);
    Walker top(pl, &head[0]);
    top.tasks = &tasks;
    top.taskLevel = taskLevel;
    // The pairs decided above the task level involve cells above it only,
    // and M2L pairs are same-level, so every target keeps its serial-walk
    // order when the top walk's lists come first and the tasks follow in the
    // order met.
    top.walk(0, 0);
    if (top.err) { msg = "M2L offset outside the table range"; return top.err; }
    const int64_t ntask = (int64_t)tasks.size();
    std::vector<PairLists> tl(ntask);
    std::vector<int> owner(ntask, 0);
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    std::vector<std::unique_ptr<Walker>> ws;
    for (int t = 0; t < nthr; t++) ws.emplace_back(new Walker(pl, nullptr));
    int werr = OK;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : werr)
    for (int64_t k = 0; k < ntask; k++) {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      Walker& w = *ws[tid];
      w.out = &tl[k];
      w.walk(tasks[k][0], tasks[k][1]);
      owner[k] = tid;
      werr |= w.err;
    }
    if (werr) { msg = "M2L offset outside the table range"; return ERR_TABLE; }
    // table ids: sorted (level, offset) keys, independent of thread timing
    std::vector<std::array<int, 4>> all = top.tabKey;
    for (auto& w : ws) all.insert(all.end(), w->tabKey.begin(), w->tabKey.end());
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    tabKeys = all;
    auto remap = [&](const Walker& w) {
      std::vector<int> m(w.tabKey.size());
      for (size_t k = 0; k < m.size(); k++)
        m[k] = (int)(std::lower_bound(all.begin(), all.end(), w.tabKey[k]) - all.begin());
      return m;
    };
    std::vector<int> topMap = remap(top);
    std::vector<std::vector<int>> wMap;
    for (auto& w : ws) wMap.push_back(remap(*w));
    // join at prefix-summed offsets, in parallel: head lists, then tasks
    const int64_t nseg = ntask + 1;
    auto seg = [&](int64_t k) -> const PairLists& { return k == 0 ? head[0] : tl[k - 1]; };
    auto segMap = [&](int64_t k) -> const std::vector<int>& {
      return k == 0 ? topMap : wMap[owner[k - 1]];
    };
    std::vector<int64_t> mOff(nseg + 1, 0), pOff(nseg + 1, 0);
    for (int64_t k = 0; k < nseg; k++) {
      mOff[k + 1] = mOff[k] + (int64_t)seg(k).mt.size();
      pOff[k + 1] = pOff[k] + (int64_t)seg(k).pt.size();
    }
    for (auto* v : {&pr.mt, &pr.ms, &pr.mtab, &pr.mh, &pr.mcone}) v->resize(mOff[nseg]);
    for (auto* v : {&pr.pt, &pr.ps}) v->resize(pOff[nseg]);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < nseg; k++) {
      const PairLists& s = seg(k);
      const std::vector<int>& m = segMap(k);
      const int64_t o = mOff[k], q = pOff[k];
      for (size_t e = 0; e < s.mt.size(); e++) {
        pr.mt[o + e] = s.mt[e];
        pr.ms[o + e] = s.ms[e];
        pr.mtab[o + e] = m[s.mtab[e]];
        pr.mh[o + e] = s.mh[e];
        pr.mcone[o + e] = s.mcone[e];
      }
      std::copy(s.pt.begin(), s.pt.end(), pr.pt.begin() + q);
      std::copy(s.ps.begin(), s.ps.end(), pr.ps.begin() + q);
    }
    std::vector<PairLists>().swap(tl);
    int64_t nHF = top.nHF, nLF = top.nLF, nP = top.nP2P, pP = top.pairsP2P;
    for (auto& w : ws) { nHF += w->nHF; nLF += w->nLF; nP += w->nP2P; pP += w->pairsP2P; }
    pl.st.hf_m2l = nHF; pl.st.lf_m2l = nLF; pl.st.p2p_calls = nP; pl.st.p2p_pairs = pP;
  }
  if (!pl.noHF) { propagate(pl, pl.S); propagate(pl, pl.T); }
  numberSlots(pl, pl.S);
  numberSlots(pl, pl.T);
  pl.st.src_slots = pl.S.nslots;
  pl.st.trg_slots = pl.T.nslots;
  pl.st.nlvl = nlvl; pl.st.cpt = cpt; pl.st.max_stage = ms;
  pl.st.ncells_t = pl.T.ncells(); pl.st.ncells_s = pl.S.ncells();

  pl.st.ms_walk = tick();
  // ---- tables ----
  pl.embed.resize(LL);
  for (int i = 0; i < LL; i++) {
    int a = i / (L * L), b = (i / L) % L, c = i % L;
    pl.embed[i] = (a * P + b) * P + c;
  }
  pl.perm.assign(64 * PP, 0);
  for (int p0 = 0; p0 < 8; p0++)
    for (int p1 = 0; p1 < 8; p1++) {
      int h = p0 + (p1 << 3);
      for (int pp = 0; pp < PP; pp++) {
        int tp[3] = {pp / (P * P), (pp / P) % P, pp % P}, tm[3];
        for (int k = 0; k < 3; k++) {
          tm[k] = tp[kSigma[p1][k]];
          if ((p0 >> kSigma[p1][k]) & 1)
            if (tm[k]) tm[k] = P - tm[k];
        }
        pl.perm[h * PP + pp] = (tm[0] * P + tm[1]) * P + tm[2];
      }
    }
  // circulant.hpp:73-93,133-144: kernel samples on the (2L-1)^3 difference grid
  std::vector<double> tabZ(3 * PP);
  for (int i = 0; i < PP; i++) {
    int E[3] = {i / (P * P), (i / P) % P, i % P};
    for (int k = 0; k < 3; k++) {
      int ek = E[k] < L ? E[k] : -(P - E[k]);
      int ip = ek > -1 ? ek : 0, jn = ek > -1 ? 0 : -ek;
      tabZ[3 * i + k] = gridPoint(ip, L, 1.) - gridPoint(jn, L, 1.);
    }
  }
  pl.ntab = (int)tabKeys.size();
  pl.m2lTab.assign((size_t)pl.ntab * PP, 0.);
#pragma omp parallel
  {
    std::vector<cplx> g(PP);
#pragma omp for schedule(dynamic, 4)
    for (int id = 0; id < pl.ntab; id++) {
      int l = tabKeys[id][0];
      double rad = pl.boxRad / double(1 << l);
      double v[3];
      for (int k = 0; k < 3; k++) v[k] = 2. * double(tabKeys[id][1 + k]) * rad;
      for (int i = 0; i < PP; i++) {
        double u[3];
        for (int k = 0; k < 3; k++) u[k] = v[k] + tabZ[3 * i + k] * rad;
        double R = nrm3(u);
        g[i] = std::exp(I1 * R * kappa) / (PI4 * R);
      }
      dft3(P, g.data(), pl.m2lTab.data() + (size_t)id * PP, -1);
    }
  }
  for (int id = 0; id < pl.ntab; id++) {
    if (pl.isHF[tabKeys[id][0]]) pl.st.hf_tables++; else pl.st.lf_tables++;
  }
  // planewave.hpp:43-69: phase tables of every precomputation HF level
  pl.sigRow0.assign(nlvl, -1);
  int rows = 0;
  for (int l = 0; l < nlvl; l++)
    if (pl.precomp[l] && pl.isHF[l]) { pl.sigRow0[l] = rows; rows += pl.nDir[l]; }
  pl.sig.assign((size_t)rows * LL, 0.);
  pl.sig2.assign((size_t)rows * LL, 0.);
  for (int l = 0; l < nlvl; l++) {
    if (pl.sigRow0[l] < 0) continue;
    double hr = .5 * (pl.boxRad / double(1 << l));
#pragma omp parallel for schedule(static)
    for (int d = 0; d < pl.nDir[l]; d++) {
      const double* dv = pl.dirs[pl.HFStage[l]][d].data();
      size_t r = (size_t)(pl.sigRow0[l] + d) * LL;
      for (int i = 0; i < LL; i++) {
        int I[3] = {i / (L * L), (i / L) % L, i % L};
        double dp = 0.;
        for (int k = 0; k < 3; k++) dp += dv[k] * gridPoint(I[k], L, hr);
        cplx s = std::exp(-I1 * kappa * dp);
        pl.sig[r + i] = s;
        pl.sig2[r + i] = s * s;
      }
    }
  }
  // interpolation.hpp:183-235
  for (int s = 0; s < 2; s++) {
    pl.A[s].assign(L * L, 0.);
    pl.B[s].assign(L * L, 0.);
    double sh = s ? 1. : -1.;
    for (int p = 0; p < L; p++)
      for (int k = 0; k < L; k++) {
        double v = lagrange(k, gridPoint(p, L, 1.) + sh, L, 2.);
        pl.A[s][k * L + p] = v;
        pl.B[s][p * L + k] = v;
      }
  }

  pl.st.ms_tables = tick();
  // ---- P2M / L2P leaves (upwpass.hpp:38-45, dwnpass.hpp:38-45) ----
  for (int c = 0; c < pl.S.ncells(); c++)
    if (pl.S.leaf(c) && pl.precomp[pl.S.lvl[c]]) {
      pl.p2mCell.push_back(c);
      if (pl.isHF[pl.S.lvl[c]]) pl.st.hf_leaves_p2m++; else pl.st.lf_leaves_p2m++;
    }
  for (int c = 0; c < pl.T.ncells(); c++)
    if (pl.T.leaf(c) && pl.precomp[pl.T.lvl[c]]) {
      pl.l2pCell.push_back(c);
      if (pl.isHF[pl.T.lvl[c]]) pl.st.hf_l2p++; else pl.st.lf_l2p++;
    }

  // ---- M2M (upwpass.hpp:50-61; M2M.hpp:34-124) ----
  // Output = parent slot; inputs = the children's slot for direction d/4
  // (HF parent above HF children), the child's single slot (HF stage-0
  // parent above LF children, or LF parent).
  auto phase = [&](const double* d, const double v[3], double sgn) {
    double dv = 0.;
    for (int k = 0; k < 3; k++) dv += d[k] * v[k];
    return std::exp(sgn * I1 * kappa * dv);
  };
  // Lists are built per level in two passes over the level's cells (count,
  // then fill at the prefix-summed offsets, in parallel): the same order as a
  // serial walk over cells, slots and children.
  auto cellsAt = [](const Tree& t, int l) {
    std::vector<int> cs;
    for (int c = 0; c < t.ncells(); c++)
      if (t.lvl[c] == l && !t.leaf(c)) cs.push_back(c);
    return cs;
  };
  auto sizeLevel = [](TransLevel& tl, int64_t nOut, int64_t nEnt) {
    tl.out.resize(nOut);
    tl.beg.resize(nOut + 1);
    tl.beg[0] = 0;
    tl.in.resize(nEnt);
    tl.sgn.resize(nEnt);
    tl.edv.resize(nEnt);
    tl.sigIdx.resize(nEnt);
  };
  for (int l = nlvl - 2; l >= 0; l--) {
    if (!pl.precomp[l]) continue;
    TransLevel tl;
    tl.level = l;
    tl.hf = pl.isHF[l] ? 1 : 0;
    const Tree& S = pl.S;
    const std::vector<int> cs = cellsAt(S, l);
    const int64_t nc = (int64_t)cs.size();
    std::vector<int64_t> oOff(nc + 1, 0), eOff(nc + 1, 0);
    for (int64_t i = 0; i < nc; i++) {
      const int c = cs[i];
      oOff[i + 1] = oOff[i] + S.nslot[c];
      eOff[i + 1] = eOff[i] + (int64_t)S.nslot[c] * S.nson[c];
      if (pl.isHF[l]) { if (pl.HFStage[l]) pl.st.hf_m2m += S.nson[c]; else pl.st.lf2hf_m2m += S.nson[c]; }
      else pl.st.lf_m2m += S.nson[c];
    }
    sizeLevel(tl, oOff[nc], eOff[nc]);
    int missing = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : missing)
    for (int64_t i = 0; i < nc; i++) {
      const int c = cs[i];
      int64_t o = oOff[i], e = eOff[i];
      for (int k = 0; k < S.nslot[c]; k++) {
        int os = S.slot0[c] + k;
        int d = S.slotDir[os];
        for (int a = 0; a < S.nson[c]; a++) {
          int ch = S.sons[S.son0[c] + a];
          double v[3];
          for (int kk = 0; kk < 3; kk++) v[kk] = S.ctr[3 * ch + kk] - S.ctr[3 * c + kk];
          int in;
          cplx edv = 1.;
          if (tl.hf) {
            in = pl.HFStage[l] ? S.slotOf(ch, d / 4) : S.slotOf(ch, -1);
            edv = phase(pl.dirs[pl.HFStage[l]][d].data(), v, -1.);
          } else {
            in = S.slotOf(ch, -1);
          }
          if (in < 0) missing++;
          int sg = (v[0] < 0. ? 0 : 1) | (v[1] < 0. ? 0 : 2) | (v[2] < 0. ? 0 : 4);
          tl.in[e] = in; tl.sgn[e] = sg; tl.edv[e] = edv;
          tl.sigIdx[e] = tl.hf ? pl.sigRow0[l] + d : -1;
          e++;
        }
        tl.out[o] = os;
        tl.beg[o + 1] = (int)e;
        o++;
      }
    }
    if (missing) { msg = "internal: missing child slot in M2M"; return ERR_STATE; }
    if (!tl.out.empty()) pl.m2m.push_back(std::move(tl));
  }

  // ---- L2L (dwnpass.hpp:47-58; L2L.hpp:34-127) ----
  // Output = child slot; inputs = parent slots (directions 4d..4d+3 of an HF
  // parent above HF children; all parent slots of an HF stage-0 parent above
  // an LF child; the LF parent's slot).
  for (int l = 0; l < nlvl - 1; l++) {
    if (!pl.precomp[l]) continue;
    TransLevel tl;
    tl.level = l;
    tl.hf = pl.isHF[l] ? 1 : 0;
    const Tree& T = pl.T;
    const bool byQuad = tl.hf && pl.HFStage[l];  // parent slots 4cd..4cd+3 only
    const std::vector<int> cs = cellsAt(T, l);
    const int64_t nc = (int64_t)cs.size();
    std::vector<int64_t> oOff(nc + 1, 0), eOff(nc + 1, 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nc; i++) {
      const int c = cs[i];
      int64_t no = 0, ne = 0;
      for (int a = 0; a < T.nson[c]; a++) {
        int ch = T.sons[T.son0[c] + a];
        no += T.nslot[ch];
        for (int k = 0; k < T.nslot[ch]; k++) {
          if (!byQuad) { ne += T.nslot[c]; continue; }
          int cd = T.slotDir[T.slot0[ch] + k];
          for (int pd = 4 * cd; pd < 4 * cd + 4; pd++) ne += T.slotOf(c, pd) >= 0;
        }
      }
      oOff[i + 1] = no;
      eOff[i + 1] = ne;
    }
    for (int64_t i = 0; i < nc; i++) {
      oOff[i + 1] += oOff[i];
      eOff[i + 1] += eOff[i];
      const int c = cs[i];
      if (pl.isHF[l]) { if (pl.HFStage[l]) pl.st.hf_l2l += T.nson[c]; else pl.st.hf2lf_l2l += T.nson[c]; }
      else pl.st.lf_l2l += T.nson[c];
    }
    sizeLevel(tl, oOff[nc], eOff[nc]);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nc; i++) {
      const int c = cs[i];
      int64_t o = oOff[i], e = eOff[i];
      auto entry = [&](int is, const double v[3], int sg) {
        int pd = T.slotDir[is];
        cplx edv = 1.;
        if (tl.hf) edv = phase(pl.dirs[pl.HFStage[l]][pd].data(), v, 1.);
        tl.in[e] = is; tl.sgn[e] = sg; tl.edv[e] = edv;
        tl.sigIdx[e] = tl.hf ? pl.sigRow0[l] + pd : -1;
        e++;
      };
      for (int a = 0; a < T.nson[c]; a++) {
        int ch = T.sons[T.son0[c] + a];
        double v[3];
        for (int kk = 0; kk < 3; kk++) v[kk] = T.ctr[3 * ch + kk] - T.ctr[3 * c + kk];
        int sg = (v[0] < 0. ? 0 : 1) | (v[1] < 0. ? 0 : 2) | (v[2] < 0. ? 0 : 4);
        for (int k = 0; k < T.nslot[ch]; k++) {
          int os = T.slot0[ch] + k;
          int cd = T.slotDir[os];
          if (byQuad) {
            // a cell's slots have ascending directions: same order as a scan
            for (int pd = 4 * cd; pd < 4 * cd + 4; pd++) {
              int is = T.slotOf(c, pd);
              if (is >= 0) entry(is, v, sg);
            }
          } else {
            for (int pk = 0; pk < T.nslot[c]; pk++) entry(T.slot0[c] + pk, v, sg);
          }
          tl.out[o] = os;
          tl.beg[o + 1] = (int)e;
          o++;
        }
      }
    }
    if (!tl.out.empty()) pl.l2l.push_back(std::move(tl));
  }

  pl.st.ms_trans = tick();
  // ---- M2L, CSR by target slot; M2F only for slots read by M2L ----
  if (dev) {
    int r = dev->m2lLists(pl, msg);
    if (r != OK) return r;
  } else {
    int64_t nm = (int64_t)pr.mt.size();
    std::vector<int> tslot(nm), sslot(nm);
    int missing = 0;
#pragma omp parallel for schedule(static) reduction(+ : missing)
    for (int64_t e2 = 0; e2 < nm; e2++) {
      int c = pr.mcone[e2];
      tslot[e2] = pl.T.slotOf(pr.mt[e2], c);
      sslot[e2] = pl.S.slotOf(pr.ms[e2], c);
      if (tslot[e2] < 0 || sslot[e2] < 0) missing++;
    }
    if (missing) { msg = "internal: missing M2L slot"; return ERR_STATE; }
    {
      std::vector<uint8_t> used(pl.S.nslots, 0);
#pragma omp parallel for schedule(static)
      for (int64_t e2 = 0; e2 < nm; e2++) __atomic_store_n(&used[sslot[e2]], (uint8_t)1, __ATOMIC_RELAXED);
      pl.m2fIndex.assign(pl.S.nslots, -1);
      for (int s = 0; s < pl.S.nslots; s++)
        if (used[s]) { pl.m2fIndex[s] = (int)pl.m2fSlot.size(); pl.m2fSlot.push_back(s); }
    }
    // CSR by target slot, stable (entries of a target in walk order): a
    // counting sort over contiguous chunks of the entries, one chunk per thread
    pl.m2lBeg.assign(pl.T.nslots + 1, 0);
    pl.m2lSrcF.resize(nm); pl.m2lTabId.resize(nm); pl.m2lHash.resize(nm);
    pl.m2lEntries = nm;
    {
      int nth = 1;
#ifdef _OPENMP
      nth = omp_get_max_threads();
#endif
      const int nts = pl.T.nslots;
      std::vector<std::vector<int>> cnt(nth);
#pragma omp parallel num_threads(nth)
      {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        const int64_t a = nm * t / nth, b = nm * (t + 1) / nth;
        std::vector<int>& c = cnt[t];
        c.assign(nts, 0);
        for (int64_t e2 = a; e2 < b; e2++) c[tslot[e2]]++;
#pragma omp barrier
#pragma omp for schedule(static)
        for (int s = 0; s < nts; s++) {
          int run = 0;
          for (int u = 0; u < nth; u++) {
            int v = cnt[u][s];
            cnt[u][s] = run;  // offset of chunk u within slot s
            run += v;
          }
          pl.m2lBeg[s + 1] = run;
        }
#pragma omp single
        for (int s = 0; s < nts; s++) pl.m2lBeg[s + 1] += pl.m2lBeg[s];
        // implicit barrier after single
        for (int64_t e2 = a; e2 < b; e2++) {
          const int s = tslot[e2];
          const int64_t pos = (int64_t)pl.m2lBeg[s] + c[s]++;
          pl.m2lSrcF[pos] = pl.m2fIndex[sslot[e2]];
          pl.m2lTabId[pos] = pr.mtab[e2];
          pl.m2lHash[pos] = pr.mh[e2];
        }
      }
    }
  }
  pl.st.m2f_slots = (int64_t)pl.m2fSlot.size();
  pl.st.m2l_entries = pl.m2lEntries;
  for (int s = 0; s < pl.T.nslots; s++)
    if (pl.m2lBeg[s + 1] > pl.m2lBeg[s]) pl.st.m2l_targets++;
  // Processing order of the M2L targets: by level, then direction, then
  // cell (pre-order = Morton order).  Targets of one direction in nearby
  // cells read largely the same source F rows, so running them together
  // keeps those rows in cache.  Targets without M2L only need L = 0.
  {
    // counting sort by (level, direction), stable in slot order (an LF
    // slot's direction -1 sorts first within its level)
    const int nts = pl.T.nslots;
    int maxDir = 1, maxLvl = 0;
    for (int s = 0; s < nts; s++) {
      maxDir = std::max(maxDir, pl.T.slotDir[s] + 2);
      maxLvl = std::max(maxLvl, pl.T.lvl[pl.T.slotCell[s]]);
    }
    auto keyOf = [&](int s) {
      return (int64_t)pl.T.lvl[pl.T.slotCell[s]] * maxDir + (pl.T.slotDir[s] + 1);
    };
    const int64_t nk = (int64_t)(maxLvl + 1) * maxDir;
    std::vector<int64_t> start(nk + 1, 0);
    for (int s = 0; s < nts; s++)
      if (pl.m2lBeg[s + 1] > pl.m2lBeg[s]) start[keyOf(s) + 1]++;
      else pl.m2lZero.push_back(s);
    for (int64_t k = 0; k < nk; k++) start[k + 1] += start[k];
    std::vector<int> ord(start[nk]);
    for (int s = 0; s < nts; s++)
      if (pl.m2lBeg[s + 1] > pl.m2lBeg[s]) ord[start[keyOf(s)]++] = s;
    pl.m2lOrder.swap(ord);
  }
  // Groups of up to m2lGroupSize(L) consecutive targets of one level and direction
  // (neighbouring cells): a group's entries are the union of its targets'
  // source rows, each with the mask of targets that read it and their
  // (table, permutation) codes, so that a source row is loaded once per group.
  {
    const int G = m2lGroupSize(L);
    pl.m2lG = G;
    // group boundaries in the processing order (serial, cheap)
    std::vector<int> gStart;  // index into m2lOrder of each group's first target
    {
      size_t i = 0;
      while (i < pl.m2lOrder.size()) {
        gStart.push_back((int)i);
        int t0 = pl.m2lOrder[i];
        int l0 = pl.T.lvl[pl.T.slotCell[t0]], d0 = pl.T.slotDir[t0];
        int ng = 0;
        while (ng < G && i < pl.m2lOrder.size()) {
          int t = pl.m2lOrder[i];
          if (pl.T.lvl[pl.T.slotCell[t]] != l0 || pl.T.slotDir[t] != d0) break;
          ng++;
          i++;
        }
      }
      gStart.push_back((int)pl.m2lOrder.size());
    }
    const int64_t ngroups = (int64_t)gStart.size() - 1;
    pl.m2lGroupTarget.resize(ngroups * G);
#pragma omp parallel for schedule(static)
    for (int64_t gi = 0; gi < ngroups; gi++) {
      const int ng = gStart[gi + 1] - gStart[gi];
      for (int g = 0; g < G; g++)
        pl.m2lGroupTarget[gi * G + g] = g < ng ? pl.m2lOrder[gStart[gi] + g] : -1;
    }
    if (dev) {
      // a merge of the targets' (ascending) source rows per group, on the device
      int r = dev->m2lGroups(pl, msg);
      if (r != OK) return r;
    } else {
      // entries of contiguous ranges of groups built in parallel, then joined
      // in range order (the result does not depend on the thread count)
      const int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(ngroups, 1024));
      struct Part {
        std::vector<int> src, mask, code, nent;
      };
      std::vector<Part> parts(nchunk);
#pragma omp parallel for schedule(dynamic, 1)
      for (int64_t ch = 0; ch < nchunk; ch++) {
        Part& pt = parts[ch];
        std::vector<std::array<int, 3>> tmp;  // srcF, target in group, tabHash
        const int64_t g0 = ngroups * ch / nchunk, g1 = ngroups * (ch + 1) / nchunk;
        for (int64_t gi = g0; gi < g1; gi++) {
          const int ng = gStart[gi + 1] - gStart[gi];
          tmp.clear();
          for (int g = 0; g < ng; g++) {
            const int t = pl.m2lOrder[gStart[gi] + g];
            for (int e = pl.m2lBeg[t]; e < pl.m2lBeg[t + 1]; e++)
              tmp.push_back({pl.m2lSrcF[e], g, (pl.m2lTabId[e] << 6) | pl.m2lHash[e]});
          }
          std::stable_sort(tmp.begin(), tmp.end(),
                           [](const std::array<int, 3>& a, const std::array<int, 3>& b) {
                             return a[0] < b[0];
                           });
          int cnt = 0;
          size_t k = 0;
          while (k < tmp.size()) {
            int src = tmp[k][0];
            int mask = 0;
            int codes[kM2LGroupMax] = {0};
            size_t k2 = k;
            // a target reads a source row at most once per group entry; a
            // repeat (the same pair listed twice) starts a new entry
            while (k2 < tmp.size() && tmp[k2][0] == src && !((mask >> tmp[k2][1]) & 1)) {
              mask |= 1 << tmp[k2][1];
              codes[tmp[k2][1]] = tmp[k2][2];
              k2++;
            }
            pt.src.push_back(src);
            pt.mask.push_back(mask);
            for (int g = 0; g < G; g++) pt.code.push_back(codes[g]);
            cnt++;
            k = k2;
          }
          pt.nent.push_back(cnt);
        }
      }
      int64_t total = 0;
      for (const Part& pt : parts) total += (int64_t)pt.src.size();
      if (total >= (int64_t)1 << 31) { msg = "M2L lists exceed 2^31 entries"; return ERR_MEMORY; }
      // the parts in range order: offsets first, then copied in parallel
      std::vector<int64_t> eOff(nchunk + 1, 0), gOff(nchunk + 1, 0);
      for (int64_t ch = 0; ch < nchunk; ch++) {
        eOff[ch + 1] = eOff[ch] + (int64_t)parts[ch].src.size();
        gOff[ch + 1] = gOff[ch] + (int64_t)parts[ch].nent.size();
      }
      pl.m2lGSrcF.resize(total);
      pl.m2lGMask.resize(total);
      pl.m2lGCode.resize(total * G);
      pl.m2lGroupBeg.assign(ngroups + 1, 0);
      {
        // group starts: a running sum over the parts' entry counts
        int64_t run = 0;
        for (int64_t ch = 0; ch < nchunk; ch++)
          for (size_t k = 0; k < parts[ch].nent.size(); k++) {
            run += parts[ch].nent[k];
            pl.m2lGroupBeg[gOff[ch] + (int64_t)k + 1] = (int)run;
          }
      }
#pragma omp parallel for schedule(dynamic, 1)
      for (int64_t ch = 0; ch < nchunk; ch++) {
        Part& pt = parts[ch];
        std::copy(pt.src.begin(), pt.src.end(), pl.m2lGSrcF.begin() + eOff[ch]);
        std::copy(pt.mask.begin(), pt.mask.end(), pl.m2lGMask.begin() + eOff[ch]);
        std::copy(pt.code.begin(), pt.code.end(), pl.m2lGCode.begin() + eOff[ch] * G);
        Part().src.swap(pt.src);
        Part().mask.swap(pt.mask);
        Part().code.swap(pt.code);
      }
    }
    pl.st.m2l_groups = ngroups;
    pl.st.m2l_group_entries = (int64_t)pl.m2lGSrcF.size();
  }

  pl.st.ms_m2l = tick();
  // ---- P2P, CSR by target leaf (P2P.hpp:42-79 over whole cell ranges) ----
  // A pair with a non-leaf target cell is split over its descendant leaves;
  // each leaf's source ranges are sorted and touching ranges merged.
  if (dev) {
    int r = dev->p2pLists(pl, msg);
    if (r != OK) return r;
  } else {
    const Tree& T = pl.T;
    std::vector<int> leafOfCell(T.ncells(), -1);
    std::vector<int> leaves;
    for (int c = 0; c < T.ncells(); c++)
      if (T.leaf(c)) { leafOfCell[c] = (int)leaves.size(); leaves.push_back(c); }
    // descendant leaves of a cell are a contiguous run in pre-order
    std::vector<int> firstLeaf(T.ncells()), nLeaves(T.ncells());
    for (int c = T.ncells() - 1; c >= 0; c--) {
      if (T.leaf(c)) { firstLeaf[c] = leafOfCell[c]; nLeaves[c] = 1; continue; }
      int f = 1 << 30, n = 0;
      for (int a = 0; a < T.nson[c]; a++) {
        int s = T.sons[T.son0[c] + a];
        f = std::min(f, firstLeaf[s]);
        n += nLeaves[s];
      }
      firstLeaf[c] = f; nLeaves[c] = n;
    }
    std::vector<int64_t> cnt(leaves.size() + 1, 0);
    for (size_t k = 0; k < pr.pt.size(); k++)
      for (int q = 0; q < nLeaves[pr.pt[k]]; q++) cnt[firstLeaf[pr.pt[k]] + q + 1]++;
    for (size_t k = 0; k < leaves.size(); k++) cnt[k + 1] += cnt[k];
    std::vector<int> rb(cnt.back()), re(cnt.back());
    {
      std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
      for (size_t k = 0; k < pr.pt.size(); k++) {
        int s = pr.ps[k];
        for (int q = 0; q < nLeaves[pr.pt[k]]; q++) {
          int64_t pos = fill[firstLeaf[pr.pt[k]] + q]++;
          rb[pos] = pl.S.pbeg[s];
          re[pos] = pl.S.pbeg[s] + pl.S.nprt[s];
        }
      }
    }
    // sort each leaf's ranges and merge touching ones (in parallel), then
    // concatenate in leaf order
    int64_t nl = (int64_t)leaves.size();
    std::vector<int> nOut(nl, 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < nl; k++) {
      int64_t b = cnt[k], e = cnt[k + 1];
      if (b == e) continue;
      std::vector<std::pair<int, int>> rr;
      rr.reserve(e - b);
      for (int64_t q = b; q < e; q++) rr.push_back({rb[q], re[q]});
      std::sort(rr.begin(), rr.end());
      // merge touching ranges only; overlapping ones stay separate so that
      // every listed pair is summed exactly as often as the walk lists it
      int64_t w = b;
      int cb = rr[0].first, ce = rr[0].second;
      for (size_t q = 1; q < rr.size(); q++) {
        if (rr[q].first == ce) { ce = rr[q].second; continue; }
        rb[w] = cb; re[w] = ce; w++;
        cb = rr[q].first; ce = rr[q].second;
      }
      rb[w] = cb; re[w] = ce; w++;
      nOut[k] = (int)(w - b);
    }
    pl.p2pBeg.push_back(0);
    for (int64_t k = 0; k < nl; k++) {
      if (!nOut[k]) continue;
      for (int64_t q = cnt[k]; q < cnt[k] + nOut[k]; q++) {
        pl.p2pSrcB.push_back(rb[q]);
        pl.p2pSrcE.push_back(re[q]);
      }
      pl.p2pLeaf.push_back(leaves[k]);
      pl.p2pBeg.push_back((int)pl.p2pSrcB.size());
    }
  }
  pl.st.ms_p2p = tick();
  return validatePlan(pl, msg);
}

// Every index the device kernels follow, checked on the host before upload.
int validatePlan(const Plan& pl, std::string& msg) {
  auto bad = [&](const char* what) {
    msg = std::string("internal: plan index out of range: ") + what;
    return ERR_STATE;
  };
  const int LL = pl.LL;
  const int nrows = (int)(pl.sig.size() / (size_t)LL);
  const int64_t ns = (int64_t)pl.S.x.size(), nt = (int64_t)pl.T.x.size();
  for (const Tree* t : {&pl.S, &pl.T}) {
    int64_t np = (int64_t)t->x.size();
    for (int c = 0; c < t->ncells(); c++) {
      if (t->pbeg[c] < 0 || t->nprt[c] < 1 || (int64_t)t->pbeg[c] + t->nprt[c] > np) return bad("cell range");
      if (t->slot0[c] < 0 || t->slot0[c] + t->nslot[c] > t->nslots) return bad("cell slots");
      if (t->lvl[c] < 0 || t->lvl[c] >= pl.nlvl) return bad("cell level");
      int l = t->lvl[c];
      for (int k = 0; k < t->nslot[c]; k++) {
        int d = t->slotDir[t->slot0[c] + k];
        if (pl.isHF[l] && (d < 0 || d >= pl.nDir[l] || pl.sigRow0[l] < 0)) return bad("slot direction");
      }
    }
  }
  for (int c : pl.p2mCell)
    if (c < 0 || c >= pl.S.ncells() || !pl.S.leaf(c)) return bad("p2m leaf");
  for (int c : pl.l2pCell)
    if (c < 0 || c >= pl.T.ncells() || !pl.T.leaf(c)) return bad("l2p leaf");
  for (int w = 0; w < 2; w++) {
    const auto& lst = w ? pl.l2l : pl.m2m;
    int nout = w ? pl.T.nslots : pl.S.nslots;
    for (const auto& tl : lst) {
      if (tl.beg.size() != tl.out.size() + 1 || tl.beg.back() != (int)tl.in.size()) return bad("trans csr");
      for (int o : tl.out)
        if (o < 0 || o >= nout) return bad("trans out");
      for (size_t e = 0; e < tl.in.size(); e++) {
        if (tl.in[e] < 0 || tl.in[e] >= nout) return bad("trans in");
        if (tl.sigIdx[e] >= nrows || (tl.hf && tl.sigIdx[e] < 0)) return bad("trans phase row");
        if (tl.sgn[e] < 0 || tl.sgn[e] > 7) return bad("trans sign");
      }
    }
  }
  for (int s : pl.m2fSlot)
    if (s < 0 || s >= pl.S.nslots) return bad("m2f slot");
  if ((int)pl.m2lBeg.size() != pl.T.nslots + 1 || pl.m2lBeg.back() != pl.m2lEntries) return bad("m2l csr");
  if (!pl.m2lSrcF.empty() && (int64_t)pl.m2lSrcF.size() != pl.m2lEntries) return bad("m2l entries");
  {
    // the entry checks in parallel, failures as bits
    const int64_t nm = (int64_t)pl.m2lSrcF.size();
    const int nf = (int)pl.m2fSlot.size();
    int fail = 0;
#pragma omp parallel for schedule(static) reduction(| : fail)
    for (int64_t e = 0; e < nm; e++) {
      if (pl.m2lSrcF[e] < 0 || pl.m2lSrcF[e] >= nf) fail |= 1;
      if (pl.m2lTabId[e] < 0 || pl.m2lTabId[e] >= pl.ntab || pl.m2lTabId[e] >= (1 << 25)) fail |= 2;
      if (pl.m2lHash[e] < 0 || pl.m2lHash[e] > 63) fail |= 4;
    }
    if (fail & 1) return bad("m2l source");
    if (fail & 2) return bad("m2l table");
    if (fail & 4) return bad("m2l hash");
  }
  for (int v : pl.perm)
    if (v < 0 || v >= pl.PP) return bad("permutation");
  if (pl.m2lOrder.size() + pl.m2lZero.size() != (size_t)pl.T.nslots) return bad("m2l order size");
  for (int v : pl.m2lOrder)
    if (v < 0 || v >= pl.T.nslots) return bad("m2l order");
  for (int v : pl.m2lZero)
    if (v < 0 || v >= pl.T.nslots) return bad("m2l zero list");
  {
    size_t ng = pl.m2lGroupBeg.size() - 1;
    const int GG = pl.m2lG;
    if (GG < 1 || GG > kM2LGroupMax) return bad("m2l group size");
    if (pl.m2lGroupTarget.size() != ng * GG) return bad("m2l group targets");
    if (pl.m2lGroupBeg.back() != (int)pl.m2lGSrcF.size() ||
        pl.m2lGMask.size() != pl.m2lGSrcF.size() ||
        pl.m2lGCode.size() != pl.m2lGSrcF.size() * GG)
      return bad("m2l group csr");
    std::vector<char> seen(pl.T.nslots, 0);
    std::vector<int> haveOf(ng, 0);
    for (size_t g = 0; g < ng; g++) {
      if (pl.m2lGroupBeg[g + 1] < pl.m2lGroupBeg[g]) return bad("m2l group order");
      int have = 0;
      for (int k = 0; k < GG; k++) {
        int t = pl.m2lGroupTarget[g * GG + k];
        if (t < -1 || t >= pl.T.nslots) return bad("m2l group target");
        if (t >= 0) {
          if (seen[t]) return bad("m2l target in two groups");
          seen[t] = 1;
          have |= 1 << k;
        }
      }
      haveOf[g] = have;
    }
    // the groups' entries in parallel, failures as bits
    const int nf = (int)pl.m2fSlot.size();
    int fail = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(| : fail)
    for (int64_t g = 0; g < (int64_t)ng; g++)
      for (int e = pl.m2lGroupBeg[g]; e < pl.m2lGroupBeg[g + 1]; e++) {
        if (pl.m2lGSrcF[e] < 0 || pl.m2lGSrcF[e] >= nf) fail |= 1;
        int m = pl.m2lGMask[e];
        if (m <= 0 || (m & ~haveOf[g])) fail |= 2;
        for (int k = 0; k < GG; k++) {
          if (!((m >> k) & 1)) continue;
          int c = pl.m2lGCode[(size_t)e * GG + k];
          if ((c >> 6) < 0 || (c >> 6) >= pl.ntab) fail |= 4;
        }
      }
    if (fail & 1) return bad("m2l group source");
    if (fail & 2) return bad("m2l group mask");
    if (fail & 4) return bad("m2l group table");
    for (int v : pl.m2lOrder)
      if (!seen[v]) return bad("m2l target not in a group");
  }
  if (pl.p2pBeg.size() != pl.p2pLeaf.size() + 1) return bad("p2p csr");
  for (size_t k = 0; k < pl.p2pLeaf.size(); k++) {
    int c = pl.p2pLeaf[k];
    if (c < 0 || c >= pl.T.ncells() || !pl.T.leaf(c)) return bad("p2p leaf");
    if (pl.p2pBeg[k + 1] <= pl.p2pBeg[k]) return bad("p2p empty list");
  }
  for (size_t r = 0; r < pl.p2pSrcB.size(); r++)
    if (pl.p2pSrcB[r] < 0 || pl.p2pSrcE[r] <= pl.p2pSrcB[r] || pl.p2pSrcE[r] > ns) return bad("p2p range");
  (void)nt;
  return OK;
}

}  // namespace dfm
