// A CPU model of the device set-up (defmm_amd/csrc/setup.hip), plugged into
// the plan builder through its SetupDevice interface and compared, array by
// array, with the plan the host builder makes for the same points.
//
// The model takes the device path's steps in the device path's order: the
// level-synchronous octree build (a splitting cell's points partitioned
// stably by octant, level by level, then renumbered in pre-order), the
// level-synchronous walk (the sources a target meets at level l are the
// children of the sources its parent expanded at level l - 1, visited in
// the parent's list order and child order, as a wavefront's lanes do), the
// smallest (target, source) pair as each offset's representative, the
// host's offsetCodes for table ids, hashes and cones, the marks of both
// cells of each HF pair, the slot counts, the stable placement of each
// target cell's pairs, the merge of a group's source rows, and per target
// leaf the merge of its and its ancestors' P2P source lists.  What it
// checks without a GPU is that these steps, and the host functions they
// use (m2lWindow, offsetCodes), give the host builder's plan: the claims
// the device path rests on (one target's sources met in ascending cell
// order; cones equal whichever pair of an offset they come from; strictly
// ascending rows per target slot).
//
// Prints one line per case: "case <name> OK <entries>" or the first
// mismatching array.
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../../defmm_amd/csrc/plan.hpp"

using namespace dfm;

namespace {

bool macHFModel(const Plan& pl, int i, int j) {
  const double d0 = pl.T.ctr[3 * i] - pl.S.ctr[3 * j], d1 = pl.T.ctr[3 * i + 1] - pl.S.ctr[3 * j + 1],
               d2 = pl.T.ctr[3 * i + 2] - pl.S.ctr[3 * j + 2];
  double s = 0.;
  s += d0 * d0;
  s += d1 * d1;
  s += d2 * d2;
  const double sr = pl.S.rad[j];
  const double tmp = (1. - 1.e-8) * (std::sqrt(s) - 2. * sr);
  const double tt = 2. * sr;
  const double a = tt * tt * pl.kappa;
  return tmp >= (a < tt ? tt : a);
}

bool macLFModel(int z0, int z1, int z2) {
  const int test = std::abs(z0) + std::abs(z1) + std::abs(z2);
  const int ncmp = (z0 != 0) + (z1 != 0) + (z2 != 0);
  return test > ncmp;
}

int winIndex(int z0, int z1, int z2, int W) {
  const int n = 2 * W + 1;
  return ((z0 + W) * n + (z1 + W)) * n + (z2 + W);
}

struct ModelSetup final : SetupDevice {
  struct Level {
    int l = 0, kind = 0, W = -1;
    std::vector<int> tcell, mBeg, mSrc;
  };
  std::vector<Level> lv;
  std::vector<std::vector<uint64_t>> code;
  ListVec srcF, tab, hash;
  // P2P pairs kept by the walk: per level the source cells, per target cell
  // its range [first, second) in its level's list
  std::vector<std::vector<int>> pLv;
  std::vector<std::pair<int, int>> pSeg;

  uint64_t codeOf(const Plan& pl, const Level& L, int i, int j) const {
    const int z0 = pl.T.idx[3 * i] - pl.S.idx[3 * j], z1 = pl.T.idx[3 * i + 1] - pl.S.idx[3 * j + 1],
              z2 = pl.T.idx[3 * i + 2] - pl.S.idx[3 * j + 2];
    return code[L.l][winIndex(z0, z1, z2, L.W)];
  }

  // level by level: each splitting cell's points partitioned stably by octant
  int tree(Tree& t, const double* xyz, int64_t n, const double ctr[3], double rad, int ncrit,
           int lvmax, std::string& msg) override {
    std::vector<int> perm(n);
    for (int64_t i = 0; i < n; i++) perm[i] = (int)i;
    std::vector<std::vector<LevelCell>> lev(1);
    LevelCell r;
    r.b = 0;
    r.n = n;
    for (int q = 0; q < 3; q++) r.ctr[q] = ctr[q];
    r.rad = rad;
    lev[0].push_back(r);
    for (int l = 0;; l++) {
      std::vector<LevelCell> next;
      for (LevelCell& c : lev[l]) {
        if (cellStops(c.n, l, ncrit, lvmax)) continue;
        if (l >= 30) return ERR_DEPTH;
        int64_t cnt[8] = {0};
        std::vector<int> bucket[8];
        for (int64_t i = c.b; i < c.b + c.n; i++) {
          const int64_t p = perm[i];
          const int o = (xyz[3 * p] > c.ctr[0] ? 1 : 0) | (xyz[3 * p + 1] > c.ctr[1] ? 2 : 0) |
                        (xyz[3 * p + 2] > c.ctr[2] ? 4 : 0);
          bucket[o].push_back((int)p);
          cnt[o]++;
        }
        int64_t w = c.b;
        for (int o = 0; o < 8; o++)
          for (int p : bucket[o]) perm[w++] = p;
        splitCell(c, cnt, next);
      }
      if (next.empty()) break;
      lev.push_back(std::move(next));
    }
    treeFromLevels(t, lev, perm, xyz);
    return OK;
  }

  int walk(Plan& pl, std::vector<std::array<int, 4>>& tabKeys, std::vector<int>& pt,
           std::vector<int>& ps, std::string& msg) override {
    const Tree &T = pl.T, &S = pl.S;
    std::vector<int> W(pl.nlvl, -1);
    for (int l = 0; l < pl.nlvl; l++) {
      W[l] = m2lWindow(pl, l);
      if (pl.precomp[l] && W[l] < 0) return kSetupFallback;
    }
    pt.clear();
    ps.clear();
    pLv.assign(pl.nlvl, {});
    pSeg.assign(T.ncells(), {0, 0});
    if (T.leaf(0) || S.leaf(0)) {
      pLv[0] = {0};
      pSeg[0] = {0, 1};
      pl.st.hf_m2l = pl.st.lf_m2l = 0;
      pl.st.p2p_calls = 1;
      tabKeys.clear();
      return OK;
    }
    const int nT = T.ncells();
    std::vector<int> pos(nT, 0), parent(nT, -1);
    std::vector<std::vector<int>> cells(pl.nlvl);
    for (int c = 0; c < nT; c++) {
      pos[c] = (int)cells[T.lvl[c]].size();
      cells[T.lvl[c]].push_back(c);
      for (int a = 0; a < T.nson[c]; a++) parent[T.sons[T.son0[c] + a]] = c;
    }
    std::vector<std::vector<uint64_t>> rep(pl.nlvl);
    for (int l = 0; l < pl.nlvl; l++)
      if (W[l] >= 0) rep[l].assign((size_t)(2 * W[l] + 1) * (2 * W[l] + 1) * (2 * W[l] + 1), ~0ull);
    std::vector<int> eBeg = {0, 1}, eSrc = {0};
    int64_t nHF = 0, nLF = 0, nP = 0;
    for (int l = 1; l < pl.nlvl && !eSrc.empty(); l++) {
      if (cells[l].empty()) break;
      Level L;
      L.l = l;
      L.kind = pl.precomp[l] ? ((l && pl.isHF[l]) ? 1 : 2) : 0;
      L.W = W[l];
      L.tcell = cells[l];
      L.mBeg.push_back(0);
      std::vector<int> nb = {0}, ns;
      for (int i : cells[l]) {
        const int p = pos[parent[i]];
        pSeg[i].first = (int)pLv[l].size();
        for (int e = eBeg[p]; e < eBeg[p + 1]; e++) {
          const int q = eSrc[e];
          for (int b = 0; b < S.nson[q]; b++) {
            const int j = S.sons[S.son0[q] + b];
            const int z0 = T.idx[3 * i] - S.idx[3 * j], z1 = T.idx[3 * i + 1] - S.idx[3 * j + 1],
                      z2 = T.idx[3 * i + 2] - S.idx[3 * j + 2];
            bool far = false;
            if (L.kind == 1) far = macHFModel(pl, i, j);
            else if (L.kind == 2) far = macLFModel(z0, z1, z2);
            if (far) {
              L.mSrc.push_back(j);
              if (std::abs(z0) > L.W || std::abs(z1) > L.W || std::abs(z2) > L.W) return kSetupFallback;
              uint64_t& r = rep[l][winIndex(z0, z1, z2, L.W)];
              const uint64_t v = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
              if (v < r) r = v;
            } else if (T.leaf(i) || S.leaf(j)) {
              pLv[l].push_back(j);
              nP++;
            } else {
              ns.push_back(j);
            }
          }
        }
        L.mBeg.push_back((int)L.mSrc.size());
        nb.push_back((int)ns.size());
        pSeg[i].second = (int)pLv[l].size();
      }
      if (L.kind == 1) nHF += (int64_t)L.mSrc.size();
      if (L.kind == 2) nLF += (int64_t)L.mSrc.size();
      eBeg.swap(nb);
      eSrc.swap(ns);
      if (!L.mSrc.empty()) lv.push_back(std::move(L));
    }
    int r = offsetCodes(pl, W, rep, code, tabKeys, msg);
    if (r != OK) return r;
    for (const Level& L : lv) {
      if (L.kind != 1) continue;
      for (size_t t = 0; t < L.tcell.size(); t++) {
        const int i = L.tcell[t];
        for (int e = L.mBeg[t]; e < L.mBeg[t + 1]; e++) {
          const int j = L.mSrc[e];
          const int cone = (int)(codeOf(pl, L, i, j) >> 40) - 1;
          pl.T.marks[pl.T.mark0[i] + cone] = 1;
          pl.S.marks[pl.S.mark0[j] + cone] = 1;
        }
      }
    }
    pl.st.hf_m2l = nHF;
    pl.st.lf_m2l = nLF;
    pl.st.p2p_calls = nP;
    return OK;
  }

  // per target leaf: its and its ancestors' P2P source lists merged by
  // source range, touching ranges joined
  int p2pLists(Plan& pl, std::string& msg) override {
    const Tree &T = pl.T, &S = pl.S;
    std::vector<int> parent(T.ncells(), -1);
    for (int c = 0; c < T.ncells(); c++)
      for (int k = 0; k < T.nson[c]; k++) parent[T.sons[T.son0[c] + k]] = c;
    pl.p2pLeaf.clear();
    pl.p2pBeg.assign(1, 0);
    pl.p2pSrcB.clear();
    pl.p2pSrcE.clear();
    int64_t pairs = 0;
    for (int leaf = 0; leaf < T.ncells(); leaf++) {
      if (!T.leaf(leaf)) continue;
      std::vector<std::array<int, 3>> heads;  // level, next, end
      for (int c = leaf; c >= 0; c = parent[c])
        if (pSeg[c].second > pSeg[c].first) heads.push_back({T.lvl[c], pSeg[c].first, pSeg[c].second});
      int cb = 0, ce = 0, n = 0;
      bool open = false;
      for (;;) {
        int best = -1, bb = 0x7FFFFFFF, be = 0x7FFFFFFF;
        for (size_t h = 0; h < heads.size(); h++) {
          if (heads[h][1] >= heads[h][2]) continue;
          const int j = pLv[heads[h][0]][heads[h][1]];
          const int rb = S.pbeg[j], re = rb + S.nprt[j];
          if (rb < bb || (rb == bb && re < be)) { best = (int)h; bb = rb; be = re; }
        }
        if (best < 0) break;
        heads[best][1]++;
        pairs += (int64_t)(be - bb) * T.nprt[leaf];
        if (open && bb == ce) {
          ce = be;
        } else {
          if (open) { pl.p2pSrcB.push_back(cb); pl.p2pSrcE.push_back(ce); n++; }
          cb = bb;
          ce = be;
          open = true;
        }
      }
      if (open) { pl.p2pSrcB.push_back(cb); pl.p2pSrcE.push_back(ce); n++; }
      if (n) {
        pl.p2pLeaf.push_back(leaf);
        pl.p2pBeg.push_back((int)pl.p2pSrcB.size());
      }
    }
    pl.st.p2p_pairs = pairs;
    return OK;
  }

  static int slotOfModel(const Tree& t, int c, int d) {
    if (t.nslot[c] == 0) return -1;
    if (t.slotDir[t.slot0[c]] < 0) return t.slot0[c];
    return (d >= 0 && t.mark0[c] >= 0) ? t.slotIdx[t.mark0[c] + d] : -1;
  }

  int m2lLists(Plan& pl, std::string& msg) override {
    const int nts = pl.T.nslots, nss = pl.S.nslots;
    std::vector<int> cnt(nts, 0);
    std::vector<uint8_t> used(nss, 0);
    int64_t nm = 0;
    auto each = [&](auto&& fn) {
      for (const Level& L : lv)
        for (size_t t = 0; t < L.tcell.size(); t++)
          for (int e = L.mBeg[t]; e < L.mBeg[t + 1]; e++) {
            const int i = L.tcell[t], j = L.mSrc[e];
            const uint64_t c = codeOf(pl, L, i, j);
            const int cone = L.kind == 1 ? (int)(c >> 40) - 1 : -1;
            fn(slotOfModel(pl.T, i, cone), slotOfModel(pl.S, j, cone), (int)(c & 0xFFFFFFFFu),
               (int)((c >> 32) & 63));
          }
    };
    bool missing = false;
    each([&](int ts, int ss, int, int) {
      if (ts < 0 || ss < 0) { missing = true; return; }
      cnt[ts]++;
      used[ss] = 1;
      nm++;
    });
    if (missing) { msg = "internal: missing M2L slot"; return ERR_STATE; }
    pl.m2lBeg.assign((size_t)nts + 1, 0);
    for (int s = 0; s < nts; s++) pl.m2lBeg[s + 1] = pl.m2lBeg[s] + cnt[s];
    pl.m2fIndex.assign(nss, -1);
    pl.m2fSlot.clear();
    for (int s = 0; s < nss; s++)
      if (used[s]) { pl.m2fIndex[s] = (int)pl.m2fSlot.size(); pl.m2fSlot.push_back(s); }
    std::vector<int> cur(pl.m2lBeg.begin(), pl.m2lBeg.end() - 1);
    pl.m2lSrcF.assign(nm, 0);
    pl.m2lTabId.assign(nm, 0);
    pl.m2lHash.assign(nm, 0);
    each([&](int ts, int ss, int id, int h) {
      const int q = cur[ts]++;
      pl.m2lSrcF[q] = pl.m2fIndex[ss];
      pl.m2lTabId[q] = id;
      pl.m2lHash[q] = h;
    });
    pl.m2lEntries = nm;
    srcF = pl.m2lSrcF;
    tab = pl.m2lTabId;
    hash = pl.m2lHash;
    return OK;
  }

  int m2lGroups(Plan& pl, std::string& msg) override {
    const int G = pl.m2lG;
    const int64_t ng = (int64_t)pl.m2lGroupTarget.size() / G;
    pl.m2lGroupBeg.assign(ng + 1, 0);
    pl.m2lGSrcF.clear();
    pl.m2lGMask.clear();
    pl.m2lGCode.clear();
    const int kEnd = 0x7FFFFFFF;
    for (int64_t gi = 0; gi < ng; gi++) {
      std::vector<int> h(G), e(G), v(G);
      for (int g = 0; g < G; g++) {
        const int t = pl.m2lGroupTarget[gi * G + g];
        h[g] = t >= 0 ? pl.m2lBeg[t] : 0;
        e[g] = t >= 0 ? pl.m2lBeg[t + 1] : 0;
        v[g] = h[g] < e[g] ? srcF[h[g]] : kEnd;
      }
      for (;;) {
        int m = kEnd;
        for (int g = 0; g < G; g++) m = std::min(m, v[g]);
        if (m == kEnd) break;
        int mask = 0;
        for (int g = 0; g < G; g++) {
          int c = 0;
          if (v[g] == m) {
            mask |= 1 << g;
            c = (tab[h[g]] << 6) | hash[h[g]];
            h[g]++;
            const int nv = h[g] < e[g] ? srcF[h[g]] : kEnd;
            if (nv <= m) { msg = "model: a target's sources do not strictly ascend"; return ERR_STATE; }
            v[g] = nv;
          }
          pl.m2lGCode.push_back(c);
        }
        pl.m2lGSrcF.push_back(m);
        pl.m2lGMask.push_back(mask);
      }
      pl.m2lGroupBeg[gi + 1] = (int)pl.m2lGSrcF.size();
    }
    return OK;
  }
};

template <class V>
bool same(const char* name, const V& a, const V& b, std::string& bad) {
  if (a == b) return true;
  if (bad.empty()) bad = name;
  return false;
}

std::string comparePlans(const Plan& A, const Plan& B) {
  std::string bad;
  for (int w = 0; w < 2; w++) {
    const Tree& a = w ? A.T : A.S;
    const Tree& b = w ? B.T : B.S;
    same("perm", a.perm, b.perm, bad);
    same("x", a.x, b.x, bad);
    same("y", a.y, b.y, bad);
    same("z", a.z, b.z, bad);
    same("lvl", a.lvl, b.lvl, bad);
    same("nprt", a.nprt, b.nprt, bad);
    same("pbeg", a.pbeg, b.pbeg, bad);
    same("son0", a.son0, b.son0, bad);
    same("nson", a.nson, b.nson, bad);
    same("sons", a.sons, b.sons, bad);
    same("idx", a.idx, b.idx, bad);
    same("ctr", a.ctr, b.ctr, bad);
    same("rad", a.rad, b.rad, bad);
    same("nPerLevel", a.nPerLevel, b.nPerLevel, bad);
    same("marks", a.marks, b.marks, bad);
    same("slot0", a.slot0, b.slot0, bad);
    same("nslot", a.nslot, b.nslot, bad);
    same("slotCell", a.slotCell, b.slotCell, bad);
    same("slotDir", a.slotDir, b.slotDir, bad);
  }
  same("m2lTab", A.m2lTab, B.m2lTab, bad);
  same("m2fSlot", A.m2fSlot, B.m2fSlot, bad);
  same("m2lBeg", A.m2lBeg, B.m2lBeg, bad);
  same("m2lSrcF", A.m2lSrcF, B.m2lSrcF, bad);
  same("m2lTabId", A.m2lTabId, B.m2lTabId, bad);
  same("m2lHash", A.m2lHash, B.m2lHash, bad);
  same("m2lOrder", A.m2lOrder, B.m2lOrder, bad);
  same("m2lZero", A.m2lZero, B.m2lZero, bad);
  same("m2lGroupBeg", A.m2lGroupBeg, B.m2lGroupBeg, bad);
  same("m2lGroupTarget", A.m2lGroupTarget, B.m2lGroupTarget, bad);
  same("m2lGSrcF", A.m2lGSrcF, B.m2lGSrcF, bad);
  same("m2lGMask", A.m2lGMask, B.m2lGMask, bad);
  same("m2lGCode", A.m2lGCode, B.m2lGCode, bad);
  same("p2pLeaf", A.p2pLeaf, B.p2pLeaf, bad);
  same("p2pBeg", A.p2pBeg, B.p2pBeg, bad);
  same("p2pSrcB", A.p2pSrcB, B.p2pSrcB, bad);
  same("p2pSrcE", A.p2pSrcE, B.p2pSrcE, bad);
  if (A.m2m.size() != B.m2m.size() || A.l2l.size() != B.l2l.size()) same("trans levels", 0, 1, bad);
  for (int w = 0; w < 2 && bad.empty(); w++) {
    const auto& a = w ? A.l2l : A.m2m;
    const auto& b = w ? B.l2l : B.m2m;
    for (size_t k = 0; k < a.size(); k++) {
      same("trans out", a[k].out, b[k].out, bad);
      same("trans in", a[k].in, b[k].in, bad);
      same("trans edv", a[k].edv, b[k].edv, bad);
    }
  }
  const Stats &s = A.st, &t = B.st;
  const int64_t x[] = {s.p2p_calls, s.p2p_pairs, s.hf_m2l, s.lf_m2l, s.src_slots, s.trg_slots,
                       s.hf_tables, s.lf_tables, s.m2f_slots, s.m2l_entries, s.m2l_targets,
                       s.m2l_groups, s.m2l_group_entries};
  const int64_t y[] = {t.p2p_calls, t.p2p_pairs, t.hf_m2l, t.lf_m2l, t.src_slots, t.trg_slots,
                       t.hf_tables, t.lf_tables, t.m2f_slots, t.m2l_entries, t.m2l_targets,
                       t.m2l_groups, t.m2l_group_entries};
  for (size_t k = 0; k < sizeof(x) / sizeof(x[0]); k++)
    if (x[k] != y[k] && bad.empty()) bad = "stat " + std::to_string(k);
  return bad;
}

// points: 0 sphere, 1 eight Gaussian clusters projected to the sphere and
// stretched to a (1, 1/4, 1/4) ellipsoid, 2 uniform cube
std::vector<double> cloud(int kind, int n, uint64_t seed) {
  std::mt19937_64 g(seed);
  std::normal_distribution<double> nd;
  std::uniform_real_distribution<double> ud(-1., 1.);
  std::vector<double> p(3 * (size_t)n);
  std::vector<std::array<double, 3>> ctr(8);
  for (auto& c : ctr) {
    double v[3], r = 0;
    for (double& x : v) { x = nd(g); r += x * x; }
    for (int k = 0; k < 3; k++) c[k] = v[k] / std::sqrt(r);
  }
  for (int i = 0; i < n; i++) {
    double v[3], r = 0;
    for (int k = 0; k < 3; k++) {
      v[k] = kind == 0 ? nd(g) : kind == 1 ? ctr[i % 8][k] + 0.05 * nd(g) : ud(g);
      r += v[k] * v[k];
    }
    for (int k = 0; k < 3; k++) {
      double x = kind == 2 ? v[k] : v[k] / std::sqrt(r);
      if (kind == 1 && k) x *= 0.25;
      p[3 * (size_t)i + k] = x;
    }
  }
  return p;
}

}  // namespace

int main() {
  struct Case {
    const char* name;
    int ks, ns, kt, nt, L, ncrit;
    double nlambda;  // wavelengths across the unit-box diameter 2
    int lvmax;
  };
  const Case cases[] = {
      {"sphere_32l_L5", 0, 20000, -1, 0, 5, 33, 32, -1},
      {"ellipsoid_64l_L4", 1, 30000, -1, 0, 4, 65, 64, -1},
      {"sphere_16l_L6", 0, 15000, -1, 0, 6, 33, 16, -1},
      {"cube_laplace_L5", 2, 15000, -1, 0, 5, 40, 0, -1},
      {"clusters_to_sphere", 1, 12000, 0, 9000, 5, 33, 16, -1},
      {"cube_fixed_depth4", 2, 8000, -1, 0, 4, 1, 4, 4},
      {"sphere_8l_L8", 0, 12000, -1, 0, 8, 65, 8, -1},
      {"target_root_leaf", 0, 20000, 0, 20, 4, 33, 8, -1},
  };
  int failed = 0;
  for (const Case& c : cases) {
    std::vector<double> src = cloud(c.ks, c.ns, 11), trg;
    if (c.kt >= 0) {
      trg = cloud(c.kt, c.nt, 12);
      for (double& x : trg) x *= 0.8;
    } else {
      trg = src;
    }
    const int nt = (int)(trg.size() / 3);
    const double kappa = 2. * M_PI * c.nlambda / 2.;
    const int nc = c.lvmax >= 0 ? -c.lvmax : c.ncrit;
    Plan host, model;
    std::string m1, m2;
    ModelSetup dev;
    const int r1 = buildPlan(host, src.data(), c.ns, trg.data(), nt, c.L, nc, kappa, m1);
    const int r2 = buildPlan(model, src.data(), c.ns, trg.data(), nt, c.L, nc, kappa, m2, &dev);
    if (r1 != OK || r2 != OK) {
      std::printf("case %s FAIL build host %d (%s) model %d (%s)\n", c.name, r1, m1.c_str(), r2,
                  m2.c_str());
      failed++;
      continue;
    }
    const std::string bad = comparePlans(host, model);
    if (!bad.empty() || model.st.setup_device != 1) {
      std::printf("case %s FAIL %s\n", c.name, bad.empty() ? "setup_device" : bad.c_str());
      failed++;
    } else {
      std::printf("case %s OK %zu entries %lld hf %lld lf %zu groups\n", c.name, host.m2lSrcF.size(),
                  (long long)host.st.hf_m2l, (long long)host.st.lf_m2l,
                  host.m2lGroupBeg.size() - 1);
    }
  }
  return failed ? 1 : 0;
}
