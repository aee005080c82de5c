// ============================================================================
//  oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
//  A plain FP64 CPU restatement of defmm's directional equispaced-interpolation
//  FMM (3D Helmholtz, exp(i*kappa*r)/(4*pi*r)), written as recursive tree walks
//  in the same shape as the reference so that every discrete choice (octree,
//  frequency regimes, direction cones, admissibility, slot numbering) follows
//  the reference step by step.  It is the checker for the GPU product and the
//  CPU baseline ("port") in bench.py; nothing in the product links it.
//
//  Reference files followed (paths relative to the reference root):
//    octree ............ include/structs/sort.hpp:98-228, tree.hpp:72-102,
//                        particles.hpp:79-100
//    regimes ........... include/tables/frequency.hpp:43-184,219-240
//    directions ........ include/tables/newdirections.hpp:75-185
//    symmetries ........ include/symmetries/hypercube.hpp:108-136,
//                        permutations.hpp:87-150,208-232
//    M2L tables ........ include/calculus/circulant.hpp:46-144,
//                        include/tables/precompute.hpp:37-76,
//                        include/operators/M2L.hpp:88-200
//    slots ............. include/structs/initdevs.hpp:31-53,187-245
//    leaf matrices ..... include/tables/mapOnTrees.hpp:36-134
//    phase tables ...... include/calculus/planewave.hpp:43-101
//    sweeps ............ include/calculus/interpolation.hpp:183-235,244-513
//    operators ......... include/operators/{P2P,P2M,M2M,M2F,M2L,F2L,L2L,L2P}.hpp
//    passes ............ include/traversals/{upw,m2f,hrz,f2l,dwn}pass.hpp
//    gemv .............. include/interface.hpp:153-170
//    direct check ...... include/interface.hpp:178-221
//
//  The DFT is written out here (separable, unnormalised, FFTW sign
//  convention: forward exp(-2*pi*i*jk/P), backward exp(+2*pi*i*jk/P)), since
//  FFTW is not part of this image.  Pinning: see tests/test_oracle_*.py.
// ============================================================================
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

typedef std::complex<double> cplx;
static const cplx I1(0.0, 1.0);
static const double PI4 = 4.0 * M_PI;

// ---------------------------------------------------------------- keys ----
// thetis/thetis_keys.hpp:35-53 : last index fastest.
static inline int I2i(const int I[3], int L) { return (I[0] * L + I[1]) * L + I[2]; }
static inline void i2I(int i, int L, int I[3]) {
  I[2] = i % L; i /= L;
  I[1] = i % L; i /= L;
  I[0] = i;
}

// ------------------------------------------------------ Lagrange basis ----
// thetis/thetis_polynomials.hpp:81-94 and :124-127.
static inline double point1D(int i, int L, double alpha) {
  return alpha * (-1. + 2. * ((double)i / (double)(L - 1)));
}
static inline double S1D(int i, double x, int P, double alpha) {
  double res = 1.;
  double nx = x / alpha;
  double pi = -1. + 2. * ((double)i / (double)(P - 1));
  for (int ii = 0; ii < i; ii++) {
    double pii = -1. + 2. * ((double)ii / (double)(P - 1));
    res *= (nx - pii) / (pi - pii);
  }
  for (int ii = i + 1; ii < P; ii++) {
    double pii = -1. + 2. * ((double)ii / (double)(P - 1));
    res *= (nx - pii) / (pi - pii);
  }
  return res;
}

// Norms and normalisations.  Built normally, the oracle uses strict IEEE
// double arithmetic (no contraction): norm = sqrt((x*x + y*y) + z*z), and
// v /= norm.  Built with ORACLE_REFERENCE_ROUNDING and the reference's own
// compiler flags (makefile:4: g++ -O3 -ffast-math -mfma ...), the same loops
// are left to that compiler's contraction and reassociation, which is how the
// reference's norm2 (gaia_small_vectors.hpp:66) and in-place normalisation
// (newdirections.hpp:116, M2L.hpp:192) round.  g_var selects alternative
// explicit orders for rounding experiments (0 = the default above).
static int g_var[3] = {0, 0, 0};
static inline double norm3(const double v[3]) {
  double t = 0.;
  for (int k = 0; k < 3; k++) t += v[k] * v[k];
  return std::sqrt(t);
}
static inline void normalize3(double v[3]) {
  double n = norm3(v);
  for (int k = 0; k < 3; k++) v[k] /= n;
}
#ifdef ORACLE_REFERENCE_ROUNDING
static inline double norm3v(const double v[3], int) { return norm3(v); }
static inline void normalize3v(double v[3], int) { normalize3(v); }
#else
static inline double norm3v(const double v[3], int var) {
  switch (var & 7) {
    case 1: return std::sqrt(std::fma(v[2], v[2], std::fma(v[0], v[0], v[1] * v[1])));
    case 2: return std::sqrt(std::fma(v[1], v[1], std::fma(v[0], v[0], v[2] * v[2])));
    case 3: return std::sqrt(std::fma(v[2], v[2], std::fma(v[1], v[1], v[0] * v[0])));
    case 4: return std::sqrt(std::fma(v[0], v[0], std::fma(v[1], v[1], v[2] * v[2])));
    case 5: return std::sqrt(std::fma(v[0], v[0], std::fma(v[2], v[2], v[1] * v[1])));
    case 6: return std::sqrt(std::fma(v[1], v[1], std::fma(v[2], v[2], v[0] * v[0])));
    default: return norm3(v);
  }
}
static inline void normalize3v(double v[3], int var) {
  double n = norm3v(v, var);
  if (var & 8) {
    double r = 1. / n;
    for (int k = 0; k < 3; k++) v[k] *= r;
  } else {
    for (int k = 0; k < 3; k++) v[k] /= n;
  }
}
#endif

// -------------------------------------------------------------- types ----
struct Particle {
  double pos[3];
  int ind;
};

struct Cell {
  int idx[3];
  int lvl;
  int nprt;
  double ctr[3];
  double rad;
  std::vector<int> son;
  int pbeg;                 // first particle (tree order)
  std::vector<int> effDir;  // 0/1 marks, then 1-based slot numbers
  int ndir = 0;
  std::vector<int> dirs;
  std::vector<cplx> M, F, L, J;  // ndir (or 1) slots of LL / PP
  bool hasDevs = false;
};

struct Tree {
  std::vector<Cell> cells;
  std::vector<Particle> parts;
  double ctr[3];
  double rad;
  std::vector<int> nCellPerStage;
};

struct Stats {
  long long v[32];
};
enum {
  ST_P2P_CALLS = 0, ST_P2P_PAIRS, ST_HFM2L, ST_LFM2L, ST_SRC_SLOTS, ST_TRG_SLOTS,
  ST_HF_TABLES, ST_LF_TABLES, ST_PHFP2M, ST_PLFP2M, ST_HFM2MBLOCK, ST_LF2HFM2M,
  ST_LFM2M, ST_PHFL2P, ST_PLFL2P, ST_HFL2LBLOCK, ST_HF2LFL2L, ST_LFL2L, ST_NLVL,
  ST_CPT, ST_MAXSTAGE, ST_NCELLS_T, ST_NCELLS_S, ST_SRC_CELLS_WITH_SLOTS,
  ST_TRG_CELLS_WITH_SLOTS
};

// ---------------------------------------------------------- the solver ----
struct Oracle {
  // parameters (global.hpp:93-102)
  int L, P, LL, PP, Ncrit;
  double kappa;
  Tree T, S;
  std::vector<cplx> q, p;  // tree order
  // frequency (frequency.hpp)
  int nlvl = 0, cpt = 0, maxStage = 0;
  bool noHF = true;
  std::vector<int> isHFStage, HFStage, nDirPerStage, nCellPerDirMax, nCellPerDirMin;
  std::vector<char> precomp;
  // directions (newdirections.hpp): dirs[e][i][k]
  std::vector<std::vector<std::array<double, 3>>> dtab;
  // symmetries
  int sigma[8][3];
  std::vector<std::vector<int>> perm;  // 64 x PP
  // circulant
  std::vector<int> itilda;
  std::vector<double> tabZ;
  std::vector<cplx> twf, twb;  // P x P twiddles
  // M2L tables: keyed by (level, fundamental triple)
  std::map<std::tuple<int, int, int, int>, std::vector<cplx>> m2lTab;
  std::map<std::tuple<int, int, int, int, int>, int> dirTable;  // (lvl, triple, hsh) -> cone
  // phase tables per level: Sig[l][dir][I], Sig2 = Sig^2
  std::vector<std::vector<std::vector<cplx>>> Sig, Sig2;
  // M2M / L2L 1D matrices (interpolation.hpp:183-235): A[s][k][p]
  std::vector<double> Am[2], Bl[2];
  Stats st;
  int mode = 0;  // 0: full, 1: near field only, 2: far field only

  // ------------------------------------------------------------ tree ----
  // particles.hpp:79-100
  void getBounds(std::vector<Particle>& tp, std::vector<Particle>& sp) {
    double mn[3], mx[3];
    for (int k = 0; k < 3; k++) mn[k] = mx[k] = tp[0].pos[k];
    for (size_t i = 1; i < tp.size(); i++)
      for (int k = 0; k < 3; k++) {
        mn[k] = tp[i].pos[k] < mn[k] ? tp[i].pos[k] : mn[k];
        mx[k] = tp[i].pos[k] > mx[k] ? tp[i].pos[k] : mx[k];
      }
    for (size_t i = 0; i < sp.size(); i++)
      for (int k = 0; k < 3; k++) {
        mn[k] = sp[i].pos[k] < mn[k] ? sp[i].pos[k] : mn[k];
        mx[k] = sp[i].pos[k] > mx[k] ? sp[i].pos[k] : mx[k];
      }
    double c[3], a = 0, b = 0;
    for (int k = 0; k < 3; k++) c[k] = (mn[k] + mx[k]) / 2;
    double d1[3], d2[3];
    for (int k = 0; k < 3; k++) { d1[k] = c[k] - mn[k]; d2[k] = mx[k] - c[k]; }
    a = d1[0]; for (int k = 1; k < 3; k++) a = a > d1[k] ? a : d1[k];
    b = d2[0]; for (int k = 1; k < 3; k++) b = b > d2[k] ? b : d2[k];
    double r = std::max(a, b);
    for (int k = 0; k < 3; k++) { T.ctr[k] = c[k]; S.ctr[k] = c[k]; }
    T.rad = r; S.rad = r;
  }

  // sort.hpp:98-162 (stable octant partition) and :187-220 (pre-order cells),
  // done in one recursion: the particle order and the cell numbering are the
  // same as the two-pass reference build.
  void build(Tree& tr, int b, int n, const double ctr[3], double rad, const int idx[3], int lvl,
             std::vector<Particle>& buf) {
    int me = (int)tr.cells.size();
    tr.cells.emplace_back();
    {
      Cell& c = tr.cells[me];
      for (int k = 0; k < 3; k++) { c.idx[k] = idx[k]; c.ctr[k] = ctr[k]; }
      c.lvl = lvl; c.nprt = n; c.rad = rad; c.pbeg = b;
    }
    if ((int)tr.nCellPerStage.size() <= lvl) tr.nCellPerStage.push_back(0);
    tr.nCellPerStage[lvl]++;
    if (n < Ncrit) return;  // sort.hpp:49
    int histo[8] = {0}, off[8], add[8] = {0};
    std::vector<int> code(n);
    for (int i = 0; i < n; i++) {
      int bc = 0;
      for (int k = 0; k < 3; k++) bc += ((tr.parts[b + i].pos[k] > ctr[k]) ? 1 : 0) << k;
      code[i] = bc;
      histo[bc]++;
    }
    off[0] = 0;
    for (int i = 1; i < 8; i++) off[i] = off[i - 1] + histo[i - 1];
    for (int i = 0; i < n; i++) buf[b + off[code[i]] + (add[code[i]]++)] = tr.parts[b + i];
    for (int i = 0; i < n; i++) tr.parts[b + i] = buf[b + i];
    double nr = rad / 2.;
    for (int i = 0; i < 8; i++) {
      if (!histo[i]) continue;
      double nc[3];
      int ni[3];
      for (int k = 0; k < 3; k++) {
        nc[k] = ctr[k] + nr * (((i & (1 << k)) >> k) * 2 - 1);
        ni[k] = 2 * idx[k] + ((i & (1 << k)) ? 1 : 0);
      }
      int sonId = (int)tr.cells.size();
      tr.cells[me].son.push_back(sonId);
      build(tr, b + off[i], histo[i], nc, nr, ni, lvl + 1, buf);
    }
  }

  void buildTree(Tree& tr, const double* xyz, int n) {
    tr.parts.resize(n);
    for (int i = 0; i < n; i++) {
      for (int k = 0; k < 3; k++) tr.parts[i].pos[k] = xyz[3 * i + k];
      tr.parts[i].ind = i;
    }
    std::vector<Particle> buf(n);
    int idx0[3] = {0, 0, 0};
    tr.cells.clear();
    tr.cells.reserve(2 * n / std::max(1, Ncrit / 4) + 16);
    build(tr, 0, n, tr.ctr, tr.rad, idx0, 0, buf);
  }

  // ------------------------------------------------------- directions ----
  // newdirections.hpp:75-171
  void getDirections(int nStages) {
    dtab.assign(nStages, {});
    for (int i = 0; i < nStages; i++) dtab[i].resize((size_t)(1 << (2 * i)) * 6);
    for (int k = 0; k < 3; k++) {
      dtab[0][2 * k] = {0., 0., 0.};
      dtab[0][2 * k][k] = 1.;
      dtab[0][2 * k + 1] = {0., 0., 0.};
      dtab[0][2 * k + 1][k] = -1.;
    }
    double hyp[4][3];
    for (int k = 0; k < 4; k++) {
      hyp[k][0] = 0.; hyp[k][1] = 1.; hyp[k][2] = 1.;
      for (int q = 0; q < 2; q++)
        if (k & (1 << q)) hyp[k][1 + q] = -1.;
    }
    int npe = 1;
    for (int e = 1; e < nStages; e++) {
      double radius = 1. / double(1 << (e + 1));
      for (int i = 0; i < npe; i++)
        for (int q = 0; q < 4; q++) {
          auto d = dtab[e - 1][i];
          for (int k = 0; k < 3; k++) d[k] += hyp[q][k] * radius;
          normalize3v(d.data(), g_var[0]);
          dtab[e][i * 4 + q] = d;
        }
      npe *= 4;
    }
    npe = 1;
    for (int e = 1; e < nStages; e++) {
      npe *= 4;
      for (int i = 0; i < npe; i++) {
        auto d = dtab[e][i];
        d[0] *= -1.;
        dtab[e][npe + i] = d;
      }
    }
    npe = 1;
    for (int e = 1; e < nStages; e++) {
      npe *= 4;
      int fs2 = 2 * npe;
      for (int j = 1; j < 3; j++)
        for (int ii = 0; ii < fs2; ii++)
          for (int k = 0; k < 3; k++) dtab[e][j * fs2 + ii][k] = dtab[e][ii][(k + j) % 3];
    }
  }

  // newdirections.hpp:174-185.  The reference keeps the first index among
  // equal distances (strict '>').  Distances of mirror-image cones to a
  // symmetric offset are equal in exact arithmetic but may round either way,
  // so a distance counts as smaller only below the best one by more than
  // 1e-12 (distances lie in [0, 2]).
  int bestCone(int stage, const double dir[3]) {
    double tmp = 1.e30;
    int ind = -1;
    for (size_t i = 0; i < dtab[stage].size(); i++) {
      double d[3];
      for (int k = 0; k < 3; k++) d[k] = dtab[stage][i][k] - dir[k];
      double nn = norm3v(d, g_var[1]);
#ifdef ORACLE_REFERENCE_ROUNDING
      if (tmp > nn) { tmp = nn; ind = (int)i; }
#else
      if (nn < tmp - 1.e-12) { tmp = nn; ind = (int)i; }
#endif
    }
    return ind;
  }

  // -------------------------------------------------------- frequency ----
  void initFrequency() {
    nlvl = std::max((int)T.nCellPerStage.size(), (int)S.nCellPerStage.size());
    isHFStage.assign(nlvl, 0); HFStage.assign(nlvl, 0); nDirPerStage.assign(nlvl, 1);
    nCellPerDirMax.assign(nlvl, -1); nCellPerDirMin.assign(nlvl, 2); precomp.assign(nlvl, 0);
    // frequency.hpp:43-82 (eta = 1, :221)
    double eta = 1.;
    double sd = 2. * T.rad;
    cpt = 0;
    if (sd * kappa > eta) while (sd * kappa > eta) { sd /= 2.; cpt++; }
    if (cpt) getDirections(cpt);
    noHF = !(cpt > 1);
    int hf = 1;
    for (int l = cpt; l > -1; l--) {
      if (l < nlvl) { isHFStage[l] = hf; HFStage[l] = hf - 1; }
      hf++;
    }
    for (int l = nlvl - 1; l > cpt; l--) { isHFStage[l] = 0; HFStage[l] = 0; }
    // :89-98
    for (int l = nlvl - 1; l > -1; l--)
      nDirPerStage[l] = (isHFStage[l] && l != 0) ? (int)dtab[HFStage[l]].size() : 1;
    // :108-129
    int ms = 0;
    double BD = 2. * T.rad;
    while (1. < kappa * BD / double(1 << ms) / double(1 << ms)) {
      ms++;
      if (ms >= nlvl) { fprintf(stderr, "oracle: no precomputation level\n"); throw 1; }
    }
    ms = std::max(ms, 2);
    maxStage = ms;
    for (int l = 0; l < nlvl; l++) precomp[l] = (l >= ms);
    // :158-184
    for (int l = 0; l < nlvl; l++) {
      if (precomp[l]) {
        if (isHFStage[l]) nCellPerDirMax[l] = 1 << l;
        else nCellPerDirMax[l] = isHFStage[l - 1] ? 4 * std::max(2, int(eta)) : 4;
      } else nCellPerDirMax[l] = -1;
    }
  }

  // ------------------------------------------------------- symmetries ----
  static int hashSigma(const int I[3]) {
    int h = 0, u = 0;
    for (int k = 0; k < 3; k++)
      for (int l = k + 1; l < 3; l++) h += ((std::abs(I[k]) < std::abs(I[l])) ? 1 : 0) << u++;
    return h;
  }
  static int hashZ(const int I[3]) {
    int h = 0;
    for (int k = 0; k < 3; k++) h += ((I[k] < 0) ? 1 : 0) << k;
    return h;
  }
  void initSymmetries() {
    static const int sg[8][3] = {{0, 1, 2}, {1, 0, 2}, {0, 1, 2}, {1, 2, 0},
                                 {0, 2, 1}, {0, 1, 2}, {2, 0, 1}, {2, 1, 0}};
    memcpy(sigma, sg, sizeof(sg));
    perm.assign(64, std::vector<int>(PP));
    for (int p0 = 0; p0 < 8; p0++)
      for (int p1 = 0; p1 < 8; p1++) {
        int h = p0 + (p1 << 3);
        for (int pp = 0; pp < PP; pp++) {
          int tp[3], tm[3];
          i2I(pp, P, tp);
          for (int k = 0; k < 3; k++) {
            tm[k] = tp[sigma[p1][k]];
            if (p0 & (1 << sigma[p1][k]))
              if (tm[k]) tm[k] = P - tm[k];
          }
          perm[h][pp] = I2i(tm, P);
        }
      }
  }
  // permutations.hpp:208-232
  void fundamental(const int X[3], int R[3], int& hsh) {
    int h0 = hashZ(X), h1 = hashSigma(X);
    hsh = h0 + (h1 << 3);
    for (int k = 0; k < 3; k++) R[k] = std::abs(X[sigma[h1][k]]);
  }

  // -------------------------------------------------------- circulant ----
  void initCirculant() {
    itilda.assign(LL, 0);
    for (int i = 0; i < LL; i++) {
      int I[3];
      i2I(i, L, I);
      itilda[i] = I2i(I, P);
    }
    tabZ.assign(3 * PP, 0.);
    for (int i = 0; i < PP; i++) {
      int E[3];
      i2I(i, P, E);
      for (int k = 0; k < 3; k++) {
        E[k] = (E[k] < L) ? E[k] : -(2 * L - 1 - E[k]);
        int Ip = E[k] > -1 ? E[k] : 0;
        int Jn = E[k] > -1 ? 0 : -E[k];
        tabZ[3 * i + k] = point1D(Ip, L, 1.) - point1D(Jn, L, 1.);
      }
    }
    twf.resize(P * P); twb.resize(P * P);
    for (int j = 0; j < P; j++)
      for (int k = 0; k < P; k++) {
        int m = (j * k) % P;
        double a = 2. * M_PI * (double)m / (double)P;
        twf[j * P + k] = cplx(std::cos(a), -std::sin(a));
        twb[j * P + k] = cplx(std::cos(a), std::sin(a));
      }
  }
  // unnormalised separable 3D DFT of a P^3 row-major array
  void dft3(const cplx* in, cplx* out, bool fwd) {
    const std::vector<cplx>& tw = fwd ? twf : twb;
    std::vector<cplx> a(in, in + PP), b(PP);
    for (int ax = 2; ax >= 0; ax--) {
      int stride = (ax == 2) ? 1 : (ax == 1 ? P : P * P);
      for (int i = 0; i < PP; i++) {
        int I[3];
        i2I(i, P, I);
        int k = I[ax];
        int base = i - k * stride;
        cplx s = 0.;
        for (int j = 0; j < P; j++) s += a[base + j * stride] * tw[j * P + k];
        b[i] = s;
      }
      a.swap(b);
    }
    std::copy(a.begin(), a.end(), out);
  }
  // circulant.hpp:133-144
  std::vector<cplx> fillCircCoeffs(double rad, const double v[3]) {
    std::vector<cplx> in(PP), out(PP);
    for (int i = 0; i < PP; i++) {
      double Z[3], w[3];
      for (int k = 0; k < 3; k++) { Z[k] = tabZ[3 * i + k] * rad; w[k] = v[k] + Z[k]; }
      double R = norm3(w);
      in[i] = std::exp(I1 * R * kappa) / (PI4 * R);
    }
    dft3(in.data(), out.data(), true);
    return out;
  }
  // precompute.hpp:37-76
  void precomputeLFTables() {
    for (int l = 2; l < nlvl; l++) {
      if (!(precomp[l] && !isHFStage[l])) continue;
      double locrad = T.rad / double(1 << l);
      int mx = nCellPerDirMax[l];
      for (int a = 0; a < mx; a++)
        for (int b = 0; b < mx; b++)
          for (int c = 0; c < mx; c++) {
            int iV[3] = {a, b, c};
            bool ok = false;
            for (int k = 0; k < 3; k++) ok = ok || (iV[k] >= nCellPerDirMin[l]);
            for (int k = 0; k < 2; k++) ok = ok && (iV[k] >= iV[k + 1]);
            if (!ok) continue;
            double fV[3];
            for (int k = 0; k < 3; k++) fV[k] = 2. * double(iV[k]) * locrad;
            m2lTab[std::make_tuple(l, a, b, c)] = fillCircCoeffs(locrad, fV);
            st.v[ST_LF_TABLES]++;
          }
    }
  }

  // -------------------------------------------------------- planewave ----
  // planewave.hpp:43-69 (tables stored per level instead of per HF index)
  void computeSignatures() {
    Sig.assign(nlvl, {}); Sig2.assign(nlvl, {});
    for (int l = 2; l < nlvl; l++) {
      if (!(precomp[l] && isHFStage[l])) continue;
      double locrad = T.rad / double(1 << l);
      Sig[l].resize(nDirPerStage[l]); Sig2[l].resize(nDirPerStage[l]);
      for (int i = 0; i < nDirPerStage[l]; i++) {
        const double* d = dtab[HFStage[l]][i].data();
        Sig[l][i].resize(LL); Sig2[l][i].resize(LL);
        for (int k = 0; k < LL; k++) {
          int I[3];
          i2I(k, L, I);
          double x[3], dp = 0.;
          for (int kk = 0; kk < 3; kk++) x[kk] = point1D(I[kk], L, .5 * locrad);
          for (int kk = 0; kk < 3; kk++) dp += d[kk] * x[kk];
          Sig[l][i][k] = std::exp(-I1 * kappa * dp);
          Sig2[l][i][k] = Sig[l][i][k] * Sig[l][i][k];
        }
      }
    }
  }

  // ---------------------------------------------------- interpolation ----
  // interpolation.hpp:183-235.  Am[s][k*L+p] = S_k((xi_p -/+ 1)/2): parent k,
  // child p.  Bl[s][p*L+k] = S_k((xi_p -/+ 1)/2): child p, parent k.
  void initInterp() {
    for (int s = 0; s < 2; s++) {
      Am[s].assign(L * L, 0.); Bl[s].assign(L * L, 0.);
      double sh = s ? 1. : -1.;
      for (int p = 0; p < L; p++)
        for (int k = 0; k < L; k++) {
          double v = S1D(k, point1D(p, L, 1.) + sh, L, 2.);
          Am[s][k * L + p] = v;
          Bl[s][p * L + k] = v;
        }
    }
  }
  // out[I] = sum_J prod_k A^{s_k}[I_k, J_k] in[J], s_k = (v_k < 0 ? 0 : 1);
  // A indexed [out*L+in].
  void sweep3(const std::vector<double>* A, const double v[3], const cplx* in, cplx* out) {
    std::vector<cplx> a(in, in + LL), b(LL);
    for (int ax = 2; ax >= 0; ax--) {
      const std::vector<double>& M = A[v[ax] < 0. ? 0 : 1];
      int stride = (ax == 2) ? 1 : (ax == 1 ? L : L * L);
      for (int i = 0; i < LL; i++) {
        int I[3];
        i2I(i, L, I);
        int k = I[ax];
        int base = i - k * stride;
        cplx s = 0.;
        for (int j = 0; j < L; j++) s += M[k * L + j] * a[base + j * stride];
        b[i] = s;
      }
      a.swap(b);
    }
    std::copy(a.begin(), a.end(), out);
  }

  // ------------------------------------------------------- setup walks ----
  // hrzpass.hpp:51-56
  bool MACgeomhf(const Cell& X, const Cell& Y) {
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = X.ctr[k] - Y.ctr[k];
    double tmp = (1. - 1.e-8) * (norm3(d) - 2. * Y.rad);
    double tt = 2. * Y.rad;
    return tmp >= std::max(tt * tt * kappa, tt);
  }
  // hrzpass.hpp:39-49
  static bool MAClist(const Cell& X, const Cell& Y) {
    int test = 0, ncmp = 3;
    for (int k = 0; k < 3; k++) {
      int d = X.idx[k] - Y.idx[k];
      ncmp -= (d ? 0 : 1);
      test += std::abs(d);
    }
    return test > ncmp;
  }
  // M2L.hpp:163-200
  void blankM2Lhf(Cell& t, Cell& s) {
    int iZ[3], nI[3], hsh;
    for (int k = 0; k < 3; k++) iZ[k] = t.idx[k] - s.idx[k];
    fundamental(iZ, nI, hsh);
    auto key = std::make_tuple(t.lvl, nI[0], nI[1], nI[2]);
    if (!m2lTab.count(key)) {
      double fV[3];
      for (int k = 0; k < 3; k++) fV[k] = 2. * double(nI[k]) * t.rad;
      m2lTab[key] = fillCircCoeffs(t.rad, fV);
      st.v[ST_HF_TABLES]++;
    }
    auto dkey = std::make_tuple(t.lvl, nI[0], nI[1], nI[2], hsh);
    auto it = dirTable.find(dkey);
    int cone;
    if (it == dirTable.end()) {
      double fZ[3];
      for (int k = 0; k < 3; k++) fZ[k] = t.ctr[k] - s.ctr[k];
      normalize3v(fZ, g_var[2]);
      cone = bestCone(HFStage[t.lvl], fZ);
      dirTable[dkey] = cone;
    } else cone = it->second;
    t.effDir[cone] = 1;
    s.effDir[cone] = 1;
  }
  // hrzpass.hpp:111-138
  void blankHorizontal(int i, int j) {
    Cell& X = T.cells[i];
    Cell& Y = S.cells[j];
    if ((!isHFStage[X.lvl] && !isHFStage[Y.lvl]) || (!isHFStage[X.lvl] && Y.son.empty()) ||
        (X.son.empty() && !isHFStage[Y.lvl]))
      return;
    if (X.lvl == Y.lvl && MACgeomhf(X, Y) && precomp[X.lvl]) {
      blankM2Lhf(X, Y);
      return;
    }
    if (!X.son.empty() && !Y.son.empty())
      for (int a : T.cells[i].son)
        for (int b : S.cells[j].son) blankHorizontal(a, b);
  }
  // dwnpass.hpp:80-104
  void blankDownward(Tree& tr, int c) {
    Cell& C = tr.cells[c];
    if (C.son.empty()) return;
    int ns = 4;  // directions.nsons[*] (newdirections.hpp:168-170)
    if (isHFStage[C.lvl] && precomp[C.lvl]) {
      for (int sId : C.son) {
        if (isHFStage[C.lvl + 1]) {
          Cell& Sn = tr.cells[sId];
          for (int dt = 0; dt < nDirPerStage[C.lvl + 1]; dt++)
            for (int k = 0; k < ns; k++)
              if (tr.cells[c].effDir[ns * dt + k]) Sn.effDir[dt] = 1;
          blankDownward(tr, sId);
        }
      }
    } else {
      for (int sId : C.son) blankDownward(tr, sId);
    }
  }
  // initdevs.hpp:187-245
  void initDevs(Tree& tr, bool src) {
    for (auto& c : tr.cells) {
      if (!precomp[c.lvl]) continue;
      c.hasDevs = true;
      if (isHFStage[c.lvl]) {
        int cnt = 0;
        for (int k = 0; k < nDirPerStage[c.lvl]; k++)
          if (c.effDir[k]) c.effDir[k] = ++cnt;
        c.ndir = cnt;
        c.dirs.clear();
        for (int u = 0; u < nDirPerStage[c.lvl]; u++)
          if (c.effDir[u]) c.dirs.push_back(u);
      } else {
        c.ndir = 1;
      }
      if (src) {
        c.M.assign((size_t)c.ndir * LL, 0.);
        c.F.assign((size_t)c.ndir * PP, 0.);
        st.v[ST_SRC_SLOTS] += c.ndir;
        if (c.ndir) st.v[ST_SRC_CELLS_WITH_SLOTS]++;
      } else {
        c.L.assign((size_t)c.ndir * LL, 0.);
        c.J.assign((size_t)c.ndir * PP, 0.);
        st.v[ST_TRG_SLOTS] += c.ndir;
        if (c.ndir) st.v[ST_TRG_CELLS_WITH_SLOTS]++;
      }
    }
  }

  // ---------------------------------------------------------- prcmpt ----
  // interface.hpp:100-151
  void prcmpt(const double* src, int ns, const double* trg, int nt) {
    memset(&st, 0, sizeof(st));
    P = 2 * L - 1; LL = L * L * L; PP = P * P * P;
    std::vector<Particle> tp(nt), sp(ns);
    for (int i = 0; i < nt; i++) for (int k = 0; k < 3; k++) tp[i].pos[k] = trg[3 * i + k];
    for (int i = 0; i < ns; i++) for (int k = 0; k < 3; k++) sp[i].pos[k] = src[3 * i + k];
    getBounds(tp, sp);
    buildTree(T, trg, nt);
    buildTree(S, src, ns);
    initCirculant();
    initFrequency();
    computeSignatures();
    initSymmetries();
    precomputeLFTables();
    initInterp();
    for (auto* tr : {&S, &T})
      for (auto& c : tr->cells)
        if (precomp[c.lvl] && isHFStage[c.lvl]) c.effDir.assign(nDirPerStage[c.lvl], 0);
    blankHorizontal(0, 0);
    if (!noHF) {
      blankDownward(S, 0);
      blankDownward(T, 0);
    }
    initDevs(S, true);
    initDevs(T, false);
    st.v[ST_NLVL] = nlvl;
    st.v[ST_CPT] = cpt;
    st.v[ST_MAXSTAGE] = maxStage;
    st.v[ST_NCELLS_T] = (long long)T.cells.size();
    st.v[ST_NCELLS_S] = (long long)S.cells.size();
  }

  // ------------------------------------------------------- operators ----
  // Lagrange weights of particle positions inside a cell: Z[k][l][j]
  void leafS(const Tree& tr, const Cell& c, std::vector<double>& Z) {
    Z.assign((size_t)3 * L * c.nprt, 0.);
    for (int k = 0; k < 3; k++)
      for (int l = 0; l < L; l++)
        for (int j = 0; j < c.nprt; j++)
          Z[((size_t)k * L + l) * c.nprt + j] =
              S1D(l, tr.parts[c.pbeg + j].pos[k] - c.ctr[k], L, c.rad);
  }
  // mapOnTrees.hpp:57-67 (Zmat) + P2M.hpp:113-136
  void P2M(Cell& c) {
    std::vector<double> Z;
    leafS(S, c, Z);
    int n = c.nprt;
    std::vector<double> Zm((size_t)LL * n);
    for (int l = 0; l < LL; l++) {
      int I[3];
      i2I(l, L, I);
      for (int j = 0; j < n; j++) {
        double v = 1.;
        for (int k = 0; k < 3; k++) v *= Z[((size_t)k * L + I[k]) * n + j];
        Zm[(size_t)l * n + j] = v;
      }
    }
    const cplx* qq = q.data() + c.pbeg;
    if (isHFStage[c.lvl]) {
      st.v[ST_PHFP2M]++;
      for (int k = 0; k < c.ndir; k++) {
        const double* d = dtab[HFStage[c.lvl]][c.dirs[k]].data();
        std::vector<cplx> bq(n);
        for (int j = 0; j < n; j++) {
          double diff[3], dp = 0.;
          for (int kk = 0; kk < 3; kk++) diff[kk] = S.parts[c.pbeg + j].pos[kk] - c.ctr[kk];
          for (int kk = 0; kk < 3; kk++) dp += d[kk] * diff[kk];
          bq[j] = std::exp(-I1 * kappa * dp) * qq[j];
        }
        const std::vector<cplx>& e2 = Sig2[c.lvl][c.dirs[k]];
        for (int l = 0; l < LL; l++) {
          cplx s = 0.;
          for (int j = 0; j < n; j++) s += Zm[(size_t)l * n + j] * bq[j];
          c.M[(size_t)k * LL + l] = std::conj(e2[l]) * s;
        }
      }
    } else {
      st.v[ST_PLFP2M]++;
      for (int l = 0; l < LL; l++) {
        cplx s = 0.;
        for (int j = 0; j < n; j++) s += Zm[(size_t)l * n + j] * qq[j];
        c.M[l] = s;
      }
    }
  }
  // M2M.hpp:34-41
  void lfM2M(Cell& t, Cell& s) {
    st.v[ST_LFM2M]++;
    double v[3];
    for (int k = 0; k < 3; k++) v[k] = s.ctr[k] - t.ctr[k];
    std::vector<cplx> b(LL);
    sweep3(Am, v, s.M.data(), b.data());
    for (int i = 0; i < LL; i++) t.M[i] += b[i];
  }
  // M2M.hpp:75-105
  void hfM2Mblock(Cell& t, Cell& s) {
    st.v[ST_HFM2MBLOCK]++;
    double v[3];
    for (int k = 0; k < 3; k++) v[k] = s.ctr[k] - t.ctr[k];
    int ns = 4;
    std::vector<cplx> a(LL), b(LL);
    for (int k = 0; k < t.ndir; k++) {
      int dir = t.dirs[k];
      const cplx* sm = s.M.data() + (size_t)(s.effDir[dir / ns] - 1) * LL;
      const std::vector<cplx>& e = Sig[t.lvl][dir];
      for (int i = 0; i < LL; i++) a[i] = e[i] * sm[i];
      sweep3(Am, v, a.data(), b.data());
      const std::vector<cplx>& e2 = Sig2[t.lvl][dir];
      const double* d = dtab[HFStage[t.lvl]][dir].data();
      double dv = 0.;
      for (int kk = 0; kk < 3; kk++) dv += d[kk] * v[kk];
      cplx edv = std::exp(-I1 * kappa * dv);
      cplx* res = t.M.data() + (size_t)(t.effDir[dir] - 1) * LL;
      for (int i = 0; i < LL; i++) res[i] += (std::conj(e2[i]) * b[i]) * edv;
    }
  }
  // M2M.hpp:107-124
  void lf2hfM2M(Cell& t, Cell& s) {
    st.v[ST_LF2HFM2M]++;
    double v[3];
    for (int k = 0; k < 3; k++) v[k] = s.ctr[k] - t.ctr[k];
    std::vector<cplx> a(LL), b(LL);
    for (int dt = 0; dt < nDirPerStage[t.lvl]; dt++) {
      if (!t.effDir[dt]) continue;
      const double* d = dtab[HFStage[t.lvl]][dt].data();
      const std::vector<cplx>& e = Sig[t.lvl][dt];
      double dv = 0.;
      for (int kk = 0; kk < 3; kk++) dv += d[kk] * v[kk];
      cplx edv = std::exp(-I1 * kappa * dv);
      for (int i = 0; i < LL; i++) a[i] = (e[i] * s.M[i]) * edv;
      sweep3(Am, v, a.data(), b.data());
      const std::vector<cplx>& e2 = Sig2[t.lvl][dt];
      cplx* res = t.M.data() + (size_t)(t.effDir[dt] - 1) * LL;
      for (int i = 0; i < LL; i++) res[i] += std::conj(e2[i]) * b[i];
    }
  }
  // upwpass.hpp:36-64
  void upward(int c) {
    Cell& C = S.cells[c];
    if (C.son.empty()) {
      if (precomp[C.lvl]) P2M(C);
      return;
    }
    for (int s : C.son) upward(s);
    if (!precomp[C.lvl]) return;
    for (int s : C.son) {
      Cell& Cc = S.cells[c];
      Cell& Sn = S.cells[s];
      if (isHFStage[Cc.lvl]) {
        if (HFStage[Cc.lvl]) hfM2Mblock(Cc, Sn);
        else lf2hfM2M(Cc, Sn);
      } else lfM2M(Cc, Sn);
    }
  }
  // M2F.hpp:35-46, circulant.hpp:95-106
  void m2f() {
    std::vector<cplx> tmp(PP);
    for (auto& c : S.cells) {
      if (!precomp[c.lvl]) continue;
      for (int k = 0; k < c.ndir; k++) {
        std::fill(tmp.begin(), tmp.end(), cplx(0.));
        for (int i = 0; i < LL; i++) tmp[itilda[i]] = c.M[(size_t)k * LL + i];
        dft3(tmp.data(), c.F.data() + (size_t)k * PP, true);
      }
    }
  }
  // M2L.hpp:114-126
  void hfM2L(Cell& t, Cell& s) {
    st.v[ST_HFM2L]++;
    int iZ[3], nI[3], hsh;
    for (int k = 0; k < 3; k++) iZ[k] = t.idx[k] - s.idx[k];
    fundamental(iZ, nI, hsh);
    const std::vector<cplx>& u = m2lTab.at(std::make_tuple(t.lvl, nI[0], nI[1], nI[2]));
    int cone = dirTable.at(std::make_tuple(t.lvl, nI[0], nI[1], nI[2], hsh));
    const std::vector<int>& pm = perm[hsh];
    const cplx* y = s.F.data() + (size_t)(s.effDir[cone] - 1) * PP;
    cplx* x = t.J.data() + (size_t)(t.effDir[cone] - 1) * PP;
    for (int i = 0; i < PP; i++) x[i] += u[pm[i]] * y[i];
  }
  // M2L.hpp:128-157 (M2L_LF_NO_INDIRECTIONS off)
  void lfM2L(Cell& t, Cell& s) {
    st.v[ST_LFM2L]++;
    int iZ[3], nI[3], hsh;
    for (int k = 0; k < 3; k++) iZ[k] = t.idx[k] - s.idx[k];
    fundamental(iZ, nI, hsh);
    const std::vector<cplx>& u = m2lTab.at(std::make_tuple(t.lvl, nI[0], nI[1], nI[2]));
    const std::vector<int>& pm = perm[hsh];
    for (int i = 0; i < PP; i++) t.J[i] += u[pm[i]] * s.F[i];
  }
  // P2P.hpp:42-79
  void P2P(const Cell& t, const Cell& s) {
    st.v[ST_P2P_CALLS]++;
    st.v[ST_P2P_PAIRS] += (long long)t.nprt * s.nprt;
    if (mode == 2) return;
    double pi4m1 = 1. / PI4;
    for (int i = 0; i < t.nprt; i++) {
      const double* pt = T.parts[t.pbeg + i].pos;
      double pr = 0., pim = 0.;
      for (int j = 0; j < s.nprt; j++) {
        const double* ps = S.parts[s.pbeg + j].pos;
        double qr = q[s.pbeg + j].real(), qi = q[s.pbeg + j].imag();
        double x = pt[0] - ps[0], y = pt[1] - ps[1], z = pt[2] - ps[2];
        double R = x * x + y * y + z * z;
        double K = 1. / std::sqrt(R);
        if (R < 1.e-16) K = 0.;
        double r = K * R;
        r *= kappa;
        K *= pi4m1;
        double co = std::cos(r), si = std::sin(r);
        co *= K;
        si *= K;
        pr += co * qr - si * qi;
        pim += co * qi + si * qr;
      }
      p[t.pbeg + i] += cplx(pr, pim);
    }
  }
  // hrzpass.hpp:61-91
  void horizontal(int i, int j) {
    Cell& X = T.cells[i];
    Cell& Y = S.cells[j];
    if (precomp[X.lvl] && X.lvl == Y.lvl) {
      if (X.lvl && isHFStage[X.lvl]) {
        if (MACgeomhf(X, Y)) {
          if (mode != 1) hfM2L(X, Y); else st.v[ST_HFM2L]++;
          return;
        }
      } else if (MAClist(X, Y)) {
        if (mode != 1) lfM2L(X, Y); else st.v[ST_LFM2L]++;
        return;
      }
    }
    if (X.son.empty() || Y.son.empty()) {
      P2P(X, Y);
      return;
    }
    for (int a : T.cells[i].son)
      for (int b : S.cells[j].son) horizontal(a, b);
  }
  // F2L.hpp:35-46, circulant.hpp:108-118
  void f2l() {
    std::vector<cplx> tmp(PP);
    for (auto& c : T.cells) {
      if (!precomp[c.lvl]) continue;
      for (int k = 0; k < c.ndir; k++) {
        dft3(c.J.data() + (size_t)k * PP, tmp.data(), false);
        for (int i = 0; i < LL; i++) c.L[(size_t)k * LL + i] += tmp[itilda[i]] / double(PP);
      }
    }
  }
  // L2L.hpp:34-41
  void lfL2L(Cell& t, Cell& s) {
    st.v[ST_LFL2L]++;
    double v[3];
    for (int k = 0; k < 3; k++) v[k] = t.ctr[k] - s.ctr[k];
    std::vector<cplx> b(LL);
    sweep3(Bl, v, s.L.data(), b.data());
    for (int i = 0; i < LL; i++) t.L[i] += b[i];
  }
  // L2L.hpp:77-108
  void hfL2Lblock(Cell& t, Cell& s) {
    st.v[ST_HFL2LBLOCK]++;
    double v[3];
    for (int k = 0; k < 3; k++) v[k] = t.ctr[k] - s.ctr[k];
    int ns = 4;
    std::vector<cplx> a(LL), b(LL);
    for (int k = 0; k < s.ndir; k++) {
      int dir = s.dirs[k];
      const std::vector<cplx>& e2 = Sig2[s.lvl][dir];
      const cplx* sl = s.L.data() + (size_t)(s.effDir[dir] - 1) * LL;
      for (int i = 0; i < LL; i++) a[i] = e2[i] * sl[i];
      sweep3(Bl, v, a.data(), b.data());
      const std::vector<cplx>& e = Sig[s.lvl][dir];
      const double* d = dtab[HFStage[s.lvl]][dir].data();
      double dv = 0.;
      for (int kk = 0; kk < 3; kk++) dv += v[kk] * d[kk];
      cplx edv = std::exp(I1 * kappa * dv);
      cplx* res = t.L.data() + (size_t)(t.effDir[dir / ns] - 1) * LL;
      for (int i = 0; i < LL; i++) res[i] += (std::conj(e[i]) * b[i]) * edv;
    }
  }
  // L2L.hpp:110-127
  void hf2lfL2L(Cell& t, Cell& s) {
    st.v[ST_HF2LFL2L]++;
    double v[3];
    for (int k = 0; k < 3; k++) v[k] = t.ctr[k] - s.ctr[k];
    std::vector<cplx> a(LL), b(LL);
    for (int ds = 0; ds < nDirPerStage[s.lvl]; ds++) {
      if (!s.effDir[ds]) continue;
      const double* d = dtab[HFStage[s.lvl]][ds].data();
      const std::vector<cplx>& e2 = Sig2[s.lvl][ds];
      const cplx* sl = s.L.data() + (size_t)(s.effDir[ds] - 1) * LL;
      for (int i = 0; i < LL; i++) a[i] = e2[i] * sl[i];
      sweep3(Bl, v, a.data(), b.data());
      const std::vector<cplx>& e = Sig[s.lvl][ds];
      double dv = 0.;
      for (int kk = 0; kk < 3; kk++) dv += v[kk] * d[kk];
      cplx ev = std::exp(I1 * kappa * dv);
      for (int i = 0; i < LL; i++) t.L[i] += (std::conj(e[i]) * b[i]) * ev;
    }
  }
  // mapOnTrees.hpp:107-118 (Zmat) + L2P.hpp:121-142,164-173
  void L2P(Cell& c) {
    std::vector<double> Z;
    leafS(T, c, Z);
    int n = c.nprt;
    std::vector<double> Zm((size_t)n * LL);
    for (int l = 0; l < LL; l++) {
      int I[3];
      i2I(l, L, I);
      for (int j = 0; j < n; j++) {
        double v = 1.;
        for (int k = 0; k < 3; k++) v *= Z[((size_t)k * L + I[k]) * n + j];
        Zm[(size_t)l * n + j] = v;
      }
    }
    cplx* pp = p.data() + c.pbeg;
    if (isHFStage[c.lvl]) {
      st.v[ST_PHFL2P]++;
      for (int iD = 0; iD < nDirPerStage[c.lvl]; iD++) {
        if (!c.effDir[iD]) continue;
        const std::vector<cplx>& e2 = Sig2[c.lvl][iD];
        const cplx* Ls = c.L.data() + (size_t)(c.effDir[iD] - 1) * LL;
        std::vector<cplx> buf(LL);
        for (int i = 0; i < LL; i++) buf[i] = e2[i] * Ls[i];
        const double* d = dtab[HFStage[c.lvl]][iD].data();
        for (int i = 0; i < n; i++) {
          cplx s = 0.;
          for (int l = 0; l < LL; l++) s += Zm[(size_t)l * n + i] * buf[l];
          double diff[3], dp = 0.;
          for (int kk = 0; kk < 3; kk++) diff[kk] = T.parts[c.pbeg + i].pos[kk] - c.ctr[kk];
          for (int kk = 0; kk < 3; kk++) dp += d[kk] * diff[kk];
          pp[i] += s * std::exp(I1 * kappa * dp);
        }
      }
    } else {
      st.v[ST_PLFL2P]++;
      for (int i = 0; i < n; i++) {
        cplx s = 0.;
        for (int l = 0; l < LL; l++) s += Zm[(size_t)l * n + i] * c.L[l];
        pp[i] += s;
      }
    }
  }
  // dwnpass.hpp:36-62
  void downward(int c) {
    Cell& C = T.cells[c];
    if (C.son.empty()) {
      if (precomp[C.lvl]) L2P(C);
      return;
    }
    for (int s : T.cells[c].son) {
      Cell& Cc = T.cells[c];
      if (precomp[Cc.lvl]) {
        if (isHFStage[Cc.lvl]) {
          if (HFStage[Cc.lvl]) hfL2Lblock(T.cells[s], Cc);
          else hf2lfL2L(T.cells[s], Cc);
        } else lfL2L(T.cells[s], Cc);
      }
      downward(s);
    }
  }

  // ------------------------------------------------------------ gemv ----
  // interface.hpp:153-170
  void gemv(const cplx* In, cplx* Out) {
    q.resize(S.parts.size());
    p.assign(T.parts.size(), cplx(0.));
    for (size_t i = 0; i < S.parts.size(); i++) q[i] = In[S.parts[i].ind];
    if (mode != 1) {
      upward(0);
      m2f();
    }
    horizontal(0, 0);
    if (mode != 1) {
      f2l();
      downward(0);
    }
    for (size_t i = 0; i < T.parts.size(); i++) Out[T.parts[i].ind] = p[i];
  }
};

}  // namespace orc

// ============================================================== C ABI ====
extern "C" {

// Full FMM evaluation: prcmpt + one gemv.  q and p are interleaved complex in
// input order.  mode: 0 full, 1 near field (P2P) only, 2 far field only.
// stats: at least 32 int64 slots.  Returns 0 on success.
int oracle_fmm(const double* src, int ns, const double* trg, int nt, const double* q, int L,
               int ncrit, double kappa, int mode, double* p, long long* stats,
               double* t_setup, double* t_apply) {
  try {
    orc::Oracle o;
    o.L = L;
    o.Ncrit = ncrit;
    o.kappa = kappa;
    o.mode = mode;
    struct timespec a, b, c;
    clock_gettime(CLOCK_MONOTONIC, &a);
    o.prcmpt(src, ns, trg, nt);
    clock_gettime(CLOCK_MONOTONIC, &b);
    o.gemv(reinterpret_cast<const orc::cplx*>(q), reinterpret_cast<orc::cplx*>(p));
    clock_gettime(CLOCK_MONOTONIC, &c);
    if (t_setup) *t_setup = (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
    if (t_apply) *t_apply = (c.tv_sec - b.tv_sec) + 1e-9 * (c.tv_nsec - b.tv_nsec);
    if (stats) memcpy(stats, o.st.v, sizeof(o.st.v));
    return 0;
  } catch (...) {
    return 1;
  }
}

// FP64 direct sum (interface.hpp:178-221 kernel) for a list of targets, with
// compensated (Kahan) accumulation and OpenMP over targets.
void oracle_direct(const double* src, int ns, const double* q, const double* trg, int nt,
                   double kappa, double* out) {
  const double pi4m1 = 1. / (4. * M_PI);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; i++) {
    double pr = 0., pim = 0., cr = 0., ci = 0.;
    for (int j = 0; j < ns; j++) {
      double x = trg[3 * i] - src[3 * j], y = trg[3 * i + 1] - src[3 * j + 1],
             z = trg[3 * i + 2] - src[3 * j + 2];
      double R = x * x + y * y + z * z;
      if (R < 1.e-16) continue;
      double r = std::sqrt(R);
      double K = pi4m1 / r;
      double co = std::cos(kappa * r) * K, si = std::sin(kappa * r) * K;
      double qr = q[2 * j], qi = q[2 * j + 1];
      double tr = co * qr - si * qi - cr;
      double s1 = pr + tr;
      cr = (s1 - pr) - tr;
      pr = s1;
      double ti = co * qi + si * qr - ci;
      double s2 = pim + ti;
      ci = (s2 - pim) - ti;
      pim = s2;
    }
    out[2 * i] = pr;
    out[2 * i + 1] = pim;
  }
}

// The reference test's charges (test/testinterface.cpp:23-26): real parts
// rand()/RAND_MAX from glibc's generator at its default seed.
void oracle_glibc_rand_charges(int n, double* out) {
  srand(1);
  for (int i = 0; i < n; i++) {
    out[2 * i] = rand() / double(RAND_MAX);
    out[2 * i + 1] = 0.;
  }
}

// Rounding variants of norms (see g_var).
void oracle_set_rounding(int tab, int cone, int fz) {
  orc::g_var[0] = tab; orc::g_var[1] = cone; orc::g_var[2] = fz;
}

// Setup only; writes the best-cone cache entries (lvl, a, b, c, hsh, cone),
// ordered by key.  Returns the number of entries (or -1 if > maxn).
int oracle_cones(const double* src, int ns, const double* trg, int nt, int L, int ncrit,
                 double kappa, int* out, int maxn) {
  orc::Oracle o;
  o.L = L; o.Ncrit = ncrit; o.kappa = kappa;
  try { o.prcmpt(src, ns, trg, nt); } catch (...) { return -2; }
  int n = 0;
  for (auto& kv : o.dirTable) {
    if (n >= maxn) return -1;
    out[6 * n] = std::get<0>(kv.first); out[6 * n + 1] = std::get<1>(kv.first);
    out[6 * n + 2] = std::get<2>(kv.first); out[6 * n + 3] = std::get<3>(kv.first);
    out[6 * n + 4] = std::get<4>(kv.first); out[6 * n + 5] = kv.second;
    n++;
  }
  return n;
}

// Exposed building blocks for closed-form checks.
double oracle_S1D(int i, double x, int L, double alpha) { return orc::S1D(i, x, L, alpha); }

// Direction table of one HF stage: writes 6*4^stage*3 doubles.
int oracle_directions(int nStages, int stage, double* out) {
  orc::Oracle o;
  o.getDirections(nStages);
  if (stage < 0 || stage >= nStages) return -1;
  int n = (int)o.dtab[stage].size();
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) out[3 * i + k] = o.dtab[stage][i][k];
  return n;
}

// Unnormalised 3D DFT of a P^3 complex array (fwd != 0: exp(-2 pi i jk/P)).
void oracle_dft3(int P, const double* in, double* out, int fwd) {
  orc::Oracle o;
  o.L = (P + 1) / 2;
  o.P = P;
  o.LL = o.L * o.L * o.L;
  o.PP = P * P * P;
  o.initCirculant();
  o.dft3(reinterpret_cast<const orc::cplx*>(in), reinterpret_cast<orc::cplx*>(out), fwd != 0);
}

// M2L kernel table of one offset: FFT of the circulant embedding of
// G(|v + rad*tabZ|).  out: P^3 complex.
void oracle_m2l_table(int L, double kappa, double rad, const double* v, double* out) {
  orc::Oracle o;
  o.L = L;
  o.P = 2 * L - 1;
  o.LL = L * L * L;
  o.PP = o.P * o.P * o.P;
  o.kappa = kappa;
  o.initCirculant();
  auto t = o.fillCircCoeffs(rad, v);
  memcpy(out, t.data(), sizeof(double) * 2 * o.PP);
}

// Fourier-index permutation for hashcode h (0..63): out[PP].
void oracle_perm(int L, int h, int* out) {
  orc::Oracle o;
  o.L = L;
  o.P = 2 * L - 1;
  o.LL = L * L * L;
  o.PP = o.P * o.P * o.P;
  o.initSymmetries();
  memcpy(out, o.perm[h].data(), sizeof(int) * o.PP);
}

// 1D M2M matrices A^-, A^+ (L*L each, [parent*L + child]).
void oracle_m2m_1d(int L, double* out) {
  orc::Oracle o;
  o.L = L;
  o.initInterp();
  memcpy(out, o.Am[0].data(), sizeof(double) * L * L);
  memcpy(out + L * L, o.Am[1].data(), sizeof(double) * L * L);
}

}  // extern "C"
