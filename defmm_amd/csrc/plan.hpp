// Host-side evaluation plan of the directional FMM: octree, frequency regimes,
// direction cones, effective-direction slots, flattened per-stage work lists
// and the operator tables the device kernels read.  Built once per prcmpt on
// the host (in double precision, strict IEEE: this file must be compiled
// without floating-point contraction so that the discrete choices are the
// same on every compiler), then uploaded by the executor.
#pragma once
#include <array>
#include <complex>
#include <cstdint>
#include <map>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace dfm {

typedef std::complex<double> cplx;

// Allocator whose construction without a value leaves the element
// default-initialised (an int: uninitialised), so that resize() of the large
// M2L lists does not zero-fill memory that is overwritten right after (the
// device set-up copies them back into host vectors of hundreds of millions
// of entries).
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    typedef NoInitAlloc<U> other;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
    ::new ((void*)p) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};
typedef std::vector<int, NoInitAlloc<int>> ListVec;

// Target slots per M2L work group: a thread holds one accumulator per group
// target and frequency it owns (2 targets at L = 8, where a thread owns 4
// frequencies, so that the FP64 kernel fits its registers; 3 at L = 7, where
// the float kernel with two right-hand sides then fits two 768-thread blocks
// per CU: 1M points / 16 wavelengths, 4 right-hand sides 182.8 -> 159.8 ms,
// FP64 89.1 -> 91.4 ms; 1 from L = 8, where a thread owns a whole column of
// 2L - 1 frequencies: at 20,000 points / 32 wavelengths / L = 8 the column
// kernel took M2L from 3.41 to 2.44 ms in float (16 right-hand sides: 38.8 ->
// 34.8 ms) and from 4.93 to 5.15 ms in double, against groups of 2).
constexpr int kM2LGroupMax = 8;
constexpr int m2lGroupSize(int L) { return L >= 8 ? 1 : L == 7 ? 3 : 4; }

// Error codes returned through the C ABI.
enum {
  OK = 0,
  ERR_ARG = 1,          // invalid argument (null pointer, bad size, bad order)
  ERR_NO_PRECOMP = 2,   // tree too shallow for any M2L level (frequency.hpp:112-121)
  ERR_DEPTH = 3,        // octree deeper than 30 levels (>= Ncrit coincident points)
  ERR_HIP = 4,          // a HIP runtime call failed
  ERR_MEMORY = 5,       // device or host allocation failed
  ERR_TABLE = 6,        // an M2L offset outside the precomputed table range
  ERR_NO_DEVICE = 7,    // no GPU visible
  ERR_STATE = 8,        // call out of sequence (e.g. execute before create)
};

// One octree (sources or targets), cells in pre-order.
struct Tree {
  std::vector<double> x, y, z;  // points in tree order
  std::vector<int> perm;        // tree position -> input index
  std::vector<int> lvl, nprt, pbeg, son0, nson, sons;
  std::vector<int> idx;         // 3 per cell
  std::vector<double> ctr;      // 3 per cell
  std::vector<double> rad;
  std::vector<int> nPerLevel;
  // effective directions: marks are per cell over all directions of its level
  std::vector<int64_t> mark0;   // offset into marks, -1 if the cell has none
  std::vector<uint8_t> marks;
  std::vector<int> slotIdx;     // per (cell, direction) of marks: slot or -1
  // slots: a slot is (cell, effective direction); LF cells have one slot
  std::vector<int> slot0, nslot;  // per cell
  std::vector<int> slotCell, slotDir;  // per slot (dir -1 for LF)
  int nslots = 0;
  int ncells() const { return (int)lvl.size(); }
  bool leaf(int c) const { return nson[c] == 0; }
  // slot of cell c for direction d (HF) or its single slot (LF); -1 if none
  int slotOf(int c, int d) const;
};

struct Stats {
  int64_t p2p_calls = 0, p2p_pairs = 0, hf_m2l = 0, lf_m2l = 0, src_slots = 0, trg_slots = 0;
  int64_t hf_tables = 0, lf_tables = 0, hf_leaves_p2m = 0, lf_leaves_p2m = 0;
  int64_t hf_m2m = 0, lf2hf_m2m = 0, lf_m2m = 0, hf_l2p = 0, lf_l2p = 0;
  int64_t hf_l2l = 0, hf2lf_l2l = 0, lf_l2l = 0, nlvl = 0, cpt = 0, max_stage = 0;
  int64_t ncells_t = 0, ncells_s = 0;
  int64_t p2p_items = 0, m2f_slots = 0, m2l_entries = 0, m2l_targets = 0;
  int64_t m2l_groups = 0, m2l_group_entries = 0;
  int64_t m2l_batches = 0, m2l_batch_rows = 0;  // batched M2L (set at upload)
  // host set-up time per phase (ms): trees, dual walk, kernel tables and
  // phase rows, M2M/L2L lists, M2L lists and groups, P2P lists
  int64_t ms_tree = 0, ms_walk = 0, ms_tables = 0, ms_trans = 0, ms_m2l = 0, ms_p2p = 0;
  int64_t ms_upload = 0;  // device upload (set by the executor)
  int64_t m2l_table_rows = 0;  // rows of the materialised M2L table store (set at upload)
  int64_t setup_device = 0;    // 1: the dual walk and M2L lists were built on the GPU
};

// A per-level translation list (M2M or L2L): for every output slot, the
// contributing input slots.  For M2M the output is a parent source slot and
// the inputs are child slots; for L2L the output is a child target slot and
// the inputs are parent slots.
struct TransLevel {
  int level = 0;          // level of the parent cells
  int hf = 0;             // 1: the parent level is HF (phases apply)
  std::vector<int> out;   // output slot
  std::vector<int> beg;   // CSR into entries (size out.size()+1)
  std::vector<int> in;    // input slot
  std::vector<int> sgn;   // bit k set: v_k >= 0 (selects the '+' 1D matrix)
  std::vector<int> sigIdx;  // phase table row (parent level, direction), -1 for LF
  std::vector<cplx> edv;  // scalar phase factor
};

struct Plan {
  // parameters
  int L = 0, P = 0, LL = 0, PP = 0, ncrit = 0;
  double kappa = 0;
  int nlvl = 0, cpt = 0, maxStage = 0;
  bool noHF = true;
  std::vector<int> isHF, HFStage, nDir, nCellPerDirMax;
  std::vector<char> precomp;
  std::vector<std::vector<std::array<double, 3>>> dirs;  // per HF stage
  double boxCtr[3] = {0, 0, 0}, boxRad = 0;
  Tree S, T;

  // phase tables: for every (precomputation HF level, direction) row r,
  // sig[r*LL + I] = exp(-i kappa d.(r_l/2) xi_I), sig2 = sig^2
  std::vector<int> sigRow0;  // per level: first row (or -1)
  std::vector<cplx> sig, sig2;
  // 1D interpolation matrices: A[s][k*L+p] (M2M: parent k <- child p) and
  // B[s][p*L+k] (L2L: child p <- parent k), s = 0 for negative offset.
  std::vector<double> A[2], B[2];
  // M2L kernel tables (Fourier domain) and index permutations
  std::vector<cplx> m2lTab;     // ntab * PP
  int ntab = 0;
  std::vector<int> perm;        // 64 * PP
  std::vector<int> embed;       // LL: I2i(I, P)

  // ---- work lists ----
  // P2M / L2P leaves: cell, direction rows
  std::vector<int> p2mCell;     // source leaves at precomputation levels
  std::vector<int> l2pCell;     // target leaves at precomputation levels
  std::vector<TransLevel> m2m;  // bottom-up order
  std::vector<TransLevel> l2l;  // top-down order
  std::vector<int> m2fSlot;     // source slots whose F is read by M2L
  std::vector<int> m2fIndex;    // per source slot: row in the F store, -1 if none
  // M2L, CSR by target slot (every target slot appears, possibly empty);
  // the entries' rows are filled by the host builder only (the device
  // builder keeps them on the device for the grouped lists, which are what
  // the M2L kernels read)
  std::vector<int> m2lBeg;      // nslots(T)+1
  int64_t m2lEntries = 0;       // entries of the CSR
  ListVec m2lSrcF;              // F row of the source slot
  ListVec m2lTabId;             // kernel table
  ListVec m2lHash;              // permutation
  std::vector<int> m2lOrder;    // target slots with M2L entries, in processing order
  std::vector<int> m2lZero;     // target slots without M2L entries (L = 0)
  // grouped M2L: per group m2lG target slots (-1 padded); per entry a source
  // F row, the mask of the group's targets reading it and their codes
  // (table id << 6 | permutation hash), m2lG per entry
  int m2lG = 4;
  std::vector<int> m2lGroupBeg, m2lGroupTarget;
  ListVec m2lGSrcF, m2lGMask, m2lGCode;
  // P2P, CSR by target leaf: source particle ranges [b, e)
  std::vector<int> p2pLeaf;     // target leaf cells that receive near field
  std::vector<int> p2pBeg;      // CSR
  std::vector<int> p2pSrcB, p2pSrcE;

  Stats st;
};

// A cell of a level-synchronous octree build: point range [b, b + n) of the
// final point order, centre, half-width, integer coordinates, and its
// children as a range of the next level's cells.
struct LevelCell {
  int64_t b = 0, n = 0;
  double ctr[3] = {0, 0, 0}, rad = 0;
  int idx[3] = {0, 0, 0};
  int kid0 = 0, nkid = 0;
};
// Whether a cell of n points at level lvl is a leaf (sort.hpp:49-51).
bool cellStops(int64_t n, int lvl, int ncrit, int lvmax);
// Appends the children of a splitting cell (octant counts cnt) to next.
void splitCell(LevelCell& c, const int64_t cnt[8], std::vector<LevelCell>& next);
// Fills t (pre-order cells, son lists, points in tree order) from the
// levels of a level-synchronous build and its final point order (consumed).
void treeFromLevels(Tree& t, const std::vector<std::vector<LevelCell>>& lv, std::vector<int>& perm,
                    const double* xyz);

// Set-up steps that a device can take over from the host builder.  tree()
// builds an octree where the host would.  walk()
// runs where the host dual walk would: it sets the trees' effective-direction
// marks and the walk statistics, returns the sorted table keys (level,
// fundamental offset) and the P2P cell pairs, and keeps the M2L pairs.
// m2lLists() runs where the host builds the M2L lists: it fills m2fIndex,
// m2fSlot, m2lBeg and m2lEntries, and keeps the entries (the host builder's
// m2lSrcF / m2lTabId / m2lHash, a target slot's in ascending source cell
// order, the host walk's order) on the device.  m2lGroups() builds the
// grouped M2L lists from them.  Each returns
// OK, an error code, or kSetupFallback when the device cannot take this plan
// (the caller then builds it on the host).
constexpr int kSetupFallback = -1;
struct SetupDevice {
  virtual ~SetupDevice() {}
  // Builds one octree as the host builder does (same cells in pre-order,
  // same point order, same centres): OK, ERR_DEPTH or kSetupFallback.
  virtual int tree(Tree& t, const double* xyz, int64_t n, const double ctr[3], double rad,
                   int ncrit, int lvmax, std::string& msg) = 0;
  virtual int walk(Plan& pl, std::vector<std::array<int, 4>>& tabKeys, std::vector<int>& pt,
                   std::vector<int>& ps, std::string& msg) = 0;
  virtual int m2lLists(Plan& pl, std::string& msg) = 0;
  // After m2lLists and the host's grouping of targets (m2lGroupTarget):
  // fills m2lGroupBeg, m2lGSrcF, m2lGMask and m2lGCode as the host merge of
  // a group's source rows does.
  virtual int m2lGroups(Plan& pl, std::string& msg) = 0;
  // Where the host builds the P2P lists: fills p2pLeaf, p2pBeg, p2pSrcB and
  // p2pSrcE from the P2P pairs the walk kept (walk() then returns no pairs
  // to the host) and sets the P2P statistics.
  virtual int p2pLists(Plan& pl, std::string& msg) = 0;
};

// Build the plan.  Returns an error code (OK on success); msg gets details.
// With dev, the dual walk and the M2L lists are built by it.
int buildPlan(Plan& pl, const double* src, int64_t ns, const double* trg, int64_t nt, int L,
              int ncrit, double kappa, std::string& msg, SetupDevice* dev = nullptr);

// Half-width W of the offset window |iZ_k| <= W that holds every M2L pair of
// level l, or -1 (no M2L at l, or more than 2^26 offsets).
int m2lWindow(const Plan& pl, int l);

// Table keys and per-offset codes (table id | hash << 32 | (cone + 1) << 40)
// from the representative pair of each offset met at each level; the cones
// of the HF levels are searched here unless cones (per level and offset,
// found with bestCone's operations elsewhere) is given.
int offsetCodes(const Plan& pl, const std::vector<int>& W,
                const std::vector<std::vector<uint64_t>>& rep,
                std::vector<std::vector<uint64_t>>& code,
                std::vector<std::array<int, 4>>& tabKeys, std::string& msg,
                const std::vector<std::vector<int>>* cones = nullptr);

// The device set-up (setup.hip): nullptr when no device is usable.
SetupDevice* newSetupDevice();

// Separable unnormalised DFT of a P^3 array (sign -1: forward).
void dft3(int P, const cplx* in, cplx* out, int sign);

}  // namespace dfm
