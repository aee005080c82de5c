/*
 * defmm_amd.h -- C ABI of the MI355X evaluation pass of defmm's directional
 * equispaced-interpolation FMM for the 3D Helmholtz kernel
 * exp(i*kappa*r)/(4*pi*r).
 *
 * The reference exposes this path as the C++ class defmm::IBFMM_Mat<3>
 * (include/interface.hpp:45-224 of the reference): particles are added,
 * prcmpt() builds trees and tables, gemv() runs one evaluation.  Its state is
 * process-global and only one gemv per prcmpt is valid; here a plan is an
 * object, owns its device memory, and can be executed any number of times.
 * Only plain pointers and sizes cross this boundary.  Every entry point
 * returns 0 on success or a DEFMM_E* code; defmm_last_error() gives a message
 * (thread-local).  A plan is thread-compatible (one caller at a time).
 * Executions of one plan share its device stores, so they are serialised
 * whatever streams they are enqueued on: an execution on another stream than
 * the previous one waits (on the device) for the previous one to finish, and
 * the reallocating setters and defmm_plan_destroy wait on the host.
 */
#ifndef DEFMM_AMD_H
#define DEFMM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DEFMM_OK = 0,
  DEFMM_EARG = 1,        /* invalid argument */
  DEFMM_ENOPRECOMP = 2,  /* no M2L level (reference: frequency.hpp:112-121 exits) */
  DEFMM_EDEPTH = 3,      /* octree deeper than 30 levels (>= Ncrit coincident points) */
  DEFMM_EHIP = 4,        /* HIP runtime failure */
  DEFMM_EMEMORY = 5,     /* allocation failure */
  DEFMM_ETABLE = 6,      /* M2L offset outside the precomputed tables */
  DEFMM_ENODEVICE = 7,   /* no GPU visible */
  DEFMM_ESTATE = 8       /* call out of sequence */
};

/* Arithmetic of an execution. */
enum {
  DEFMM_FP64 = 0,  /* complex128 everywhere: the drop-in of the reference */
  DEFMM_MIXED = 1  /* complex64 expansions and near field, float64 phase
                      arguments and kernel tables built in float64 */
};

typedef struct defmm_plan_s* defmm_plan;

/*
 * Setup: replaces IBFMM_Mat::add{Source,Target}Particle* + prcmpt(L, Ncrit,
 * kappa, lvMAX) (reference include/interface.hpp:59-151).  src_xyz / trg_xyz
 * hold n*3 doubles (x, y, z per point, input order).  order = L (>= 2),
 * ncrit: a cell is split while it holds >= ncrit points; lvmax >= 0 instead
 * splits every cell down to depth lvmax (the reference's fixed-depth build,
 * tree.hpp:104-132), -1 to use ncrit.  kappa >= 0: the wavenumber; 0 gives
 * the Laplace kernel 1/(4*pi*r) (the reference's thetis::kernels::Laplace)
 * with every level low-frequency.  precision: DEFMM_FP64 or DEFMM_MIXED
 * (device storage is allocated for that arithmetic).  Device: the current
 * HIP device.
 */
int defmm_plan_create(const double* src_xyz, int64_t n_src, const double* trg_xyz,
                      int64_t n_trg, int order, int ncrit, double kappa, int lvmax,
                      int precision, defmm_plan* out);

/*
 * Where defmm_plan_create builds the octrees (the reference's
 * include/structs/sort.hpp:98-162), runs the dual walk (its set-up
 * traversal, include/traversals/hrzpass.hpp:111-138, and its M2L list
 * building, include/interface.hpp:100-151) and builds the M2L lists:
 * DEFMM_SETUP_DEVICE (the default) on the current GPU, DEFMM_SETUP_HOST in
 * host C++ (OpenMP).  Both build the same plan, entry for entry.  The device
 * path hands a plan it cannot take (an M2L offset window above 2^26
 * offsets, device memory exhausted) to the host path.  No counterpart in the
 * reference (its set-up is host code); for comparing the two builders.
 * Process-wide; returns the previous mode, or DEFMM_EARG negated for an
 * unknown mode.
 */
#define DEFMM_SETUP_DEVICE 0
#define DEFMM_SETUP_HOST 1
int defmm_set_setup_mode(int mode);

/*
 * Evaluation: replaces gemv(A, In, Out) (reference include/interface.hpp:
 * 153-170).  q: n_src complex charges (interleaved re, im doubles, input
 * order, host memory); p: n_trg complex potentials, overwritten (host).
 * Synchronous.
 */
int defmm_execute(defmm_plan plan, const double* q, double* p);

/*
 * Same with device-resident data: q_dev / p_dev point to HIP device memory
 * holding interleaved complex128 values in input order.  stream: a
 * hipStream_t (NULL = default stream); the call enqueues work and returns.
 */
int defmm_execute_device(defmm_plan plan, const void* q_dev, void* p_dev, void* stream);

/*
 * Several right-hand sides at once (the reference evaluates one per gemv):
 * q holds nrhs x n_src complex charges, right-hand side after right-hand
 * side (q[2*(r*n_src + i)] = Re q_r,i), p likewise nrhs x n_trg.  The pass
 * handles up to defmm_plan_set_rhs() right-hand sides at a time, M2L and P2P
 * up to 8 per kernel (one table or kernel value for all of them).  Host
 * arrays, synchronous.
 */
int defmm_execute_batch(defmm_plan plan, const double* q, double* p, int nrhs);

/* The same with device arrays, enqueued on stream (NULL = default stream). */
int defmm_execute_batch_device(defmm_plan plan, const void* q_dev, void* p_dev, int nrhs,
                               void* stream);

/*
 * Width of the expansion stores in right-hand sides (1..64, default 1):
 * reallocates them; if the M2L variant was not set explicitly, re-chooses
 * it for the new width (stored F when it fits, else batched).
 */
int defmm_plan_set_rhs(defmm_plan plan, int nrhs);

/*
 * Evaluation of a part of the interaction, for verification: parts bit 0 =
 * near field (P2P), bit 1 = far field (P2M ... L2P).  parts = 3 is
 * defmm_execute.  Host memory, synchronous.
 */
int defmm_execute_parts(defmm_plan plan, const double* q, double* p, int parts);

/*
 * Plan statistics (list sizes and call counts of the reference's passes);
 * writes min(n, DEFMM_NSTATS) int64 values, see DEFMM_STAT_* below.
 */
int defmm_plan_stats(defmm_plan plan, int64_t* out, int n);

/*
 * Milliseconds spent in each stage by the last execution (HIP events on the
 * execution stream): writes min(n, DEFMM_NSTAGES) doubles in the order
 * gather, P2M, M2M, M2F, M2L+F2L, L2L, L2P, P2P, scatter.  All stages run
 * in this order on the execution stream; the M2L+F2L stage holds the fused
 * M2L kernel alone (target slots with no M2L are cleared in the L2L stage).
 */
int defmm_stage_times(defmm_plan plan, double* ms, int n);

/*
 * Host-only setup, no device work: builds the octrees and lists of
 * defmm_plan_create and writes their statistics (as defmm_plan_stats).  For
 * inspecting list sizes before committing device memory.
 */
int defmm_plan_build_stats(const double* src_xyz, int64_t n_src, const double* trg_xyz,
                           int64_t n_trg, int order, int ncrit, double kappa, int lvmax,
                           int64_t* out, int n);

/*
 * FP64 direct sum p_i = sum_j exp(i kappa r_ij)/(4 pi r_ij) q_j (terms with
 * r_ij^2 < 1e-16 skipped) on the current device, compensated accumulation:
 * the check of the reference's CheckFULL (include/interface.hpp:178-221),
 * for sampled targets.  Host arrays, synchronous; allocates device memory.
 */
int defmm_direct(const double* src_xyz, int64_t n_src, const double* q, const double* trg_xyz,
                 int64_t n_trg, double kappa, double* p);

/*
 * The same direct sum for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Device arrays in the layouts
 * of defmm_direct (src_dev n_src x 3 doubles, q_dev n_src interleaved
 * complex doubles, trg_dev n_trg x 3, p_dev n_trg complex, overwritten),
 * enqueued on stream (NULL = default stream) of the current device; returns
 * after enqueueing.  Results are bit-identical to defmm_direct's.  Calls no
 * allocating routine: the partial sums live in a fixed 4 MiB array of the
 * library in each GPU's memory (held until the process ends), so calls that
 * overlap on one GPU (other streams, other host threads) race and must not
 * be issued.  n_src, n_trg in 1 .. 2^31-1 (larger sizes are rejected with
 * DEFMM_EARG); targets are processed in tiles of 4,096 that reuse the array.
 */
int defmm_direct_noalloc(const void* src_dev, int64_t n_src, const void* q_dev,
                         const void* trg_dev, int64_t n_trg, double kappa, void* p_dev,
                         void* stream);

/*
 * Switch the arithmetic of an existing plan (DEFMM_FP64 / DEFMM_MIXED): the
 * lists stay, the expansion stores and tables are reallocated in the new
 * precision.  Synchronises the plan's work first.
 */
int defmm_plan_set_precision(defmm_plan plan, int precision);

/*
 * M2L variants (results agree to roundoff; the switch reallocates the
 * expansion stores).  defmm_plan_create picks STORED when its stores fit in
 * 90% of the free device memory, else BATCHED.
 *  DEFMM_M2L_STORED: the F rows of every source slot read by M2L are kept
 *    in HBM (one M2F launch, one M2L launch streaming F); M and L share one
 *    store.  Memory: (2L-1)^3 complex per source slot on top of M.  Fastest.
 *  DEFMM_M2L_BATCHED: the M2L targets, in (level, direction, cell) order,
 *    are cut into batches whose source Fourier rows fit a workspace of
 *    `rows` rows (defmm_plan_set_m2l_batch_rows); per batch and per chunk
 *    of right-hand sides of one M2L launch (1 - 8), M2F fills the workspace
 *    and the M2L launch reads it while it is still cache-resident.  Memory:
 *    M and L stores plus the workspace of one chunk.
 *  DEFMM_M2L_ON_THE_FLY: M2L recomputes each F row from M inside the
 *    kernel; no F store.  Orders up to 7.
 */
#define DEFMM_M2L_ON_THE_FLY 0
#define DEFMM_M2L_STORED 1
#define DEFMM_M2L_BATCHED 2
int defmm_plan_set_m2l(defmm_plan plan, int mode);
/* The plan's current M2L variant. */
int defmm_plan_get_m2l(defmm_plan plan, int* mode);

/*
 * Workspace rows of the batched M2L (default 16384).  Raised to the largest
 * number of source rows a single M2L work group reads if smaller; takes
 * effect at once when the plan is in batched mode, else at the next switch.
 */
int defmm_plan_set_m2l_batch_rows(defmm_plan plan, int rows);

/*
 * Device memory budget of a plan in bytes (0 = none, the default).  Later
 * reallocations (defmm_plan_set_rhs, _set_precision, _set_m2l,
 * _set_m2l_batch_rows) that would exceed it fail with DEFMM_EMEMORY, and the
 * automatic M2L variant choice counts only the memory left under it.  The
 * current allocation is not shrunk.  A reallocation that fails for any reason
 * restores the previous settings; if even those cannot be allocated again,
 * the plan stays unusable and executions return DEFMM_ESTATE.
 */
int defmm_plan_set_memory_limit(defmm_plan plan, int64_t bytes);

/*
 * Execution options of a plan (no reallocation; waits for the plan's last
 * execution).  DEFMM_OPT_P2P_MFMA: 1 (default) evaluates the near field of
 * the mixed-precision path 16 right-hand sides at a time on the matrix cores
 * when the plan's width allows it; 0 keeps it on the vector units (up to 8
 * right-hand sides per kernel value).  DEFMM_OPT_L2P_MFMA: the same for L2P
 * (1 by default: 16 right-hand sides per pass over a leaf's slots on the
 * matrix cores; 0: one right-hand side per pass on the vector units).
 * DEFMM_OPT_P2M_MFMA: the same for P2M (leaves of at most 64 points; larger
 * leaves stay on the vector units).  DEFMM_OPT_M2L_SLICED (orders <= 7): 1
 * sweeps each M2L work group's sources once per slice of its frequencies
 * (a work group then reads a third of each row at a time, so rows shared
 * with neighbouring work groups stay in L2 longer); 0 (default) sweeps them
 * once.  Results are identical.
 */
#define DEFMM_OPT_P2P_MFMA 1
#define DEFMM_OPT_L2P_MFMA 2
#define DEFMM_OPT_P2M_MFMA 3
#define DEFMM_OPT_M2L_SLICED 4
int defmm_plan_set_option(defmm_plan plan, int option, int64_t value);

/*
 * Autotuner of a plan's execution knobs for passes of nrhs right-hand sides
 * (1 .. the plan's width), in the plan's current precision.  The pass is run
 * on the plan's own staging buffers with unit charges, `reps` timed passes
 * per candidate (median), one knob at a time with the others fixed, and the
 * fastest value is kept:
 *   1. M2L storage: stored Fourier rows (skipped when they do not fit) or
 *      the batched variant with a workspace of 2,048 to 65,536 rows
 *      (powers of two);
 *   2. orders <= 7: the sliced M2L sweep on or off (DEFMM_OPT_M2L_SLICED);
 *   3. in mixed precision with nrhs >= 16: matrix cores on or off for P2P,
 *      L2P and P2M (DEFMM_OPT_*_MFMA).
 * The kept M2L variant is then the plan's explicit choice: a later width
 * change (defmm_plan_set_rhs) keeps it, and fails with DEFMM_EMEMORY (the
 * plan keeping its settings) if it does not fit at the new width.
 * Writes a text report of every candidate's time and the kept values to
 * report (at most report_len bytes, NUL-terminated; report may be NULL).
 * Waits for the plan's previous executions; synchronous.
 */
int defmm_plan_autotune(defmm_plan plan, int nrhs, int reps, char* report, int report_len);

/* Bytes of device memory held by the plan. */
int defmm_plan_device_bytes(defmm_plan plan, int64_t* bytes);

/*
 * Instance batches.  The reference holds one FMM instance per process
 * (process-global state, include/global.hpp:45-90); here many independent
 * plans (sizes, geometries, orders, precisions and right-hand-side counts
 * may all differ) are evaluated by one HIP graph.  defmm_batch_create
 * captures every plan's whole pass -- nrhs[i] right-hand sides, charges
 * q_dev[i] (nrhs[i] x n_src complex128, device) to potentials p_dev[i]
 * (nrhs[i] x n_trg, device) -- as the graph's branches: the instances are
 * dealt to nbranch concurrent branches (0 = min(n, 8)), largest work first
 * to the least loaded branch.  The arrays' addresses are bound into the
 * graph.  defmm_batch_execute launches the graph on stream (NULL = default
 * stream) after every plan's previous execution, and returns.  A plan may
 * appear once per batch; reallocating one of its plans (set_rhs,
 * set_precision, set_m2l, batch rows) makes the batch stale (DEFMM_ESTATE:
 * create it again).  Destroy a batch before its plans.  Stage times are not
 * recorded for batched executions.
 */
typedef struct defmm_batch_s* defmm_batch;
int defmm_batch_create(const defmm_plan* plans, int n, const void* const* q_dev,
                       void* const* p_dev, const int* nrhs, int nbranch, defmm_batch* out);
int defmm_batch_execute(defmm_batch batch, void* stream);
int defmm_batch_destroy(defmm_batch batch);

int defmm_plan_destroy(defmm_plan plan);
const char* defmm_strerror(int code);
const char* defmm_last_error(void);

#define DEFMM_NSTAGES 9

enum {
  DEFMM_STAT_P2P_CALLS = 0, DEFMM_STAT_P2P_PAIRS, DEFMM_STAT_HF_M2L, DEFMM_STAT_LF_M2L,
  DEFMM_STAT_SRC_SLOTS, DEFMM_STAT_TRG_SLOTS, DEFMM_STAT_HF_TABLES, DEFMM_STAT_LF_TABLES,
  DEFMM_STAT_HF_LEAVES, DEFMM_STAT_LF_LEAVES, DEFMM_STAT_HF_M2M, DEFMM_STAT_LF2HF_M2M,
  DEFMM_STAT_LF_M2M, DEFMM_STAT_HF_L2P, DEFMM_STAT_LF_L2P, DEFMM_STAT_HF_L2L,
  DEFMM_STAT_HF2LF_L2L, DEFMM_STAT_LF_L2L, DEFMM_STAT_NLVL, DEFMM_STAT_CPT,
  DEFMM_STAT_MAX_STAGE, DEFMM_STAT_NCELLS_T, DEFMM_STAT_NCELLS_S, DEFMM_STAT_P2P_ITEMS,
  DEFMM_STAT_M2F_SLOTS, DEFMM_STAT_M2L_ENTRIES, DEFMM_STAT_M2L_TARGETS,
  DEFMM_STAT_M2L_GROUPS, DEFMM_STAT_M2L_GROUP_ENTRIES, DEFMM_STAT_M2L_BATCHES,
  DEFMM_STAT_M2L_BATCH_ROWS,
  /* set-up time in milliseconds: octrees, dual walk, kernel tables and phase
     rows, M2M/L2L lists, M2L lists, P2P lists, device upload */
  DEFMM_STAT_MS_TREE, DEFMM_STAT_MS_WALK, DEFMM_STAT_MS_TABLES, DEFMM_STAT_MS_TRANS,
  DEFMM_STAT_MS_M2L, DEFMM_STAT_MS_P2P, DEFMM_STAT_MS_UPLOAD,
  /* rows of the M2L kernel-table store on the device (one per table, axis
     order and sign flips in use) */
  DEFMM_STAT_M2L_TABLE_ROWS,
  /* 1 when the dual walk and the M2L lists were built on the GPU */
  DEFMM_STAT_SETUP_DEVICE,
  DEFMM_NSTATS
};

#ifdef __cplusplus
}
#endif
#endif /* DEFMM_AMD_H */
