"""defmm_amd -- MI355X evaluation pass of defmm's directional FMM (Helmholtz).

Host-side mirror of the reference's interface (``defmm::IBFMM_Mat<3>``,
include/interface.hpp:45-224 of the reference) over the C ABI of
``libdefmm_amd.so`` (declared in include/defmm_amd.h):

    A = IBFMM_Mat()
    A.addSourceParticlesINP("mesh.inp", n); A.addTargetParticlesINP(...)
    A.prcmpt(L, Ncrit, kappa)          # trees, lists, tables, device upload
    out = gemv(A, q)                   # one evaluation (complex128 arrays)

Differences from the reference, all deliberate: the plan may be executed any
number of times (the reference accumulates into never-cleared buffers, so a
second gemv is wrong), errors raise ``DefmmError`` instead of exit(1), and the
input particle lists are kept after prcmpt.  There is no CPU fallback: without
the built library or a GPU the calls raise.
"""
import ctypes
import os

import numpy as np

__all__ = ["IBFMM_Mat", "Batch", "gemv", "DefmmError", "lib", "FP64", "MIXED", "STAT_NAMES",
           "STAGE_NAMES", "build_plan_stats", "direct", "set_setup_mode", "SETUP_DEVICE",
           "SETUP_HOST"]

FP64 = 0
MIXED = 1
# M2L variants (defmm_amd.h: DEFMM_M2L_*)
M2L_ON_THE_FLY = 0
M2L_STORED = 1
M2L_BATCHED = 2
# where plan creation builds the dual walk and the M2L lists (defmm_amd.h: DEFMM_SETUP_*)
SETUP_DEVICE = 0
SETUP_HOST = 1
# execution options (defmm_amd.h: DEFMM_OPT_*)
OPT_P2P_MFMA = 1
OPT_L2P_MFMA = 2
OPT_P2M_MFMA = 3
OPT_M2L_SLICED = 4

STAT_NAMES = [
    "p2p_calls", "p2p_pairs", "hf_m2l", "lf_m2l", "src_slots", "trg_slots", "hf_tables",
    "lf_tables", "hf_leaves", "lf_leaves", "hf_m2m", "lf2hf_m2m", "lf_m2m", "hf_l2p",
    "lf_l2p", "hf_l2l", "hf2lf_l2l", "lf_l2l", "nlvl", "cpt", "max_stage", "ncells_t",
    "ncells_s", "p2p_items", "m2f_slots", "m2l_entries", "m2l_targets", "m2l_groups",
    "m2l_group_entries", "m2l_batches", "m2l_batch_rows", "ms_tree", "ms_walk", "ms_tables",
    "ms_trans", "ms_m2l", "ms_p2p", "ms_upload", "m2l_table_rows", "setup_device",
]
STAGE_NAMES = ["gather", "p2m", "m2m", "m2f", "m2l_f2l", "l2l", "l2p", "p2p", "scatter"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdefmm_amd.so")


class DefmmError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libdefmm_amd.so (built by defmm_amd/build.py); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DefmmError(f"{LIB_PATH} is not built; run python defmm_amd/build.py")
    # PyTorch ships its own HIP runtime.  Loaded first, it is the one the
    # library binds to, so that both share one runtime; loaded after the
    # library (which would then bring the system's), torch finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    D = ctypes.POINTER(ctypes.c_double)
    I64 = ctypes.POINTER(ctypes.c_int64)
    P = ctypes.c_void_p
    L.defmm_plan_create.argtypes = [D, ctypes.c_int64, D, ctypes.c_int64, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(P)]
    L.defmm_execute.argtypes = [P, D, D]
    L.defmm_execute_device.argtypes = [P, P, P, P]
    L.defmm_execute_parts.argtypes = [P, D, D, ctypes.c_int]
    L.defmm_execute_batch.argtypes = [P, D, D, ctypes.c_int]
    L.defmm_execute_batch_device.argtypes = [P, P, P, ctypes.c_int, P]
    L.defmm_plan_set_rhs.argtypes = [P, ctypes.c_int]
    L.defmm_plan_stats.argtypes = [P, I64, ctypes.c_int]
    L.defmm_stage_times.argtypes = [P, D, ctypes.c_int]
    L.defmm_plan_device_bytes.argtypes = [P, I64]
    L.defmm_plan_destroy.argtypes = [P]
    L.defmm_strerror.restype = ctypes.c_char_p
    L.defmm_strerror.argtypes = [ctypes.c_int]
    L.defmm_last_error.restype = ctypes.c_char_p
    L.defmm_plan_build_stats.argtypes = [D, ctypes.c_int64, D, ctypes.c_int64, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_int, I64,
                                         ctypes.c_int]
    L.defmm_plan_set_precision.argtypes = [P, ctypes.c_int]
    L.defmm_plan_set_m2l.argtypes = [P, ctypes.c_int]
    L.defmm_plan_set_m2l_batch_rows.argtypes = [P, ctypes.c_int]
    L.defmm_plan_get_m2l.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    L.defmm_plan_set_memory_limit.argtypes = [P, ctypes.c_int64]
    L.defmm_plan_set_option.argtypes = [P, ctypes.c_int, ctypes.c_int64]
    L.defmm_plan_autotune.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    L.defmm_batch_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.POINTER(P),
                                     ctypes.POINTER(P), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                     ctypes.POINTER(P)]
    L.defmm_batch_execute.argtypes = [P, P]
    L.defmm_batch_destroy.argtypes = [P]
    L.defmm_direct.argtypes = [D, ctypes.c_int64, D, D, ctypes.c_int64, ctypes.c_double, D]
    L.defmm_set_setup_mode.argtypes = [ctypes.c_int]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        L = lib()
        raise DefmmError(f"{L.defmm_strerror(rc).decode()} ({rc}): {L.defmm_last_error().decode()}")


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _points(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, 3))
    return a


def set_setup_mode(mode):
    """Where plan creation builds the octrees, runs the dual walk and builds
    the M2L lists: SETUP_DEVICE (the default, the current GPU) or SETUP_HOST
    (host C++).  Both build the same plan.  Process-wide; returns the
    previous mode."""
    rc = lib().defmm_set_setup_mode(int(mode))
    if rc < 0:
        _check(-rc)
    return rc


def build_plan_stats(src, trg, L, Ncrit, kappa, lvMAX=-1):
    """Host-only plan construction (no GPU): returns the list statistics."""
    s, t = _points(src), _points(trg)
    out = np.zeros(len(STAT_NAMES), dtype=np.int64)
    _check(lib().defmm_plan_build_stats(_dptr(s), len(s), _dptr(t), len(t), int(L), int(Ncrit),
                                        float(kappa), int(lvMAX),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        len(out)))
    return dict(zip(STAT_NAMES, out.tolist()))


def direct(src, q, trg, kappa):
    """FP64 direct sum on the GPU at the given targets (sampled-target check)."""
    s, t = _points(src), _points(trg)
    q = np.ascontiguousarray(np.asarray(q, dtype=np.complex128).reshape(-1))
    if q.size != len(s):
        raise DefmmError("charges and sources differ in length")
    out = np.zeros(len(t), dtype=np.complex128)
    _check(lib().defmm_direct(_dptr(s), len(s), _dptr(q.view(np.float64)), _dptr(t), len(t),
                              float(kappa), _dptr(out.view(np.float64))))
    return out


def _read_inp(name, nparts):
    with open(name) as f:
        vals = f.read().split()
    a = np.array([float(v) for v in vals[:3 * nparts]], dtype=np.float64)
    if a.size != 3 * nparts:
        raise DefmmError(f"{name}: expected {nparts} points")
    return a.reshape(nparts, 3)


class IBFMM_Mat:
    """Mirror of defmm::IBFMM_Mat<3> (reference include/interface.hpp:45-224)."""

    def __init__(self, precision=FP64):
        self._src = []
        self._trg = []
        self._plan = ctypes.c_void_p()
        self.precision = precision
        self.L = self.Ncrit = None
        self.kappa = None

    # interface.hpp:59-77
    def addSourceParticle(self, part):
        self._src = list(self._src) + [np.asarray(part, dtype=np.float64).reshape(3)]

    def addTargetParticle(self, part):
        self._trg = list(self._trg) + [np.asarray(part, dtype=np.float64).reshape(3)]

    # interface.hpp:79-98 (whitespace-separated "x y z" text)
    def addSourceParticlesINP(self, name, nparts):
        self._src = list(_read_inp(name, nparts))

    def addTargetParticlesINP(self, name, nparts):
        self._trg = list(_read_inp(name, nparts))

    def setSourceParticles(self, xyz):
        self._src = _points(xyz)

    def setTargetParticles(self, xyz):
        self._trg = _points(xyz)

    @property
    def n_src(self):
        return len(self._src)

    @property
    def n_trg(self):
        return len(self._trg)

    # interface.hpp:100-151
    def prcmpt(self, L, Ncrit, kappa, lvMAX=-1):
        self.destroy()
        self.L, self.Ncrit, self.kappa = int(L), int(Ncrit), float(kappa)
        s = _points(np.array(self._src)) if len(self._src) else np.zeros((0, 3))
        t = _points(np.array(self._trg)) if len(self._trg) else np.zeros((0, 3))
        if len(s) == 0 or len(t) == 0:
            raise DefmmError("no particles")
        plan = ctypes.c_void_p()
        _check(lib().defmm_plan_create(_dptr(s), len(s), _dptr(t), len(t), int(L), int(Ncrit),
                                       float(kappa), int(lvMAX), int(self.precision),
                                       ctypes.byref(plan)))
        self._plan = plan
        self._width = 1

    def _need_plan(self):
        if not self._plan:
            raise DefmmError("prcmpt() has not been called")

    def apply(self, q, parts=3):
        """Evaluate p = A q (q: n_src complex values, input order)."""
        self._need_plan()
        q = np.ascontiguousarray(np.asarray(q, dtype=np.complex128).reshape(-1))
        if q.size != self.n_src:
            raise DefmmError(f"expected {self.n_src} charges, got {q.size}")
        p = np.zeros(self.n_trg, dtype=np.complex128)
        qd, pd = q.view(np.float64), p.view(np.float64)
        if parts == 3:
            _check(lib().defmm_execute(self._plan, _dptr(qd), _dptr(pd)))
        else:
            _check(lib().defmm_execute_parts(self._plan, _dptr(qd), _dptr(pd), int(parts)))
        return p

    def apply_device(self, q_ptr, p_ptr, stream=None):
        """Evaluate with device pointers (complex128, input order)."""
        self._need_plan()
        _check(lib().defmm_execute_device(self._plan, ctypes.c_void_p(q_ptr),
                                          ctypes.c_void_p(p_ptr), ctypes.c_void_p(stream or 0)))

    def set_rhs(self, nrhs):
        """Width of the expansion stores in right-hand sides (reallocates)."""
        self._need_plan()
        _check(lib().defmm_plan_set_rhs(self._plan, int(nrhs)))
        self._width = int(nrhs)

    def width(self):
        """Right-hand sides the plan's stores hold (set_rhs; 1 after prcmpt)."""
        return getattr(self, "_width", 1)

    def apply_batch(self, Q):
        """Evaluate P = A Q for several right-hand sides: Q (nrhs, n_src)
        complex -> P (nrhs, n_trg)."""
        self._need_plan()
        Q = np.ascontiguousarray(np.asarray(Q, dtype=np.complex128))
        if Q.ndim != 2 or Q.shape[1] != self.n_src:
            raise DefmmError(f"expected (nrhs, {self.n_src}) charges, got {Q.shape}")
        P = np.zeros((Q.shape[0], self.n_trg), dtype=np.complex128)
        _check(lib().defmm_execute_batch(self._plan, _dptr(Q.view(np.float64)),
                                         _dptr(P.view(np.float64)), Q.shape[0]))
        return P

    def apply_batch_device(self, q_ptr, p_ptr, nrhs, stream=None):
        """Several right-hand sides with device pointers (nrhs x n complex128)."""
        self._need_plan()
        _check(lib().defmm_execute_batch_device(self._plan, ctypes.c_void_p(q_ptr),
                                                ctypes.c_void_p(p_ptr), int(nrhs),
                                                ctypes.c_void_p(stream or 0)))

    def set_precision(self, precision):
        """Switch an existing plan to FP64 or MIXED (lists are kept)."""
        self._need_plan()
        _check(lib().defmm_plan_set_precision(self._plan, int(precision)))
        self.precision = precision

    def set_m2l(self, mode, batch_rows=None):
        """M2L variant: M2L_STORED (chosen at creation when it fits in
        memory), M2L_BATCHED or M2L_ON_THE_FLY; batch_rows sets the batched
        variant's workspace rows."""
        self._need_plan()
        if batch_rows is not None:
            _check(lib().defmm_plan_set_m2l_batch_rows(self._plan, int(batch_rows)))
        _check(lib().defmm_plan_set_m2l(self._plan, int(mode)))

    def m2l_mode(self):
        """The plan's current M2L variant (M2L_STORED, M2L_BATCHED, M2L_ON_THE_FLY)."""
        self._need_plan()
        m = ctypes.c_int()
        _check(lib().defmm_plan_get_m2l(self._plan, ctypes.byref(m)))
        return m.value

    def stats(self):
        self._need_plan()
        out = np.zeros(len(STAT_NAMES), dtype=np.int64)
        _check(lib().defmm_plan_stats(self._plan, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                      len(out)))
        return dict(zip(STAT_NAMES, out.tolist()))

    def stage_times(self):
        self._need_plan()
        ms = np.zeros(len(STAGE_NAMES))
        _check(lib().defmm_stage_times(self._plan, _dptr(ms), len(ms)))
        return dict(zip(STAGE_NAMES, ms.tolist()))

    def set_option(self, option, value):
        """Execution option: OPT_P2P_MFMA, OPT_L2P_MFMA, OPT_P2M_MFMA (0 or 1:
        the stage of the mixed path on the matrix cores, 16 right-hand sides
        at a time; 1 by default)."""
        self._need_plan()
        _check(lib().defmm_plan_set_option(self._plan, int(option), int(value)))

    def autotune(self, nrhs=None, reps=3):
        """Pick the M2L storage variant and the matrix-core options by timing
        passes of nrhs right-hand sides (default: the plan's width); returns
        the autotuner's report."""
        self._need_plan()
        buf = ctypes.create_string_buffer(4096)
        _check(lib().defmm_plan_autotune(self._plan, int(nrhs or self.width()), int(reps), buf,
                                         len(buf)))
        return buf.value.decode()

    def set_memory_limit(self, nbytes):
        """Device memory budget of later reallocations (0: none)."""
        self._need_plan()
        _check(lib().defmm_plan_set_memory_limit(self._plan, int(nbytes)))

    def device_bytes(self):
        self._need_plan()
        b = ctypes.c_int64()
        _check(lib().defmm_plan_device_bytes(self._plan, ctypes.byref(b)))
        return b.value

    def destroy(self):
        if self._plan:
            lib().defmm_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Batch:
    """Independent plans evaluated by one HIP graph (defmm_batch_*): mats is
    a list of IBFMM_Mat with plans, q_ptrs / p_ptrs device addresses of each
    instance's charges (nrhs x n_src complex128) and potentials, nrhs the
    right-hand sides per instance.  The plans must outlive the batch."""

    def __init__(self, mats, q_ptrs, p_ptrs, nrhs, nbranch=0):
        n = len(mats)
        for A in mats:
            A._need_plan()
        self._mats = list(mats)
        plans = (ctypes.c_void_p * n)(*[A._plan.value for A in mats])
        qs = (ctypes.c_void_p * n)(*q_ptrs)
        ps = (ctypes.c_void_p * n)(*p_ptrs)
        nr = (ctypes.c_int * n)(*[int(r) for r in nrhs])
        self._b = ctypes.c_void_p()
        _check(lib().defmm_batch_create(plans, n, qs, ps, nr, int(nbranch), ctypes.byref(self._b)))

    def execute(self, stream=None):
        _check(lib().defmm_batch_execute(self._b, ctypes.c_void_p(stream or 0)))

    def destroy(self):
        if self._b:
            lib().defmm_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def gemv(A, In):
    """gemv(A, In) -> Out (reference include/interface.hpp:153-170)."""
    return A.apply(In)
