"""ctypes access to the oracle (oracle/liboracle.so): test infrastructure."""
import ctypes
import gzip
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")

STATS = ["p2p_calls", "p2p_pairs", "hf_m2l", "lf_m2l", "src_slots", "trg_slots", "hf_tables",
         "lf_tables", "hf_leaves", "lf_leaves", "hf_m2m", "lf2hf_m2m", "lf_m2m", "hf_l2p",
         "lf_l2p", "hf_l2l", "hf2lf_l2l", "lf_l2l", "nlvl", "cpt", "max_stage", "ncells_t",
         "ncells_s", "src_cells_with_slots", "trg_cells_with_slots"]

_D = ctypes.POINTER(ctypes.c_double)
_libs = {}


def load(path=None):
    path = path or os.path.join(ORACLE_DIR, "liboracle.so")
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        import importlib.util
        spec = importlib.util.spec_from_file_location("oracle_build", os.path.join(ORACLE_DIR, "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build(refround=path.endswith("refround.so"))
    lib = ctypes.CDLL(path)
    lib.oracle_fmm.argtypes = [_D, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int,
                               ctypes.c_double, ctypes.c_int, _D,
                               ctypes.POINTER(ctypes.c_longlong), _D, _D]
    lib.oracle_direct.argtypes = [_D, ctypes.c_int, _D, _D, ctypes.c_int, ctypes.c_double, _D]
    lib.oracle_glibc_rand_charges.argtypes = [ctypes.c_int, _D]
    lib.oracle_S1D.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    lib.oracle_S1D.restype = ctypes.c_double
    lib.oracle_directions.argtypes = [ctypes.c_int, ctypes.c_int, _D]
    lib.oracle_dft3.argtypes = [ctypes.c_int, _D, _D, ctypes.c_int]
    lib.oracle_m2l_table.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, _D]
    lib.oracle_perm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.oracle_m2m_1d.argtypes = [ctypes.c_int, _D]
    _libs[path] = lib
    return lib


def _p(a):
    return a.ctypes.data_as(_D)


def fmm(src, trg, q, L, ncrit, kappa, mode=0, lib=None):
    """Oracle FMM (prcmpt + gemv); returns (p, stats, t_setup, t_apply)."""
    lib = lib or load()
    src = np.ascontiguousarray(src, dtype=np.float64)
    trg = np.ascontiguousarray(trg, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.complex128)
    p = np.zeros(len(trg), dtype=np.complex128)
    st = np.zeros(32, dtype=np.int64)
    ts, ta = ctypes.c_double(), ctypes.c_double()
    rc = lib.oracle_fmm(_p(src), len(src), _p(trg), len(trg), _p(q.view(np.float64)), L, ncrit,
                        kappa, mode, _p(p.view(np.float64)),
                        st.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), ctypes.byref(ts),
                        ctypes.byref(ta))
    if rc:
        raise RuntimeError("oracle failed")
    return p, dict(zip(STATS, st.tolist())), ts.value, ta.value


def direct(src, q, trg, kappa, lib=None):
    """FP64 direct sum (compensated) at the given targets."""
    lib = lib or load()
    src = np.ascontiguousarray(src, dtype=np.float64)
    trg = np.ascontiguousarray(trg, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.complex128)
    out = np.zeros(len(trg), dtype=np.complex128)
    lib.oracle_direct(_p(src), len(src), _p(q.view(np.float64)), _p(trg), len(trg), kappa,
                      _p(out.view(np.float64)))
    return out


def glibc_rand_charges(n, lib=None):
    lib = lib or load()
    q = np.zeros(n, dtype=np.complex128)
    lib.oracle_glibc_rand_charges(n, _p(q.view(np.float64)))
    return q


def refined_cube():
    """The reference's test mesh (mesh/refinedcube.inp), 41,334 points."""
    with gzip.open(os.path.join(GOLDEN, "refinedcube.inp.gz"), "rt") as f:
        vals = f.read().split()
    return np.array([float(v) for v in vals], dtype=np.float64).reshape(-1, 3)


def rel_errors(exact, approx):
    """Relative L-inf, L1, L2 norms of the reference's CheckFULL
    (include/gaia/gaia_vectors.hpp:119-172), exact vector first."""
    d = np.abs(exact - approx)
    a = np.abs(exact)
    return d.max() / a.max(), d.sum() / a.sum(), np.sqrt((d ** 2).sum() / (a ** 2).sum())
