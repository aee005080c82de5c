"""Host logic of the product on the CPU (no device work): the plan builder's
octree, regimes, best cones, slots and lists against the oracle's counts,
and the C ABI library's exported symbols."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_api as oa

ROOT = oa.ROOT


def _lib_or_skip():
    import defmm_amd as dm
    try:
        return dm, dm.lib()
    except (dm.DefmmError, OSError) as e:
        pytest.skip(f"libdefmm_amd.so not built: {e}")


KEYS = ["p2p_calls", "p2p_pairs", "hf_m2l", "lf_m2l", "src_slots", "trg_slots", "hf_leaves",
        "lf_leaves", "hf_m2m", "lf2hf_m2m", "lf_m2m", "hf_l2p", "lf_l2p", "hf_l2l", "hf2lf_l2l",
        "lf_l2l", "nlvl", "cpt", "max_stage", "ncells_t", "ncells_s"]


def test_header_symbols_exported():
    _, lib = _lib_or_skip()
    hdr = open(os.path.join(ROOT, "include", "defmm_amd.h")).read()
    names = re.findall(r"^\s*(?:int|const char\*)\s+(defmm_\w+)\s*\(", hdr, re.M)
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n


def test_refined_cube_plan_counts(cube):
    dm, _ = _lib_or_skip()
    s = dm.build_plan_stats(cube, cube, 6, 32, 1.0)
    assert (s["p2p_calls"], s["hf_m2l"], s["lf_m2l"], s["hf_leaves"]) == (129928, 336776, 17334, 2193)
    assert (s["hf_m2m"], s["lf2hf_m2m"], s["lf_m2m"]) == (2508, 733, 96)
    assert s["hf_tables"] == 615  # HF offsets met; the LF tables are built on demand (28 of 132)


@pytest.mark.parametrize("case", range(6))
def test_plan_counts_match_oracle(case):
    dm, _ = _lib_or_skip()
    from defmm_amd import workloads as wl
    rng = np.random.default_rng(case)
    n = int(rng.integers(500, 2500))
    geo = case % 3
    if geo == 0:
        src = rng.random((n, 3))
    elif geo == 1:
        src = wl.sphere(n, seed=case + 1)
    else:
        src = wl.cube_edges(n, seed=case + 1)
    trg = src if case % 2 == 0 else rng.random((n // 2, 3)) * 0.9
    L = int(rng.integers(2, 7))
    ncrit = int(rng.integers(8, 40))
    kappa = float(rng.choice([0.3, 3.0, 20.0, 60.0]))
    q = np.ones(len(src), dtype=complex)
    try:
        _, so, _, _ = oa.fmm(src, trg, q, L, ncrit, kappa)
    except RuntimeError:
        with pytest.raises(dm.DefmmError):
            dm.build_plan_stats(src, trg, L, ncrit, kappa)
        return
    sp = dm.build_plan_stats(src, trg, L, ncrit, kappa)
    for k in KEYS:
        assert sp[k] == so[k], (k, sp[k], so[k])


def test_plan_errors():
    dm, _ = _lib_or_skip()
    pts = np.random.default_rng(0).random((30, 3))
    with pytest.raises(dm.DefmmError, match="no M2L level"):
        dm.build_plan_stats(pts, pts, 4, 64, 1.0)
    with pytest.raises(dm.DefmmError, match="invalid argument"):
        dm.build_plan_stats(pts, pts, 1, 8, 1.0)
    same = np.zeros((100, 3))
    same[:, 0] = np.arange(100) * 1e-3
    same[50:] = 0.0  # 50 coincident points with Ncrit 8: unbounded split
    with pytest.raises(dm.DefmmError, match="deeper than 30"):
        dm.build_plan_stats(same, same, 3, 8, 1.0)


def test_fixed_depth_build():
    """lvMAX >= 0: every cell split to that depth (tree.hpp:104-132)."""
    dm, _ = _lib_or_skip()
    pts = np.random.default_rng(1).random((4000, 3))
    s = dm.build_plan_stats(pts, pts, 3, 0, 5.0, lvMAX=4)
    assert s["nlvl"] == 5


def test_degenerate_inputs():
    """Inputs the reference does not guard against give errors, not crashes:
    no points, a single point (all points coincide), non-finite coordinates
    or wavenumber, and a tree too shallow for any M2L level (two points)."""
    dm, _ = _lib_or_skip()
    rng = np.random.default_rng(3)
    pts = rng.random((500, 3))
    with pytest.raises(dm.DefmmError, match="n_src and n_trg"):
        dm.build_plan_stats(np.zeros((0, 3)), pts, 4, 8, 1.0)
    with pytest.raises(dm.DefmmError, match="n_src and n_trg"):
        dm.build_plan_stats(pts, np.zeros((0, 3)), 4, 8, 1.0)
    with pytest.raises(dm.DefmmError, match="coincide"):
        dm.build_plan_stats(pts[:1], pts[:1], 4, 8, 1.0)
    bad = pts.copy()
    bad[17, 1] = np.nan
    with pytest.raises(dm.DefmmError, match="non-finite"):
        dm.build_plan_stats(bad, pts, 4, 8, 1.0)
    bad[17, 1] = np.inf
    with pytest.raises(dm.DefmmError, match="non-finite"):
        dm.build_plan_stats(pts, bad, 4, 8, 1.0)
    with pytest.raises(dm.DefmmError, match="invalid argument"):
        dm.build_plan_stats(pts, pts, 4, 8, float("nan"))
    with pytest.raises(dm.DefmmError, match="no M2L level"):
        dm.build_plan_stats(pts[:2], pts[:2], 4, 8, 1.0)


def test_setup_mode_switch():
    """defmm_set_setup_mode (no device work): returns the previous mode and
    refuses an unknown one, leaving the mode unchanged."""
    dm, _ = _lib_or_skip()
    old = dm.set_setup_mode(dm.SETUP_HOST)
    try:
        assert dm.set_setup_mode(dm.SETUP_DEVICE) == dm.SETUP_HOST
        with pytest.raises(dm.DefmmError, match="unknown set-up mode"):
            dm.set_setup_mode(7)
        assert dm.set_setup_mode(dm.SETUP_DEVICE) == dm.SETUP_DEVICE
    finally:
        dm.set_setup_mode(old)
