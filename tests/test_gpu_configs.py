"""GPU parity on reduced-size versions of the benchmark configurations the
headline run does not cover: the clustered 4:1:1 ellipsoid in FP64 and mixed
precision (L = 5, 6, 7), several right-hand sides on the ellipsoid (L = 6,
leaf <= 96; L = 7), and a batch of uneven instances (geometry, size,
wavelengths, order and right-hand sides drawn per instance).  Results are
compared with the oracle on the same inputs (FP64 to relative L2 1e-12 / max
1e-10, mixed to relative L2 3e-6), except the L = 7 right-hand-side test,
which compares batches with single evaluations (the L = 7 single evaluation
is the one pinned to the oracle above).  Errors against the FP64 direct sum on
sampled targets are printed.
"""
import sys

import numpy as np
import pytest

import oracle_api as oa

pytestmark = pytest.mark.gpu

FP64_L2 = 1e-12
FP64_MAX = 1e-10
MIXED_L2 = 3e-6


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _diff(a, b):
    d = np.abs(a - b)
    return np.sqrt((d ** 2).sum() / (np.abs(b) ** 2).sum()), d.max() / np.abs(b).max()


def _plan(dm, pts, L, ncrit, kappa, prec):
    A = dm.IBFMM_Mat(precision=prec)
    A.setSourceParticles(pts)
    A.setTargetParticles(pts)
    A.prcmpt(L, ncrit, kappa)
    return A


@pytest.mark.parametrize("L", [5, 6, 7])
def test_ellipsoid_fp64_and_mixed(L):
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    n = 8000
    pts = wl.ellipsoid_clusters(n, seed=3)
    kappa = wl.kappa_for(16, pts)
    q = wl.charges(n, seed=5).astype(np.complex64).astype(np.complex128)  # complex64 charges
    _say(f"case: clustered ellipsoid N={n} 16 wavelengths L={L} Ncrit=65")
    po, _, _, _ = oa.fmm(pts, pts, q, L, 65, kappa)
    idx = np.arange(0, n, n // 400)
    ex = oa.direct(pts, q, pts[idx], kappa)
    for prec, tol in ((dm.FP64, FP64_L2), (dm.MIXED, MIXED_L2)):
        A = _plan(dm, pts, L, 65, kappa, prec)
        p = A.apply(q)
        A.destroy()
        l2, mx = _diff(p, po)
        e = np.sqrt((np.abs(p[idx] - ex) ** 2).sum() / (np.abs(ex) ** 2).sum())
        _say(f"  precision {prec}: vs oracle rel L2 {l2:.3e} max {mx:.3e}; vs direct sum {e:.3e}")
        assert l2 <= tol
        if prec == dm.FP64:
            assert mx <= FP64_MAX


def test_ellipsoid_multi_rhs_vs_oracle():
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    n, nr = 8000, 3
    pts = wl.ellipsoid_clusters(n, seed=4)
    kappa = wl.kappa_for(32, pts)
    Q = np.stack([wl.charges(n, seed=60 + r) for r in range(nr)])
    _say(f"case: clustered ellipsoid N={n} 32 wavelengths L=6 Ncrit=97, {nr} right-hand sides")
    ref = np.stack([oa.fmm(pts, pts, Q[r], 6, 97, kappa)[0] for r in range(nr)])
    for prec, tol in ((dm.FP64, FP64_L2), (dm.MIXED, MIXED_L2)):
        A = _plan(dm, pts, 6, 97, kappa, prec)
        A.set_rhs(nr)
        P = A.apply_batch(Q)
        A.destroy()
        for r in range(nr):
            l2, mx = _diff(P[r], ref[r])
            _say(f"  precision {prec} rhs {r}: vs oracle rel L2 {l2:.3e} max {mx:.3e}")
            assert l2 <= tol


@pytest.mark.parametrize("mode", [1, 2])
def test_ellipsoid_multi_rhs_L7_matches_single(mode):
    """Order 7 with several right-hand sides: the float M2L kernels take 2
    right-hand sides per launch there (groups of 3 targets, compiled for 6
    waves per SIMD), FP64 one; 3 right-hand sides run as 2 + 1 and 1 + 1 + 1.
    Each must equal its single-right-hand-side evaluation, with the stored
    (mode 1) and the batched (mode 2) M2L variant."""
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    n, nr = 8000, 3
    pts = wl.ellipsoid_clusters(n, seed=3)
    kappa = wl.kappa_for(16, pts)
    Q = np.stack([wl.charges(n, seed=80 + r) for r in range(nr)])
    for prec, tol in ((dm.FP64, 1e-13), (dm.MIXED, 1e-6)):
        _say(f"case: clustered ellipsoid N={n} 16 wavelengths L=7, {nr} right-hand sides, "
             f"M2L mode {mode}, precision {prec}")
        A = _plan(dm, pts, 7, 65, kappa, prec)
        A.set_m2l(mode)
        single = np.stack([A.apply(Q[r]) for r in range(nr)])
        A.set_rhs(nr)
        P = A.apply_batch(Q)
        A.destroy()
        l2, mx = _diff(P, single)
        _say(f"  batch vs single rel L2 {l2:.3e} max {mx:.3e}")
        assert l2 <= tol and mx <= 10 * tol


def test_uneven_instance_batch():
    """Instances of different geometry, size (log-uniform), wavelengths,
    order and number of right-hand sides, evaluated one after another in
    mixed precision and in FP64."""
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    u = wl.uniform(777, 64)
    geos = ["sphere", "cube_edges", "ellipsoid", "clusters"]
    for i in range(6):
        n = int(10 ** (3.3 + 0.7 * u[6 * i]))
        geo = geos[i % 4]
        nl = 2 ** (2 + 2 * u[6 * i + 2])
        L = 4 + int(u[6 * i + 3] * 3) % 3
        nr = 1 + int(u[6 * i + 4] * 3) % 3
        pts = wl.GEOMETRIES[geo](n, seed=100 + i)
        kappa = wl.kappa_for(nl, pts)
        Q = np.stack([wl.charges(n, seed=200 + 10 * i + r) for r in range(nr)])
        _say(f"case: instance {i}: {geo} N={n} {nl:.1f} wavelengths L={L} nrhs={nr}")
        try:
            ref = np.stack([oa.fmm(pts, pts, Q[r], L, 33, kappa)[0] for r in range(nr)])
        except RuntimeError:
            _say("  oracle: no M2L level for this instance (tree too shallow); skipped")
            continue
        for prec, tol in ((dm.FP64, FP64_L2), (dm.MIXED, MIXED_L2)):
            A = _plan(dm, pts, L, 33, kappa, prec)
            A.set_rhs(nr)
            P = A.apply_batch(Q)
            A.destroy()
            l2 = max(_diff(P[r], ref[r])[0] for r in range(nr))
            _say(f"  precision {prec}: worst rel L2 vs oracle {l2:.3e}")
            assert l2 <= tol
