"""GPU parity in the regimes of the benchmark configurations, at sizes the
oracle finishes in seconds:

  - the unit sphere at 128 wavelengths, order 5 (the 1M-point sphere's
    frequency regime: every level high-frequency, 384+ cones at the finest
    level), FP64 against the oracle to roundoff;
  - the clustered 4:1:1 ellipsoid at 128 wavelengths, order 5, FP64 to
    roundoff and mixed precision within 3e-6 of the oracle;
  - the clustered ellipsoid at 256 wavelengths, order 6, leaves <= 96
    points, 16 complex64 right-hand sides in one batched evaluation: each
    equals its single evaluation, and the first equals the oracle;
  - the 20,000-point edge-refined cube at 32 wavelengths with order 8, the
    order at which this method reaches relative error <= 1e-6 there: FP64
    and mixed precision against the FP64 direct sum.

Oracle and GPU run on the same inputs; FP64 tolerance relative L2 1e-12 and
relative max 1e-10 (different rounding, same operations).
"""
import sys
from concurrent.futures import ThreadPoolExecutor

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


def _err(p, exact):
    return np.sqrt((np.abs(p - exact) ** 2).sum() / (np.abs(exact) ** 2).sum())


def _plan(dm, pts, L, ncrit, kappa, prec):
    A = dm.IBFMM_Mat(precision=prec)
    A.setSourceParticles(pts)
    A.setTargetParticles(pts)
    A.prcmpt(L, ncrit, kappa)
    return A


STAT_KEYS = ["p2p_calls", "p2p_pairs", "hf_m2l", "lf_m2l", "src_slots", "trg_slots", "hf_leaves",
             "hf_m2m", "hf_l2l", "nlvl", "cpt", "max_stage"]


def test_sphere_128_wavelengths_fp64():
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    n, L, ncrit = 3000, 5, 6
    pts = wl.sphere(n, seed=5)
    kappa = wl.kappa_for(128, pts)
    q = wl.charges(n, seed=6)
    _say(f"case: sphere N={n} 128 wavelengths kappa={kappa:.3f} L={L} Ncrit={ncrit} fp64")
    po, so, _, ta = oa.fmm(pts, pts, q, L, ncrit, kappa)
    A = _plan(dm, pts, L, ncrit, kappa, dm.FP64)
    p = A.apply(q)
    sg = A.stats()
    A.destroy()
    l2, mx = _diff(p, po)
    e = _err(p, oa.direct(pts, q, pts, kappa))
    _say(f"  oracle {ta:.1f} s; hf_m2l {sg['hf_m2l']} slots {sg['src_slots']}; vs oracle rel L2 "
         f"{l2:.3e} max {mx:.3e}; vs direct sum {e:.3e}")
    for k in STAT_KEYS:
        assert sg[k] == so[k], k
    assert sg["cpt"] >= sg["nlvl"] - 1 and sg["hf_m2l"] > 50000
    assert l2 <= FP64_L2 and mx <= FP64_MAX


def test_ellipsoid_128_wavelengths_fp64_and_mixed():
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    n, L, ncrit = 20000, 5, 17
    pts = wl.ellipsoid_clusters(n, seed=3)
    kappa = wl.kappa_for(128, pts)
    q = wl.charges(n, seed=5).astype(np.complex64).astype(np.complex128)
    _say(f"case: clustered ellipsoid N={n} 128 wavelengths L={L} Ncrit={ncrit}")
    po, so, _, ta = oa.fmm(pts, pts, q, L, ncrit, kappa)
    idx = np.arange(0, n, n // 300)
    ex = oa.direct(pts, q, pts[idx], kappa)
    for prec, tol in ((dm.FP64, FP64_L2), (dm.MIXED, MIXED_L2)):
        A = _plan(dm, pts, L, ncrit, kappa, prec)
        p = A.apply(q)
        sg = A.stats()
        A.destroy()
        l2, mx = _diff(p, po)
        _say(f"  precision {prec}: vs oracle rel L2 {l2:.3e} max {mx:.3e}; vs direct sum "
             f"{_err(p[idx], ex):.3e} (hf_m2l {sg['hf_m2l']}, oracle {ta:.1f} s)")
        for k in STAT_KEYS:
            assert sg[k] == so[k], k
        assert l2 <= tol
        if prec == dm.FP64:
            assert mx <= FP64_MAX


def test_ellipsoid_256_wavelengths_16_rhs():
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    n, L, ncrit, nr = 20000, 6, 97, 16
    pts = wl.ellipsoid_clusters(n, seed=4)
    kappa = wl.kappa_for(256, pts)
    Q = np.stack([wl.charges(n, seed=300 + r) for r in range(nr)]).astype(np.complex64)
    Q = Q.astype(np.complex128)
    _say(f"case: clustered ellipsoid N={n} 256 wavelengths L={L} Ncrit={ncrit}, {nr} "
         f"right-hand sides")
    po, so, _, ta = oa.fmm(pts, pts, Q[0], L, ncrit, kappa)
    for prec, tol_single, tol_oracle in ((dm.FP64, 1e-13, FP64_L2), (dm.MIXED, 1e-6, MIXED_L2)):
        A = _plan(dm, pts, L, ncrit, kappa, prec)
        single = np.stack([A.apply(Q[r]) for r in range(nr)])
        A.set_rhs(nr)
        P = A.apply_batch(Q)
        sg = A.stats()
        A.destroy()
        l2, mx = _diff(P, single)
        lo, mo = _diff(P[0], po)
        _say(f"  precision {prec}: batch of {nr} vs single rel L2 {l2:.3e} max {mx:.3e}; "
             f"first vs oracle rel L2 {lo:.3e} (hf_m2l {sg['hf_m2l']}, oracle {ta:.1f} s)")
        assert sg["hf_m2l"] == so["hf_m2l"] and sg["hf_m2l"] > 10000
        assert l2 <= tol_single and mx <= 10 * tol_single
        assert lo <= tol_oracle


def test_cube20k_order8_reaches_1e6():
    """The 20,000-point edge-refined cube at 32 wavelengths (leaves <= 64):
    at order 8 the oracle's relative L2 error against the FP64 direct sum on
    1,000 sampled targets is 7.214e-7 (orders 6, 7, 9: 4.196e-5, 5.760e-6,
    8.902e-8, computed by the oracle on this input).  The FP64 path must
    give the same error, and the mixed path must stay below 1e-6."""
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts, kappa, _, ncrit = wl.make("cube20k_32l", seed=1)
    n = len(pts)
    q = wl.charges(n, seed=7)
    idx = np.linspace(0, n - 1, 1000).astype(np.int64)
    ex = oa.direct(pts, q, pts[idx], kappa)
    _say(f"case: edge-refined cube N={n} 32 wavelengths L=8 Ncrit={ncrit}")
    errs = {}
    for prec in (dm.FP64, dm.MIXED):
        A = _plan(dm, pts, 8, ncrit, kappa, prec)
        p = A.apply(q)
        A.destroy()
        errs[prec] = _err(p[idx], ex)
        _say(f"  precision {prec}: rel L2 vs direct sum {errs[prec]:.4e}")
    assert abs(errs[dm.FP64] - 7.214e-7) / 7.214e-7 < 2e-3
    assert errs[dm.MIXED] <= 1e-6


@pytest.mark.parametrize("L", [4, 6])
def test_laplace_kernel_at_zero_wavenumber(L):
    """The reference's Laplace kernel 1/(4 pi R) (include/thetis/
    thetis_kernels.hpp:35) is its Helmholtz kernel at kappa = 0: every level
    low-frequency, M2L from level 2 on the same Fourier-domain tables.  FP64
    against the oracle to roundoff, mixed within 3e-6, both against the
    direct sum (the oracle gives 1.390e-4 at L = 4 and 2.425e-6 at L = 6 on
    this input)."""
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts = wl.sphere(6000, seed=3)
    q = wl.charges(6000, seed=4)
    _say(f"case: Laplace (kappa = 0), sphere N=6000 L={L} Ncrit=33")
    po, so, _, _ = oa.fmm(pts, pts, q, L, 33, 0.0)
    idx = np.arange(0, 6000, 12)
    ex = oa.direct(pts, q, pts[idx], 0.0)
    for prec, tol in ((dm.FP64, FP64_L2), (dm.MIXED, MIXED_L2)):
        A = _plan(dm, pts, L, 33, 0.0, prec)
        p = A.apply(q)
        sg = A.stats()
        A.destroy()
        l2, mx = _diff(p, po)
        _say(f"  precision {prec}: vs oracle rel L2 {l2:.3e}; vs direct sum {_err(p[idx], ex):.3e} "
             f"(lf_m2l {sg['lf_m2l']}, hf_m2l {sg['hf_m2l']})")
        assert sg["lf_m2l"] == so["lf_m2l"] > 0 and sg["hf_m2l"] == 0
        assert l2 <= tol
