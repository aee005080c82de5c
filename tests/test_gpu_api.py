"""GPU tests of the C ABI's execution paths and plan life cycle:

  - device-pointer executions (single and batched), which the benchmark
    times, equal the host-array executions;
  - switching the arithmetic of an existing plan FP64 -> MIXED -> FP64 gives
    the results of fresh plans of each arithmetic;
  - mixed precision at order 8 (the column M2L kernel, 2 right-hand sides per
    launch in float) with 5 right-hand sides against FP64;
  - a reallocation that cannot be satisfied (over the plan's memory limit)
    returns an error and leaves the plan usable with its previous settings;
    when even those cannot be restored, executions return an error instead
    of running;
  - executions on two streams are serialised by the plan.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _diff(a, b):
    d = np.abs(a - b)
    return np.sqrt((d ** 2).sum() / (np.abs(b) ** 2).sum()), d.max() / np.abs(b).max()


def _cloud(n=6000, seed=12):
    from defmm_amd import workloads as wl
    pts = wl.ellipsoid_clusters(n, seed=seed)
    return pts, wl.kappa_for(16, pts)


def _plan(dm, pts, L, ncrit, kappa, prec):
    A = dm.IBFMM_Mat(precision=prec)
    A.setSourceParticles(pts)
    A.setTargetParticles(pts)
    A.prcmpt(L, ncrit, kappa)
    return A


def test_device_execute_matches_host():
    import torch
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts, kappa = _cloud()
    n, nr = len(pts), 3
    Q = np.stack([wl.charges(n, seed=500 + r) for r in range(nr)])
    dev = torch.device("cuda:0")
    for prec in (dm.FP64, dm.MIXED):
        _say(f"case: device-pointer execution vs host arrays, ellipsoid N={n} L=5, precision {prec}")
        A = _plan(dm, pts, 5, 65, kappa, prec)
        ph = A.apply(Q[0])
        qd = torch.from_numpy(Q[0]).to(dev)
        pd = torch.full((n,), 7 + 7j, dtype=torch.complex128, device=dev)
        s = torch.cuda.current_stream(dev)
        A.apply_device(qd.data_ptr(), pd.data_ptr(), s.cuda_stream)
        torch.cuda.synchronize()
        assert np.array_equal(pd.cpu().numpy(), ph)
        A.set_rhs(nr)
        Ph = A.apply_batch(Q)
        Qd = torch.from_numpy(Q).to(dev)
        Pd = torch.zeros((nr, n), dtype=torch.complex128, device=dev)
        A.apply_batch_device(Qd.data_ptr(), Pd.data_ptr(), nr, s.cuda_stream)
        torch.cuda.synchronize()
        A.destroy()
        assert np.array_equal(Pd.cpu().numpy(), Ph)
        l2, _ = _diff(Ph[0], ph)
        _say(f"  batched first right-hand side vs single rel L2 {l2:.3e}")
        assert l2 <= (1e-13 if prec == dm.FP64 else 1e-6)


def test_set_precision_roundtrip_equals_fresh_plans():
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts, kappa = _cloud()
    q = wl.charges(len(pts), seed=41)
    _say("case: FP64 -> MIXED -> FP64 on one plan vs fresh plans, L=6")
    fresh = {}
    for prec in (dm.FP64, dm.MIXED):
        B = _plan(dm, pts, 6, 65, kappa, prec)
        fresh[prec] = B.apply(q)
        B.destroy()
    A = _plan(dm, pts, 6, 65, kappa, dm.FP64)
    p0 = A.apply(q)
    A.set_precision(dm.MIXED)
    p1 = A.apply(q)
    A.set_precision(dm.FP64)
    p2 = A.apply(q)
    A.destroy()
    assert np.array_equal(p0, fresh[dm.FP64])
    assert np.array_equal(p1, fresh[dm.MIXED])
    assert np.array_equal(p2, fresh[dm.FP64])
    l2, _ = _diff(p1, p0)
    _say(f"  mixed vs fp64 rel L2 {l2:.3e}")
    assert l2 <= 3e-6


def test_mixed_order8_five_rhs():
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts, kappa = _cloud(5000, seed=13)
    n, nr = len(pts), 5
    Q = np.stack([wl.charges(n, seed=600 + r) for r in range(nr)])
    _say(f"case: order 8, {nr} right-hand sides (float M2L 2 + 2 + 1 per launch) vs FP64")
    A = _plan(dm, pts, 8, 65, kappa, dm.FP64)
    st = A.stats()
    assert st["hf_m2l"] + st["lf_m2l"] > 0
    ref = np.stack([A.apply(Q[r]) for r in range(nr)])
    A.set_precision(dm.MIXED)
    single = np.stack([A.apply(Q[r]) for r in range(nr)])
    A.set_rhs(nr)
    P = A.apply_batch(Q)
    A.destroy()
    l2s, _ = _diff(P, single)
    l2r = max(_diff(P[r], ref[r])[0] for r in range(nr))
    _say(f"  batch vs single (mixed) rel L2 {l2s:.3e}; worst mixed vs FP64 {l2r:.3e}")
    assert l2s <= 1e-6
    assert l2r <= 3e-6


def test_failed_reallocation_keeps_or_invalidates_plan():
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts, kappa = _cloud(4000, seed=14)
    q = wl.charges(len(pts), seed=42)
    _say("case: reallocation over the plan's memory limit")
    A = _plan(dm, pts, 5, 65, kappa, dm.FP64)
    p0 = A.apply(q)
    used = A.device_bytes()
    A.set_memory_limit(used + 4096)
    with pytest.raises(dm.DefmmError, match="out of memory"):
        A.set_rhs(16)
    # previous settings restored: same width, same results
    assert A.device_bytes() == used
    assert np.array_equal(A.apply(q), p0)
    # a limit below what the plan already holds: the restore fails as well,
    # and executions refuse to run
    A.set_memory_limit(1)
    with pytest.raises(dm.DefmmError, match="out of memory"):
        A.set_precision(dm.MIXED)
    with pytest.raises(dm.DefmmError, match="call out of sequence"):
        A.apply(q)
    # lifting the limit and asking again makes it usable
    A.set_memory_limit(0)
    A.set_precision(dm.FP64)
    assert np.array_equal(A.apply(q), p0)
    A.destroy()


def test_executions_on_two_streams_are_serialised():
    import torch
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts, kappa = _cloud()
    n = len(pts)
    q1, q2 = wl.charges(n, seed=43), wl.charges(n, seed=44)
    _say("case: executions of one plan on two streams")
    A = _plan(dm, pts, 5, 65, kappa, dm.FP64)
    ref1, ref2 = A.apply(q1), A.apply(q2)
    dev = torch.device("cuda:0")
    sa, sb = torch.cuda.Stream(dev), torch.cuda.Stream(dev)
    q1d, q2d = torch.from_numpy(q1).to(dev), torch.from_numpy(q2).to(dev)
    p1d = torch.zeros(n, dtype=torch.complex128, device=dev)
    p2d = torch.zeros(n, dtype=torch.complex128, device=dev)
    torch.cuda.synchronize()
    for _ in range(3):
        A.apply_device(q1d.data_ptr(), p1d.data_ptr(), sa.cuda_stream)
        A.apply_device(q2d.data_ptr(), p2d.data_ptr(), sb.cuda_stream)
    torch.cuda.synchronize()
    A.destroy()
    assert np.array_equal(p1d.cpu().numpy(), ref1)
    assert np.array_equal(p2d.cpu().numpy(), ref2)


def test_p2p_matrix_cores_match_vector_units():
    """The mixed-precision near field on the matrix cores (16 right-hand
    sides per kernel value, V_MFMA_F32_16X16X4_F32) against the vector-unit
    kernel (8 per kernel value) and FP64: 32 right-hand sides (two launches of
    16) on a leaf-heavy clustered ellipsoid, plan width 16."""
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts = wl.ellipsoid_clusters(12000, seed=15)
    kappa = wl.kappa_for(64, pts)
    n, nr = len(pts), 32
    Q = np.stack([wl.charges(n, seed=800 + r) for r in range(nr)])
    _say(f"case: near field on matrix cores, ellipsoid N={n} 64 wavelengths L=5 Ncrit=97, "
         f"{nr} right-hand sides")
    A = _plan(dm, pts, 5, 97, kappa, dm.FP64)
    st = A.stats()
    assert st["p2p_pairs"] > 10_000_000
    A.set_rhs(16)
    ref = A.apply_batch(Q)
    A.set_precision(dm.MIXED)
    A.set_option(dm.OPT_P2P_MFMA, 0)
    valu = A.apply_batch(Q)
    A.set_option(dm.OPT_P2P_MFMA, 1)
    mfma = A.apply_batch(Q)
    mfma2 = A.apply_batch(Q)
    A.destroy()
    assert np.array_equal(mfma, mfma2)
    d_valu = max(_diff(mfma[r], valu[r])[0] for r in range(nr))
    d_ref = max(_diff(mfma[r], ref[r])[0] for r in range(nr))
    _say(f"  matrix cores vs vector units: worst rel L2 {d_valu:.3e}; vs FP64 {d_ref:.3e}")
    assert d_valu <= 1e-6
    assert d_ref <= 3e-6


@pytest.mark.parametrize("L", [4, 5, 8, 9])
def test_l2p_matrix_cores_match_vector_units(L):
    """L2P of the mixed path on the matrix cores (16 right-hand sides per pass
    over a leaf's slots, tiles of 32 points up to L^3 = 512, else 16) against
    the vector-unit kernel, and the whole pass against FP64: 16 right-hand
    sides on the edge-refined cube (leaves of up to 64 points: several point
    tiles per leaf), plan width 16."""
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts = wl.cube_edges(8000, seed=16)
    kappa = wl.kappa_for(16, pts)
    n, nr = len(pts), 16
    Q = np.stack([wl.charges(n, seed=900 + r) for r in range(nr)])
    _say(f"case: L2P on matrix cores, edge-refined cube N={n} 16 wavelengths L={L} Ncrit=65, "
         f"{nr} right-hand sides")
    A = _plan(dm, pts, L, 65, kappa, dm.FP64)
    assert A.stats()["hf_l2p"] + A.stats()["lf_l2p"] > 0
    A.set_rhs(nr)
    ref = A.apply_batch(Q)
    A.set_precision(dm.MIXED)
    A.set_option(dm.OPT_L2P_MFMA, 0)
    valu = A.apply_batch(Q)
    A.set_option(dm.OPT_L2P_MFMA, 1)
    mfma = A.apply_batch(Q)
    A.destroy()
    d_valu = max(_diff(mfma[r], valu[r])[0] for r in range(nr))
    d_ref = max(_diff(mfma[r], ref[r])[0] for r in range(nr))
    _say(f"  matrix cores vs vector units: worst rel L2 {d_valu:.3e}; vs FP64 {d_ref:.3e}")
    assert d_valu <= 1e-6
    assert d_ref <= 3e-6


@pytest.mark.parametrize("L,ncrit", [(5, 97), (8, 65)])
def test_p2m_matrix_cores_match_vector_units(L, ncrit):
    """P2M of the mixed path on the matrix cores (16 right-hand sides, leaves
    of at most 64 points; with Ncrit 97 the larger leaves stay on the vector
    units in the same pass) against the vector-unit kernel, and the whole
    pass against FP64."""
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts = wl.ellipsoid_clusters(10000, seed=17)
    kappa = wl.kappa_for(32, pts)
    n, nr = len(pts), 16
    Q = np.stack([wl.charges(n, seed=950 + r) for r in range(nr)])
    _say(f"case: P2M on matrix cores, ellipsoid N={n} 32 wavelengths L={L} Ncrit={ncrit}, "
         f"{nr} right-hand sides")
    A = _plan(dm, pts, L, ncrit, kappa, dm.FP64)
    A.set_rhs(nr)
    ref = A.apply_batch(Q)
    A.set_precision(dm.MIXED)
    A.set_option(dm.OPT_P2M_MFMA, 0)
    valu = A.apply_batch(Q)
    A.set_option(dm.OPT_P2M_MFMA, 1)
    mfma = A.apply_batch(Q)
    A.destroy()
    d_valu = max(_diff(mfma[r], valu[r])[0] for r in range(nr))
    d_ref = max(_diff(mfma[r], ref[r])[0] for r in range(nr))
    _say(f"  matrix cores vs vector units: worst rel L2 {d_valu:.3e}; vs FP64 {d_ref:.3e}")
    assert d_valu <= 1e-6
    assert d_ref <= 3e-6


def test_autotune_keeps_results():
    """The autotuner times the M2L storage variants and the matrix-core
    options on the plan's own buffers, keeps the fastest, and the plan's
    results stay those of the mixed path (within float roundoff of the
    untuned plan, which runs other kernels for some stages)."""
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts = wl.cube_edges(6000, seed=18)
    kappa = wl.kappa_for(16, pts)
    n, nr = len(pts), 16
    Q = np.stack([wl.charges(n, seed=990 + r) for r in range(nr)])
    _say(f"case: autotune, edge-refined cube N={n} 16 wavelengths L=6, {nr} right-hand sides")
    A = _plan(dm, pts, 6, 65, kappa, dm.MIXED)
    A.set_rhs(nr)
    before = A.apply_batch(Q)
    rep = A.autotune(nr, reps=2)
    _say("  " + rep.replace("\n", "\n  "))
    after = A.apply_batch(Q)
    A.destroy()
    assert "kept m2l" in rep and "p2p_mfma" in rep and "l2p_mfma" in rep and "p2m_mfma" in rep
    d = max(_diff(after[r], before[r])[0] for r in range(nr))
    _say(f"  after vs before: worst rel L2 {d:.3e}")
    assert d <= 1e-6


@pytest.mark.parametrize("L,prec,nr", [(4, "mixed", 5), (5, "fp64", 3), (5, "mixed", 3),
                                       (7, "fp64", 2), (7, "mixed", 3)])
def test_m2l_sliced_sweep_identical(L, prec, nr):
    """The grouped M2L kernel sweeping its entries once per frequency slice
    (DEFMM_OPT_M2L_SLICED) equals the single sweep bit for bit: each
    accumulator sums the same terms in the same order.  Several right-hand
    sides (launches of NR = 1 .. 4) and the stored and batched variants."""
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts = wl.sphere(12000, seed=23)
    kappa = wl.kappa_for(32, pts)
    n = len(pts)
    Q = np.stack([wl.charges(n, seed=1030 + r) for r in range(nr)])
    p = dm.FP64 if prec == "fp64" else dm.MIXED
    _say(f"case: sliced M2L sweep, sphere N={n} 32 wavelengths L={L}, {prec}, {nr} RHS")
    A = _plan(dm, pts, L, 33, kappa, p)
    A.set_rhs(nr)
    for mode in (dm.M2L_STORED, dm.M2L_BATCHED):
        A.set_m2l(mode, batch_rows=1024)
        A.set_option(dm.OPT_M2L_SLICED, 0)
        ref = A.apply_batch(Q)
        A.set_option(dm.OPT_M2L_SLICED, 1)
        out = A.apply_batch(Q)
        same = bool(np.array_equal(out, ref))
        _say(f"  m2l mode {mode}: identical {same}")
        assert same
    st = A.stats()
    A.destroy()
    assert st["hf_m2l"] > 0


def test_device_out_of_memory_keeps_plan():
    """A reallocation that the device cannot satisfy (hipMalloc fails, as
    opposed to the plan's own memory limit) returns DEFMM_EMEMORY and the plan
    keeps its previous settings and results: the failed allocation's error is
    not reported again by the restored upload or the next execution.  A 2M-point
    ellipsoid at order 6 in mixed precision with stored Fourier rows takes
    about 39 GB per right-hand side: 16 do not fit in one MI355X."""
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts, kappa, L, ncrit = wl.make("ellipsoid2m_256l_L6")
    n = len(pts)
    q = wl.charges(n, seed=1077)
    _say(f"case: device out of memory, ellipsoid N={n} L={L} mixed stored, set_rhs(16)")
    A = _plan(dm, pts, L, ncrit, kappa, dm.MIXED)
    A.set_m2l(dm.M2L_STORED)
    before = A.apply(q)
    with pytest.raises(dm.DefmmError, match="out of memory.*keeps its previous settings"):
        A.set_rhs(16)
    after = A.apply(q)
    A.destroy()
    assert np.array_equal(after, before)


def test_sixty_four_rhs_in_one_pass():
    """The widest pass (64 right-hand sides, the largest width): four chunks
    of 16 on the matrix cores and M2L launches of 2 (order 8) in float.  Every
    fifth right-hand side against its own FP64 evaluation (within 3e-6), and
    the pass equals the same right-hand sides evaluated 16 at a time."""
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts = wl.cube_edges(6000, seed=31)
    kappa = wl.kappa_for(16, pts)
    n, nr = len(pts), 64
    Q = np.stack([wl.charges(n, seed=2000 + r) for r in range(nr)])
    _say(f"case: 64 right-hand sides, edge-refined cube N={n} 16 wavelengths L=8, mixed")
    A = _plan(dm, pts, 8, 65, kappa, dm.FP64)
    ref = {r: A.apply(Q[r]) for r in range(0, nr, 5)}
    A.set_precision(dm.MIXED)
    A.set_rhs(nr)
    P = A.apply_batch(Q)
    A.set_rhs(16)
    P16 = np.concatenate([A.apply_batch(Q[k:k + 16]) for k in range(0, nr, 16)])
    A.destroy()
    worst = max(_diff(P[r], ref[r])[0] for r in ref)
    _say(f"  worst rel L2 vs FP64 over {len(ref)} right-hand sides {worst:.3e}")
    assert worst <= 3e-6
    assert np.array_equal(P, P16)
