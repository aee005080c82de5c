"""The plan builder's octrees, dual walk, M2L and P2P lists on the GPU
(setup.hip) against the host builder (plan.cpp): for the same points and
parameters both give the same plan -- every list statistic equal (pair
counts, slots, tables, M2F rows, M2L entries and groups, P2P items,
table-store rows) and executions bit-identical, which needs the M2L entries
of every target slot in the same order.  Workloads cover HF and LF levels,
a 200,000-point sphere at 64 wavelengths, Laplace (kappa = 0), a
fixed-depth tree, different source and target clouds, a leaf target root,
and a tree too shallow for any M2L level (both builders refuse it alike)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SKIP = ("ms_tree", "ms_walk", "ms_tables", "ms_trans", "ms_m2l", "ms_p2p", "ms_upload",
        "setup_device")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _build(dm, mode, src, trg, L, ncrit, kappa, lvmax=-1, prec=None):
    old = dm.set_setup_mode(mode)
    try:
        A = dm.IBFMM_Mat(precision=dm.FP64 if prec is None else prec)
        A.setSourceParticles(src)
        A.setTargetParticles(trg)
        A.prcmpt(L, ncrit, kappa, lvmax)
    finally:
        dm.set_setup_mode(old)
    return A


def _compare(name, src, trg, L, ncrit, kappa, lvmax=-1, prec=None, nrhs=1):
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    _say(f"case: {name}: host vs device set-up, Ns={len(src)} Nt={len(trg)} L={L} "
         f"Ncrit={ncrit} kappa={kappa:.4f} lvmax={lvmax}")
    H = _build(dm, dm.SETUP_HOST, src, trg, L, ncrit, kappa, lvmax, prec)
    D = _build(dm, dm.SETUP_DEVICE, src, trg, L, ncrit, kappa, lvmax, prec)
    sh, sd = H.stats(), D.stats()
    assert sh["setup_device"] == 0 and sd["setup_device"] == 1
    diff = {k: (sh[k], sd[k]) for k in sh if k not in SKIP and sh[k] != sd[k]}
    _say(f"  walk ms host {sh['ms_walk']} device {sd['ms_walk']}, M2L lists ms host "
         f"{sh['ms_m2l']} device {sd['ms_m2l']}; {sd['hf_m2l']} HF + {sd['lf_m2l']} LF M2L, "
         f"{sd['p2p_calls']} P2P; differing stats {diff}")
    assert not diff
    Q = np.stack([wl.charges(len(src), seed=900 + r) for r in range(nrhs)])
    if nrhs > 1:
        H.set_rhs(nrhs)
        D.set_rhs(nrhs)
        ph, pd = H.apply_batch(Q), D.apply_batch(Q)
    else:
        ph, pd = H.apply(Q[0]), D.apply(Q[0])
    same = bool(np.array_equal(ph, pd))
    _say(f"  results identical: {same}")
    H.destroy()
    D.destroy()
    assert same
    return sd


def test_refined_cube_known_answer_plan(cube):
    """The reference's own test case (41,334 points, kappa 1, Ncrit 32, L 6):
    the device walk reproduces the reference's call counts."""
    sd = _compare("refined cube", cube, cube, 6, 32, 1.0)
    assert (sd["p2p_calls"], sd["hf_m2l"], sd["lf_m2l"]) == (129928, 336776, 17334)


@pytest.mark.parametrize("geo,n,nl,L,ncrit", [
    ("sphere", 20000, 32, 5, 33),
    ("ellipsoid", 30000, 64, 4, 65),
    ("cube_edges", 20000, 8, 6, 40),
])
def test_hf_workloads(geo, n, nl, L, ncrit):
    from defmm_amd import workloads as wl
    pts = {"sphere": lambda: wl.sphere(n, seed=31),
           "ellipsoid": lambda: wl.ellipsoid_clusters(n, seed=32),
           "cube_edges": lambda: wl.cube_edges(n, seed=33)}[geo]()
    sd = _compare(f"{geo} {nl} wavelengths", pts, pts, L, ncrit, wl.kappa_for(nl, pts))
    assert sd["hf_m2l"] > 0


def test_sphere_200k_64_wavelengths():
    """A larger plan in the benchmark's all-HF regime (200,000 points on the
    sphere, 64 wavelengths, order 5, leaf <= 64: about 20M M2L entries)."""
    from defmm_amd import workloads as wl
    pts = wl.sphere(200000, seed=43)
    sd = _compare("sphere 200k, 64 wavelengths", pts, pts, 5, 65, wl.kappa_for(64, pts))
    assert sd["hf_m2l"] > 10_000_000


def test_mixed_precision_several_rhs():
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts = wl.sphere(15000, seed=34)
    _compare("sphere, mixed precision, 4 RHS", pts, pts, 6, 33, wl.kappa_for(16, pts),
             prec=dm.MIXED, nrhs=4)


def test_distinct_source_and_target_clouds():
    from defmm_amd import workloads as wl
    src = wl.gaussian_clusters(12000, seed=35)
    trg = wl.sphere(9000, seed=36) * 0.8
    _compare("clusters -> sphere", src, trg, 5, 33, wl.kappa_for(16, src, trg))


def test_laplace_and_fixed_depth():
    from defmm_amd import workloads as wl
    pts = wl.cube_edges(15000, seed=37)
    sd = _compare("Laplace (kappa 0)", pts, pts, 5, 40, 0.0)
    assert sd["lf_m2l"] > 0 and sd["hf_m2l"] == 0
    u = wl.uniform(38, 3 * 8000).reshape(-1, 3)
    _compare("uniform cube, fixed depth 4", u, u, 4, 1, wl.kappa_for(4, u), lvmax=4)


def test_target_root_leaf():
    """A target cloud of fewer than Ncrit points: the target root is a leaf,
    so the walk ends at the root pair (all near field)."""
    from defmm_amd import workloads as wl
    src = wl.sphere(20000, seed=40)
    trg = wl.sphere(20, seed=41) * 0.5
    sd = _compare("20 targets, target root leaf", src, trg, 4, 33, wl.kappa_for(8, src))
    assert sd["p2p_calls"] == 1 and sd["hf_m2l"] + sd["lf_m2l"] == 0


def test_too_shallow_refused_alike():
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts = wl.sphere(300, seed=39)
    _say("case: root leaf (Ncrit above N): both builders refuse the plan")
    msgs = []
    for mode in (dm.SETUP_HOST, dm.SETUP_DEVICE):
        with pytest.raises(dm.DefmmError) as ei:
            _build(dm, mode, pts, pts, 4, 1000, 1.0)
        msgs.append(str(ei.value))
    assert msgs[0] == msgs[1]
