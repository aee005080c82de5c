"""Instance batches (defmm_batch_*): many independent FMM problems evaluated
by one HIP graph.  40 instances with the uneven mix of the benchmark's
instance-batch configuration, scaled down: sizes log-uniform in 2,000 -
20,000 points, geometry from {sphere, edge-refined cube, clustered 4:1:1
ellipsoid, 16 Gaussian clusters}, 4 - 32 wavelengths, orders 4 - 6, 1 - 4
right-hand sides, mixed precision for most and FP64 for every fifth.

Each instance's batched result must equal (bit for bit) its own evaluation
by defmm_execute_batch_device, on three launches of the graph with an eager
execution of one of the plans between them; three instances are compared
with the oracle as well (FP64 to roundoff, mixed within 3e-6)."""
import sys

import numpy as np
import pytest

import oracle_api as oa

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _diff(a, b):
    d = np.abs(a - b)
    return np.sqrt((d ** 2).sum() / (np.abs(b) ** 2).sum())


def _instances(count, seed=4242):
    from defmm_amd import workloads as wl
    u = wl.uniform(seed, 8 * count)
    geos = ["sphere", "cube_edges", "ellipsoid", "clusters"]
    out = []
    for i in range(count):
        n = int(10 ** (np.log10(2000) + (np.log10(20000) - np.log10(2000)) * u[8 * i]))
        geo = geos[min(3, int(u[8 * i + 1] * 4))]
        nl = 2 ** (2 + 3 * u[8 * i + 2])
        L = 4 + min(2, int(u[8 * i + 3] * 3))
        nr = 1 + min(3, int(u[8 * i + 4] * 4))
        prec = 0 if i % 5 == 0 else 1
        out.append((i, geo, n, nl, L, nr, prec))
    return out


def test_uneven_instance_batch_in_one_graph():
    import torch
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    dev = torch.device("cuda:0")
    mats, Qs, Pref, Pbat, meta = [], [], [], [], []
    for (i, geo, n, nl, L, nr, prec) in _instances(40):
        pts = wl.GEOMETRIES[geo](n, seed=300 + i)
        kappa = wl.kappa_for(nl, pts)
        A = dm.IBFMM_Mat(precision=prec)
        A.setSourceParticles(pts)
        A.setTargetParticles(pts)
        try:
            A.prcmpt(L, 33, kappa)
        except dm.DefmmError as e:
            _say(f"  instance {i} ({geo} N={n} {nl:.1f} wavelengths): {e}; not batched")
            continue
        A.set_rhs(nr)
        Q = np.stack([wl.charges(n, seed=700 + 10 * i + r) for r in range(nr)])
        mats.append(A)
        meta.append((i, geo, n, nl, L, nr, prec, pts, kappa, Q))
        Qs.append(torch.from_numpy(Q).to(dev))
        Pref.append(torch.zeros((nr, n), dtype=torch.complex128, device=dev))
        Pbat.append(torch.full((nr, n), 5 + 5j, dtype=torch.complex128, device=dev))
    _say(f"case: batch of {len(mats)} instances in one graph "
         f"({sum(m[2] for m in meta):,} points, {sum(m[5] for m in meta)} right-hand sides)")
    assert len(mats) >= 32
    s = torch.cuda.current_stream(dev)
    for A, m, q, p in zip(mats, meta, Qs, Pref):
        A.apply_batch_device(q.data_ptr(), p.data_ptr(), m[5], s.cuda_stream)
    torch.cuda.synchronize()
    B = dm.Batch(mats, [q.data_ptr() for q in Qs], [p.data_ptr() for p in Pbat],
                 [m[5] for m in meta], nbranch=8)
    scratch = torch.zeros_like(Pref[0])
    for launch in range(3):
        for p in Pbat:
            p.fill_(9 + 9j)
        # the plans' own eager work between the replays
        mats[0].apply_batch_device(Qs[0].data_ptr(), scratch.data_ptr(), meta[0][5], s.cuda_stream)
        torch.cuda.synchronize()
        B.execute(s.cuda_stream)
        torch.cuda.synchronize()
        for m, a, b in zip(meta, Pref, Pbat):
            assert torch.equal(a, b), f"launch {launch}: instance {m[0]} differs"
        assert torch.equal(scratch, Pref[0])
    _say("  every instance equals its own evaluation, on three launches with eager passes between")
    # three instances against the oracle (FP64 and mixed)
    # the first FP64 instance and the first two mixed ones
    picks = [k for k, m in enumerate(meta) if m[6] == 0][:1] + \
        [k for k, m in enumerate(meta) if m[6] == 1][:2]
    checked = 0
    for k in picks:
        i, geo, n, nl, L, nr, prec, pts, kappa, Q = meta[k]
        b = Pbat[k]
        ref = oa.fmm(pts, pts, Q[0], L, 33, kappa)[0]
        e = _diff(b[0].cpu().numpy(), ref)
        _say(f"  instance {i} ({geo} N={n} {nl:.1f} wavelengths L={L} precision {prec}): "
             f"vs oracle rel L2 {e:.3e}")
        assert e <= (1e-12 if prec == 0 else 3e-6)
        checked += 1
    assert checked == 3
    B.destroy()
    # reallocating a plan makes the batch stale
    B = dm.Batch(mats[:2], [q.data_ptr() for q in Qs[:2]], [p.data_ptr() for p in Pbat[:2]],
                 [m[5] for m in meta[:2]])
    mats[0].set_rhs(meta[0][5] + 1)
    with pytest.raises(dm.DefmmError, match="call out of sequence"):
        B.execute(s.cuda_stream)
    B.destroy()
    for A in mats:
        A.destroy()
