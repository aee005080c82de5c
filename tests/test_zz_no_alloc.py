"""The direct sum entry point that allocates nothing (defmm_direct_noalloc,
partial sums in a fixed array of the library) against the allocating one
(defmm_direct): bit-identical results on two calls in a row, the second with
a smaller input (fewer sources and targets, one target tile instead of two),
so that the second call reuses the array the first one filled."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_direct_noalloc_matches_direct():
    import torch
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    lib = dm.lib()
    fn = lib.defmm_direct_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    for ns, nt, seed in ((30000, 5000, 1), (7000, 1500, 2)):
        src = wl.sphere(ns, seed=seed)
        trg = np.stack([wl.uniform(40 + seed + k, nt) for k in range(3)], 1) * 2 - 1
        q = wl.charges(ns, seed=50 + seed)
        kappa = wl.kappa_for(8, src)
        _say(f"case: direct sum without allocation, {ns} sources, {nt} targets")
        ref = dm.direct(src, q, trg, kappa)
        sd = torch.from_numpy(np.ascontiguousarray(src)).to(dev)
        qd = torch.from_numpy(np.ascontiguousarray(q)).to(dev)
        td = torch.from_numpy(np.ascontiguousarray(trg)).to(dev)
        pd = torch.full((nt,), 3 + 3j, dtype=torch.complex128, device=dev)
        torch.cuda.synchronize()
        rc = fn(sd.data_ptr(), ns, qd.data_ptr(), td.data_ptr(), nt, kappa, pd.data_ptr(),
                stream.cuda_stream)
        assert rc == 0, lib.defmm_last_error()
        torch.cuda.synchronize()
        got = pd.cpu().numpy()
        d = np.abs(got - ref).max() / np.abs(ref).max()
        _say(f"  max rel difference {d:.3e}")
        assert np.array_equal(got, ref)
    # sizes beyond the maximum are rejected before anything is enqueued
    assert fn(sd.data_ptr(), 1 << 31, qd.data_ptr(), td.data_ptr(), 10, 1.0, pd.data_ptr(),
              stream.cuda_stream) == 1
