"""vine_solve_device_static, the entry point that allocates no device memory
(every buffer of the solve is cut from the library's static scratch array):
its answers against vine_solve_device's on the same inputs, two calls in a
row with the second one smaller, and configurations it refuses."""
import ctypes

import numpy as np
import pytest
import torch

from vine_amd import learn as L
from vine_amd import solver as S
from vine_amd import workload as W


def _static(dg, opts, mlp, src, dst, bnd):
    dev = torch.device("cuda:0")
    nq = len(src)
    ts = [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (src, dst, bnd)]
    c = torch.full((nq,), -1.0, dtype=torch.float32, device=dev)
    d = torch.full((nq,), -1.0, dtype=torch.float32, device=dev)
    f = torch.full((nq,), -1, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    st = S.Stats()
    rc = S._L().vine_solve_device_static(dg.h, ctypes.byref(opts), None if mlp is None else mlp.ctypes.data,
                                         0 if mlp is None else len(mlp), nq, ts[0].data_ptr(), ts[1].data_ptr(),
                                         ts[2].data_ptr(), c.data_ptr(), d.data_ptr(), f.data_ptr(), None,
                                         ctypes.byref(st))
    torch.cuda.synchronize()
    return rc, c.cpu().numpy(), d.cpu().numpy(), f.cpu().numpy(), st


@pytest.mark.gpu
@pytest.mark.parametrize("policy", [S.DELTA_MLP, S.DELTA_FIXED])
def test_static_matches_device(gpu, policy):
    g, q = W.preset("grid64")
    dg = S.DeviceGraph(g)
    mlp = np.load(L.DEFAULT_MODEL).astype(np.float32) if policy == S.DELTA_MLP else None
    kw = dict(max_batch=16, prune=S.PRUNE_BOUNDS, delta_policy=policy,
              delta_fixed=4.0 if policy == S.DELTA_MLP else 300.0, rounds_per_sync=8)
    print("static scratch solve, policy %d" % policy, flush=True)
    with S.Solver(dg, **kw) as s:
        if mlp is not None:
            s.set_mlp(mlp)
        ref = s.solve(q.src, q.dst, q.bound, paths=False)
    opts = S.default_options(**kw)
    for n in (100, 37):
        rc, c, d, f, st = _static(dg, opts, mlp, q.src[:n], q.dst[:n], q.bound[:n])
        assert rc == S.VINE_OK, S._L().vine_last_error()
        assert np.array_equal(f.astype(bool), ref.feasible[:n]), n
        assert np.array_equal(c, ref.cost[:n]) and np.array_equal(d, ref.delay[:n]), n
        assert st.batches == (n + 15) // 16 and st.expanded > 0


@pytest.mark.gpu
def test_static_refuses(gpu):
    g, q = W.preset("grid64")
    dg = S.DeviceGraph(g)
    L_ = S._L()
    for kw in ({"max_batch": 100000}, {"max_batch": 16, "concurrency": 2}, {"max_batch": 16, "use_graph": 1},
               {"max_batch": 16, "time_kernels": 1}):
        opts = S.default_options(prune=S.PRUNE_BOUNDS, **kw)
        rc, c, d, f, st = _static(dg, opts, None, q.src[:4], q.dst[:4], q.bound[:4])
        assert rc == S.VINE_E_ARG, kw
        assert np.all(f == -1)  # nothing written
    assert b"static scratch" in L_.vine_last_error() or b"allocations" in L_.vine_last_error()
    opts = S.default_options(prune=S.PRUNE_BOUNDS, max_batch=100000)
    _static(dg, opts, None, q.src[:4], q.dst[:4], q.bound[:4])
    assert b"static scratch" in L_.vine_last_error()
