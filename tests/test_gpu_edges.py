"""Edge cases at the C ABI on the GPU: empty query sets, s == t, unreachable
targets, zero bounds, bounds and ids out of range, a path buffer too small,
fewer queries than internal parts; for the main solver and the family
solver."""
import numpy as np
import pytest
import torch

from vine_amd import solver as S
from vine_amd import workload as W


def _line_graph():
    # 0 -> 1 -> 2 -> 3, and 4 isolated; (len, delay) = (2, 3) per arc
    row = np.array([0, 1, 2, 3, 3, 3], np.int32)
    dst = np.array([1, 2, 3], np.int32)
    return W.from_csr(row, dst, np.full(3, 2, np.float32), np.full(3, 3, np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("conc", [0, 3])
def test_edges_main_solver(gpu, conc):
    g = _line_graph()
    print("edge cases, concurrency %d" % conc, flush=True)
    with S.Solver(g, max_batch=4, prune=S.PRUNE_BOUNDS, concurrency=conc) as s:
        a = s.solve(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
        assert a.cost.shape == (0,) and a.path_off.tolist() == [0]
        src = np.array([2, 0, 0, 0, 3], np.int32)
        dst = np.array([2, 3, 3, 4, 0], np.int32)
        bnd = np.array([0, 9, 8, 100, 100], np.float32)
        a = s.solve(src, dst, bnd)
        assert a.feasible.tolist() == [True, True, False, False, False]
        assert (a.cost[0], a.delay[0]) == (0, 0) and a.path_of(0).tolist() == [2]
        assert (a.cost[1], a.delay[1]) == (6, 9) and a.path_of(1).tolist() == [0, 1, 2, 3]
        assert np.all(np.isinf(a.cost[2:]))
        # a path buffer too small: answers written, VINE_E_OVERFLOW with the needed size
        rc_cost = np.zeros(5, np.float32)
        rc_delay = np.zeros(5, np.float32)
        feas = np.zeros(5, np.int32)
        off = np.zeros(6, np.int64)
        path = np.zeros(2, np.int32)
        L = S._L()
        st = S.Stats()
        rc = L.vine_solve(s.h, 5, src.ctypes.data, dst.ctypes.data, bnd.ctypes.data, rc_cost.ctypes.data,
                          rc_delay.ctypes.data, feas.ctypes.data, off.ctypes.data, path.ctypes.data, 2, st)
        assert rc == S.VINE_E_OVERFLOW and off[-1] == 5
        assert np.array_equal(rc_cost, a.cost) and np.array_equal(feas.astype(bool), a.feasible)
        with pytest.raises(S.VineError):
            s.solve(np.array([0], np.int32), np.array([5], np.int32), np.array([1], np.float32))
        with pytest.raises(S.VineError):
            s.solve(np.array([0], np.int32), np.array([3], np.int32), np.array([2.0 ** 24], np.float32))
        # device entry with fewer queries than parts
        dev = torch.device("cuda:0")
        ts = [torch.from_numpy(x).to(dev) for x in (src[:2], dst[:2], bnd[:2])]
        c = torch.empty(2, dtype=torch.float32, device=dev)
        d = torch.empty(2, dtype=torch.float32, device=dev)
        f = torch.empty(2, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        s.solve_device(ts[0].data_ptr(), ts[1].data_ptr(), ts[2].data_ptr(), c.data_ptr(), d.data_ptr(),
                       f.data_ptr(), 2)
        assert c.cpu().tolist() == [0.0, 6.0] and f.cpu().tolist() == [1, 1]


@pytest.mark.gpu
def test_edges_family_solver(gpu):
    g = _line_graph()
    print("edge cases, family solver", flush=True)
    with S.FamilySolver(g, S.FAMILY_HOP, max_batch=2) as fs:
        a = fs.solve(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
        assert a.label.shape == (0, 4)
        a = fs.solve(np.array([1, 0, 0], np.int32), np.array([1, 3, 3], np.int32), np.array([0, 9, 9], np.float32),
                     hop_L=np.array([0, 3, 2], np.float32))
        assert a.feasible.tolist() == [True, True, False]
        assert a.label[0].tolist() == [0, 0, 0, 0] and a.label[1].tolist() == [6, 9, 3, 0]
        assert a.path_of(1).tolist() == [0, 1, 2, 3]


@pytest.mark.gpu
@pytest.mark.parametrize("conc", [0, 3])
def test_paths_without_resolving(gpu, conc):
    """A path buffer too small: vine_solve returns VINE_E_OVERFLOW with the size
    it needs, and vine_solver_paths reads the same batch's paths from the
    device without solving again; a query count other than the last solve's
    is refused."""
    g, q = W.preset("grid64")
    print("vine_solver_paths, concurrency %d" % conc, flush=True)
    with S.Solver(g, max_batch=100, prune=S.PRUNE_BOUNDS, concurrency=conc) as s:
        full = s.solve(q.src, q.dst, q.bound)
        L = S._L()
        nq = len(q)
        cost = np.zeros(nq, np.float32)
        delay = np.zeros(nq, np.float32)
        feas = np.zeros(nq, np.int32)
        off = np.zeros(nq + 1, np.int64)
        small = np.zeros(8, np.int32)
        rc = L.vine_solve(s.h, nq, q.src.ctypes.data, q.dst.ctypes.data, q.bound.ctypes.data, cost.ctypes.data,
                          delay.ctypes.data, feas.ctypes.data, off.ctypes.data, small.ctypes.data, 8, None)
        assert rc == S.VINE_E_OVERFLOW and off[-1] == full.path_off[-1]
        assert np.array_equal(cost, full.cost)
        # still too small: the offsets, then VINE_E_OVERFLOW again
        off2 = np.zeros(nq + 1, np.int64)
        assert L.vine_solver_paths(s.h, nq, off2.ctypes.data, small.ctypes.data, 8) == S.VINE_E_OVERFLOW
        assert np.array_equal(off2, full.path_off)
        path = np.zeros(int(off[-1]), np.int32)
        assert L.vine_solver_paths(s.h, nq, off2.ctypes.data, path.ctypes.data, int(off[-1])) == S.VINE_OK
        assert np.array_equal(path, full.path) and np.array_equal(off2, full.path_off)
        assert L.vine_solver_paths(s.h, nq - 1, off2.ctypes.data, path.ctypes.data, int(off[-1])) == S.VINE_E_ARG
        # the wrapper takes that route by itself when its first buffer is short
        a = s.solve(q.src, q.dst, q.bound, path_cap=16)
        assert np.array_equal(a.path, full.path) and np.array_equal(a.path_off, full.path_off)
