"""Shared checks for the parity tests."""
import numpy as np


def check_path(g, path, s, t, cost, delay, bound):
    """A returned path must be an s->t arc sequence whose len sums to cost
    (bit-exact in float32 order-independent integer arithmetic) and whose
    delay sums to delay <= bound."""
    assert len(path) >= 1 and path[0] == s and path[-1] == t, (path[:5], s, t)
    tl, td = 0.0, 0.0
    for a, b in zip(path[:-1], path[1:]):
        lo, hi = g.row_ptr[a], g.row_ptr[a + 1]
        cand = np.nonzero(g.dst[lo:hi] == b)[0]
        assert cand.size > 0, "no arc %d->%d" % (a, b)
        # parallel arcs: take the one that is lexicographically best is not
        # needed; any arc works as long as the sums match one choice
        best = None
        for k in cand:
            l, d = float(g.len[lo + k]), float(g.delay[lo + k])
            if best is None or (l, d) < best:
                best = (l, d)
        tl += best[0]
        td += best[1]
    assert np.float32(tl) == np.float32(cost), (tl, cost)
    assert np.float32(td) == np.float32(delay), (td, delay)
    assert td <= bound
