"""Shared checks for the parity tests."""
import numpy as np


def check_path(g, path, s, t, cost, delay, bound):
    """A returned path must be an s->t vertex sequence along existing arcs such
    that some choice among parallel arcs sums to exactly (cost, delay), with
    delay <= bound.  Sums of integer-valued weights are exact."""
    path = np.asarray(path)
    assert len(path) >= 1 and path[0] == s and path[-1] == t, (path[:5], s, t)
    sums = {(0.0, 0.0)}
    for a, b in zip(path[:-1], path[1:]):
        lo, hi = g.row_ptr[a], g.row_ptr[a + 1]
        ks = np.nonzero(g.dst[lo:hi] == b)[0]
        assert ks.size > 0, "no arc %d->%d" % (a, b)
        arcs = {(float(g.len[lo + k]), float(g.delay[lo + k])) for k in ks}
        sums = {(x + l, y + d) for (x, y) in sums for (l, d) in arcs if y + d <= bound}
        assert sums, "path exceeds the bound"
    assert (float(cost), float(delay)) in sums, (cost, delay, sorted(sums)[:4])
    assert float(delay) <= float(bound)
