"""Writes tests/golden/ar_golden.json: input series and the coefficients the
reference's own AR::AutoRegression returns for them (Burg, method 0, and least
squares, method 1), computed by the reference's include/AR.cpp compiled from
source into oracle/_ref/libar_ref.so (make -C oracle ref).

The series are log2 frontier-size traces of the oracle on the 64x64 grid
workload plus seeded random walks.  The DOP sequences are the reference rule
(algorithm/cspp.cuh:254-284, calculate_p :33-41) evaluated with the reference's
coefficients.  Run: python tests/golden/make_golden.py
"""
import json
import math
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from vine_amd import workload as W  # noqa: E402


def dop_sequence(sizes, max_labels, ar):
    """The reference DOP rule, fed one new frontier size per round."""
    dop = int(max_labels / 1000)
    hist = []
    out = []
    for n in sizes:
        lb = math.log2(n)
        k = len(hist)
        if k >= 3 and lb < hist[-1] < hist[-2] < hist[-3]:
            dop = int(min(max_labels, dop + dop // 2))
            hist.append(lb)
        else:
            hist.append(lb)
            if len(hist) > 3:
                c = ar(hist)
                est = sum(c[j] * hist[len(hist) - j - 1] for j in range(3))
                est = 2.0 ** est
                if est > n:
                    p = max_labels / 12 / (est / n - 0.95)
                    if p < 0:
                        p = max_labels
                    p = min(p, max_labels)
                    dop = int(round(p))
                    dop = max(dop, 1)
        out.append(dop)
    return out


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    series = []
    g, q = W.preset("grid64")
    for i in range(8):
        r = O.cspp(g, q.src[i], q.dst[i], q.bound[i], trace_cap=4096)
        tr = [int(x) for x in r.trace if x > 0]
        series.append([math.log2(x) for x in tr])
    rng = np.random.default_rng(2024)
    for length in (4, 5, 7, 12, 30, 64, 150):
        series.append([float(x) for x in np.cumsum(rng.normal(0.3, 1.0, length)) + 5.0])
    cases = []
    for s in series:
        if len(s) < 4:
            continue
        # the reference exits the process on a singular least-squares system and
        # divides by zero on a constant series: keep only well-posed series
        try:
            ok = np.all(np.isfinite(O.ar_burg(s, 3))) and np.all(np.isfinite(O.ar_lsq(s, 3)))
        except ValueError:
            ok = False
        if not ok:
            continue
        burg = O.ref_ar(s, 3, 0)
        lsq = O.ref_ar(s, 3, 1)
        cases.append({"series": s, "burg": [float(x) for x in burg], "lsq": [float(x) for x in lsq]})
    dops = []
    for i in range(8):
        r = O.cspp(g, q.src[i], q.dst[i], q.bound[i], trace_cap=4096)
        sizes = [int(x) for x in r.trace if x > 0]
        for mx in (1e4, 1e6, 2.4e10):
            seq = dop_sequence(sizes, mx, lambda h: O.ref_ar(h, 3, 0))
            dops.append({"sizes": sizes, "max_labels": mx, "dop": seq})
    out = {"source": "include/AR.cpp compiled from the reference (oracle/_ref/libar_ref.so)",
           "ar": cases, "dop": dops}
    with open(os.path.join(HERE, "ar_golden.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote %d AR cases, %d DOP sequences" % (len(cases), len(dops)))


if __name__ == "__main__":
    main()
