"""Writes expected_counts.json: neighbour counts of the reference's own test files.

test100.txt / test128.txt are the only data the reference ships with its join
(18-D rows, read with INPUT_DATA_DIM 18).  Squared distances are computed here
exactly, in rational arithmetic over the FP64 values the reader produces, so the
counts do not depend on any floating-point summation order.  An eps is kept only
if no pair lies within a relative 1e-9 of the boundary, where the FP64 direct form
could round across it.  "f32" holds the same for eps rounded to binary32, the value
the reference's main.cu:20 passes on (ACCUM_TYPE epsilon = atof(...)) and the
drop-in GPUJoinMainBruteForce takes.  Run: python tests/golden/make_golden.py
"""
import json
import os
import re
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DIM = 18
EPS = [0.001, 0.02, 0.03, 0.04, 0.05, 0.07, 0.1]
EPS_F32 = [0.05, 0.1]  # as the reference's command line would pass them (binary32)


def read_rows(path):
    vals = [float(t) for t in re.split(r"[,\s]+", open(path).read()) if t]
    m = len(vals) // DIM
    return [vals[i * DIM:(i + 1) * DIM] for i in range(m)]


def main():
    out = {}
    for name in ("test100.txt", "test128.txt"):
        rows = read_rows(os.path.join(HERE, name))
        fr = [[Fraction(v) for v in r] for r in rows]
        n = len(rows)
        d2 = [[sum((a - b) * (a - b) for a, b in zip(fr[i], fr[j])) for j in range(n)] for i in range(n)]
        def table(eps_list):
            keep, counts, totals = [], [], []
            for e in eps_list:
                e2 = Fraction(e) * Fraction(e)
                close = min(abs(d2[i][j] - e2) / e2 for i in range(n) for j in range(n))
                if close < Fraction(1, 10**9):
                    continue
                c = [sum(1 for j in range(n) if j != i and d2[i][j] <= e2) for i in range(n)]
                keep.append(e)
                counts.append(c)
                totals.append(sum(c) + n)
            return keep, counts, totals

        keep, counts, totals = table(EPS)
        kf, cf, tf = table([float(np.float32(e)) for e in EPS_F32])
        out[name] = {"dim": DIM, "n": n, "eps": keep, "totals": totals, "counts": counts,
                     "f32": {"eps_arg": EPS_F32, "eps": kf, "totals": tf, "counts": cf}}
        print(name, n, list(zip(keep, totals)), list(zip(kf, tf)))
    with open(os.path.join(HERE, "expected_counts.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
