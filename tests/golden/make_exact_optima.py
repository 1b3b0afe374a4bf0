"""Write tests/golden/exact_optima.json: exact optimal objective values of
seeded instances, from the HiGHS LP solver that scipy bundles
(scipy.optimize.linprog(method="highs"), scipy 1.15).

HiGHS is the reference's own correctness oracle for PDHG
(tools/oracle/install_highs.sh builds HiGHS v1.11.0 from the network and
tools/check_pdhg_vs_oracle.py compares PDHG objectives with it); scipy's
bundled copy is the same solver and needs no network.  These values pin the
objective of every PDHG solve the tests run at sizes where the CPU oracle's
own PDHG would take minutes: the 2,000 x 5,000 LP, the 2,000 x 2,000
transportation LP (4M columns) and smaller family members.

The fixture is data: per instance its generator arguments, a digest of the
generated arrays (so a change to a generator is caught), HiGHS's status and
objective, and the seconds HiGHS took.

Run in the build container:  python tests/golden/make_exact_optima.py
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "sovereign-solver_amd"))
sys.path.insert(0, HERE)

from make_golden import digest  # noqa: E402


def highs_optimum(a: dict) -> tuple[int, float, float]:
    """(status, objective, seconds) of min/max c x, row_lo <= A x <= row_hi,
    col_lo <= x <= col_hi, with the model's objective constant."""
    from scipy.optimize import linprog
    from scipy.sparse import csr_matrix, vstack
    A = csr_matrix((a["val"], a["col_idx"], a["row_ptr"]), shape=(a["m"], a["n"]))
    lo, hi = np.asarray(a["row_lo"]), np.asarray(a["row_hi"])
    eq = lo == hi
    ub_rows, ub_rhs = [], []
    fin_hi = ~eq & np.isfinite(hi)
    fin_lo = ~eq & np.isfinite(lo)
    if fin_hi.any():
        ub_rows.append(A[np.flatnonzero(fin_hi)])
        ub_rhs.append(hi[fin_hi])
    if fin_lo.any():
        ub_rows.append(-A[np.flatnonzero(fin_lo)])
        ub_rhs.append(-lo[fin_lo])
    sign = -1.0 if a["maximize"] else 1.0
    bounds = [(None if not np.isfinite(l) else l, None if not np.isfinite(h) else h)
              for l, h in zip(a["col_lo"], a["col_hi"])]
    kw = {}
    if eq.any():
        kw.update(A_eq=A[np.flatnonzero(eq)], b_eq=lo[eq])
    if ub_rows:
        kw.update(A_ub=vstack(ub_rows).tocsr(), b_ub=np.concatenate(ub_rhs))
    t0 = time.perf_counter()
    r = linprog(sign * np.asarray(a["c"]), bounds=bounds, method="highs", **kw)
    dt = time.perf_counter() - t0
    obj = sign * r.fun + a["obj_const"] if r.status == 0 else float("nan")
    return int(r.status), float(obj), dt


def instances():
    import sovereign_amd as S
    yield "lp_200x500_s0", dict(kind="random", m=200, n=500, k=5, seed=0), S.LP.random(200, 500, 5, 0)
    yield "lp_200x500_s1", dict(kind="random", m=200, n=500, k=5, seed=1), S.LP.random(200, 500, 5, 1)
    yield "lp_300x800_s3", dict(kind="random", m=300, n=800, k=5, seed=3), S.LP.random(300, 800, 5, 3)
    yield "lp_2000x5000_s0", dict(kind="random", m=2000, n=5000, k=5, seed=0), S.LP.random(2000, 5000, 5, 0)
    yield "lp_2000x5000_s1", dict(kind="random", m=2000, n=5000, k=5, seed=1), S.LP.random(2000, 5000, 5, 1)
    yield "transport_20x30_s0", dict(kind="transportation", sources=20, sinks=30, seed=0), \
        S.LP.transportation(20, 30, 0)
    yield "transport_200x200_s0", dict(kind="transportation", sources=200, sinks=200, seed=0), \
        S.LP.transportation(200, 200, 0)
    yield "zipf_3000x7000_s2", dict(kind="zipf", m=3000, n=7000, seed=2), S.LP.random_zipf(3000, 7000, 2)
    yield "transport_2000x2000_s0", dict(kind="transportation", sources=2000, sinks=2000, seed=0), \
        S.LP.transportation(2000, 2000, 0)


def main() -> None:
    out = {"note": "exact optima from HiGHS (scipy.optimize.linprog, method='highs'); "
                   "written by tests/golden/make_exact_optima.py", "instances": {}}
    for name, args, lp in instances():
        a = lp.arrays()
        st, obj, dt = highs_optimum(a)
        out["instances"][name] = dict(args=args, m=a["m"], n=a["n"], nnz=a["nnz"],
                                      digest=digest(a["row_ptr"], a["col_idx"], a["val"], a["c"], a["row_lo"]),
                                      highs_status=st, objective=obj, highs_seconds=round(dt, 2))
        print(name, st, repr(obj), f"{dt:.1f}s", flush=True)
    with open(os.path.join(HERE, "exact_optima.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
