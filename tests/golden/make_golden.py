"""Write the golden fixtures of tests/golden/ from the reference's own host
PDHG, compiled in place by `make -C oracle ref` (oracle/_ref/libref.so).

Run in a container that has the reference tree:  python tests/golden/make_golden.py
The fixtures are data only: model arrays, and the reference's results on them
(status, objective, iteration and restart counts, operator-norm estimate,
per-check residual trace, final point), plus scaling factors and generator
pins.  The models are:
  * the LPs of the reference's tests/test_pdhg.cpp (:29-147) and the four
    smoke instances of benchmarks/smoke/*.mps, parsed by the reference's own
    readers;
  * seeded instances of this project's random-LP generator (which the
    generator pins below freeze).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "sovereign-solver_amd"))

import oracle_ffi as O  # noqa: E402

REF_TREE = "/root/reference"

# Model texts of the reference's tests/test_pdhg.cpp, as inputs.
TEST_PDHG_LPS = {
    "two_variable_optimum": "Maximize\n obj: x + y\nSubject To\n c1: x + 2 y <= 4\n c2: 4 x + 2 y <= 12\nEnd\n",
    "binding_lower_bound": "Minimize\n obj: 3 x + 2 y\nSubject To\n c1: x + y >= 4\nEnd\n",
    "equality_row": "Minimize\n obj: x + y\nSubject To\n c1: x + y = 10\nBounds\n x <= 6\nEnd\n",
    "column_bound": "Minimize\n obj: -x\nSubject To\n c1: x + y <= 100\nBounds\n x <= 3\nEnd\n",
    "degenerate": "Minimize\n obj: -x1 - x2\nSubject To\n c1: x1 <= 4\n c2: x1 + x2 <= 4\n c3: x2 <= 2\nEnd\n",
}


def arrays_of(M: O.Model) -> dict:
    return dict(m=M.m, n=M.n, row_ptr=M.row_ptr.tolist(), col_idx=M.col_idx[: M.nnz].tolist(),
                val=M.val[: M.nnz].tolist(), c=M.c.tolist(), col_lo=[_f(v) for v in M.col_lo],
                col_hi=[_f(v) for v in M.col_hi], row_lo=[_f(v) for v in M.row_lo],
                row_hi=[_f(v) for v in M.row_hi], maximize=M.maximize, obj_const=M.obj_const)


def _f(v: float):
    if np.isinf(v):
        return "inf" if v > 0 else "-inf"
    return float(v)


def result_of(r: dict) -> dict:
    out = {k: r[k] for k in ("status", "objective", "primal_infeasibility", "dual_infeasibility", "duality_gap",
                             "iterations", "restarts", "operator_norm")}
    out = {k: (_f(v) if isinstance(v, float) else v) for k, v in out.items()}
    out["trace"] = [float(v) for v in r["trace"]]
    out["trace_it"] = [int(v) for v in r["trace_it"]]
    out["x"] = [float(v) for v in r["x"]]
    out["y"] = [float(v) for v in r["y"]]
    return out


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main() -> None:
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    import sovereign_amd as S

    out: dict = {"note": "written by tests/golden/make_golden.py from the reference host PDHG"}
    # -- hand-written LPs and smoke instances
    lps = {}
    for name, text in TEST_PDHG_LPS.items():
        M = O.Model(O.ref_parse(text))
        lps[name] = dict(model=arrays_of(M), result=result_of(O.ref_solve(M, trace_cap=100000)))
        if name == "two_variable_optimum":
            lps[name + "_period5"] = dict(model=arrays_of(M), period=5,
                                          result=result_of(O.ref_solve(M, period=5, trace_cap=100000)))
    for fname in ("tiny_lp", "tiny_degenerate", "tiny_infeasible", "tiny_unbounded"):
        with open(os.path.join(REF_TREE, "benchmarks", "smoke", fname + ".mps")) as f:
            M = O.Model(O.ref_parse(f.read(), mps=True))
        lps["smoke_" + fname] = dict(model=arrays_of(M),
                                     result=result_of(O.ref_solve(M, max_iterations=20000, trace_cap=100000)))
    out["lps"] = lps

    # -- seeded random LPs (reduced C0 shape 200 x 500 and full C0 2000 x 5000)
    rand = {}
    for (m, n, seed, tol, cap) in ((200, 500, 0, 1e-4, 200000), (200, 500, 1, 1e-4, 200000),
                                   (200, 500, 2, 1e-4, 200000), (200, 500, 1, 1e-6, 400000),
                                   (2000, 5000, 0, 1e-4, 20000)):
        L = S.LP.random(m, n, 5, seed)
        M = O.Model(L.arrays())
        r = result_of(O.ref_solve(M, tolerance=tol, max_iterations=cap, trace_cap=100000))
        if m > 500:
            r.pop("x")
            r.pop("y")
        rand[f"c0_{m}x{n}_s{seed}_tol{tol:g}"] = dict(m=m, n=n, k=5, seed=seed, tolerance=tol, max_iterations=cap,
                                                     digest=digest(M.row_ptr, M.col_idx[: M.nnz], M.val[: M.nnz],
                                                                   M.c, M.row_lo),
                                                     result=r)
        print("done", m, n, seed, tol, r["status"], r["iterations"], flush=True)
    out["random"] = rand

    # -- scaling factors of the reference (Scaling::equilibrate)
    ill = O.Model(dict(m=4, n=4, row_ptr=[0, 2, 4, 7, 9], col_idx=[0, 1, 1, 2, 0, 2, 3, 1, 3],
                       val=[1e6, 1e-3, 1e-6, 1e3, 1e-4, 1e2, 1.0, 1e4, 1e-5], c=[0] * 4, col_lo=[0] * 4,
                       col_hi=[0] * 4, row_lo=[0] * 4, row_hi=[0] * 4))
    L = S.LP.random(200, 500, 5, 0)
    Mr = O.Model(L.arrays())
    sc = {}
    for name, M in (("ill_scaled_4x4", ill), ("c0_200x500_s0", Mr)):
        for ruiz, pc in ((1, 1), (1, 0), (0, 1)):
            rs, cs = O.ref_scaling(M, ruiz, pc)
            sc[f"{name}_r{ruiz}p{pc}"] = dict(row_scale=rs.tolist(), col_scale=cs.tolist())
    sc["ill_scaled_4x4_model"] = arrays_of(ill)
    out["scaling"] = sc

    # -- generator pins (so a change to the generators is caught)
    gens = {}
    for name, lp in (("random_200x500_k5_s0", S.LP.random(200, 500, 5, 0)),
                     ("random_2000x5000_k5_s0", S.LP.random(2000, 5000, 5, 0)),
                     ("zipf_5000x10000_s0", S.LP.random_zipf(5000, 10000, 0)),
                     ("transport_20x30_s0", S.LP.transportation(20, 30, 0))):
        a = lp.arrays()
        gens[name] = dict(m=a["m"], n=a["n"], nnz=a["nnz"],
                          digest=digest(a["row_ptr"], a["col_idx"], a["val"], a["c"], a["row_lo"]),
                          val_sum=float(np.sum(a["val"])), c_sum=float(np.sum(a["c"])),
                          row_lo_sum=float(np.sum(a["row_lo"])))
    gens["batch_sizes_256_s0"] = S.batch_sizes(256, 200, 20000, 0).tolist()
    out["generators"] = gens

    with open(os.path.join(HERE, "reference_pdhg.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", os.path.join(HERE, "reference_pdhg.json"))


if __name__ == "__main__":
    main()
