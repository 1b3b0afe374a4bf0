"""Write tests/golden/reference_checkpoints.json: the checkpoint hand-off of
the reference's own host PDHG (PdhgOptions::checkpoint, reference
include/sovereign/pdhg.hpp:121-133, called at src/l1/pdhg.cpp:373-381), from
oracle/_ref (the reference compiled from its sources by `make -C oracle ref`).

Per case: the model's generator arguments and digest, the solve's status and
number of checkpoints, and the first checkpoints' iteration, original-space
residual and unscaled candidate (x, y) -- what a crossover consumer
receives.  Data only.

Run in a container that has the reference tree:
    python tests/golden/make_checkpoints.py
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "sovereign-solver_amd"))

import oracle_ffi as O  # noqa: E402
from make_golden import TEST_PDHG_LPS, arrays_of, digest  # noqa: E402


def main() -> None:
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    import sovereign_amd as S
    out = {"note": "checkpoint hand-off of the reference host PDHG; written by tests/golden/make_checkpoints.py",
           "cases": {}}
    M = O.Model(O.ref_parse(TEST_PDHG_LPS["two_variable_optimum"]))
    cases = [("two_variable_optimum", M, dict(model=arrays_of(M)), 1e-4, 16)]
    lp = S.LP.random(100, 250, 5, 1)
    Mr = O.Model(lp.arrays())
    cases.append(("c0_100x250_s1", Mr, dict(m=100, n=250, k=5, seed=1,
                                            digest=digest(Mr.row_ptr, Mr.col_idx[: Mr.nnz], Mr.val[: Mr.nnz],
                                                          Mr.c, Mr.row_lo)), 1e-4, 12))
    for name, model, desc, tol, cap in cases:
        st, total, its, res, xs, ys = O.ref_checkpoints(model, tolerance=tol, cap=cap)
        out["cases"][name] = dict(desc, tolerance=tol, status=st, checkpoints=total, iterations=its.tolist(),
                                  residuals=res.tolist(), x=xs.tolist(), y=ys.tolist())
        print(name, st, total, "checkpoints;", len(its), "kept", flush=True)
    with open(os.path.join(HERE, "reference_checkpoints.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
