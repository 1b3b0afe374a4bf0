"""Generated LPs shared by the CPU and GPU tests."""
import numpy as np

import sovereign_amd as S


def infeasible_and_unbounded(m=200, n=500, seed=4):
    """A random feasible LP made (a) infeasible by one extra row
    sum_{j<50} x_j <= -1 over columns with x >= 0, and (b) unbounded by one
    extra free-above column with cost -1 and no matrix entries."""
    a = S.LP.random(m, n, 5, seed).arrays()
    rp, ci, v = list(a["row_ptr"]), list(a["col_idx"][: a["nnz"]]), list(a["val"][: a["nnz"]])
    bad = S.LP.from_csr(m + 1, n, rp + [rp[-1] + 50], ci + list(range(50)), v + [1.0] * 50, a["c"], a["col_lo"],
                        a["col_hi"], list(a["row_lo"]) + [-np.inf], list(a["row_hi"]) + [-1.0])
    unb = S.LP.from_csr(m, n + 1, rp, ci, v, list(a["c"]) + [-1.0], list(a["col_lo"]) + [0.0],
                        list(a["col_hi"]) + [np.inf], a["row_lo"], a["row_hi"])
    return bad, unb
