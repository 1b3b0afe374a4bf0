"""Write tests/golden/reader_cases.json: edge cases of the MPS format as the
reference's own reader (src/io/mps_reader.cpp, compiled into oracle/_ref by
`make -C oracle ref`) reads them -- the model arrays, or the fact that it
rejects the text.  Data only: each case is an MPS text and the reference's
verdict.

Run in a container that has the reference tree:
    python tests/golden/make_reader_cases.py
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, HERE)

import oracle_ffi as O  # noqa: E402
from make_golden import arrays_of  # noqa: E402

HEAD = "NAME t\nROWS\n N obj\n L c1\n G c2\n E c3\nCOLUMNS\n x obj 1 c1 1\n x c2 1 c3 1\n y obj -2 c1 1\n y c3 1\n"
RHS = "RHS\n RHS c1 4 c2 1\n RHS c3 2\n"

CASES = {
    # bound cards with and without a set name, infinite spellings
    "bounds_forms": HEAD + RHS + "BOUNDS\n UP BND x 4\n LO y -3\n MI BND y\n UP y 1e30\nENDATA\n",
    "infinite_tokens": HEAD + RHS + "BOUNDS\n UP BND x INF\n LO BND y -Infinity\nENDATA\n",
    # a valued bound with no value
    "bound_without_value": HEAD + RHS + "BOUNDS\n UP BND x\nENDATA\n",
    # numbers that do not parse
    "bad_coefficient": HEAD.replace("x c2 1 c3 1", "x c2 1.5.2 c3 1") + RHS + "ENDATA\n",
    "bad_rhs": HEAD + "RHS\n RHS c1 four\nENDATA\n",
    "bad_bound": HEAD + RHS + "BOUNDS\n UP BND x 4x\nENDATA\n",
    # negative upper bounds: UP moves the default lower bound to -inf, UI does not
    "negative_up": HEAD + RHS + "BOUNDS\n UP BND x -2\nENDATA\n",
    "negative_ui": HEAD + RHS + "BOUNDS\n UI BND x -2\nENDATA\n",
    # a bound on a column without coefficients is added (with a warning)
    "bound_on_new_column": HEAD + RHS + "BOUNDS\n UP BND z 7\nENDATA\n",
    # ranges on every row type, negative on the equality row; objective constant
    "ranges_and_constant": HEAD + "RHS\n RHS c1 4 c2 1\n RHS c3 2 obj 3\nRANGES\n RNG c1 2 c2 -5\n RNG c3 -1\nENDATA\n",
    # a second N row (kept as a free row) with coefficients and a right-hand side
    "free_row": HEAD.replace(" E c3\n", " E c3\n N spare\n").replace("y c3 1", "y c3 1 spare 4") + RHS
                + "RHS\n RHS spare 9\nENDATA\n",
    # integer markers and objective sense
    "markers_and_sense": "NAME t\nOBJSENSE\n MAX\nROWS\n N obj\n L c1\nCOLUMNS\n M1 'MARKER' 'INTORG'\n"
                         " x obj 1 c1 1\n M2 'MARKER' 'INTEND'\n y obj 1 c1 2\nRHS\n RHS c1 3\nENDATA\n",
    "unknown_row": HEAD.replace("y c3 1", "y c9 1") + RHS + "ENDATA\n",
}


def main() -> None:
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    out = {"note": "MPS edge cases as the reference reader reads them; written by tests/golden/make_reader_cases.py",
           "cases": {}}
    for name, text in CASES.items():
        try:
            model = arrays_of(O.Model(O.ref_parse(text, mps=True)))
            out["cases"][name] = dict(text=text, rejected=False, model=model)
        except RuntimeError:
            out["cases"][name] = dict(text=text, rejected=True)
        print(name, "rejected" if out["cases"][name]["rejected"] else "read", flush=True)
    with open(os.path.join(HERE, "reader_cases.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
