"""Writes lm_recorded_baseline.json: the reference LM fitter's six recorded
cases (my-version/src/fit.rs:4654-4832) as data — each case's pre-order tree,
rows, targets, weights and the recorded constants/SSE IEEE bits
(fit.rs:4805-4823).  The inputs are regenerated from the formulas of that
test; the expected bits are the reference's recorded values, transcribed.
Run: python tests/golden/make_lm_baseline.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "the-rule-and-compass_amd"))

from test_oracle_golden import EXPECTED, _cases  # noqa: E402


def main():
    out = {"source": "my-version/src/fit.rs:4654-4832 (recorded bits 4805-4823)",
           "iters": 8, "cases": []}
    for (name, tree, x, y, w), (cb, sb) in zip(_cases(), EXPECTED):
        ops, vals = tree.preorder()
        out["cases"].append({
            "name": name,
            "ops": [int(o) for o in ops],
            "vals": [float(v) for v in vals],
            "x": [r[0] for r in x],
            "y": list(y),
            "w": w,
            "consts_bits": [hex(b) for b in cb],
            "sse_bits": hex(sb),
        })
    with open(os.path.join(HERE, "lm_recorded_baseline.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
