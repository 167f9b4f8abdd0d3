"""Pins the CPU f64 oracle against the reference's own recorded vectors.

The six cases and their expected IEEE bits are the recorded baseline of the
reference LM fitter (my-version/src/fit.rs:4654-4832, expected bits at
4805-4823).  A restatement that drops a term, swaps an index or changes the
damping schedule changes these bits.
"""
import math

import numpy as np
import pytest

import oracle
from rcsr_hip.expr import ADD, DIV, EXP, MUL, SIN, SUB, const, node, var


def _bits(v: float) -> int:
    return int(np.float64(v).view(np.uint64))


def _cases():
    lin_x = [[-4.0 + 0.2 * i] for i in range(41)]
    lin_y = [2.0 * r[0] + 1.0 + 0.05 * math.sin(3.0 * r[0]) for r in lin_x]
    quad_x = [[-3.0 + 0.2 * i] for i in range(31)]
    quad_y = [0.5 * r[0] * r[0] - 2.0 * r[0] + 1.0 + 0.02 * math.cos(5.0 * r[0]) for r in quad_x]
    exp_x = [[-1.0 + 0.05 * i] for i in range(41)]
    exp_y = [math.exp(1.5 * r[0]) + 0.001 * math.sin(10.0 * r[0]) for r in exp_x]
    rat_x = [[0.5 + 0.25 * i] for i in range(19)]
    rat_y = [2.0 / (r[0] - 0.2) + 0.5 for r in rat_x]
    sin_x = [[-3.1 + 0.1 * i] for i in range(63)]
    sin_y = [2.5 * math.sin(0.7 * r[0]) + 0.3 for r in sin_x]
    wts = [0.0 if i % 5 == 0 else (0.25 if i % 5 in (1, 2) else 2.0) for i in range(41)]
    line = lambda: node(ADD, node(MUL, const(0.0), var(0)), const(0.0))
    return [
        ("linear", line(), lin_x, lin_y, None),
        ("quadratic",
         node(ADD, node(MUL, const(0.0), node(MUL, var(0), var(0))),
              node(ADD, node(MUL, const(0.0), var(0)), const(0.0))),
         quad_x, quad_y, None),
        ("exp_bad_scale", node(ADD, node(EXP, node(MUL, const(3.0), var(0))), const(0.5)),
         exp_x, exp_y, None),
        ("rational", node(ADD, node(DIV, const(1.0), node(SUB, var(0), const(0.0))), const(0.0)),
         rat_x, rat_y, None),
        ("damped_sine",
         node(ADD, node(MUL, const(1.0), node(SIN, node(MUL, const(1.0), var(0)))), const(0.0)),
         sin_x, sin_y, None),
        ("weighted_linear", line(), lin_x, lin_y, wts),
    ]


# fit.rs:4805-4823
EXPECTED = [
    ([0x3ffff3cde812a095, 0x3fefffffffffffff], 0x3fa9c6058c8f6e1a),
    ([0x3fe000441da12855, 0xc000000000000000, 0x3ff000b65a9771ca], 0x3f78fdaf8bfef98f),
    ([0x3ff8004d303015c9, 0xbf097785fa307cfa], 0x3ef4074508c02711),
    ([0x3ffffffffffffffe, 0x3fc999999999999d, 0x3fe0000000000001], 0x39b8400000000000),
    ([0x4004000000000000, 0x3fe6666666666666, 0x3fd3333333333333], 0x0000000000000000),
    ([0x3ffff2ce37c38d91, 0x3feff88cc309f8c7], 0x3fab8577d7aab41d),
]


@pytest.mark.parametrize("idx", range(6))
def test_lm_matches_recorded_baseline_bit_for_bit(idx):
    name, tree, x, y, w = _cases()[idx]
    sse, c = oracle.fit_weighted(tree, x, y, w, 8)
    got = ([_bits(v) for v in c], _bits(sse))
    exp_c, exp_sse = EXPECTED[idx]
    assert got == (exp_c, exp_sse), (
        f"{name}: consts={[hex(b) for b in got[0]]} sse={hex(got[1])}; "
        f"expected {[hex(b) for b in exp_c]} {hex(exp_sse)}")


def test_golden_fixture_file_matches_inline_table():
    """The committed fixture (used by the device parity tests) holds the same bits."""
    import json, os
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "lm_recorded_baseline.json")) as fh:
        fx = json.load(fh)
    assert [ (c["consts_bits"], c["sse_bits"]) for c in fx["cases"] ] == \
        [([hex(b) for b in e[0]], hex(e[1])) for e in EXPECTED]
