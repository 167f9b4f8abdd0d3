"""CPU checks of the oracle's order, rank and maximum statistics
(metrics.rs:1193-1386 and the maximum-error scan, eval.cu:4505-4600).

Pinned by the reference's own fixture (metrics.rs:2191-2201), by brute-force
definitions that share no code with the oracle (average rank as a count of
smaller and equal values, the window mean as a Python loop) and by closed
forms (a permutation's ranks, sum A = 0, Cauchy-Schwarz, interval ends that
solve |r + b d| = t).
"""
import math
import sys

import numpy as np

import oracle


def brute_doubled_ranks(v):
    v = list(v)
    n = len(v)
    out = []
    for x in v:
        less = sum(1 for y in v if y < x)
        eq = sum(1 for y in v if y == x)
        avg2 = 2 * less + eq + 1          # twice the average rank
        out.append(avg2 - (n + 1))
    return np.asarray(out, np.int64)


def test_reference_fixture_negative_zero_ties():
    a = [-0.0, 0.0, 1.0, 1.0, 1.0, 2.0]
    assert oracle.centred_ranks_doubled(a).tolist() == [-4, -4, 1, 1, 1, 5]
    ab, aa, bb = oracle.rank_sums(oracle.centred_ranks_doubled(a), oracle.centred_ranks_doubled(a))
    assert ab == aa == bb == 16 + 16 + 1 + 1 + 1 + 25


def test_ranks_match_the_counting_definition():
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 17, 200):
        v = np.round(rng.normal(size=n), 1)   # many ties
        got = oracle.centred_ranks_doubled(v)
        assert np.array_equal(got, brute_doubled_ranks(v)), n
        assert int(got.sum()) == 0
    perm = rng.permutation(50).astype(float)
    assert np.array_equal(oracle.centred_ranks_doubled(perm), (2 * (perm + 1) - 51).astype(np.int64))
    assert oracle.centred_ranks_doubled([1.0, np.nan]) is None
    assert oracle.centred_ranks_doubled([np.inf, 1.0]) is None


def test_rank_sums_are_exact_and_bounded():
    rng = np.random.default_rng(4)
    n = 4001
    a = oracle.centred_ranks_doubled(rng.normal(size=n))
    b = oracle.centred_ranks_doubled(np.round(rng.normal(size=n), 2))
    ab, aa, bb = oracle.rank_sums(a, b)
    assert ab == int(sum(int(x) * int(y) for x, y in zip(a, b)))
    assert aa == int(sum(int(x) * int(x) for x in a))
    assert ab * ab <= aa * bb
    # without ties sum A^2 has the closed form (n^3 - n) / 3
    assert aa == (n ** 3 - n) // 3
    # 128-bit values beyond 64 bits survive the (hi, lo) round trip
    big = np.full(3, 3_000_000_000, np.int64)
    _, aa_big, _ = oracle.rank_sums(big, big)
    assert aa_big == 3 * 3_000_000_000 ** 2 and aa_big > 2 ** 64


def test_median_and_interquartile_mean():
    assert oracle.median_abs([3.0, -1.0, 2.0]) == 2.0
    assert oracle.median_abs([1.0, -4.0, 2.0, 3.0]) == 2.5
    # the window is chosen by rank: sorted 1,2,3,3 averages 2,3 (eval.cu:4453-4456)
    assert oracle.iqr_mean_abs([3.0, 1.0, 3.0, 2.0]) == 2.5
    assert oracle.iqr_mean_abs([5.0, -1.0, 2.0]) == (1.0 + 2.0 + 5.0) / 3
    rng = np.random.default_rng(5)
    for n in (4, 7, 100, 1001):
        r = rng.normal(size=n)
        s = sorted(abs(x) for x in r)
        assert oracle.median_abs(r) == (s[n // 2] if n % 2 else 0.5 * (s[n // 2 - 1] + s[n // 2]))
        lo, hi = n // 4, (3 * n) // 4
        acc = 0.0
        for x in s[lo:hi]:
            acc += x
        assert oracle.iqr_mean_abs(r) == acc / (hi - lo)


def test_max_scan_maximum_row_and_level_interval():
    r = np.asarray([1.0, -3.0, 3.0, 2.0])
    vm, row, lo, hi, st = oracle.max_scan(r)
    assert (vm, row, st) == (3.0, 1, 0) and lo == -math.inf and hi == math.inf
    # |r_i + b_i d| <= t on every row; the ends solve one row's bound exactly
    rng = np.random.default_rng(6)
    r = rng.normal(size=50)
    b = rng.normal(size=50)
    t = 2.5
    vm, row, lo, hi, st = oracle.max_scan(r, b, t)
    assert st == 0 and lo <= hi
    for d in (lo, hi, 0.5 * (lo + hi)):
        assert np.all(np.abs(r + b * d) <= t * (1 + 1e-12))
    for d in (lo - 1e-6, hi + 1e-6):
        assert np.any(np.abs(r + b * d) > t)
    # a row the constant cannot move vetoes an unreachable level
    _, _, lo, hi, st = oracle.max_scan([0.5, 4.0], [1.0, 0.0], 1.0)
    assert st == 0 and lo == math.inf and hi == -math.inf
    vm, row, _, _, st = oracle.max_scan([1.0, np.nan])
    assert st == 2 and row == -1 and math.isnan(vm)


def test_flat_predictions_equal_the_tree_evaluator_on_fit_programs():
    from rcsr_hip import workloads
    wl = workloads.config3(n=300, ncand=60, iters=8)
    for t in wl.trees:
        a = oracle.predict_flat(t, wl.x)
        b = oracle.predict(t, wl.x)
        assert np.array_equal(a, b, equal_nan=True), str(t)
