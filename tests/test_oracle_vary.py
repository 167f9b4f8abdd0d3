"""Pins of the oracle's counter-based generator and flat postfix variation.

Philox4x32-10 (genome.rs:214-300) is held to the Random123 known-answer
vectors (the first is the reference's own test, genome.rs:440-448).  The
variation engine (postfix.rs:500-893) has no recorded child vectors, so it is
held to the reference's invariants (postfix.rs:1019-1089): every child is a
bounded, valid program; a child is a pure function of its position; an
island's children do not depend on how many islands are evolved; elitism
keeps each island's best parent; restarted islands are fresh random
programs; and, with dedup, no accepted child repeats a parent or an earlier
accepted child.
"""
import numpy as np
import pytest

import oracle

HALT = 255


def test_philox_matches_the_random123_known_answers():
    assert oracle.philox([0, 0], [0, 0, 0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF, 0xFFFFFFFF], [0xFFFFFFFF] * 4) == \
        [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox([0xA4093822, 0x299F31D0],
                         [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def _arity(w):
    if w in (6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 21, 22, 23, 24, 25, 28, 29, 33,
             42):
        return 1
    if w in (2, 3, 4, 5, 20, 26, 27, 30, 31, 34, 35, 36, 37, 38, 39, 40, 41):
        return 2
    return 3 if w == 32 else None


def _parse(slot):
    """Independent postfix parse (postfix.rs:268-312): node count or None."""
    depth = nodes = i = 0
    while i < len(slot) and slot[i] != HALT:
        w = int(slot[i])
        if w in (0, 1):
            if i + 1 >= len(slot) or slot[i + 1] == HALT:
                return None
            depth += 1
            nodes += 1
            i += 2
        else:
            a = _arity(w)
            if a is None or depth < a:
                return None
            depth -= a - 1
            nodes += 1
            i += 1
    return nodes if depth == 1 and i > 0 and i < len(slot) else None


def _fnv(slot):
    h = 0xCBF29CE484222325
    for w in slot:
        h ^= int(w)
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
        if w == HALT:
            break
    return h


def _parents(islands, pop, width=64, seed=7):
    """A random parent generation: every island restarted (fill_random)."""
    blank = np.full((islands * pop, width), HALT, np.uint32)
    blank[:, 0], blank[:, 1], blank[:, 2] = 1, 0, HALT
    g = oracle.vary(blank, islands, pop, np.full(islands * pop, np.nan), seed, 0,
                    restart=[1] * islands)
    err = np.array([abs(np.sin(0.5 * i)) for i in range(islands * pop)])
    return g["words"], err


def test_children_are_bounded_valid_programs_with_true_metadata():
    parents, err = _parents(3, 32)
    nxt = oracle.vary(parents, 3, 32, err, 7, 1)
    for s in range(96):
        slot = nxt["words"][s]
        nodes = _parse(slot)
        assert nodes is not None and nodes <= 30, s
        ln = int(np.argmax(slot == HALT)) + 1
        assert nxt["len"][s] == ln <= 64
        assert nxt["cost"][s] == nodes
        assert int(nxt["hash"][s]) == _fnv(slot)
        assert np.all(slot[ln:] == HALT)


def test_a_child_is_a_pure_function_of_its_position():
    parents, err = _parents(4, 40, seed=9)
    a = oracle.vary(parents, 4, 40, err, 9, 3)
    b = oracle.vary(parents, 4, 40, err, 9, 3)
    assert np.array_equal(a["words"], b["words"])
    assert np.array_equal(a["hash"], b["hash"])
    assert np.array_equal(a["err"].view(np.uint64), b["err"].view(np.uint64))
    c = oracle.vary(parents, 4, 40, err, 9, 4)
    assert not np.array_equal(a["words"], c["words"])


def test_islands_do_not_depend_on_how_many_are_evolved():
    # No shared stream or slot cursor: without dedup island 0's children are
    # the same whether it is evolved alone or among four islands.
    parents, err = _parents(4, 24, seed=4)
    four = oracle.vary(parents, 4, 24, err, 4, 2, dedup=False)
    solo = oracle.vary(parents[:24], 1, 24, err[:24], 4, 2, dedup=False)
    assert np.array_equal(four["words"][:24], solo["words"])
    assert np.array_equal(four["err"][:24].view(np.uint64), solo["err"].view(np.uint64))
    # The dedup set spans the generation (every parent's hash is in it,
    # postfix.rs:556-561), so with dedup a child may differ only where the solo
    # run accepted a program that repeats another island's parent.
    four = oracle.vary(parents, 4, 24, err, 4, 2)
    solo = oracle.vary(parents[:24], 1, 24, err[:24], 4, 2)
    others = {_fnv(q) for q in parents[24:]}
    for j in range(24):
        if not np.array_equal(four["words"][j], solo["words"][j]):
            assert _fnv(solo["words"][j]) in others, j


def test_elitism_restart_and_dedup():
    parents, err = _parents(3, 32, seed=11)
    nxt = oracle.vary(parents, 3, 32, err, 11, 5, restart=[0, 1, 0])
    for i in (0, 2):
        best = min(range(32 * i, 32 * i + 32),
                   key=lambda s: (np.isnan(err[s]), err[s], _parse(parents[s]), s))
        assert np.array_equal(nxt["words"][32 * i], parents[best])
        assert nxt["err"][32 * i] == err[best]
        assert np.all(np.isnan(nxt["err"][32 * i + 1:32 * i + 32]))
    # A restarted island's slots are fresh random programs keyed by its own
    # position (fill_random with salt 0): the same in a two-island arena.
    two = oracle.vary(parents[:64], 2, 32, err[:64], 11, 5, restart=[0, 1])
    assert np.array_equal(two["words"][32:64], nxt["words"][32:64])
    assert not np.array_equal(nxt["words"][32:64], parents[32:64])
    # Dedup: accepted children (not elites, not restarted islands) never
    # repeat a parent or an earlier accepted child.
    seen = {_fnv(p) for p in parents}
    for i in (0, 2):
        for j in range(1, 32):
            h = _fnv(nxt["words"][32 * i + j])
            assert h not in seen
            seen.add(h)
    # Without dedup the same positions may repeat, and every child is still
    # valid.
    free = oracle.vary(parents, 3, 32, err, 11, 5, dedup=False)
    assert all(_parse(s) is not None for s in free["words"])
