"""Seeded synthetic workloads for the five benchmark configurations.

Every stream comes from a Philox4x32-10 counter-based generator (the family
the reference uses for its genome pool, genome.rs:216-300) seeded per
configuration (0x5EED0001 ... 0x5EED0005), so the same seed yields the same
rows, programs and constants on every machine.

  C1 plain CPU-runnable: x ~ U(-3,3), N=1000, y = 2x+1+N(0,0.1); 4 islands x
     200 random programs ({+,-,*,/,exp,log,sin,cos}, depth <= 4, 0-4 constants
     ~ N(0,1)); eval + LM fit, squared.
  C2 evaluation: 4096 programs of 3..60 nodes, 5 columns ~ N(0,1),
     N = 1048576, y ~ N(0,1).
  C3 LM fit: 1024 candidates, k ~ U{1..20}, 3 inputs ~ U(-2,2), N = 262144,
     y = random depth-5 expression + N(0, 0.05).
  C4 uneven: 8 islands x 512, per-island N log-uniform in [512, 4194304],
     k ~ U{0..20}, squared / absolute 50:50, inputs ~ N(0,1), 1% Cauchy y.
  C5 implicit: (x1, x2) on the unit circle with N(0, 0.01) radial noise,
     y ~ U(-1.5, 1.5) appended as input 3, 2048 candidates with <= 8 constants.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .expr import (ADD, COS, DIV, EXP, LOG, MUL, SIN, SUB, SQUARE, Node, const, node, var)

SEEDS = {1: 0x5EED0001, 2: 0x5EED0002, 3: 0x5EED0003, 4: 0x5EED0004, 5: 0x5EED0005}

BASE_BINARY = [ADD, SUB, MUL, DIV]
BASE_UNARY = [EXP, LOG, SIN, COS]


def rng_for(config: int, stream: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=SEEDS[config] + (stream << 32)))


def random_tree(rng, nvars: int, depth: int, p_const: float = 0.35, p_leaf: float = 0.3) -> Node:
    """Grow a random tree of the base operator set with depth <= `depth`."""
    if depth <= 0 or (depth < 4 and rng.random() < p_leaf):
        if rng.random() < p_const:
            return const(float(rng.standard_normal()))
        return var(int(rng.integers(nvars)))
    if rng.random() < 0.3:
        op = BASE_UNARY[int(rng.integers(4))]
        return node(op, random_tree(rng, nvars, depth - 1, p_const, p_leaf))
    op = BASE_BINARY[int(rng.integers(4))]
    return node(op, random_tree(rng, nvars, depth - 1, p_const, p_leaf),
                random_tree(rng, nvars, depth - 1, p_const, p_leaf))


def sized_tree(rng, nvars: int, size: int, p_const: float = 0.4) -> Node:
    """A random tree of exactly `size` nodes (size >= 1)."""
    if size <= 1:
        if rng.random() < p_const:
            return const(float(rng.standard_normal()))
        return var(int(rng.integers(nvars)))
    if size == 2 or rng.random() < 0.3:
        op = BASE_UNARY[int(rng.integers(4))]
        return node(op, sized_tree(rng, nvars, size - 1, p_const))
    left = int(rng.integers(1, size - 1))
    op = BASE_BINARY[int(rng.integers(4))]
    return node(op, sized_tree(rng, nvars, left, p_const),
                sized_tree(rng, nvars, size - 1 - left, p_const))


def _count_consts(t: Node) -> int:
    return len(t.constants())


@dataclass
class Workload:
    name: str
    x: np.ndarray                       # row-major (n, v) f64
    y: np.ndarray                       # f64
    trees: List[Node]
    w: Optional[np.ndarray] = None
    metrics: Optional[np.ndarray] = None
    rowsets: Optional[List[np.ndarray]] = None   # C4: per-island row indices
    island_of: Optional[np.ndarray] = None
    iters: int = 8
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.x.shape[0]

    @property
    def v(self) -> int:
        return self.x.shape[1]

    def cols32(self, x: Optional[np.ndarray] = None) -> np.ndarray:
        x = self.x if x is None else x
        return np.ascontiguousarray(x.T.astype(np.float32)).reshape(-1)

    def cols64(self, x: Optional[np.ndarray] = None) -> np.ndarray:
        x = self.x if x is None else x
        return np.ascontiguousarray(x.T).reshape(-1)


def config1() -> Workload:
    rng = rng_for(1)
    n = 1000
    x = rng.uniform(-3.0, 3.0, size=(n, 1))
    y = 2.0 * x[:, 0] + 1.0 + rng.normal(0.0, 0.1, size=n)
    trees = []
    while len(trees) < 800:
        t = random_tree(rng, 1, int(rng.integers(1, 5)))
        if _count_consts(t) <= 4:
            trees.append(t)
    return Workload("plain-cpu-runnable: N=1000 x1 f64, 4x200 programs depth<=4, squared",
                    x, y, trees, iters=8)


def config2(n: int = 1048576, nprog: int = 4096) -> Workload:
    rng = rng_for(2)
    x = rng.standard_normal(size=(n, 5)).astype(np.float32).astype(np.float64)
    y = rng.standard_normal(size=n).astype(np.float32).astype(np.float64)
    trees = [sized_tree(rng, 5, int(rng.integers(3, 61))) for _ in range(nprog)]
    return Workload(f"batched-eval: {nprog} programs (3-60 nodes) x N={n}, 5 cols fp32", x, y,
                    trees)


def _target_expr(rng, nvars: int, x: np.ndarray, depth: int = 5) -> Node:
    """A random depth-`depth` expression that is finite and varied on x."""
    from . import expr as E
    for _ in range(10000):
        t = random_tree(rng, nvars, depth, p_const=0.3, p_leaf=0.15)
        vals = eval_tree_np(t, x)
        if vals is None or not np.all(np.isfinite(vals)):
            continue
        s = float(np.std(vals))
        if 0.1 < s < 50.0 and float(np.max(np.abs(vals))) < 1e3:
            return t
    raise RuntimeError("no admissible target expression")


def eval_tree_np(t: Node, x: np.ndarray) -> Optional[np.ndarray]:
    """Vectorised f64 evaluation with the evaluator's guards (NaN on a domain
    error); used only to screen synthetic targets."""
    with np.errstate(all="ignore"):
        def ev(n: Node) -> np.ndarray:
            if n.op == 0:
                return np.full(x.shape[0], n.val)
            if n.op == 1:
                return x[:, int(n.val)].astype(np.float64)
            a = ev(n.kids[0])
            if n.op == EXP:
                return np.where(a > 700.0, np.inf, np.exp(np.minimum(a, 700.0)))
            if n.op == LOG:
                return np.where(a > 0, np.log(np.where(a > 0, a, 1.0)), np.nan)
            if n.op == SIN:
                return np.sin(a)
            if n.op == COS:
                return np.cos(a)
            if n.op == SQUARE:
                return a * a
            b = ev(n.kids[1])
            if n.op == ADD:
                return a + b
            if n.op == SUB:
                return a - b
            if n.op == MUL:
                return a * b
            if n.op == DIV:
                return np.where(np.abs(b) < 1e-12, np.nan, a / np.where(b == 0, 1.0, b))
            raise ValueError(n.op)
        return ev(t)


def _term(rng, nvars: int, budget: int) -> Node:
    """One additive term using at most `budget` (>= 1) constants."""
    xi = var(int(rng.integers(nvars)))
    kind = int(rng.integers(6)) if budget >= 3 else (int(rng.integers(3)) if budget >= 2 else 0)
    c = lambda s=1.0: const(float(rng.normal(0.0, s)))
    if kind == 0:           # c * basis(x)
        basis = [xi, node(MUL, xi, var(int(rng.integers(nvars)))), node(SIN, xi), node(COS, xi),
                 node(MUL, xi, xi)][int(rng.integers(5))]
        return node(MUL, c(), basis)
    if kind == 1:           # c * sin(c' x)
        return node(MUL, c(), node(SIN, node(MUL, c(0.7), xi)))
    if kind == 2:           # c * exp(c' x)
        return node(MUL, c(), node(EXP, node(MUL, c(0.3), xi)))
    if kind == 3:           # c * cos(c' x + c'')
        return node(MUL, c(), node(COS, node(ADD, node(MUL, c(0.7), xi), c())))
    if kind == 4:           # c * x / (c' + x^2)... as c * x / (1.5 + c'' * x * x)
        return node(MUL, c(), node(DIV, xi, node(ADD, const(1.5 + abs(float(rng.normal()))),
                                                   node(MUL, node(MUL, c(0.1), xi), xi))))
    # c * exp(c' x^2): a Gaussian-like bump (one inner constant, so every
    # constant of the term is identified).
    return node(MUL, c(), node(EXP, node(MUL, node(MUL, c(0.2), xi), xi)))


def _shape(t: Node) -> str:
    """The term with its constants blanked: two terms of the same shape are
    the same basis, which a search would have merged."""
    return str(t.with_constants([0.0] * len(t.constants())))


def structured_candidate(rng, nvars: int, k: int) -> Node:
    """A random candidate with exactly k constants (k >= 1) and no two terms
    of the same shape."""
    terms, shapes = [], set()
    left = k
    tries = 0
    while left > 0:
        t = _term(rng, nvars, left)
        used = _count_consts(t)
        tries += 1
        if used > left or (_shape(t) in shapes and tries < 200):
            continue
        shapes.add(_shape(t))
        terms.append(t)
        left -= used
    tree = terms[0]
    for t in terms[1:]:
        tree = node(ADD, tree, t)
    return tree


def config3(n: int = 262144, ncand: int = 1024, iters: int = 30) -> Workload:
    rng = rng_for(3)
    x = rng.uniform(-2.0, 2.0, size=(n, 3))
    tgt = _target_expr(rng, 3, x[:2048])
    y = eval_tree_np(tgt, x) + rng.normal(0.0, 0.05, size=n)
    ks = rng.integers(1, 21, size=ncand)
    trees = [structured_candidate(rng, 3, int(k)) for k in ks]
    wl = Workload(f"batched-LM: {ncand} candidates k~U{{1..20}}, 3 inputs, N={n}, "
                  f"{iters} LM iterations", x, y, trees, iters=iters)
    wl.extra["target"] = tgt
    return wl


def config4(nisland: int = 8, pop: int = 512, nmin: int = 512, nmax: int = 4194304,
            iters: int = 8) -> Workload:
    rng = rng_for(4)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=nisland)).astype(np.int64)
    total = int(sizes.sum())
    x = rng.standard_normal(size=(total, 3))
    tgt = _target_expr(rng, 3, x[:2048], depth=4)
    y = eval_tree_np(tgt, x) + rng.normal(0.0, 0.05, size=total)
    out = rng.random(total) < 0.01
    y[out] += np.tan(np.pi * (rng.random(int(out.sum())) - 0.5))
    rowsets, start = [], 0
    for s in sizes:
        rowsets.append(np.arange(start, start + int(s)))
        start += int(s)
    trees, metrics, island = [], [], []
    for i in range(nisland):
        for _ in range(pop):
            k = int(rng.integers(0, 21))
            trees.append(structured_candidate(rng, 3, k) if k > 0 else
                         node(SIN, var(int(rng.integers(3)))))
            metrics.append(0 if rng.random() < 0.5 else 4)
            island.append(i)
    return Workload("uneven: 8 islands x 512, N log-uniform [512, 4194304], mixed metrics",
                    x, y, trees, metrics=np.asarray(metrics, np.int32), rowsets=rowsets,
                    island_of=np.asarray(island, np.int32), iters=iters)


def config5(n: int = 1048576, ncand: int = 2048, iters: int = 8) -> Workload:
    rng = rng_for(5)
    th = rng.uniform(0.0, 2.0 * np.pi, size=n)
    r = 1.0 + rng.normal(0.0, 0.01, size=n)
    x1, x2 = r * np.cos(th), r * np.sin(th)
    yy = rng.uniform(-1.5, 1.5, size=n)
    x = np.stack([x1, x2, yy], axis=1)
    basis = [lambda: node(SQUARE, var(0)), lambda: node(SQUARE, var(1)),
             lambda: node(MUL, var(0), var(1)), lambda: var(0), lambda: var(1), lambda: var(2),
             lambda: node(SIN, var(0)), lambda: node(COS, var(1))]
    trees = []
    for i in range(ncand):
        k = int(rng.integers(2, 9))
        idx = rng.choice(len(basis), size=k - 1, replace=False)
        t = None
        for j in idx:
            term = node(MUL, const(float(rng.normal())), basis[j]())
            t = term if t is None else node(ADD, t, term)
        t = node(SUB, t, const(float(rng.normal())))
        trees.append(t)
    return Workload("implicit: unit circle N=1048576, f(x1,x2,y), <=8 constants", x,
                    np.zeros(n), trees, iters=iters)
