"""Seeded synthetic instances of the benchmark configurations — shared by the
oracle tests, the HIP parity tests and the benchmark (pure data generation).

  * `ising_lattice`  — ferromagnetic open LxL lattice (J=1, h=0), edges in
                       the row-major order of examples/double_layer_ising.jl:110-117;
  * `spin_glass_3reg`— random 3-regular graph by the pairing model (seeded),
                       w uniform +-1, h ~ N(0, 0.1^2), edges sorted (i<j);
  * `batch_instances`— +-J square-lattice spin glasses, L uniform in {8,...,24},
                       beta uniform in [0.2, 1.2], h = 0;
  * `dense_lattice`  — the dense-regime network: an open LxL grid of tensors
                       with bond D on every edge (legs up, down, left, right
                       where present), entries uniform [0, 1) or N(0, 1).
"""
from __future__ import annotations

import numpy as np


def square_edges(L):
    e = []
    for r in range(1, L + 1):
        for c in range(1, L + 1):
            gid = (r - 1) * L + c
            if c < L:
                e.append((gid, gid + 1))
            if r < L:
                e.append((gid, gid + L))
    return e


def ising_lattice(L, J=1.0, h=0.0):
    edges = square_edges(L)
    n = L * L
    return n, edges, np.full(len(edges), float(J)), np.full(n, float(h))


def random_regular_edges(n, d, rng):
    """Pairing model: shuffle n*d stubs, pair neighbours, reject self-loops and
    multi-edges, retry until simple."""
    assert (n * d) % 2 == 0
    while True:
        stubs = np.repeat(np.arange(1, n + 1), d)
        rng.shuffle(stubs)
        pairs = stubs.reshape(-1, 2)
        ok = True
        seen = set()
        for a, b in pairs:
            a, b = int(a), int(b)
            if a == b:
                ok = False
                break
            key = (min(a, b), max(a, b))
            if key in seen:
                ok = False
                break
            seen.add(key)
        if ok:
            return sorted(seen)


def spin_glass_3reg(n=100, seed=0, hstd=0.1):
    rng = np.random.default_rng(seed)
    edges = random_regular_edges(n, 3, rng)
    w = rng.choice(np.array([-1.0, 1.0]), size=len(edges))
    h = rng.normal(0.0, hstd, size=n)
    return n, edges, w, h


def batch_instances(count=64, seed=0, Ls=(8, 10, 12, 14, 16, 18, 20, 22, 24),
                    beta_range=(0.2, 1.2)):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        L = int(rng.choice(np.array(Ls)))
        edges = square_edges(L)
        w = rng.choice(np.array([-1.0, 1.0]), size=len(edges))
        beta = float(rng.uniform(*beta_range))
        out.append((L, L * L, edges, w, np.zeros(L * L), beta))
    return out


def dense_lattice(L, D=4, seed=0, signed=False):
    """(tensors, ixs) of the open L x L dense-regime grid, row-major sites,
    each with its present legs in the order up, down, left, right."""
    rng = np.random.default_rng(seed)
    ts, ixs = [], []
    for r in range(L):
        for c in range(L):
            ix = []
            if r > 0:
                ix.append(f"v{r - 1}_{c}")
            if r < L - 1:
                ix.append(f"v{r}_{c}")
            if c > 0:
                ix.append(f"h{r}_{c - 1}")
            if c < L - 1:
                ix.append(f"h{r}_{c}")
            shape = [D] * len(ix)
            ts.append(rng.standard_normal(shape) if signed else rng.uniform(0.0, 1.0, shape))
            ixs.append(ix)
    return ts, ixs
