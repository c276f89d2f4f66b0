"""Exact (closed-form / brute-force) partition functions — TEST INFRASTRUCTURE ONLY.

Independent of the CATN algorithm; these pin the oracle and the HIP path:
  * `brute_lnZ`      — enumeration over 2^n spin states (test/test_ising.jl:88-100,
                       examples/double_layer_ising.jl:120-135);
  * `chain_lnZ`      — open chain via the 2x2 transfer matrix (test/test_ising.jl:4-10);
  * `lattice_lnZ`    — open LxL square lattice via the row transfer matrix with
                       arbitrary per-edge couplings and fields (L <= 24);
  * `elimination_lnZ`— variable elimination with a min-fill order on any graph;
  * `exact_contract` — dense einsum of a labelled network (test/exact.jl:4-19);
  * `brute_observables` — <s_k> and <s_i s_j> by enumeration (test/test_ising.jl:107-118);
  * `grid_lnZ`       — exact contraction of an open LxL grid of tensors
                       (configs.dense_lattice layout) by row transfer, D^(L+1)
                       entries, for any real or complex entries.
All sums are done in scaled form (log-sum-exp) so large lattices do not overflow.
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence, Tuple

import numpy as np


def square_edges(L: int) -> List[Tuple[int, int]]:
    """examples/double_layer_ising.jl:110-117: row-major, E then S."""
    e = []
    for r in range(1, L + 1):
        for c in range(1, L + 1):
            gid = (r - 1) * L + c
            if c < L:
                e.append((gid, gid + 1))
            if r < L:
                e.append((gid, gid + L))
    return e


def _spins(n):
    idx = np.arange(2 ** n, dtype=np.int64)
    return 1 - 2 * ((idx[:, None] >> np.arange(n)[None, :]) & 1)  # bit k -> s_{k+1}; 0 -> +1


def brute_lnZ(n, edges, w, h, beta):
    """Enumerate all 2^n states (n <= 22)."""
    assert n <= 22
    s = _spins(n).astype(np.float64)
    E = np.zeros(s.shape[0])
    for e, (i, j) in enumerate(edges):
        E += w[e] * s[:, i - 1] * s[:, j - 1]
    E += s @ np.asarray(h, dtype=np.float64)
    x = beta * E
    m = x.max()
    return float(m + math.log(np.exp(x - m).sum()))


def brute_observables(n, edges, w, h, beta):
    s = _spins(n).astype(np.float64)
    E = np.zeros(s.shape[0])
    for e, (i, j) in enumerate(edges):
        E += w[e] * s[:, i - 1] * s[:, j - 1]
    E += s @ np.asarray(h, dtype=np.float64)
    x = beta * E
    p = np.exp(x - x.max())
    p /= p.sum()
    mag = p @ s
    cor = np.array([p @ (s[:, i - 1] * s[:, j - 1]) for (i, j) in edges])
    return mag, cor


def chain_lnZ(L, J, beta):
    """test/test_ising.jl:4-10."""
    T = np.array([[math.exp(beta * J), math.exp(-beta * J)],
                  [math.exp(-beta * J), math.exp(beta * J)]])
    v = np.ones(2)
    return float(math.log(v @ np.linalg.matrix_power(T, L - 1) @ v))


def lattice_lnZ(L, w, h, beta, edges=None):
    """Open LxL lattice, spins numbered (r-1)*L+c, couplings `w` in the order
    of `edges` (default `square_edges(L)`), fields `h`. Row transfer matrix:
    psi over the 2^L states of the current row, bit (c-1) = column c is -1."""
    if edges is None:
        edges = square_edges(L)
    wd = {}
    for e, (i, j) in enumerate(edges):
        wd[(min(i, j), max(i, j))] = float(w[e])
    s = _spins(L).astype(np.float64)            # (2^L, L)
    h = np.asarray(h, dtype=np.float64)
    logscale = 0.0
    psi = None
    for r in range(1, L + 1):
        ids = [(r - 1) * L + c for c in range(1, L + 1)]
        x = beta * (s @ h[np.array(ids) - 1])
        for c in range(1, L):
            x = x + beta * wd.get((ids[c - 1], ids[c]), 0.0) * s[:, c - 1] * s[:, c]
        diag = np.exp(x - x.max())
        logscale += x.max()
        psi = diag if psi is None else psi * diag
        if r == L:
            break
        # vertical bonds to the next row: apply a 2x2 matrix on each column axis
        t = psi.reshape([2] * L, order="F")
        for c in range(1, L + 1):
            wv = wd.get((ids[c - 1], ids[c - 1] + L), 0.0)
            K = np.array([[math.exp(beta * wv), math.exp(-beta * wv)],
                          [math.exp(-beta * wv), math.exp(beta * wv)]])
            t = np.moveaxis(np.tensordot(K, t, axes=([0], [c - 1])), 0, c - 1)
        psi = t.reshape(-1, order="F")
        mx = psi.max()
        psi = psi / mx
        logscale += math.log(mx)
    return float(logscale + math.log(psi.sum()))


def elimination_lnZ(n, edges, w, h, beta, return_width=False):
    """Variable elimination with a greedy min-fill order (ties: min degree,
    then smallest id). Factors are stored in log-scaled form."""
    factors: List[Tuple[Tuple[int, ...], np.ndarray]] = []
    for k in range(1, n + 1):
        factors.append(((k,), np.array([math.exp(beta * h[k - 1]), math.exp(-beta * h[k - 1])])))
    for e, (i, j) in enumerate(edges):
        ww = float(w[e])
        factors.append(((i, j), np.array([[math.exp(beta * ww), math.exp(-beta * ww)],
                                          [math.exp(-beta * ww), math.exp(beta * ww)]])))
    adj: Dict[int, set] = {k: set() for k in range(1, n + 1)}
    for i, j in edges:
        adj[i].add(j)
        adj[j].add(i)
    remaining = set(range(1, n + 1))
    logscale = 0.0
    width = 0
    while remaining:
        best = None
        for v in remaining:
            nb = adj[v]
            fill = sum(1 for a in nb for b in nb if a < b and b not in adj[a])
            key = (fill, len(nb), v)
            if best is None or key < best:
                best = key
        v = best[2]
        nb = sorted(adj[v])
        width = max(width, len(nb))
        inv = [f for f in factors if v in f[0]]
        factors = [f for f in factors if v not in f[0]]
        scope = tuple(nb + [v])
        letters = {var: chr(97 + q) if q < 26 else chr(65 + q - 26) for q, var in enumerate(scope)}
        subs = ",".join("".join(letters[x] for x in f[0]) for f in inv)
        out = "".join(letters[x] for x in nb)
        newf = np.einsum(subs + "->" + out, *[f[1] for f in inv], optimize=True)
        mx = float(np.max(np.abs(newf)))
        newf = newf / mx
        logscale += math.log(mx)
        factors.append((tuple(nb), np.asarray(newf)))
        for a in nb:
            adj[a].discard(v)
            for b in nb:
                if a != b:
                    adj[a].add(b)
        del adj[v]
        remaining.discard(v)
    total = 1.0
    for f in factors:
        total *= float(f[1])
    res = float(logscale + math.log(total))
    return (res, width) if return_width else res


def exact_contract(tensors, ixs):
    """test/exact.jl:4-19: output labels are those appearing once, first-seen order."""
    counts: Dict = {}
    orderl = []
    for ix in ixs:
        for l in ix:
            if l not in counts:
                orderl.append(l)
            counts[l] = counts.get(l, 0) + 1
    letters = {}
    for q, l in enumerate(orderl):
        letters[l] = chr(97 + q) if q < 26 else chr(65 + q - 26)
    subs = ",".join("".join(letters[l] for l in ix) for ix in ixs)
    out = "".join(letters[l] for l in orderl if counts[l] == 1)
    return np.einsum(subs + "->" + out, *tensors, optimize="greedy")


def grid_lnZ(ts, L):
    """Exact contraction of an open L x L grid (configs.dense_lattice: site
    r * L + c, present legs up, down, left, right) by row transfer: the state
    carries the open vertical bonds below the finished rows and one
    horizontal bond inside the current row, rescaled per row. Returns
    (ln|Z|, Z / |Z|)."""
    psi, lnscale = np.ones(()), 0.0
    for r in range(L):
        for c in range(L):
            ax = [a for a, on in (("u", r > 0), ("d", r < L - 1), ("l", c > 0), ("r", c < L - 1)) if on]
            nd = c if r < L - 1 else 0          # finished down legs of this row
            pa, ta = [], []
            if c > 0:
                pa.append(nd)
                ta.append(ax.index("l"))
            if r > 0:
                pa.append(nd + (1 if c > 0 else 0))
                ta.append(ax.index("u"))
            out = np.tensordot(psi, ts[r * L + c], axes=(pa, ta))
            rest = [a for i, a in enumerate(ax) if i not in ta]
            nu = (L - c - 1) if r > 0 else 0    # up legs still open
            order = list(range(nd))
            order += [nd + nu + rest.index(a) for a in ("d", "r") if a in rest]
            order += list(range(nd, nd + nu))
            psi = np.transpose(out, order)
        s = float(np.max(np.abs(psi)))
        psi = psi / s
        lnscale += math.log(s)
    v = complex(psi.item()) if np.iscomplexobj(psi) else float(psi)
    return lnscale + math.log(abs(v)), v / abs(v)
