"""FP64 CPU restatement of CATN.jl's contraction engine — TEST INFRASTRUCTURE ONLY.

This module is the oracle: a line-by-line NumPy restatement of the reference
algorithm (xuanzhaogao/CATN.jl), used by `tests/`, by `__graft_entry__.smoke()`
and by `bench.py`'s `cpu_baseline` leg as the checker. The product path
(`catn_amd`) never imports it and fails loudly when its HIP library is missing.

Conventions kept from the reference:
  * MPS sites are rank-3 arrays `(dl, d, dr)` with Julia column-major reshape
    semantics (every reshape here is `order="F"`);
  * node ids and neighbour values are the reference's 1-based ids; positions
    inside `neighbor` lists are 0-based Python indices, and every place where
    the reference does 1-based index arithmetic (`cld`, `idx1+1`, ...) is
    translated explicitly with a comment;
  * LAPACK is called the way Julia's LinearAlgebra does: `svd` -> gesdd thin
    ('S'), fallback gesvd; `qr` -> Householder QR (R diagonal signs as dlarfg).
    NumPy bundles OpenBLAS 0.3.29, the version the reference's Manifest pins.

Every function cites the reference file:line it restates.

Parity status of each function (what pins it) is listed in `PINS` at the end
of this file and checked by `tests/test_oracle.py`.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# ---------------------------------------------------------------------------
# schedule logger: records every truncation decision of the sweep
# ---------------------------------------------------------------------------


class Schedule:
    """Per-primitive log: kind, matrix shape, kept count and the margin of the
    closest singular value to the absolute cutoff (log10 ratio)."""

    def __init__(self):
        self.events: List[tuple] = []
        self.edges: List[Tuple[int, int]] = []
        # structural trace of contraction!: per step (sel_i, sel_j, i, j,
        # reversed_i, reversed_j, idx_j_in_i, idx_i_in_j); per duplicate bond
        # (step, k, cross)
        self.steps: List[tuple] = []
        self.dups: List[tuple] = []
        self.enabled = True

    def log(self, kind, shape, nkeep, S, cutoff):
        if not self.enabled:
            return
        margin = float("inf")
        if S is not None and len(S) and cutoff > 0:
            Sp = np.abs(np.asarray(S, dtype=np.float64))
            with np.errstate(divide="ignore"):
                r = np.abs(np.log10(np.maximum(Sp, 1e-300) / cutoff))
            margin = float(r.min())
        self.events.append((kind, tuple(int(x) for x in shape), int(nkeep), margin))


_SCHED: Optional[Schedule] = None


def set_schedule(s: Optional[Schedule]):
    global _SCHED
    _SCHED = s


def _log(kind, shape, nkeep, S, cutoff):
    if _SCHED is not None:
        _SCHED.log(kind, shape, nkeep, S, cutoff)


# ---------------------------------------------------------------------------
# src/linalg_utils.jl
# ---------------------------------------------------------------------------

class SVDFailure(Exception):
    pass


# host-settable fault flag: the primary SVD "fails" so the fallback runs
# (the reference has no injection hook; src/linalg_utils.jl:16-21)
FORCE_GESDD_FAILURE = False


def _svd(A):
    """Julia `svd(A)` thin -> (U, S, V) with A = U diag(S) V'
    (src/linalg_utils.jl:17; gesdd 'S'); on failure gesvd (:20)."""
    try:
        if FORCE_GESDD_FAILURE:
            raise np.linalg.LinAlgError("forced")
        U, S, Vh = np.linalg.svd(A, full_matrices=False)
    except np.linalg.LinAlgError:
        import scipy.linalg as sl
        U, S, Vh = sl.svd(A, full_matrices=False, lapack_driver="gesvd")
    return U, S, Vh.conj().T


def tsvd(A, cutoff=1e-15, maxdim=None, kind="tsvd"):
    """src/linalg_utils.jl:15-28. Keep min(count(S>cutoff), maxdim, nfull),
    at least 1; discarded = sum of dropped S."""
    U, S, V = _svd(A)
    nfull = len(S)
    nkeep = int(np.count_nonzero(S > cutoff))
    if nkeep == 0:
        nkeep = 1
    else:
        nkeep = min(nkeep, nfull if maxdim is None else maxdim, nfull)
    discarded = float(np.sum(S[nkeep:])) if nkeep < nfull else 0.0
    _log(kind, A.shape, nkeep, S, cutoff)
    return U[:, :nkeep], S[:nkeep], V[:, :nkeep], discarded


def thin_q(A):
    """src/linalg_utils.jl:32-37 (`_thin_q`) together with `qr(A)`: the thin
    Householder Q (m x n) and R (n x n) of a tall A."""
    Q, R = np.linalg.qr(A, mode="reduced")
    return Q, R


def rsvd(A, k, oversample=10, power=10, rng=None):
    """src/linalg_utils.jl:44-58. The reference draws G from Julia's global RNG
    inside swap! (src/mps_node.jl:155); here G comes from a seeded NumPy
    generator (the stream itself is not reproducible outside Julia)."""
    if rng is None:
        rng = np.random.default_rng(0)
    m, n = A.shape
    p = min(n, oversample * k)
    G = rng.standard_normal((n, p))
    if np.iscomplexobj(A):
        G = G.astype(A.dtype)
    Y = A @ G
    for _ in range(power):
        M = A @ (A.conj().T @ Y)
        Y, _ = np.linalg.qr(M, mode="reduced")
    Q, _ = np.linalg.qr(Y, mode="reduced")
    B = Q.conj().T @ A
    U, S, V = _svd(B)
    kk = min(k, U.shape[1])
    _log("rsvd", A.shape, kk, None, 0.0)
    return (Q @ U)[:, :kk], S[:kk], V[:, :kk]


# ---------------------------------------------------------------------------
# src/mps_node.jl
# ---------------------------------------------------------------------------

def _r(a, *shape):
    return np.reshape(a, shape, order="F")


@dataclass
class MPSNode:
    """src/mps_node.jl:1-29. `cano` is 1-based like the reference."""
    mps: List[np.ndarray]
    neighbor: List[int]
    cano: int = 0
    chi: int = 32
    cutoff: float = 1e-15
    norm_method: int = 1
    svdopt: bool = True
    swapopt: bool = True

    @staticmethod
    def from_tensor(tensor, neighbor, chi=32, cutoff=1e-15, norm_method=1,
                    svdopt=True, swapopt=True):
        """src/mps_node.jl:23-29."""
        mps = raw2mps(np.asarray(tensor), chi, cutoff)
        return MPSNode(mps, list(neighbor), len(mps), chi, cutoff, norm_method,
                       svdopt, swapopt)


def raw2mps(tensor, chi, cutoff):
    """src/mps_node.jl:31-50."""
    nd = tensor.ndim
    if nd == 0:
        return []
    dims = tensor.shape
    if nd == 1:
        return [_r(tensor, 1, dims[0], 1)]
    mps = []
    R = _r(tensor, 1, *dims)
    dleft = 1
    for i in range(nd - 1):
        M = _r(R, dleft * dims[i], -1)
        U, S, V, _ = tsvd(M, cutoff=cutoff, maxdim=chi, kind="raw2mps")
        x = len(S)
        mps.append(_r(U, dleft, dims[i], x))
        R = _r(np.diag(S) @ V.conj().T, x, *dims[i + 1:])
        dleft = x
    mps.append(_r(R, dleft, dims[-1], 1))
    return mps


def mps2raw(node: MPSNode):
    """src/mps_node.jl:52-70."""
    mps = node.mps
    if not mps:
        return np.ones(())
    A = mps[0]
    cur = _r(A, A.shape[1], A.shape[2])
    pdim = A.shape[1]
    phys = [A.shape[1]]
    for B in mps[1:]:
        r, p, r2 = B.shape
        res = cur @ _r(B, r, p * r2)
        pdim *= p
        cur = _r(_r(res, pdim // p, p, r2), pdim, r2)
        phys.append(p)
    return _r(cur, *phys)


def cano_to(node: MPSNode, idx: int):
    """src/mps_node.jl:72-104. `idx` is 1-based; 0 means the last site."""
    if idx == 0:
        idx = len(node.mps)
    if node.cano == idx:
        return node
    if node.cano < idx:
        for i in range(node.cano, idx):            # i = cano .. idx-1 (1-based)
            t = node.mps[i - 1]
            dl, d, dr = t.shape
            U, S, V, _ = tsvd(_r(t, dl * d, dr), cutoff=0.0, kind="cano_r")
            nkeep = int(np.count_nonzero(S > node.cutoff))
            nkeep = len(S) if nkeep == 0 else nkeep
            _log("cano_r_keep", (dl * d, dr), nkeep, S, node.cutoff)
            U, S, V = U[:, :nkeep], S[:nkeep], V[:, :nkeep]
            node.mps[i - 1] = _r(U, dl, d, nkeep)
            R = np.diag(S) @ V.conj().T
            node.mps[i] = np.einsum("ij,jab->iab", R, node.mps[i])
            node.cano = i + 1
    else:
        for i in range(node.cano, idx, -1):        # i = cano .. idx+1
            t = node.mps[i - 1]
            dl, d, dr = t.shape
            Mt = _r(np.transpose(t, (1, 2, 0)), d * dr, dl)
            U, S, V, _ = tsvd(Mt, cutoff=0.0, kind="cano_l")
            nkeep = int(np.count_nonzero(S > node.cutoff))
            nkeep = len(S) if nkeep == 0 else nkeep
            _log("cano_l_keep", (d * dr, dl), nkeep, S, node.cutoff)
            U, S, V = U[:, :nkeep], S[:nkeep], V[:, :nkeep]
            node.mps[i - 1] = np.transpose(_r(U, d, dr, nkeep), (2, 0, 1))
            R = np.diag(S) @ V.conj().T             # (nkeep, dl)
            node.mps[i - 2] = np.einsum("abc,cd->abd", node.mps[i - 2], R.T)
            node.cano = i - 1
    return node


def left_canonical(node: MPSNode):
    """src/mps_node.jl:106."""
    node.cano = 1
    return cano_to(node, len(node.mps))


def order(node):
    return len(node.mps)


def shape(node):
    return [t.shape[1] for t in node.mps]


# thresholds of the randomized-SVD switch (src/mps_node.jl:154)
def rsvd_switch(swapopt, rows, cols):
    return swapopt and ((rows > 7000 and cols > 7000) or rows > 20000 or cols > 20000)


_RSVD_RNG = np.random.default_rng(12345)


def swap(node: MPSNode, i: int, j: int) -> float:
    """src/mps_node.jl:126-178. i, j are 1-based adjacent positions."""
    assert abs(i - j) == 1, "swap!: i and j must be consecutive indices"
    assert 1 <= i <= len(node.mps) and 1 <= j <= len(node.mps)
    if node.cano != i and node.cano != j:
        target = i if abs(node.cano - i) < abs(node.cano - j) else j
        cano_to(node, target)
    mps = node.mps
    if i < j:
        tl, tr = mps[i - 1], mps[j - 1]
    else:
        tl, tr = mps[j - 1], mps[i - 1]
    d0, d1, d2, d3 = tl.shape[0], tr.shape[1], tl.shape[1], tr.shape[2]
    W = np.einsum("ijk,kab->iajb", tl, tr)
    M = _r(W, d0 * d1, d2 * d3)
    rows, cols = M.shape
    if rsvd_switch(node.swapopt, rows, cols):
        U, S, V = rsvd(M, node.chi, 10, 10, rng=_RSVD_RNG)
        err = 0.0
    else:
        U, S, V, err = tsvd(M, cutoff=node.cutoff, maxdim=node.chi, kind="swap")
    myd = len(S)
    if i < j:
        node.mps[i - 1] = _r(U, d0, d1, myd)
        node.mps[j - 1] = _r(np.diag(S) @ V.conj().T, myd, d2, d3)
    else:
        node.mps[j - 1] = _r(U @ np.diag(S), d0, d1, myd)
        node.mps[i - 1] = _r(V.conj().T.copy(), myd, d2, d3)
    node.cano = j
    nb = node.neighbor
    nb[i - 1], nb[j - 1] = nb[j - 1], nb[i - 1]
    return err


def move(node, a, b):
    """src/mps_node.jl:186-199 (1-based a, b)."""
    err = 0.0
    if a == b:
        return err
    if b > a:
        for idx in range(a, b):
            err += swap(node, idx, idx + 1)
    else:
        for idx in range(a, b, -1):
            err += swap(node, idx, idx - 1)
    return err


def move2tail(node, idx):
    """src/mps_node.jl:208-217."""
    err = 0.0
    n = len(node.mps)
    if idx == n:
        cano_to(node, n)
        return err
    for k in range(idx, n):
        err += swap(node, k, k + 1)
    cano_to(node, n)
    return err


def move2head(node, idx):
    """src/mps_node.jl:225-229."""
    err = move(node, idx, 1)
    cano_to(node, 1)
    return err


def reverse(node):
    """src/mps_node.jl:238-244."""
    if len(node.mps) <= 1:
        return node
    node.mps = [np.transpose(t, (2, 1, 0)).copy() for t in node.mps[::-1]]
    node.neighbor.reverse()
    node.cano = len(node.mps) + 1 - node.cano
    return node


def find_neighbor(node, j):
    """src/mps_node.jl:257-260: 1-based position, 0 when absent."""
    try:
        return node.neighbor.index(j) + 1
    except ValueError:
        return 0


def add_neighbor(node, n, pos=0):
    """src/mps_node.jl:267-274 (pos 1-based; 0 appends)."""
    if pos == 0:
        node.neighbor.append(n)
    else:
        node.neighbor.insert(pos - 1, n)
    return node


def delete_neighbor(node, n):
    """src/mps_node.jl:282-287: returns the former 1-based position."""
    idx = find_neighbor(node, n)
    if idx == 0:
        raise KeyError(f"delete_neighbor!: neighbor {n} not found")
    del node.neighbor[idx - 1]
    return idx


def logdim(node, idx=None):
    """src/mps_node.jl:299-311. Left-to-right Float64 sum of log2 dims."""
    if idx is not None:
        return math.log2(float(node.mps[idx - 1].shape[1]))
    if not node.mps:
        return 0.0
    s = 0.0
    for t in node.mps:
        s += math.log2(float(t.shape[1]))
    return s


def lognorm(node):
    """src/mps_node.jl:319-328."""
    if not node.mps:
        return 0.0, 1.0
    if len(node.mps) == 1 and node.mps[0].shape[1] == 1:
        z = node.mps[0].sum()
        return math.log(abs(z)), (z / abs(z) if z != 0 else 0.0 * z)
    raise ValueError("lognorm: computing norm of a non-scalar MPS is not supported in contraction")


def clear(node):
    """src/mps_node.jl:335-339."""
    node.mps.clear()
    node.neighbor.clear()
    return node


def compress(node):
    """src/mps_node.jl:356-376."""
    err = 0.0
    if not node.mps:
        return err
    left_canonical(node)
    mps = node.mps
    for j in range(len(mps), 1, -1):
        i = j - 1
        tl, tr = mps[i - 1], mps[j - 1]
        d0, d1 = tl.shape[0], tl.shape[1]
        d2, d3 = tr.shape[1], tr.shape[2]
        mat = _r(np.einsum("ijk,kab->ijab", tl, tr), d0 * d1, d2 * d3)
        U, S, V, e = tsvd(mat, cutoff=node.cutoff, maxdim=node.chi, kind="compress")
        err += e
        myd = len(S)
        mps[i - 1] = _r(U @ np.diag(S), d0, d1, myd)
        mps[j - 1] = _r(V.conj().T.copy(), myd, d2, d3)
    node.cano = 1
    return err


def compress_opt(node):
    """src/mps_node.jl:394-447."""
    err = 0.0
    if not node.mps:
        return err
    left_canonical(node)
    mps = node.mps
    for j in range(len(mps), 1, -1):
        i = j - 1
        tl, tr = mps[i - 1], mps[j - 1]
        d0, d1 = tl.shape[0], tl.shape[1]
        d2, d3 = tr.shape[1], tr.shape[2]
        dd = tl.shape[2]
        matl = _r(tl, d0 * d1, dd)
        matr = _r(tr, dd, d2 * d3)
        fl = matl.shape[0] > matl.shape[1]
        fr = matr.shape[0] < matr.shape[1]
        if fl:
            Ql, Rl = thin_q(matl)
        else:
            Rl = matl
        if fr:
            Qr, Rr = thin_q(matr.conj().T)
        else:
            Rr = matr.conj().T
        mat = Rl @ Rr.conj().T
        U, S, V, e = tsvd(mat, cutoff=node.cutoff, maxdim=node.chi, kind="compress_opt")
        err += e
        myd = len(S)
        U = U @ np.diag(S)
        if fl:
            U = Ql @ U
        if fr:
            V = Qr @ V
        mps[i - 1] = _r(U, d0, d1, myd)
        mps[j - 1] = _r(V.conj().T.copy(), myd, d2, d3)
    node.cano = 1
    return err


def _norm(t, method):
    if method == 1:
        return float(np.linalg.norm(t.ravel()))
    if method == 2:
        return float(np.max(np.abs(t)))
    return 1.0


def eat(node, nodej, idx, idxi):
    """src/mps_node.jl:463-575. idx, idxi are 1-based leg positions.
    Returns (lognorm, error, phase)."""
    err = 0.0
    one = node.mps[0].dtype.type(1) if node.mps else 1.0
    if len(node.mps) == 1:                                        # case (a)
        assert len(nodej.mps) == 1
        vi = node.mps[0].ravel(order="F")
        vj = nodej.mps[0].ravel(order="F")
        r = np.sum(vi * vj)
        absr = abs(r)
        if absr <= node.cutoff:
            node.mps.clear()
            del node.neighbor[idx - 1]
            return 0.0, 0.0, one
        node.mps.clear()
        del node.neighbor[idx - 1]
        return math.log(absr), 0.0, r / absr
    err += move2tail(node, idx)
    last = node.mps[-1]
    mati = _r(last, last.shape[0], last.shape[1])
    if len(nodej.mps) == 1:                                       # case (b)
        tj = nodej.mps[0]
        matj = _r(tj, tj.shape[1], 1)
        mat = mati @ matj
        new = np.einsum("ijk,ka->ija", node.mps[-2], mat)
        nrm = _norm(new, node.norm_method)
        node.mps[-2] = new
        node.cano -= 1
        node.mps.pop()
        del node.neighbor[-1]
        if nrm <= node.cutoff:
            return 0.0, err, one
        node.mps[-1] = node.mps[-1] / nrm
        return math.log(nrm), err, one
    err += move2head(nodej, idxi)                                 # case (c)
    h = nodej.mps[0]
    matj = _r(h, h.shape[1], h.shape[2])
    mat = mati @ matj
    node.mps[-2] = np.einsum("ijk,ka->ija", node.mps[-2], mat)
    node.mps.pop()
    node.cano = len(node.mps)
    node.mps.extend(nodej.mps[1:])
    del node.neighbor[-1]
    node.neighbor.extend(nodej.neighbor[1:])
    cano_to(node, len(node.mps))
    center = node.mps[node.cano - 1]
    nrm = _norm(center, node.norm_method)
    if nrm <= node.cutoff:
        return 0.0, err, one
    node.mps[node.cano - 1] = center / nrm
    return math.log(nrm), err, one


def merge(node, j, cross=False):
    """src/mps_node.jl:595-622."""
    pos = [k + 1 for k, v in enumerate(node.neighbor) if v == j]
    if len(pos) != 2:
        raise ValueError(f"merge!: expected exactly 2 sites with neighbor {j}, found {len(pos)}")
    idx1, idx2 = pos
    target = idx1 if cross else idx1 + 1
    err = move(node, idx2, target)
    del node.neighbor[target - 1]
    cano_to(node, idx1)
    a, b = node.mps[idx1 - 1], node.mps[idx1]
    dl, dr2 = a.shape[0], b.shape[2]
    node.mps[idx1 - 1] = _r(np.einsum("ijk,kab->ijab", a, b), dl, -1, dr2)
    del node.mps[idx1]
    cano_to(node, idx1)
    return err


# ---------------------------------------------------------------------------
# src/tensor_network.jl
# ---------------------------------------------------------------------------

@dataclass
class TensorNetwork:
    """src/tensor_network.jl:1-22. `tensors` is insertion-ordered by id; the
    reference's Dict iteration order is replaced by ascending id order."""
    tensors: Dict[int, MPSNode]
    Dmax: int = 32
    chi: int = 32
    cutoff: float = 1e-15
    norm_method: int = 1
    select: int = 1
    reverse: bool = True
    svdopt: bool = True
    swapopt: bool = True
    compress: bool = False
    cut_bond: bool = False
    edge_count: Dict[int, List[List[int]]] = field(default_factory=dict)
    lnZ: float = 0.0
    sign: float = 1.0
    psi: float = 1.0
    maxdim_intermediate: int = -1
    num_isolated: int = 0
    n: int = 0
    beta: float = 1.0


def tensor_network(tensors, ixs, Dmax=32, chi=32, cutoff=1e-15, norm_method=1,
                   select=1, reverse=True, svdopt=True, swapopt=True,
                   compress=False, cut_bond=False, seed=1):
    """src/tensor_network.jl:41-140."""
    n = len(tensors)
    assert len(ixs) == n
    for t in range(n):
        if len(set(ixs[t])) != len(ixs[t]):
            raise ValueError(f"Self-loop detected in tensor {t + 1}")
    label_map: Dict = {}
    for t in range(n):
        for k, lbl in enumerate(ixs[t]):
            label_map.setdefault(lbl, []).append((t + 1, k))
    for lbl, occ in label_map.items():
        if len(occ) > 2:
            raise ValueError(f"Label {lbl} appears in {len(occ)} tensors (max 2 allowed)")
    sentinel = [0]
    nbv = [[0] * len(ixs[t]) for t in range(n)]
    for lbl, occ in label_map.items():
        if len(occ) == 2:
            (t1, k1), (t2, k2) = occ
            nbv[t1 - 1][k1] = t2
            nbv[t2 - 1][k2] = t1
        else:
            sentinel[0] -= 1
            t, k = occ[0]
            nbv[t - 1][k] = sentinel[0]
    nodes = {}
    for t in range(n):
        nodes[t + 1] = MPSNode.from_tensor(tensors[t], nbv[t], chi=chi, cutoff=cutoff,
                                           norm_method=norm_method, svdopt=svdopt,
                                           swapopt=swapopt)
    num_iso = sum(1 for t in range(n) if all(nb < 0 for nb in nbv[t]))
    return TensorNetwork(nodes, Dmax, chi, cutoff, norm_method, select, reverse, svdopt,
                         swapopt, compress, cut_bond, {}, 0.0, 1.0, 1.0, -1, num_iso, n, 1.0)


def dim_after_merge(tn, i, j):
    """src/tensor_network.jl:152-155 (Julia round: ties to even == Python round)."""
    ni = tn.tensors[i]
    idx = find_neighbor(ni, j)
    return int(round(logdim(ni) + logdim(tn.tensors[j]) - 2 * logdim(ni, idx)))


def count_add_edges(tn, edges):
    """src/tensor_network.jl:167-176."""
    for i, j in edges:
        c = dim_after_merge(tn, i, j)
        tn.edge_count.setdefault(c, []).append([i, j])


def count_add_nodes(tn, nodes):
    """src/tensor_network.jl:185-197."""
    edges = set()
    for i in nodes:
        if i <= 0:
            continue
        for j in tn.tensors[i].neighbor:
            if j <= 0:
                continue
            edges.add((min(i, j), max(i, j)))
    count_add_edges(tn, sorted(edges))


def count_remove_nodes(tn, nodes):
    """src/tensor_network.jl:213-225."""
    for nid in nodes:
        if nid <= 0:
            continue
        for nb in list(tn.tensors[nid].neighbor):
            if nb <= 0:
                continue
            i, j = min(nid, nb), max(nid, nb)
            c = dim_after_merge(tn, i, j)
            if c in tn.edge_count:
                tn.edge_count[c] = [p for p in tn.edge_count[c] if p != [i, j]]


def select_edge_init(tn):
    """src/tensor_network.jl:233-236."""
    tn.edge_count.clear()
    count_add_nodes(tn, list(tn.tensors.keys()))


def select_edge_min_dim(tn):
    """src/tensor_network.jl:248-253."""
    cost = min(k for k, v in tn.edge_count.items() if v)
    tn.maxdim_intermediate = max(tn.maxdim_intermediate, cost)
    p = tn.edge_count[cost][0]
    return p[0], p[1]


def select_edge_min_dim_triangle(tn):
    """src/tensor_network.jl:262-288."""
    cost = min(k for k, v in tn.edge_count.items() if v)
    tn.maxdim_intermediate = max(tn.maxdim_intermediate, cost)
    cands = tn.edge_count[cost]
    best, bestc = cands[0], -1
    for p in cands:
        i, j = p
        neigh_j = {nb for nb in tn.tensors[j].neighbor if nb > 0 and nb != i}
        both = sum(1 for k in neigh_j if find_neighbor(tn.tensors[k], i) > 0)
        if both > bestc:
            bestc, best = both, p
    return best[0], best[1]


def select_edge_sequentially(tn):
    """src/tensor_network.jl:300-316 (node iteration in ascending id order)."""
    best, bs = None, None
    for i in tn.tensors:
        for nb in tn.tensors[i].neighbor:
            if nb <= 0 or not i < nb:
                continue
            if bs is None or i + nb < bs:
                bs, best = i + nb, (i, nb)
    if best is None:
        raise ValueError("select_edge_sequentially: no edges remain")
    return best


def _cut_common(tn, i, idx_j_in_i, opt):
    node_i = tn.tensors[i]
    j = node_i.neighbor[idx_j_in_i - 1]
    if j <= 0:
        raise ValueError(f"cut_bondim: leg {idx_j_in_i} of node {i} is an open leg (no neighbor)")
    node_j = tn.tensors[j]
    idx_i_in_j = find_neighbor(node_j, i)
    if idx_i_in_j == 0:
        raise ValueError(f"cut_bondim: node {i} not found as neighbor of node {j}")
    return cut_legs(node_i, idx_j_in_i, node_j, idx_i_in_j, opt, tn.cutoff, tn.Dmax)


def cut_legs(node_i, idx_j_in_i, node_j, idx_i_in_j, opt, cutoff, Dmax):
    """The cut of src/tensor_network.jl:342-388 / :399-477 on an explicit pair
    of legs that carry one bond (leg idx_j_in_i of node_i, leg idx_i_in_j of
    node_j)."""
    if opt:
        cano_to(node_i, idx_j_in_i)
        cano_to(node_j, idx_i_in_j)
    Ai, Aj = node_i.mps[idx_j_in_i - 1], node_j.mps[idx_i_in_j - 1]
    dal, d, dar = Ai.shape
    dbl, _, dbr = Aj.shape
    mati = _r(np.transpose(Ai, (0, 2, 1)), dal * dar, d)
    matj = _r(np.transpose(Aj, (0, 2, 1)), dbl * dbr, d)
    fl = fr = False
    if opt and mati.shape[0] > mati.shape[1]:
        qi, ri = thin_q(mati)
        fl = True
    else:
        ri = mati
    if opt and matj.shape[0] > matj.shape[1]:
        qj, rj = thin_q(matj)
        fr = True
    else:
        rj = matj
    merged = ri @ rj.T
    U, S, V, disc = tsvd(merged, cutoff=cutoff, kind="cut")
    myd = min(len(S), Dmax)
    if myd == 0:
        myd = 1
    extra = float(np.sum(S[myd:])) if myd < len(S) else 0.0
    err = disc + extra
    _log("cut_keep", merged.shape, myd, S, cutoff)
    S, U, V = S[:myd], U[:, :myd], V[:, :myd]
    sq = np.sqrt(S)
    mi = U * sq[None, :]
    mj = V.conj() * sq[None, :]
    if fl:
        mi = qi @ mi
    if fr:
        mj = qj @ mj
    node_i.mps[idx_j_in_i - 1] = np.transpose(_r(mi, dal, dar, myd), (0, 2, 1)).copy()
    node_j.mps[idx_i_in_j - 1] = np.transpose(_r(mj, dbl, dbr, myd), (0, 2, 1)).copy()
    return err


def cut_bondim(tn, i, idx_j_in_i):
    """src/tensor_network.jl:342-388."""
    if tn.Dmax < 0:
        return 0.0
    return _cut_common(tn, i, idx_j_in_i, False)


def cut_bondim_opt(tn, i, idx_j_in_i):
    """src/tensor_network.jl:399-477."""
    if tn.Dmax < 0:
        return 0.0
    return _cut_common(tn, i, idx_j_in_i, True)


def network_lognorm(tn):
    """src/tensor_network.jl:489-501 (ascending-id summation order)."""
    total, sgn = 0.0, 1.0
    for node in tn.tensors.values():
        if not all(s.shape[1] == 1 for s in node.mps):
            continue
        ln, sg = lognorm(node)
        total += ln
        sgn *= sg
    return total, sgn


def has_real_edges(tn):
    """src/tensor_network.jl:535-542."""
    return any(nb > 0 for node in tn.tensors.values() for nb in node.neighbor)


def contraction(tn: TensorNetwork, sched: Optional[Schedule] = None):
    """src/tensor_network.jl:556-717. Returns (lnZ, error, psi)."""
    prev = _SCHED
    if sched is not None:
        set_schedule(sched)
    try:
        return _contraction(tn, sched)
    finally:
        set_schedule(prev)


def _cld(a, b):
    return -(-a // b)


def _contraction(tn, sched):
    err = 0.0
    tn.psi = 1.0
    tn.lnZ = 0.0
    if tn.select in (0, 1) and not tn.edge_count:
        select_edge_init(tn)
    while has_real_edges(tn):
        if tn.select == 0:
            i, j = select_edge_min_dim(tn)
        elif tn.select == 1:
            i, j = select_edge_min_dim_triangle(tn)
        else:
            i, j = select_edge_sequentially(tn)
        if sched is not None:
            sched.edges.append((i, j))
        sel = (i, j)
        ni, nj = tn.tensors[i], tn.tensors[j]
        if order(nj) > order(ni):
            i, j = j, i
            ni, nj = nj, ni
        affected = [i, j] + list(ni.neighbor) + list(nj.neighbor)
        if tn.select in (0, 1):
            count_remove_nodes(tn, affected)
        neigh_j = list(nj.neighbor)
        idx_j_in_i = find_neighbor(ni, j)
        idx_i_in_j = find_neighbor(nj, i)
        rev_i = rev_j = False
        if tn.reverse:
            # 1-based comparisons of src/tensor_network.jl:609 and :614
            if idx_j_in_i < _cld(len(ni.neighbor), 2):
                reverse(ni)
                idx_j_in_i = find_neighbor(ni, j)
                rev_i = True
            if idx_i_in_j >= _cld(len(nj.neighbor), 2):
                reverse(nj)
                neigh_j = list(nj.neighbor)
                idx_i_in_j = find_neighbor(nj, i)
                rev_j = True
        if sched is not None:
            sched.steps.append(sel + (i, j, int(rev_i), int(rev_j), idx_j_in_i, idx_i_in_j))
        duplicate = []
        for l in range(1, len(neigh_j) + 1):
            if l == idx_i_in_j:
                continue
            k = neigh_j[l - 1]
            if k <= 0:
                continue
            idx_k_in_i = find_neighbor(ni, k)
            nk = tn.tensors[k]
            idx_i_in_k = find_neighbor(nk, i)
            idx_j_in_k = delete_neighbor(nk, j)
            add_neighbor(nk, i, idx_j_in_k)
            if idx_k_in_i > 0:
                duplicate.append(k)
                cross = idx_i_in_k > idx_j_in_k
                if sched is not None:
                    sched.dups.append((len(sched.steps) - 1, k, int(cross)))
                err += merge(nk, i, cross=cross)
        ln, e, phase = eat(ni, nj, idx_j_in_i, idx_i_in_j)
        err += e
        tn.psi *= phase
        tn.lnZ += ln
        for k in duplicate:
            merge(ni, k, cross=False)
            idx_k_in_i = find_neighbor(ni, k)
            if tn.cut_bond:
                err += cut_bondim_opt(tn, i, idx_k_in_i)
            elif tn.Dmax > 0 and ni.mps[idx_k_in_i - 1].shape[1] > tn.Dmax:
                err += cut_bondim_opt(tn, i, idx_k_in_i)
        clear(nj)
        if tn.compress:
            if tn.svdopt:
                compress_opt(ni)
            else:
                compress(ni)
        if tn.select in (0, 1):
            count_add_nodes(tn, [i] + list(ni.neighbor))
        for t in ni.mps:
            if t.shape[1] > tn.maxdim_intermediate:
                tn.maxdim_intermediate = t.shape[1]
    ln, sg = network_lognorm(tn)
    tn.sign = sg
    tn.lnZ += ln
    return tn.lnZ, err, tn.psi


# ---------------------------------------------------------------------------
# src/ising.jl
# ---------------------------------------------------------------------------

def ising_network(n, edges, weights, fields, beta, pos1=0, val1=0, pos2=0, val2=0,
                  Dmax=32, chi=32, cutoff=1e-15, norm_method=1, select=1, reverse=True,
                  svdopt=True, swapopt=True, compress=False, cut_bond=False, seed=1,
                  dtype=np.float64):
    """src/ising.jl:20-227. Site order of each spin = order of its edges in
    the edge list."""
    degree = [0] * (n + 1)
    for i, j in edges:
        degree[i] += 1
        degree[j] += 1
    nodes = {k: MPSNode([], [], 0, chi, cutoff, norm_method, svdopt, swapopt)
             for k in range(1, n + 1)}
    spin = np.array([1.0, -1.0])
    for e, (i, j) in enumerate(edges):
        si, sj = spin.copy(), spin.copy()
        if i == pos1:
            si[val1 - 1] = 0.0
        elif i == pos2:
            si[val2 - 1] = 0.0
        if j == pos1:
            sj[val1 - 1] = 0.0
        elif j == pos2:
            sj[val2 - 1] = 0.0
        Mij = np.outer(si, sj)
        w = float(weights[e])
        B = np.exp(beta * w * Mij)
        Q, R = B, np.eye(2)
        for node_id, other, s, mat in ((i, j, si, Q), (j, i, sj, R)):
            nd = nodes[node_id]
            h = float(fields[node_id - 1])
            fld = np.diag(np.exp(beta * h * s))
            if degree[node_id] == 1:
                sv = (fld @ mat).sum(axis=0)
                site = _r(sv, 1, 2, 1)
            elif not nd.neighbor:
                site = _r((fld @ mat).T, 1, 2, 2)
            elif len(nd.neighbor) == degree[node_id] - 1:
                site = _r(mat.copy(), 2, 2, 1)
            else:
                site = np.zeros((2, 2, 2))
                site[:, 0, :] = np.diag(mat[:, 0])
                site[:, 1, :] = np.diag(mat[:, 1])
            nd.mps.append(site)
            nd.neighbor.append(other)
    for k in range(1, n + 1):
        if degree[k] == 0:
            h = float(fields[k - 1])
            nodes[k].mps.append(_r(np.array([math.exp(beta * h) + math.exp(-beta * h)]), 1, 1, 1))
    for node in nodes.values():
        node.cano = len(node.mps)
        node.mps = [m.astype(dtype) for m in node.mps]
    num_iso = sum(1 for k in range(1, n + 1) if not nodes[k].neighbor)
    return TensorNetwork(nodes, Dmax, chi, cutoff, norm_method, select, reverse, svdopt,
                         swapopt, compress, cut_bond, {}, 0.0, 1.0, 1.0, -1, num_iso, n,
                         float(beta))


def free_energy(tn):
    """src/ising.jl:239-244."""
    lnZ, _, _ = contraction(tn)
    lz = lnZ / tn.n
    return lz, -lz / tn.beta


def magnetization(n, edges, w, h, beta, **kw):
    """src/ising.jl:263-290 (pinning)."""
    lnZ0, _, _ = contraction(ising_network(n, edges, w, h, beta, **kw))
    mag = np.zeros(n)
    for k in range(1, n + 1):
        lz = [contraction(ising_network(n, edges, w, h, beta, pos1=k, val1=v, **kw))[0]
              for v in (1, 2)]
        p = np.exp(np.array(lz) - lnZ0)
        mag[k - 1] = p[1] - p[0]
    return mag


def correlation(n, edges, w, h, beta, **kw):
    """src/ising.jl:309-339 (pinning)."""
    lnZ0, _, _ = contraction(ising_network(n, edges, w, h, beta, **kw))
    cor = np.zeros(len(edges))
    for e, (m, nn) in enumerate(edges):
        lz = np.zeros((2, 2))
        for v1 in (1, 2):
            for v2 in (1, 2):
                lz[v1 - 1, v2 - 1] = contraction(ising_network(
                    n, edges, w, h, beta, pos1=m, val1=v1, pos2=nn, val2=v2, **kw))[0]
        p = np.exp(lz - lnZ0)
        cor[e] = p[0, 0] - p[0, 1] - p[1, 0] + p[1, 1]
    return cor
