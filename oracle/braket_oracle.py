"""FP64 NumPy restatement of the reference's double-layer (bra-ket) engine,
src/braket.jl, for the parity tests of the device BraKet path.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker, never by the
product path.

It reuses the MPS primitives of oracle/catn_oracle.py (tsvd, cano_to, swap,
compress_opt, compress) where src/braket.jl restates them for BraKetNode; a
BraKetNode here is an MPSNode plus a `layer` list that the moves permute in
lockstep with `neighbor`.

Fixed choice where the reference depends on Julia Dict order: the paired
edges of braket_network are taken in the order in which their labels first
appear when scanning ixs[0], ixs[1], ... (src/braket.jl:457-468 iterates a
Dict{Any,Int}).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np

from . import catn_oracle as O

_r = O._r


@dataclass
class BraKetNode(O.MPSNode):
    """src/braket.jl:20-34: MPS [ket v1..v_deg, bra v1..v_deg] of
    E_i = sum_p T_i (x) conj(T_i); layer True = ket leg."""
    layer: List[bool] = field(default_factory=list)


def braket_node(Ti, ket_neighbors, phys_pos, chi=32, cutoff=1e-15, norm_method=1,
                svdopt=True, swapopt=False):
    """src/braket.jl:75-185. phys_pos is 1-based."""
    Ti = np.asarray(Ti)
    nd = Ti.ndim
    deg = nd - 1
    if deg < 1:
        raise ValueError("braket_node: degree-0 site (no virtual legs) is not supported.")
    if len(ket_neighbors) != deg:
        raise ValueError("braket_node: wrong number of ket neighbours")
    virt = [a for a in range(nd) if a != phys_pos - 1]
    Tk = np.transpose(Ti, virt + [phys_pos - 1])          # (D1, ..., D_deg, d)
    dims = Tk.shape
    d = dims[deg]
    # ket chain: left-to-right tsvd over D1 .. D_{deg-1}; the last ket site keeps d
    # as its right bond (src/braket.jl:116-130)
    ket = []
    R = _r(Tk, 1, *dims)
    dleft = 1
    for i in range(deg - 1):
        M = _r(R, dleft * dims[i], -1)
        U, S, V, _ = O.tsvd(M, cutoff=cutoff, maxdim=chi, kind="braket_ket")
        x = len(S)
        ket.append(_r(U, dleft, dims[i], x))
        R = _r(np.diag(S) @ V.conj().T, x, *dims[i + 1:])
        dleft = x
    ket.append(_r(R, dleft, dims[deg - 1], d))
    # bra chain from conj(Tk) permuted to (d, D1, ..., D_deg): d is the left bond
    # of the first bra site (src/braket.jl:155-172)
    Tb = np.transpose(Tk.conj(), [deg] + list(range(deg)))
    db = Tb.shape
    bra = []
    R = Tb
    dleft = d
    for i in range(deg - 1):
        M = _r(R, dleft * db[i + 1], -1)
        U, S, V, _ = O.tsvd(M, cutoff=cutoff, maxdim=chi, kind="braket_bra")
        x = len(S)
        bra.append(_r(U, dleft, db[i + 1], x))
        R = _r(np.diag(S) @ V.conj().T, x, *db[i + 2:])
        dleft = x
    bra.append(_r(R, dleft, db[deg], 1))
    node = BraKetNode(ket + bra, list(ket_neighbors) + list(ket_neighbors), 1, chi, cutoff,
                      norm_method, svdopt, swapopt, [True] * deg + [False] * deg)
    O.left_canonical(node)                                 # src/braket.jl:183
    return node


def find_leg(node, j, isket):
    """src/braket.jl:287-294; 1-based, 0 if absent."""
    for k in range(len(node.mps)):
        if node.neighbor[k] == j and node.layer[k] == isket:
            return k + 1
    return 0


def swap(node, i, j):
    """src/braket.jl:510-564: MPSNode swap! plus the layer in lockstep."""
    err = O.swap(node, i, j)
    node.layer[i - 1], node.layer[j - 1] = node.layer[j - 1], node.layer[i - 1]
    return err


def move(node, a, b):
    """src/braket.jl:570-584."""
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


def eat(node_i, node_j, j_id, i_id):
    """src/braket.jl:638-798. Returns (lognorm, error, phase)."""
    err = 0.0
    n_i, n_j = len(node_i.mps), len(node_j.mps)
    bi = find_leg(node_i, j_id, False)
    if bi <= 0:
        raise ValueError("eat!: bra leg not found in node_i")
    err += move(node_i, bi, n_i)
    ki = find_leg(node_i, j_id, True)
    if ki <= 0:
        raise ValueError("eat!: ket leg not found in node_i")
    err += move(node_i, ki, n_i - 1)
    O.cano_to(node_i, n_i)
    kj = find_leg(node_j, i_id, True)
    if kj <= 0:
        raise ValueError("eat!: ket leg not found in node_j")
    err += move(node_j, kj, 1)
    bj = find_leg(node_j, i_id, False)
    if bj <= 0:
        raise ValueError("eat!: bra leg not found in node_j")
    err += move(node_j, bj, 2)
    A_ik, A_ib = node_i.mps[n_i - 2], node_i.mps[n_i - 1]
    A_jk, A_jb = node_j.mps[0], node_j.mps[1]
    dl, Dk, Db, dr = A_ik.shape[0], A_ik.shape[1], A_ib.shape[1], A_jb.shape[2]
    W_i = _r(np.einsum("ijk,kab->ijab", A_ik, A_ib), dl, Dk, Db)
    W_j = _r(np.einsum("ijk,kab->ijab", A_jk, A_jb), Dk, Db, dr)
    mat = np.einsum("iab,abj->ij", W_i, W_j)
    one = mat.dtype.type(1)
    if n_i == 2 and n_j == 2:                                        # case A
        z = np.sum(mat)
        absz = abs(z)
        node_i.mps.clear(); node_i.neighbor.clear(); node_i.layer.clear()
        if absz <= node_i.cutoff:
            return 0.0, err, one
        return math.log(absz), err, z / absz
    if n_i == 2:                                                     # case B
        raise ValueError("eat!: case B (order(i) < order(j)) is not reached by contraction!")
    new = np.einsum("ijk,ka->ija", node_i.mps[n_i - 3], mat)
    node_i.mps[n_i - 3] = new
    node_i.cano = n_i - 2
    del node_i.mps[-2:], node_i.neighbor[-2:], node_i.layer[-2:]
    if n_j > 2:                                                      # case D
        node_i.mps.extend(node_j.mps[2:])
        node_i.neighbor.extend(node_j.neighbor[2:])
        node_i.layer.extend(node_j.layer[2:])
    O.cano_to(node_i, len(node_i.mps))                               # cases C and D
    center = node_i.mps[node_i.cano - 1]
    nrm = O._norm(center, node_i.norm_method)
    if nrm <= node_i.cutoff:
        return 0.0, err, one
    node_i.mps[-1] = node_i.mps[-1] / nrm
    return math.log(nrm), err, one


def merge_pair(node, target, isket, cross):
    """Loopy extension, no reference counterpart (src/braket.jl:405-420 rejects
    cycles): merge! (src/mps_node.jl:595-622) of the two legs of one layer that
    point to `target`, the layer list following the moves; cross moves the
    second leg in front of the first, so that the bond older than the eat is
    the fast index of the fused leg on both of its ends."""
    pos = [k + 1 for k in range(len(node.mps)) if node.neighbor[k] == target and node.layer[k] == isket]
    if len(pos) != 2:
        raise ValueError("merge: expected two legs of one layer to the shared neighbour")
    idx1, idx2 = pos
    err = move(node, idx2, idx1 if cross else idx1 + 1)
    del node.neighbor[idx1], node.layer[idx1]        # equal entries: keep one
    O.cano_to(node, idx1)
    a, b = node.mps[idx1 - 1], node.mps[idx1]
    node.mps[idx1 - 1] = _r(np.einsum("ijk,kab->ijab", a, b), a.shape[0], -1, b.shape[2])
    del node.mps[idx1]
    O.cano_to(node, idx1)
    return err


@dataclass
class BraKetNetwork:
    tensors: Dict[int, BraKetNode]
    edges: List[tuple]
    chi: int
    cutoff: float
    norm_method: int
    svdopt: bool
    swapopt: bool
    maxdim_intermediate: int = 0
    edge_log: List[tuple] = field(default_factory=list)   # (i, j) eaten, in order
    loops: bool = False
    Dmax: int = -1     # loops only: cut of merged bonds wider than Dmax


def braket_network(tensors, ixs, chi=64, cutoff=1e-15, norm_method=1, svdopt=True,
                   swapopt=True, loops=False, Dmax=-1):
    """src/braket.jl:369-496 (edge order: first appearance of the label).
    loops=True accepts cycles (no reference counterpart); two bonds between
    one pair of tensors are rejected either way."""
    n = len(tensors)
    count, owners, order_ = {}, {}, []
    for i in range(n):
        for l in ixs[i]:
            if l not in count:
                order_.append(l)
            count[l] = count.get(l, 0) + 1
            owners.setdefault(l, []).append(i + 1)
    for l, c in count.items():
        if c not in (1, 2):
            raise ValueError(f"braket_network: label {l} appears {c} times (expected 1 or 2)")
    for i in range(n):
        if sum(1 for l in ixs[i] if count[l] == 1) != 1:
            raise ValueError(f"braket_network: tensor {i + 1} needs exactly one physical label")
    parent = list(range(n + 1))

    def root(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x
    pairs = set()
    for l in order_:
        if count[l] == 2:
            a, b = owners[l]
            if (min(a, b), max(a, b)) in pairs:
                raise ValueError("braket_network: two bonds between one pair of tensors")
            pairs.add((min(a, b), max(a, b)))
            ra, rb = root(a), root(b)
            if ra == rb and not loops:
                raise ValueError("braket_network: virtual-bond graph contains a cycle")
            parent[ra] = rb
    nodes = {}
    for i in range(n):
        kn, pp = [], 0
        for ax, l in enumerate(ixs[i]):
            if count[l] == 1:
                pp = ax + 1
            else:
                a, b = owners[l]
                kn.append(b if a == i + 1 else a)
        nodes[i + 1] = braket_node(tensors[i], kn, pp, chi, cutoff, norm_method, svdopt, swapopt)
    edges = [tuple(owners[l]) for l in order_ if count[l] == 2]
    return BraKetNetwork(nodes, edges, chi, cutoff, norm_method, svdopt, swapopt, loops=loops,
                         Dmax=Dmax)


def braket_lognorm(bk):
    """src/braket.jl:915-950 (ascending node id)."""
    ln = 0.0
    for k in sorted(bk.tensors):
        node = bk.tensors[k]
        if not node.mps:
            continue
        O.cano_to(node, len(node.mps))
        nm = O._norm(node.mps[node.cano - 1], node.norm_method)
        if nm <= node.cutoff:
            continue
        ln += math.log(nm)
    return ln, 1.0


def contraction(bk):
    """src/braket.jl:960-1078. Returns (lnZ, error, psi)."""
    err, psi, lnz = 0.0, 1.0, 0.0
    edges = list(bk.edges)
    while edges:
        i, j = edges[0]
        ni, nj = bk.tensors[i], bk.tensors[j]
        if len(nj.mps) > len(ni.mps):
            i, j, ni, nj = j, i, nj, ni
        bk.edge_log.append((i, j))
        lg, e, ph = eat(ni, nj, j, i)
        err += e
        psi *= ph
        lnz += lg
        err += O.compress_opt(ni) if bk.svdopt else O.compress(ni)
        for t in ni.mps:
            bk.maxdim_intermediate = max(bk.maxdim_intermediate, t.shape[1])
        edges = [x for x in edges if not (set(x) == {i, j})]
        # re-point j's edges (and the neighbour entries of j's other
        # neighbours) to i (src/braket.jl:1020-1071); with loops, a neighbour
        # already bonded to i closes a cycle: its bond is listed once and the
        # doubled legs are merged on both ends below
        new, dups = [], []
        for (a, b) in edges:
            if j in (a, b):
                other = b if a == j else a
                on = bk.tensors[other]
                if find_leg(on, i, True) > 0:
                    if not bk.loops:
                        raise ValueError("braket contraction: a cycle closes")
                    dups.append((other, find_leg(on, i, True) > find_leg(on, j, True),
                                 find_leg(on, i, False) > find_leg(on, j, False)))
                on.neighbor = [i if v == j else v for v in on.neighbor]
                if not any({i, other} == set(r) for r in edges):
                    new.append((i, other) if a == j else (other, i))
            else:
                new.append((a, b))
        edges = new
        del bk.tensors[j]
        for (k, cross_ket, cross_bra) in dups:
            nk = bk.tensors[k]
            err += merge_pair(nk, i, True, cross_ket)
            err += merge_pair(nk, i, False, cross_bra)
            err += merge_pair(ni, k, True, False)
            err += merge_pair(ni, k, False, False)
            for t in nk.mps + ni.mps:
                bk.maxdim_intermediate = max(bk.maxdim_intermediate, t.shape[1])
        # Dmax (no reference counterpart): the merged ket bond, then the bra
        # bond, cut as cut_bondim_opt! (src/tensor_network.jl:399-477) when
        # wider than Dmax
        if bk.Dmax > 0:
            for (k, _, _) in dups:
                for isket in (True, False):
                    nk = bk.tensors[k]
                    ai, ak = find_leg(ni, k, isket), find_leg(nk, i, isket)
                    if ai > 0 and ak > 0 and ni.mps[ai - 1].shape[1] > bk.Dmax:
                        err += O.cut_legs(ni, ai, nk, ak, True, bk.cutoff, bk.Dmax)
    ln, sg = braket_lognorm(bk)
    return lnz + ln, err, psi * sg


def exact_norm(tensors, ixs):
    """test/test_braket.jl:8-16: <psi|psi> of the dense contraction."""
    count = {}
    for ix in ixs:
        for l in ix:
            count[l] = count.get(l, 0) + 1
    letters, sub = {}, []
    for ix in ixs:
        s = ""
        for l in ix:
            if l not in letters:
                letters[l] = chr(ord("a") + len(letters)) if len(letters) < 26 else chr(ord("A") + len(letters) - 26)
            s += letters[l]
        sub.append(s)
    out = "".join(letters[l] for l in letters if count[l] == 1)
    psi = np.einsum(",".join(sub) + "->" + out, *tensors)
    return float(np.real(np.sum(psi * psi.conj())))
