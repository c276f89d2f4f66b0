"""Double-layer (bra-ket) engine, src/braket.jl.

The oracle's restatement (oracle/braket_oracle.py) is pinned first, by the
checks of test/test_braket.jl: braket_node rebuilds E_i = sum_p T_i (x)
conj(T_i) (:19-92), a paired-edge eat! of two sites gives <psi|psi> (:96-121),
and contraction! of chains and a star in exact mode matches the dense
<psi|psi> (:246-305, with exact_norm of :8-16). Real and complex tensors.
The device path (real tensors here; complex ones in tests/test_complex.py) is
then compared with the pinned oracle and with the dense norm, on the CPU
emulator and, marked gpu, on the MI355X.
"""
import math

import numpy as np
import pytest

from oracle import braket_oracle as B
from oracle import catn_oracle as O


def _rand(rng, shape, cplx):
    a = rng.standard_normal(shape)
    if cplx:
        a = a + 1j * rng.standard_normal(shape)
    return a


def _E(T, phys):
    """Dense E = sum_p T (x) conj(T), axes (ket virtuals..., bra virtuals...)."""
    nd = T.ndim
    virt = [a for a in range(nd) if a != phys]
    letters = "abcdefgh"
    ket = "".join(letters[a] if a != phys else "p" for a in range(nd))
    bra = "".join(letters[a].upper() if a != phys else "p" for a in range(nd))
    out = "".join(letters[a] for a in virt) + "".join(letters[a].upper() for a in virt)
    return np.einsum(f"{ket},{bra}->{out}", T, T.conj())


# ----------------------------- oracle pins ----------------------------------

@pytest.mark.parametrize("shape,phys,cplx", [((3, 2, 4), 1, True), ((3, 2, 4), 1, False),
                                             ((3, 2), 1, True), ((2, 3, 2, 2), 1, False)])
def test_oracle_braket_node_rebuilds_double_tensor(shape, phys, cplx):
    rng = np.random.default_rng(len(shape) * 7 + phys)
    T = _rand(rng, shape, cplx)
    deg = T.ndim - 1
    node = B.braket_node(T, list(range(10, 10 + deg)), phys + 1, chi=10_000)
    assert node.layer == [True] * deg + [False] * deg
    assert node.neighbor == list(range(10, 10 + deg)) * 2
    assert node.cano == len(node.mps)
    raw = O.mps2raw(node)
    assert np.allclose(raw, _E(T, phys), rtol=1e-10, atol=1e-12)
    for t in node.mps[:-1]:                      # left-canonical
        M = t.reshape(-1, t.shape[2], order="F")
        assert np.allclose(M.conj().T @ M, np.eye(M.shape[1]), atol=1e-10)


@pytest.mark.parametrize("cplx", [True, False])
def test_oracle_paired_edge_eat_two_sites(cplx):
    rng = np.random.default_rng(3)
    T1, T2 = _rand(rng, (2, 4), cplx), _rand(rng, (4, 2), cplx)
    n1 = B.braket_node(T1, [2], 1, chi=10_000)
    n2 = B.braket_node(T2, [1], 2, chi=10_000)
    lg, _, ph = B.eat(n1, n2, 2, 1)
    ref = B.exact_norm([T1, T2], [["p1", "a"], ["a", "p2"]])
    assert abs(math.exp(lg) * ph - ref) <= 1e-10 * ref
    assert ref > 0


def _chain(rng, n, D, d, cplx):
    ts, ixs = [], []
    for k in range(n):
        if k == 0:
            ts.append(_rand(rng, (d, D), cplx)); ixs.append([f"p{k}", f"b{k}"])
        elif k == n - 1:
            ts.append(_rand(rng, (D, d), cplx)); ixs.append([f"b{k - 1}", f"p{k}"])
        else:
            ts.append(_rand(rng, (D, d, D), cplx)); ixs.append([f"b{k - 1}", f"p{k}", f"b{k}"])
    return ts, ixs


def _star(rng, cplx):
    ts = [_rand(rng, (2, 2, 2, 2), cplx), _rand(rng, (2, 2), cplx), _rand(rng, (2, 2), cplx),
          _rand(rng, (2, 2), cplx)]
    ixs = [["a", "b", "c", "pc"], ["a", "p1"], ["b", "p2"], ["c", "p3"]]
    return ts, ixs


CASES = ["chain3_real", "chain4_cplx", "chain5_cplx", "star_real", "star_cplx", "chain6_D3"]


def _case(name):
    rng = np.random.default_rng(101 + CASES.index(name))
    if name == "chain3_real":
        return _chain(rng, 3, 2, 2, False)
    if name == "chain4_cplx":
        return _chain(rng, 4, 3, 2, True)
    if name == "chain5_cplx":
        return _chain(rng, 5, 3, 2, True)
    if name == "star_real":
        return _star(rng, False)
    if name == "star_cplx":
        return _star(rng, True)
    return _chain(rng, 6, 3, 2, False)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("svdopt", [True, False])
def test_oracle_norm_matches_exact(name, svdopt):
    ts, ixs = _case(name)
    ref = B.exact_norm(ts, ixs)
    bk = B.braket_network(ts, ixs, chi=10_000, svdopt=svdopt)
    lnz, err, psi = B.contraction(bk)
    assert abs(math.exp(lnz) * psi - ref) <= 1e-10 * ref
    assert err < 1e-12


def test_oracle_rejects_cycles_and_bad_labels():
    rng = np.random.default_rng(0)
    ts = [rng.standard_normal((2, 2, 2)) for _ in range(3)]
    with pytest.raises(ValueError):
        B.braket_network(ts, [["a", "b", "p1"], ["b", "c", "p2"], ["c", "a", "p3"]])
    with pytest.raises(ValueError):
        B.braket_network([rng.standard_normal((2, 2))] * 2, [["a", "b"], ["a", "b"]])


def test_oracle_plants_a_mistake():
    """A dropped conj on the bra chain (the bra built from T, not conj(T))
    must fail the pins on complex tensors."""
    rng = np.random.default_rng(5)
    T = _rand(rng, (3, 2, 4), True)
    node = B.braket_node(T.conj(), [1, 2], 2, chi=10_000)   # ket of conj(T): bra of T
    assert not np.allclose(O.mps2raw(node), _E(T, 1), atol=1e-6)


# ----------------------------- device path ----------------------------------

@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx")
    return catn, request.getfixturevalue("gpu_ctx")


def _tree(rng, n, D, d):
    """Random tree: node k > 1 hangs off a random earlier node."""
    nb = {k: [] for k in range(1, n + 1)}
    labels = {k: [] for k in range(1, n + 1)}
    for k in range(2, n + 1):
        par = int(rng.integers(1, k))
        lab = f"e{par}_{k}"
        labels[par].append(lab)
        labels[k].append(lab)
    ts, ixs = [], []
    for k in range(1, n + 1):
        ix = labels[k][:]
        ix.insert(int(rng.integers(0, len(ix) + 1)), f"p{k}")
        ts.append(rng.standard_normal([d if l.startswith("p") else D for l in ix]))
        ixs.append(ix)
    return ts, ixs


@pytest.mark.parametrize("name", ["chain3_real", "star_real", "chain6_D3"])
@pytest.mark.parametrize("svdopt", [True, False])
def test_device_norm_matches_exact(env, name, svdopt):
    catn, ctx = env
    ts, ixs = _case(name)
    ref = B.exact_norm(ts, ixs)
    tn = catn.braket_network(ts, ixs, ctx=ctx, chi=10_000, svdopt=svdopt)
    lnz, err, psi = catn.contraction(tn)
    assert abs(math.exp(lnz) * psi - ref) <= 1e-10 * ref
    assert err < 1e-12


@pytest.mark.parametrize("n,D,chi,seed", [(8, 3, 10_000, 1), (10, 3, 6, 2), (12, 4, 8, 3),
                                          (9, 2, 3, 4)])
def test_device_matches_oracle_on_random_trees(env, n, D, chi, seed):
    """Exact and truncated (chi below the exact bond) random trees: the same
    edge sequence, maxdim_intermediate, lnZ (1e-12 relative) and discarded
    weight (1e-8 relative) as the oracle."""
    catn, ctx = env
    ts, ixs = _tree(np.random.default_rng(seed), n, D, 2)
    bk = B.braket_network(ts, ixs, chi=chi)
    lo, eo, po = B.contraction(bk)
    tn = catn.braket_network(ts, ixs, ctx=ctx, chi=chi)
    lnz, err, psi = catn.contraction(tn)
    assert [tuple(e) for e in tn.edge_log()] == bk.edge_log
    assert tn.maxdim_intermediate == bk.maxdim_intermediate
    assert abs(lnz - lo) <= 1e-12 * max(1.0, abs(lo))
    assert abs(err - eo) <= 1e-8 * max(1e-12, eo) + 1e-14
    assert psi == po == 1.0
    if chi >= 10_000:
        assert abs(lnz - math.log(B.exact_norm(ts, ixs))) < 1e-10


def test_device_rejects_cycles_and_double_bonds(env):
    """Cycles without loops=1 (src/braket.jl:405-420) and two bonds between one
    pair of tensors are refused. (Complex tensors are accepted: tests/test_complex.py.)"""
    catn, ctx = env
    rng = np.random.default_rng(0)
    ts = [rng.standard_normal((2, 2, 2)) for _ in range(3)]
    with pytest.raises(catn.CatnError):
        catn.braket_network(ts, [["a", "b", "p1"], ["b", "c", "p2"], ["c", "a", "p3"]], ctx=ctx)
    with pytest.raises(catn.CatnError):
        catn.braket_network([rng.standard_normal((2, 2)), rng.standard_normal((2, 2))],
                            [["a", "b"], ["a", "b"]], ctx=ctx)


# ------------------------- loopy double layer --------------------------------
# No reference counterpart (src/braket.jl:405-420 rejects cycles): with
# loops=True an eat! that closes a cycle leaves doubled bonds between the eater
# and a shared neighbour, merged ket and bra separately as merge! does
# (src/mps_node.jl:595-622). The oracle's extension is pinned in exact mode by
# the dense <psi|psi> on lattices and on a triangle whose third tensor lists
# the bond to the eaten node first (the merge with cross = True there).

def _lattice(rng, Lx, Ly, D, d, cplx=False):
    ts, ixs = [], []
    for x in range(Lx):
        for y in range(Ly):
            ix = []
            if x > 0:
                ix.append(f"v{x - 1}_{y}")
            if y > 0:
                ix.append(f"h{x}_{y - 1}")
            ix.append(f"p{x}_{y}")
            if y < Ly - 1:
                ix.append(f"h{x}_{y}")
            if x < Lx - 1:
                ix.append(f"v{x}_{y}")
            ts.append(_rand(rng, [d if l.startswith("p") else D for l in ix], cplx))
            ixs.append(ix)
    return ts, ixs


def _triangle(rng, D):
    ts = [rng.standard_normal((D, 2, D)), rng.standard_normal((D, D, 2)),
          rng.standard_normal((D, D, 2))]
    return ts, [["a", "p1", "b"], ["a", "c", "p2"], ["c", "b", "p3"]]


LOOPY = ["triangle_D3", "square_D2", "square_D3", "lattice2x3_D2", "lattice3x3_D2", "lattice3x3_D3"]


def _loopy(name, cplx=False):
    rng = np.random.default_rng(200 + LOOPY.index(name))
    if name == "triangle_D3":
        return _triangle(rng, 3)
    Lx, Ly, D = {"square_D2": (2, 2, 2), "square_D3": (2, 2, 3), "lattice2x3_D2": (2, 3, 2),
                 "lattice3x3_D2": (3, 3, 2), "lattice3x3_D3": (3, 3, 3)}[name]
    return _lattice(rng, Lx, Ly, D, 2, cplx)


@pytest.mark.parametrize("name", LOOPY)
@pytest.mark.parametrize("cplx", [False, True])
def test_oracle_loopy_norm_matches_exact(name, cplx):
    ts, ixs = _loopy(name, cplx and name != "triangle_D3")
    ref = B.exact_norm(ts, ixs)
    bk = B.braket_network(ts, ixs, chi=10_000, loops=True)
    lnz, err, psi = B.contraction(bk)
    assert abs(math.exp(lnz) * psi - ref) <= 1e-10 * ref
    assert err < 1e-12


def test_oracle_loopy_plants_a_mistake(monkeypatch):
    """Fusing the doubled legs in opposite orders on the two ends of the bond
    (the cross flag ignored on the shared neighbour) must break the triangle."""
    ts, ixs = _loopy("triangle_D3")
    ref = B.exact_norm(ts, ixs)
    orig, calls = B.merge_pair, [0]

    def one_end(node, target, isket, cross):
        calls[0] += 1
        return orig(node, target, isket, (not cross) if (calls[0] - 1) % 4 < 2 else cross)
    monkeypatch.setattr(B, "merge_pair", one_end)
    bk = B.braket_network(ts, ixs, chi=10_000, loops=True)
    lnz, _, psi = B.contraction(bk)
    assert abs(math.exp(lnz) * psi - ref) > 1e-3 * ref


@pytest.mark.parametrize("name", LOOPY)
@pytest.mark.parametrize("chi", [10_000, 8, 3])
def test_device_loopy_matches_oracle(env, name, chi):
    """Device against the oracle: edge sequence, maxdim_intermediate, lnZ
    (1e-12 relative), discarded weight; in exact mode also the dense norm."""
    catn, ctx = env
    ts, ixs = _loopy(name)
    bk = B.braket_network(ts, ixs, chi=chi, loops=True)
    lo, eo, po = B.contraction(bk)
    tn = catn.braket_network(ts, ixs, ctx=ctx, chi=chi, loops=True)
    lnz, err, psi = catn.contraction(tn)
    assert [tuple(e) for e in tn.edge_log()] == bk.edge_log
    assert tn.maxdim_intermediate == bk.maxdim_intermediate
    assert abs(lnz - lo) <= 1e-12 * max(1.0, abs(lo))
    assert abs(err - eo) <= 1e-8 * max(1e-12, eo) + 1e-14
    # a heavily truncated loopy double layer need not stay positive: the final
    # scalar eat's phase z/|z| is then -1, on the device as in the oracle
    assert psi == po and abs(psi) == 1.0
    if chi >= 10_000:
        assert psi == 1.0
        assert abs(lnz - math.log(B.exact_norm(ts, ixs))) < 1e-10


def test_loops_flag_and_double_bonds(env):
    """Without loops=True a cycle is rejected (as in the reference); two bonds
    between one pair of tensors are rejected either way."""
    catn, ctx = env
    ts, ixs = _loopy("square_D2")
    with pytest.raises(catn.CatnError):
        catn.braket_network(ts, ixs, ctx=ctx, chi=64)
    with pytest.raises(ValueError):
        B.braket_network(ts, ixs, chi=64)
    rng = np.random.default_rng(1)
    t2 = [rng.standard_normal((2, 2, 2)), rng.standard_normal((2, 2, 2))]
    ix2 = [["a", "b", "p1"], ["a", "b", "p2"]]
    with pytest.raises(catn.CatnError):
        catn.braket_network(t2, ix2, ctx=ctx, chi=64, loops=True)
    with pytest.raises(ValueError):
        B.braket_network(t2, ix2, chi=64, loops=True)


@pytest.mark.parametrize("name,Dmax,chi", [("lattice3x3_D2", 3, 10_000), ("lattice3x3_D2", 2, 8),
                                           ("lattice2x3_D2", 3, 10_000), ("lattice3x3_D3", 5, 10_000)])
def test_device_loopy_dmax_matches_oracle(env, name, Dmax, chi):
    """Dmax on the loopy double layer (no reference counterpart): a merged ket
    or bra bond wider than Dmax is cut as cut_bondim_opt! does. Device against
    the oracle's restatement: edge sequence, maxdim_intermediate, lnZ
    (1e-12 relative), discarded weight, phase."""
    catn, ctx = env
    ts, ixs = _loopy(name)
    bk = B.braket_network(ts, ixs, chi=chi, loops=True, Dmax=Dmax)
    lo, eo, po = B.contraction(bk)
    assert eo > 1e-10                      # the cut did truncate
    tn = catn.braket_network(ts, ixs, ctx=ctx, chi=chi, loops=True, Dmax=Dmax)
    lnz, err, psi = catn.contraction(tn)
    assert [tuple(e) for e in tn.edge_log()] == bk.edge_log
    assert tn.maxdim_intermediate == bk.maxdim_intermediate
    assert abs(lnz - lo) <= 1e-12 * max(1.0, abs(lo))
    assert abs(err - eo) <= 1e-8 * max(1e-12, eo) + 1e-14
    assert psi == po


def test_loopy_dmax_above_every_bond_changes_nothing(env):
    catn, ctx = env
    ts, ixs = _loopy("lattice3x3_D2")
    a = catn.contraction(catn.braket_network(ts, ixs, ctx=ctx, chi=10_000, loops=True))
    b = catn.contraction(catn.braket_network(ts, ixs, ctx=ctx, chi=10_000, loops=True, Dmax=64))
    assert a == b
    assert abs(a[0] - math.log(B.exact_norm(ts, ixs))) < 1e-10
