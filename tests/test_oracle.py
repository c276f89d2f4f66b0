"""Pins of the FP64 oracle (oracle/catn_oracle.py) against closed forms, dense
identities and brute force, restating the reference's own tests:
test/test_linalg.jl, test/test_mps_node.jl, test/test_contraction.jl,
test/test_ising.jl, examples/double_layer_ising.jl. The reference ships no
stored golden vectors (it uses unseeded randn); inputs here are seeded."""
import math

import numpy as np
import pytest

from oracle import catn_oracle as O
from oracle import configs as Cf
from oracle import exact as X


def rng(seed=0):
    return np.random.default_rng(seed)


# ---------------- src/linalg_utils.jl (test/test_linalg.jl:4-35) ----------------

def test_tsvd_reconstruction_and_truncation():
    A = rng(1).standard_normal((8, 5))
    U, S, V, d = O.tsvd(A)
    assert np.allclose(U @ np.diag(S) @ V.T, A)
    assert np.all(np.diff(S) <= 0)
    U2, S2, V2, d2 = O.tsvd(A, maxdim=2)
    assert len(S2) == 2 and np.allclose(S2, S[:2])
    assert np.isclose(d2, S[2:].sum())
    B = U[:, :3] @ np.diag([1.0, 1e-3, 1e-20]) @ V[:, :3].T
    _, S3, _, _ = O.tsvd(B, cutoff=1e-10)
    assert len(S3) == 2
    # at least one value is kept even when none exceeds the cutoff
    _, S4, _, _ = O.tsvd(np.zeros((3, 3)))
    assert len(S4) == 1


def test_tsvd_matches_lapack_spectrum():
    A = rng(2).standard_normal((30, 12))
    _, S, _, _ = O.tsvd(A, cutoff=0.0)
    assert np.allclose(S, np.linalg.svd(A, compute_uv=False), rtol=1e-14)


def test_tsvd_fallback_is_identical_in_value():
    A = rng(3).standard_normal((9, 6))
    U, S, V, _ = O.tsvd(A)
    O.FORCE_GESDD_FAILURE = True
    try:
        U2, S2, V2, _ = O.tsvd(A)
    finally:
        O.FORCE_GESDD_FAILURE = False
    assert np.allclose(S, S2, rtol=1e-13)
    assert np.allclose(U @ np.diag(S) @ V.T, U2 @ np.diag(S2) @ V2.T)


def test_rsvd_low_rank():
    r = rng(4)
    L = r.standard_normal((50, 4)) @ r.standard_normal((4, 40))
    U, S, V = O.rsvd(L, 4, 10, 10, rng=np.random.default_rng(0))
    assert np.allclose(U @ np.diag(S) @ V.T, L, atol=1e-8)
    U2, S2, V2 = O.rsvd(L, 4, 10, 10, rng=np.random.default_rng(0))
    assert np.allclose(S, S2)


# ---------------- src/mps_node.jl (test/test_mps_node.jl) ----------------

@pytest.mark.parametrize("dims", [(4,), (3, 5), (2, 3, 4), (2, 3, 2, 3)])
def test_raw2mps_roundtrip(dims):
    T = rng(5).standard_normal(dims)
    node = O.MPSNode.from_tensor(T, list(range(1, len(dims) + 1)), chi=1000)
    assert O.order(node) == len(dims)
    assert O.shape(node) == list(dims)
    assert np.allclose(O.mps2raw(node), T)


def test_raw2mps_low_rank_chi():
    r = rng(6)
    M = r.standard_normal((6, 2)) @ r.standard_normal((2, 6))
    node = O.MPSNode.from_tensor(M, [1, 2], chi=2)
    assert np.allclose(O.mps2raw(node), M, atol=1e-10)


def test_cano_to_preserves_and_isometry():
    T = rng(7).standard_normal((2, 3, 4, 2))
    node = O.MPSNode.from_tensor(T, [1, 2, 3, 4], chi=1000)
    ref = O.mps2raw(node)
    for idx in [1, 2, 3, 4, 0]:
        O.cano_to(node, idx)
        assert np.allclose(O.mps2raw(node), ref)
    O.left_canonical(node)
    assert node.cano == 4
    for i in range(3):
        A = node.mps[i]
        M = A.reshape(A.shape[0] * A.shape[1], A.shape[2], order="F")
        assert np.allclose(M.T @ M, np.eye(M.shape[1]), atol=1e-8)


def test_cano_to_tiny_norm_rank():
    T = rng(8).standard_normal((2, 3, 4, 2)) * 1e-16
    node = O.MPSNode.from_tensor(T, [1, 2, 3, 4], chi=1000, cutoff=0.0)
    node.cutoff = 1e-15
    ref = O.mps2raw(node)
    O.cano_to(node, 1)
    assert np.allclose(O.mps2raw(node), ref, rtol=1e-8)
    O.cano_to(node, 4)
    assert np.allclose(O.mps2raw(node), ref, rtol=1e-8)


def test_swap_reverse_move2tail():
    T = rng(9).standard_normal((2, 3, 4, 5))
    node = O.MPSNode.from_tensor(T, [10, 20, 30, 40], chi=1000)
    O.swap(node, 2, 3)
    assert node.neighbor == [10, 30, 20, 40]
    assert np.allclose(O.mps2raw(node), np.transpose(T, (0, 2, 1, 3)))
    n2 = O.MPSNode.from_tensor(T, [10, 20, 30, 40], chi=1000)
    O.reverse(n2)
    assert n2.neighbor == [40, 30, 20, 10]
    assert np.allclose(O.mps2raw(n2), np.transpose(T, (3, 2, 1, 0)))
    n3 = O.MPSNode.from_tensor(T, [10, 20, 30, 40], chi=1000)
    O.move2tail(n3, 1)
    assert n3.neighbor[-1] == 10
    assert np.allclose(O.mps2raw(n3), np.transpose(T, (1, 2, 3, 0)))
    n4 = O.MPSNode.from_tensor(T, [10, 20, 30, 40], chi=1000)
    O.swap(n4, 3, 2)
    assert n4.neighbor == [10, 30, 20, 40]
    assert np.allclose(O.mps2raw(n4), np.transpose(T, (0, 2, 1, 3)))


def test_swap_reports_truncation():
    T = rng(10).standard_normal((4, 4, 4))
    node = O.MPSNode.from_tensor(T, [1, 2, 3], chi=2)
    assert O.swap(node, 1, 2) > 0


def test_compress_preserves():
    T = rng(11).standard_normal((2, 3, 4, 3, 2))
    for f in (O.compress, O.compress_opt):
        node = O.MPSNode.from_tensor(T, [1, 2, 3, 4, 5], chi=1000)
        ref = O.mps2raw(node)
        e = f(node)
        assert node.cano == 1
        assert np.allclose(O.mps2raw(node), ref)
        assert e >= 0.0


def test_merge_fuses_duplicate_legs():
    T = rng(12).standard_normal((2, 5, 3))
    node = O.MPSNode.from_tensor(T, [7, 8, 7], chi=1000)
    O.merge(node, 7)
    assert node.neighbor.count(7) == 1
    raw = O.mps2raw(node)
    expected = np.transpose(T, (0, 2, 1)).reshape(6, 5, order="F")
    assert np.allclose(raw.reshape(raw.shape[0], raw.shape[1]), expected)


def test_eat_cases():
    r = rng(13)
    Ti, Tj = r.standard_normal((2, 3, 4)), r.standard_normal((4, 5))
    ni = O.MPSNode.from_tensor(Ti, [1, 2, 3], chi=1000, norm_method=0)
    nj = O.MPSNode.from_tensor(Tj, [3, 4], chi=1000, norm_method=0)
    ln, err, ph = O.eat(ni, nj, 3, 1)
    raw = O.mps2raw(ni) * math.exp(ln)
    assert np.allclose(raw, np.einsum("abc,cd->abd", Ti, Tj), atol=1e-8)
    u, v = r.standard_normal(3), r.standard_normal(3)
    a = O.MPSNode.from_tensor(u, [2], norm_method=0)
    b = O.MPSNode.from_tensor(v, [1], norm_method=0)
    ln, err, ph = O.eat(a, b, 1, 1)
    assert not a.mps and np.isclose(math.exp(ln) * ph, u @ v)
    Ti, Tj = r.standard_normal((2, 4)), r.standard_normal(4)
    ni = O.MPSNode.from_tensor(Ti, [5, 9], chi=1000, norm_method=0)
    nj = O.MPSNode.from_tensor(Tj, [9], chi=1000, norm_method=0)
    ln, err, ph = O.eat(ni, nj, 2, 1)
    assert ni.neighbor == [5]
    assert np.allclose(O.mps2raw(ni).ravel() * math.exp(ln) * ph, Ti @ Tj, atol=1e-8)


def test_eat_case_c_left_canonical():
    r = rng(14)
    ni = O.MPSNode.from_tensor(r.standard_normal((2, 3, 4)), [1, 2, 3], chi=1000)
    nj = O.MPSNode.from_tensor(r.standard_normal((4, 5, 6)), [3, 4, 5], chi=1000)
    O.eat(ni, nj, 3, 1)
    assert ni.cano == len(ni.mps)
    for A in ni.mps[:-1]:
        M = A.reshape(A.shape[0] * A.shape[1], A.shape[2], order="F")
        assert np.allclose(M.T @ M, np.eye(M.shape[1]), atol=1e-8)


# ---------------- src/tensor_network.jl (test/test_contraction.jl) ----------------

def test_constructor_neighbors():
    r = rng(15)
    tn = O.tensor_network([r.standard_normal((2, 3)), r.standard_normal((3, 4)),
                           r.standard_normal((4, 2))], [["a", "b"], ["b", "c"], ["c", "a"]], chi=1000)
    assert tn.tensors[1].neighbor == [3, 2]
    assert tn.num_isolated == 0


def test_selection_avoids_fat_bond_and_sequential():
    r = rng(16)
    tn = O.tensor_network([r.standard_normal((2, 2)), r.standard_normal((2, 2)),
                           r.standard_normal((2, 8)), r.standard_normal((8, 2))],
                          [["a", "x"], ["x", "y"], ["y", "z"], ["z", "a"]], chi=1000)
    O.select_edge_init(tn)
    assert set(O.select_edge_min_dim(tn)) != {3, 4}
    tn = O.tensor_network([r.standard_normal((2, 2))] * 3, [["a", "b"], ["b", "c"], ["c", "a"]], chi=1000)
    assert O.select_edge_sequentially(tn) == (1, 2)


@pytest.mark.parametrize("cutter", ["cut_bondim", "cut_bondim_opt"])
def test_cut_lossless_on_low_rank_bond(cutter):
    r = rng(17)
    P = r.standard_normal((3, 2)) @ r.standard_normal((2, 4))
    Q = r.standard_normal((4, 2)) @ r.standard_normal((2, 3))
    tn = O.tensor_network([P, Q], [["a", "m"], ["m", "b"]], chi=1000, Dmax=2)
    getattr(O, cutter)(tn, 1, 2)
    assert tn.tensors[1].mps[1].shape[1] == 2
    got = O.mps2raw(tn.tensors[1]) @ O.mps2raw(tn.tensors[2])
    assert np.allclose(got, P @ Q, atol=1e-8)


def _networks(r):
    return [
        ([r.standard_normal((3, 4)), r.standard_normal((4, 5)), r.standard_normal((5, 3))],
         [["a", "b"], ["b", "c"], ["c", "a"]]),
        ([r.standard_normal((2, 3, 4)), r.standard_normal(2), r.standard_normal(3), r.standard_normal(4)],
         [["a", "b", "c"], ["a"], ["b"], ["c"]]),
        ([r.standard_normal((2, 3)), r.standard_normal((3, 2)), r.standard_normal((2, 3)), r.standard_normal((3, 2))],
         [["a", "b"], ["b", "c"], ["c", "d"], ["d", "a"]]),
    ]


def test_exact_mode_matches_einsum():
    for ts, ixs in _networks(rng(18)):
        ref = float(X.exact_contract(ts, ixs))
        for sel in (0, 1, 2):
            for rev in (False, True):
                for comp in (False, True):
                    tn = O.tensor_network(ts, ixs, Dmax=-1, chi=10000, select=sel, reverse=rev,
                                          compress=comp)
                    lnZ, _, psi = O.contraction(tn)
                    assert np.isclose(math.exp(lnZ) * psi, ref, rtol=1e-8)


def test_finite_dmax_loop():
    r = rng(19)
    ts = [r.standard_normal((4, 4)) for _ in range(4)]
    ixs = [["a", "b"], ["b", "c"], ["c", "d"], ["d", "a"]]
    ref = float(X.exact_contract(ts, ixs))
    tn = O.tensor_network(ts, ixs, Dmax=4, chi=64, select=1, compress=True)
    lnZ, _, psi = O.contraction(tn)
    assert np.isclose(math.exp(lnZ) * psi, ref, rtol=1e-6)


# ---------------- src/ising.jl (test/test_ising.jl) ----------------

def test_ising_chain_transfer_matrix():
    L, J, beta = 6, 0.7, 0.4
    edges = [(i, i + 1) for i in range(1, L)]
    tn = O.ising_network(L, edges, [J] * (L - 1), np.zeros(L), beta, Dmax=-1, chi=10000)
    lnZ, _, _ = O.contraction(tn)
    assert np.isclose(lnZ, X.chain_lnZ(L, J, beta), rtol=1e-9)


@pytest.mark.parametrize("w,h,beta", [([0.5, -0.8, 1.1], [0.2, -0.1, 0.4], 0.3),
                                      ([0.6, -0.5, 0.9], [0.15, -0.3, 0.2], 0.45)])
def test_ising_triangle_brute_force(w, h, beta):
    edges = [(1, 2), (2, 3), (3, 1)]
    tn = O.ising_network(3, edges, w, h, beta, Dmax=-1, chi=10000)
    lnZ, _, _ = O.contraction(tn)
    assert np.isclose(lnZ, X.brute_lnZ(3, edges, w, h, beta), rtol=1e-9)
    tn2 = O.ising_network(3, edges, w, h, beta, Dmax=-1, chi=10000)
    lz, F = O.free_energy(tn2)
    assert np.isclose(lz, lnZ / 3, rtol=1e-9) and np.isclose(F, -lnZ / (3 * beta), rtol=1e-9)


def test_magnetization_correlation_brute_force():
    edges = [(1, 2), (2, 3), (3, 1)]
    w, h, beta = [0.4, 0.6, -0.3], [0.1, -0.2, 0.05], 0.35
    m_ref, c_ref = X.brute_observables(3, edges, w, h, beta)
    m = O.magnetization(3, edges, w, h, beta, Dmax=-1, chi=10000)
    c = O.correlation(3, edges, w, h, beta, Dmax=-1, chi=10000)
    assert np.allclose(m, m_ref, rtol=1e-7)
    assert np.allclose(c, c_ref, rtol=1e-7)


def test_isolated_spins():
    beta = 0.6
    lnZ, _, _ = O.contraction(O.ising_network(1, [], [], [0.5], beta, Dmax=-1, chi=10000))
    assert np.isclose(lnZ, math.log(2 * math.cosh(beta * 0.5)), rtol=1e-9)
    lnZ2, _, _ = O.contraction(O.ising_network(2, [], [], [0.3, 0.7], beta, Dmax=-1, chi=10000))
    assert np.isclose(lnZ2, math.log(2 * math.cosh(beta * 0.3)) + math.log(2 * math.cosh(beta * 0.7)))


def test_double_layer_L4_brute_force():
    """examples/double_layer_ising.jl:165-181: single-layer L=4 at beta_c."""
    beta = math.log(1 + math.sqrt(2)) / 2
    n, edges, w, h = Cf.ising_lattice(4)
    lnZ, _, _ = O.contraction(O.ising_network(n, edges, w, h, beta, Dmax=-1, chi=100000))
    assert np.isclose(lnZ, X.brute_lnZ(n, edges, w, h, beta), rtol=1e-8)


def test_random_small_graph_brute_force():
    r = rng(20)
    for trial in range(3):
        n = 8
        edges = sorted({(min(a, b), max(a, b)) for a, b in r.integers(1, n + 1, size=(14, 2)) if a != b})
        w = r.uniform(-1, 1, len(edges))
        h = r.uniform(-0.5, 0.5, n)
        lnZ, _, _ = O.contraction(O.ising_network(n, edges, w, h, 0.7, Dmax=-1, chi=100000))
        assert np.isclose(lnZ, X.brute_lnZ(n, edges, w, h, 0.7), rtol=1e-9)


# ---------------- closed-form oracles themselves ----------------

def test_exact_oracles_agree():
    r = rng(21)
    L = 4
    n, edges, _, _ = Cf.ising_lattice(L)
    w = r.choice([-1.0, 1.0], len(edges))
    h = r.normal(0, 0.3, n)
    b = X.brute_lnZ(n, edges, w, h, 0.8)
    assert np.isclose(X.lattice_lnZ(L, w, h, 0.8), b, rtol=1e-12)
    assert np.isclose(X.elimination_lnZ(n, edges, w, h, 0.8), b, rtol=1e-12)


def test_config1_against_transfer_matrix():
    """Config 1 (16x16 ferromagnet, beta=0.44, Dmax=20, chi=200) vs the exact
    row transfer matrix: truncation error per spin below 1e-8."""
    n, edges, w, h = Cf.ising_lattice(16)
    tn = O.ising_network(n, edges, w, h, 0.44, Dmax=20, chi=200)
    lnZ, _, _ = O.contraction(tn)
    ex = X.lattice_lnZ(16, w, h, 0.44)
    assert abs(lnZ - ex) / n < 1e-8


# ---------------- planted mistakes: the checks above must catch them ----------------

def test_planted_mistake_transposed_theta_is_caught(monkeypatch):
    """A swap! that builds theta with a transposed operand order fails the
    dense-identity pin."""
    orig = np.einsum

    def bad_einsum(spec, *ops, **kw):
        if spec == "ijk,kab->iajb":
            spec = "ijk,kab->ijab"
        return orig(spec, *ops, **kw)

    monkeypatch.setattr(O.np, "einsum", bad_einsum)
    T = rng(22).standard_normal((2, 3, 4, 5))
    node = O.MPSNode.from_tensor(T, [10, 20, 30, 40], chi=1000)
    O.swap(node, 2, 3)
    assert not np.allclose(O.mps2raw(node), np.transpose(T, (0, 2, 1, 3)))


def test_planted_mistake_dropped_field_is_caught(monkeypatch):
    """An ising_network that drops the field term fails the brute-force pin."""
    edges = [(1, 2), (2, 3), (3, 1)]
    w, h, beta = [0.5, -0.8, 1.1], [0.2, -0.1, 0.4], 0.3
    good = O.contraction(O.ising_network(3, edges, w, h, beta, Dmax=-1, chi=10000))[0]
    bad = O.contraction(O.ising_network(3, edges, w, [0.0] * 3, beta, Dmax=-1, chi=10000))[0]
    ref = X.brute_lnZ(3, edges, w, h, beta)
    assert np.isclose(good, ref, rtol=1e-9) and not np.isclose(bad, ref, rtol=1e-6)
