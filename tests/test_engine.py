"""Parity of the contraction engine against the FP64 oracle and exact results.

Every test runs twice: on the CPU emulator build of the engine (same source,
one-thread workgroups; part of the CPU suite) and, marked `gpu`, on the HIP
library on an MI355X through the C ABI. Checks restate the reference's tests
(test/test_mps_node.jl, test/test_contraction.jl, test/test_ising.jl,
test/test_linalg.jl, test/test_gpu.jl) plus the committed golden fixtures.

Tolerances: lnZ/N within 1e-12 of the oracle in FP64 (same algorithm, both
backward stable); dense identities to 1e-9 relative; edge sequences and
Dmax-cut kept counts bit-exact.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import catn_oracle as O
from oracle import configs as Cf
from oracle import exact as X

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx")
    return catn, request.getfixturevalue("gpu_ctx")


def rng(s):
    return np.random.default_rng(s)


def mps2raw(sites):
    return O.mps2raw(O.MPSNode(list(sites), []))


def node_from_tensor(catn, ctx, T, nb, **kw):
    net = catn.MPSNet(ctx=ctx, **kw)
    net.add_tensor(1, T, nb)
    return net


# ------------------------------ tsvd ------------------------------------

@pytest.mark.parametrize("shape", [(1, 1), (2, 2), (8, 5), (5, 8), (30, 7), (7, 30), (64, 64),
                                   (100, 20), (160, 160), (300, 120), (120, 300)])
def test_tsvd_spectrum_and_reconstruction(env, shape):
    catn, ctx = env
    A = rng(1).standard_normal(shape)
    U, S, V, disc, Sfull = catn.tsvd(A, ctx=ctx)
    ref = np.linalg.svd(A, compute_uv=False)
    assert np.allclose(Sfull, ref, rtol=1e-12, atol=1e-12 * ref[0])
    assert np.allclose(U @ np.diag(S) @ V.T, A, atol=1e-11 * ref[0])
    assert np.allclose(U.T @ U, np.eye(U.shape[1]), atol=1e-12)
    assert np.allclose(V.T @ V, np.eye(V.shape[1]), atol=1e-12)


@pytest.mark.parametrize("lds,shape,rank", [(5056, (100, 100), None), (5056, (400, 128), None),
                                            (2048, (64, 40), None), (5056, (100, 100), 30),
                                            (5056, (70, 90), None)])
def test_tsvd_column_block_jacobi(env, lds, shape, rank):
    """LDS budgets under which the Jacobi of the R factor runs by column blocks
    staged in LDS: an even block count (n=100, b=18), an odd one with a padding
    block (n=128, b=15), a ragged last block (n=40, b=14), a rank-deficient
    operand (noise columns under the threshold) and a wide one (through its
    transpose). Spectrum against LAPACK, reconstruction, orthogonality."""
    catn, ctx = env
    r = rng(7)
    A = r.standard_normal(shape)
    if rank:
        A = r.standard_normal((shape[0], rank)) @ r.standard_normal((rank, shape[1]))
    ctx.set_launch(256, lds, 0, 0)
    try:
        U, S, V, disc, Sfull = catn.tsvd(A, cutoff=0.0, ctx=ctx)
    finally:
        ctx.set_launch(256, 0, 0, 0)
    ref = np.linalg.svd(A, compute_uv=False)
    assert np.allclose(Sfull, ref, rtol=1e-12, atol=1e-12 * ref[0])
    assert np.allclose(U @ np.diag(S) @ V.T, A, atol=1e-11 * ref[0])
    k = rank or min(shape)
    assert np.allclose(U[:, :k].T @ U[:, :k], np.eye(k), atol=1e-12)
    assert np.allclose(V[:, :k].T @ V[:, :k], np.eye(k), atol=1e-12)


def test_tsvd_truncation_rules(env):
    """test/test_linalg.jl:4-18 plus the keep-at-least-one rule."""
    catn, ctx = env
    A = rng(2).standard_normal((8, 5))
    U, S, V, d, Sf = catn.tsvd(A, maxdim=2, ctx=ctx)
    assert len(S) == 2 and np.isclose(d, Sf[2:].sum())
    Ub, Sb, Vb = np.linalg.svd(A, full_matrices=False)
    B = Ub[:, :3] @ np.diag([1.0, 1e-3, 1e-20]) @ Vb[:3]
    _, S3, _, _, _ = catn.tsvd(B, cutoff=1e-10, ctx=ctx)
    assert len(S3) == 2
    _, S0, _, d0, _ = catn.tsvd(np.zeros((3, 4)), ctx=ctx)
    assert len(S0) == 1 and S0[0] == 0.0 and d0 == 0.0


def test_tsvd_rank_deficient_noise_below_cutoff(env):
    catn, ctx = env
    r = rng(3)
    A = r.standard_normal((40, 6)) @ r.standard_normal((6, 40))
    A /= np.linalg.norm(A)
    U, S, V, d, Sf = catn.tsvd(A, cutoff=1e-15, ctx=ctx)
    ref = np.linalg.svd(A, compute_uv=False)
    assert len(S) == int(np.count_nonzero(ref > 1e-15)) == 6
    assert np.allclose(U @ np.diag(S) @ V.T, A, atol=1e-14)


# ------------------------------ MPSNode primitives ------------------------

@pytest.mark.parametrize("dims", [(4,), (3, 5), (2, 3, 4), (2, 3, 2, 3)])
def test_raw2mps_roundtrip(env, dims):
    catn, ctx = env
    T = rng(4).standard_normal(dims)
    net = node_from_tensor(catn, ctx, T, list(range(1, len(dims) + 1)), chi=1000)
    order, cano, _, sdims = net.node(1)
    assert order == len(dims) and cano == len(dims)
    assert [d[1] for d in sdims] == list(dims)
    assert np.allclose(mps2raw(net.sites(1)), T, atol=1e-12)


def test_cano_to_preserves_tensor_and_isometry(env):
    catn, ctx = env
    T = rng(5).standard_normal((2, 3, 4, 2))
    net = node_from_tensor(catn, ctx, T, [1, 2, 3, 4], chi=1000)
    for idx in [1, 2, 3, 4, 0]:
        net.cano_to(1, idx)
        assert np.allclose(mps2raw(net.sites(1)), T, atol=1e-12)
    net.left_canonical(1)
    assert net.node(1)[1] == 4
    for A in net.sites(1)[:3]:
        M = A.reshape(A.shape[0] * A.shape[1], A.shape[2], order="F")
        assert np.allclose(M.T @ M, np.eye(M.shape[1]), atol=1e-10)


def test_cano_to_tiny_norm_keeps_rank(env):
    """test/test_mps_node.jl:40-49: all S below the cutoff -> keep them all."""
    catn, ctx = env
    T = rng(6).standard_normal((2, 3, 4, 2)) * 1e-16
    net = catn.MPSNet(ctx=ctx, chi=1000, cutoff=0.0)
    ref_node = O.MPSNode.from_tensor(T, [1, 2, 3, 4], chi=1000, cutoff=0.0)
    net.add_node(1, ref_node.mps, [1, 2, 3, 4])
    net2 = catn.MPSNet(ctx=ctx, chi=1000, cutoff=1e-15)
    net2.add_node(1, ref_node.mps, [1, 2, 3, 4], cano=4)
    ref = O.mps2raw(ref_node)
    net2.cano_to(1, 1)
    assert np.allclose(mps2raw(net2.sites(1)), ref, rtol=1e-8, atol=1e-30)
    net2.cano_to(1, 4)
    assert np.allclose(mps2raw(net2.sites(1)), ref, rtol=1e-8, atol=1e-30)


def test_swap_both_directions_reverse_move2tail(env):
    catn, ctx = env
    T = rng(7).standard_normal((2, 3, 4, 5))
    net = node_from_tensor(catn, ctx, T, [10, 20, 30, 40], chi=1000)
    net.swap(1, 2, 3)
    assert net.neighbor(1, 4) == [10, 30, 20, 40]
    assert np.allclose(mps2raw(net.sites(1)), np.transpose(T, (0, 2, 1, 3)), atol=1e-12)
    net = node_from_tensor(catn, ctx, T, [10, 20, 30, 40], chi=1000)
    net.swap(1, 3, 2)
    assert net.neighbor(1, 4) == [10, 30, 20, 40]
    assert np.allclose(mps2raw(net.sites(1)), np.transpose(T, (0, 2, 1, 3)), atol=1e-12)
    net = node_from_tensor(catn, ctx, T, [10, 20, 30, 40], chi=1000)
    net.reverse(1)
    assert net.neighbor(1, 4) == [40, 30, 20, 10]
    assert np.allclose(mps2raw(net.sites(1)), np.transpose(T, (3, 2, 1, 0)), atol=1e-12)
    net = node_from_tensor(catn, ctx, T, [10, 20, 30, 40], chi=1000)
    net.move2tail(1, 1)
    assert net.neighbor(1, 4)[-1] == 10
    assert np.allclose(mps2raw(net.sites(1)), np.transpose(T, (1, 2, 3, 0)), atol=1e-12)
    net = node_from_tensor(catn, ctx, T, [10, 20, 30, 40], chi=1000)
    net.move2head(1, 4)
    assert net.neighbor(1, 4)[0] == 40
    assert np.allclose(mps2raw(net.sites(1)), np.transpose(T, (3, 0, 1, 2)), atol=1e-12)


def test_swap_truncation_error_matches_oracle(env):
    catn, ctx = env
    T = rng(8).standard_normal((4, 4, 4))
    net = node_from_tensor(catn, ctx, T, [1, 2, 3], chi=2)
    err = net.swap(1, 1, 2)
    ref = O.MPSNode.from_tensor(T, [1, 2, 3], chi=2)
    e_ref = O.swap(ref, 1, 2)
    assert err > 0 and np.isclose(err, e_ref, rtol=1e-10)
    assert np.allclose(mps2raw(net.sites(1)), O.mps2raw(ref), atol=1e-10)


@pytest.mark.parametrize("opt", [False, True])
def test_compress_preserves(env, opt):
    catn, ctx = env
    T = rng(9).standard_normal((2, 3, 4, 3, 2))
    net = node_from_tensor(catn, ctx, T, [1, 2, 3, 4, 5], chi=1000)
    e = net.compress(1, opt=opt)
    assert net.node(1)[1] == 1
    assert e >= 0.0
    assert np.allclose(mps2raw(net.sites(1)), T, atol=1e-11)


def test_merge_fuses_duplicate_legs(env):
    catn, ctx = env
    T = rng(10).standard_normal((2, 5, 3))
    net = node_from_tensor(catn, ctx, T, [7, 8, 7], chi=1000)
    net.merge(1, 7)
    assert net.neighbor(1, 2).count(7) == 1
    raw = mps2raw(net.sites(1))
    exp = np.transpose(T, (0, 2, 1)).reshape(6, 5, order="F")
    assert np.allclose(raw.reshape(raw.shape[0], raw.shape[1]), exp, atol=1e-12)


def test_eat_cases(env):
    catn, ctx = env
    r = rng(11)
    # (c) general
    Ti, Tj = r.standard_normal((2, 3, 4)), r.standard_normal((4, 5))
    net = catn.MPSNet(ctx=ctx, chi=1000, norm_method=0)
    net.add_tensor(1, Ti, [1, 2, 3])
    net.add_tensor(2, Tj, [3, 4])
    ln, err, ph = net.eat(1, 2, 3, 1)
    raw = mps2raw(net.sites(1)) * math.exp(ln) * ph
    assert np.allclose(raw, np.einsum("abc,cd->abd", Ti, Tj), atol=1e-10)
    # (a) both leaves, non-conjugating dot
    u, v = r.standard_normal(3), r.standard_normal(3)
    net = catn.MPSNet(ctx=ctx, norm_method=0)
    net.add_tensor(1, u, [2])
    net.add_tensor(2, v, [1])
    ln, err, ph = net.eat(1, 2, 1, 1)
    assert net.node(1)[0] == 0 and np.isclose(math.exp(ln) * ph, u @ v)
    # (b) leaf j
    Ti, Tj = r.standard_normal((2, 4)), r.standard_normal(4)
    net = catn.MPSNet(ctx=ctx, chi=1000, norm_method=0)
    net.add_tensor(1, Ti, [5, 9])
    net.add_tensor(2, Tj, [9])
    ln, err, ph = net.eat(1, 2, 2, 1)
    assert net.neighbor(1, 1) == [5]
    assert np.allclose(mps2raw(net.sites(1)).ravel() * math.exp(ln) * ph, Ti @ Tj, atol=1e-10)


def test_eat_case_c_left_canonical_normalised(env):
    catn, ctx = env
    r = rng(12)
    net = catn.MPSNet(ctx=ctx, chi=1000, norm_method=1)
    net.add_tensor(1, r.standard_normal((2, 3, 4)), [1, 2, 3])
    net.add_tensor(2, r.standard_normal((4, 5, 6)), [3, 4, 5])
    net.eat(1, 2, 3, 1)
    order, cano, _, _ = net.node(1)
    assert cano == order
    sites = net.sites(1)
    for A in sites[:-1]:
        M = A.reshape(A.shape[0] * A.shape[1], A.shape[2], order="F")
        assert np.allclose(M.T @ M, np.eye(M.shape[1]), atol=1e-10)
    assert np.isclose(np.linalg.norm(sites[-1]), 1.0)


@pytest.mark.parametrize("opt", [False, True])
def test_cut_bond_lossless_low_rank(env, opt):
    """test/test_contraction.jl:42-57."""
    catn, ctx = env
    r = rng(13)
    P = r.standard_normal((3, 2)) @ r.standard_normal((2, 4))
    Q = r.standard_normal((4, 2)) @ r.standard_normal((2, 3))
    tn_net = catn.MPSNet(ctx=ctx, chi=1000, Dmax=2)
    tn_net.add_tensor(1, P, [-1, 2])
    tn_net.add_tensor(2, Q, [1, -2])
    tn_net.cut_bondim(1, 2, opt=opt)
    assert tn_net.node(1)[3][1][1] == 2
    got = mps2raw(tn_net.sites(1)) @ mps2raw(tn_net.sites(2))
    assert np.allclose(got, P @ Q, atol=1e-10)


def test_structural_errors(env):
    catn, ctx = env
    T = rng(14).standard_normal((2, 3, 4))
    net = node_from_tensor(catn, ctx, T, [1, 2, 3])
    with pytest.raises(catn.CatnError):
        net.swap(1, 1, 3)            # not adjacent (src/mps_node.jl:128)
    with pytest.raises(catn.CatnError):
        net.merge(1, 99)             # no duplicate (src/mps_node.jl:597)
    with pytest.raises(catn.CatnError):
        catn.tensor_network([np.ones((2, 2))], [["a", "a"]], ctx=ctx)   # self-loop
    with pytest.raises(catn.CatnError):
        catn.tensor_network([np.ones(2)] * 3, [["a"], ["a"], ["a"]], ctx=ctx)
    tn = catn.ising_network(3, [(1, 2), (2, 3)], [1.0, 1.0], [0.0] * 3, 0.5, ctx=ctx)
    catn.contraction(tn)
    with pytest.raises(catn.CatnError):
        catn.contraction(tn)         # consumed like the reference


# ------------------------------ networks ----------------------------------

def _networks(r):
    return [
        ([r.standard_normal((3, 4)), r.standard_normal((4, 5)), r.standard_normal((5, 3))],
         [["a", "b"], ["b", "c"], ["c", "a"]]),
        ([r.standard_normal((2, 3, 4)), r.standard_normal(2), r.standard_normal(3), r.standard_normal(4)],
         [["a", "b", "c"], ["a"], ["b"], ["c"]]),
        ([r.standard_normal((2, 3)), r.standard_normal((3, 2)), r.standard_normal((2, 3)), r.standard_normal((3, 2))],
         [["a", "b"], ["b", "c"], ["c", "d"], ["d", "a"]]),
    ]


def test_exact_mode_matches_einsum(env):
    """test/test_contraction.jl:69-88: select x reverse x compress, rtol 1e-8."""
    catn, ctx = env
    for ts, ixs in _networks(rng(15)):
        ref = float(X.exact_contract(ts, ixs))
        for sel in (0, 1, 2):
            for rev in (False, True):
                for comp in (False, True):
                    tn = catn.tensor_network(ts, ixs, Dmax=-1, chi=10000, select=sel, reverse=rev,
                                             compress=comp, ctx=ctx)
                    lnZ, _, psi = catn.contraction(tn)
                    assert np.isclose(math.exp(lnZ) * psi, ref, rtol=1e-8)


def test_finite_dmax_loop(env):
    """test/test_contraction.jl:90-96."""
    catn, ctx = env
    r = rng(16)
    ts = [r.standard_normal((4, 4)) for _ in range(4)]
    ixs = [["a", "b"], ["b", "c"], ["c", "d"], ["d", "a"]]
    ref = float(X.exact_contract(ts, ixs))
    tn = catn.tensor_network(ts, ixs, Dmax=4, chi=64, select=1, compress=True, ctx=ctx)
    lnZ, _, psi = catn.contraction(tn)
    assert np.isclose(math.exp(lnZ) * psi, ref, rtol=1e-6)


def test_ising_closed_forms(env):
    """test/test_ising.jl: chain TM, triangle brute force, isolated spins."""
    catn, ctx = env
    L, J, beta = 6, 0.7, 0.4
    tn = catn.ising_network(L, [(i, i + 1) for i in range(1, L)], [J] * (L - 1), np.zeros(L), beta,
                            Dmax=-1, chi=10000, ctx=ctx)
    assert np.isclose(catn.contraction(tn)[0], X.chain_lnZ(L, J, beta), rtol=1e-9)
    edges = [(1, 2), (2, 3), (3, 1)]
    w, h = [0.5, -0.8, 1.1], [0.2, -0.1, 0.4]
    tn = catn.ising_network(3, edges, w, h, 0.3, Dmax=-1, chi=10000, ctx=ctx)
    assert np.isclose(catn.contraction(tn)[0], X.brute_lnZ(3, edges, w, h, 0.3), rtol=1e-9)
    tn = catn.ising_network(3, edges, w, h, 0.3, Dmax=-1, chi=10000, ctx=ctx)
    lz, F = catn.free_energy(tn)
    ref = X.brute_lnZ(3, edges, w, h, 0.3)
    assert np.isclose(lz, ref / 3, rtol=1e-9) and np.isclose(F, -ref / (3 * 0.3), rtol=1e-9)
    tn = catn.ising_network(2, [], [], [0.3, 0.7], 0.6, Dmax=-1, chi=10000, ctx=ctx)
    assert np.isclose(catn.contraction(tn)[0],
                      math.log(2 * math.cosh(0.18)) + math.log(2 * math.cosh(0.42)), rtol=1e-9)


def test_magnetization_correlation_brute_force(env):
    """test/test_ising.jl:103-124 (pinning, here as one batch)."""
    catn, ctx = env
    edges = [(1, 2), (2, 3), (3, 1)]
    w, h, beta = [0.4, 0.6, -0.3], [0.1, -0.2, 0.05], 0.35
    m_ref, c_ref = X.brute_observables(3, edges, w, h, beta)
    m = catn.magnetization(3, edges, w, h, beta, Dmax=-1, chi=10000, ctx=ctx)
    c = catn.correlation(3, edges, w, h, beta, Dmax=-1, chi=10000, ctx=ctx)
    assert np.allclose(m, m_ref, rtol=1e-7) and np.allclose(c, c_ref, rtol=1e-7)


def test_random_graph_against_oracle_bit_exact_schedule(env):
    """Planner (C++) vs oracle (Python): identical edge sequences, structural
    traces (selected pair, order swap, reversals, leg indices, duplicate bonds
    with their cross flags) and Dmax-cut kept counts; lnZ/N to 1e-12 on
    random graphs with truncation."""
    catn, ctx = env
    for seed in range(4):
        n, edges, w, h = Cf.spin_glass_3reg(16, seed=seed)
        for sel in (0, 1, 2):
            tn = catn.ising_network(n, edges, w, h, 0.9, Dmax=4, chi=8, select=sel, ctx=ctx)
            lz, err, psi = catn.contraction(tn)
            ref = O.ising_network(n, edges, w, h, 0.9, Dmax=4, chi=8, select=sel)
            s = O.Schedule()
            lo, eo, po = O.contraction(ref, s)
            assert tn.edge_log() == s.edges
            t = tn.plan_trace()
            assert t["steps"] == s.steps and t["dups"] == s.dups
            assert t["cuts"] == [e[2] for e in s.events if e[0] == "cut_keep"]
            assert abs(lz - lo) / n < 1e-12
            assert np.isclose(err, eo, rtol=1e-6, atol=1e-13)


@pytest.mark.parametrize("name", ["ising4_exact", "sg20", "config2_sg100", "config1_ising16"])
def test_golden(env, name):
    catn, ctx = env
    g = json.load(open(os.path.join(GOLDEN, name + ".json")))
    tn = catn.ising_network(g["n"], [tuple(e) for e in g["edges_in"]], g["w"], g["h"], g["beta"],
                            Dmax=g["Dmax"], chi=g["chi"], ctx=ctx)
    lz, err, psi = catn.contraction(tn)
    assert [list(e) for e in tn.edge_log()] == g["edge_sequence"]
    assert tn.cut_kept() == g["cut_kept"]
    assert abs(lz - g["lnZ"]) / g["n"] < 1e-12
    assert np.isclose(err, g["error"], rtol=1e-6, atol=1e-14)
    assert tn.maxdim_intermediate == g["maxdim_intermediate"]
    # truncation error of the engine itself against the exact value
    tol = 1e-12 if g["Dmax"] < 0 else 1e-6
    assert abs(lz - g["exact_lnZ"]) / g["n"] < tol


def test_batch_equals_single(env):
    """contraction_batch: each network's result identical to its own run."""
    catn, ctx = env
    n, edges, w, h = Cf.spin_glass_3reg(16, seed=9)
    nets = [catn.ising_network(n, edges, w, h, b, Dmax=4, chi=16, ctx=ctx) for b in (0.3, 0.6, 0.9)]
    lz, _, _ = catn.contraction_batch(nets)
    for k, b in enumerate((0.3, 0.6, 0.9)):
        single = catn.contraction(catn.ising_network(n, edges, w, h, b, Dmax=4, chi=16, ctx=ctx))[0]
        assert lz[k] == single
