"""Randomized SVD (src/linalg_utils.jl:44-58) and swap!'s rsvd branch
(src/mps_node.jl:154-156).

Pins (the reference's own tests and properties; its Gaussian stream comes from
Julia's RNG and is not reproducible, so no stored vectors exist):
  * test/test_linalg.jl:20-34: a rank-4 50x40 and a rank-3 40x30 matrix are
    reconstructed to atol 1e-8; the same seed gives the same result;
  * the top-k singular values of a matrix with a decaying spectrum equal
    LAPACK's to 1e-10 relative, and U, V are orthonormal;
  * swap! through the rsvd branch (size switch lowered for the test) on a
    theta of rank chi gives the same swapped tensor as the tsvd route and the
    oracle, and its discarded weight is 0 as in the reference.
"""
import numpy as np
import pytest

from oracle import catn_oracle as O
from oracle import configs as Cf


@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx")
    return catn, request.getfixturevalue("gpu_ctx")


def test_low_rank_reconstruction_and_determinism(env):
    catn, ctx = env
    r = np.random.default_rng(1)
    L = r.standard_normal((50, 4)) @ r.standard_normal((4, 40))
    U, S, V = catn.rsvd(L, 4, 10, 10, seed=0, ctx=ctx)
    assert np.allclose(U @ np.diag(S) @ V.T, L, rtol=0, atol=1e-8)
    r = np.random.default_rng(2)
    L = r.standard_normal((40, 3)) @ r.standard_normal((3, 30))
    U, S, V = catn.rsvd(L, 3, 10, 10, seed=7, ctx=ctx)
    assert np.allclose(U @ np.diag(S) @ V.T, L, rtol=0, atol=1e-8)
    U2, S2, V2 = catn.rsvd(L, 3, 10, 10, seed=7, ctx=ctx)
    assert np.array_equal(S, S2) and np.array_equal(U, U2)
    U3, S3, _ = catn.rsvd(L, 3, 10, 10, seed=8, ctx=ctx)     # another sketch, same answer
    assert np.allclose(S3, S, rtol=1e-10)


def test_top_singular_values_and_orthonormality(env):
    catn, ctx = env
    r = np.random.default_rng(3)
    m, n, k = 90, 70, 5
    Qa, _ = np.linalg.qr(r.standard_normal((m, n)))
    Qb, _ = np.linalg.qr(r.standard_normal((n, n)))
    s = 0.5 ** np.arange(n)
    A = Qa @ np.diag(s) @ Qb.T
    U, S, V = catn.rsvd(A, k, 10, 10, seed=3, ctx=ctx)
    ref = np.linalg.svd(A, compute_uv=False)[:k]
    assert U.shape == (m, k) and V.shape == (n, k)
    assert np.allclose(S, ref, rtol=1e-10, atol=0)
    assert np.allclose(U.T @ U, np.eye(k), atol=1e-12)
    assert np.allclose(V.T @ V, np.eye(k), atol=1e-12)
    # p = min(n, oversample k) = n: the sketch spans the whole column space
    U, S, V = catn.rsvd(A, 10, 10, 2, seed=4, ctx=ctx)
    assert np.allclose(S, np.linalg.svd(A, compute_uv=False)[:10], rtol=1e-12, atol=1e-15)


def _rank4_node_sites(seed):
    """Two sites tl (8, 8, 4), tr (4, 8, 8) whose swap theta (64 x 64) has rank
    4: every term tl[:, :, k] and tr[k, :, :] is an outer product."""
    r = np.random.default_rng(seed)
    d0 = d1 = d2 = d3 = 8
    dd = 4
    tl = np.zeros((d0, d2, dd))
    tr = np.zeros((dd, d1, d3))
    for k in range(dd):
        tl[:, :, k] = np.outer(r.standard_normal(d0), r.standard_normal(d2))
        tr[k, :, :] = np.outer(r.standard_normal(d1), r.standard_normal(d3))
    return tl, tr


@pytest.mark.parametrize("direction", ["right", "left"])
def test_swap_rsvd_branch_matches_tsvd_route(env, direction):
    """swap! on a node whose theta (64 x 64, rank 4) trips the size switch
    (lowered to 50 for the test): rsvd(M, chi = 4, 10, 10) is exact there, so
    the swapped two-site tensor equals the tsvd route's and the oracle's, and
    the returned discarded weight is 0 (src/mps_node.jl:156)."""
    catn, ctx = env
    tl, tr = _rank4_node_sites(11)
    i, j = (1, 2) if direction == "right" else (2, 1)
    out = {}
    for name, kw in (("rsvd", dict(rsvd_both=50, rsvd_any=10_000)), ("tsvd", dict())):
        net = catn.MPSNet(ctx=ctx, chi=4, Dmax=-1, **kw)
        net.add_node(1, [tl, tr], [2, 3], cano=i)
        err = net.swap(1, i, j)
        a, b = net.sites(1)
        out[name] = (np.einsum("iak,kjb->iajb", a, b), err, a.shape, b.shape)
    node = O.MPSNode([tl.copy(), tr.copy()], [2, 3], cano=i, chi=4)
    O.swap(node, i, j)
    ref = np.einsum("iak,kjb->iajb", node.mps[0], node.mps[1])
    scale = np.abs(ref).max()
    assert out["rsvd"][2] == (8, 8, 4) and out["rsvd"][3] == (4, 8, 8)
    assert np.abs(out["rsvd"][0] - ref).max() < 1e-12 * scale
    assert np.abs(out["tsvd"][0] - ref).max() < 1e-12 * scale
    assert out["rsvd"][1] == 0.0
    assert out["tsvd"][1] > 0.0          # the tsvd route reports its rounding-level tail


def test_gradient_through_rsvd_swaps_raises(env):
    """The randomized SVD has no reverse step: a gradient run whose swaps take
    the rsvd branch raises instead of returning a wrong derivative."""
    catn, ctx = env
    n, edges, w, h = Cf.spin_glass_3reg(12, seed=2)
    tn = catn.ising_network(n, edges, w, h, 1.0, ctx=ctx, Dmax=-1, chi=64, rsvd_both=1, rsvd_any=1)
    with pytest.raises(catn.CatnError):
        catn.gradient_batch([tn])
