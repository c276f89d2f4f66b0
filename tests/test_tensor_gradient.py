"""d lnZ / d T for the input tensors of a generic real tensor network.

The reverse sweep of contraction! runs back through the raw2mps splits
(src/mps_node.jl:31-50) to the tensors given to TensorNetwork
(src/tensor_network.jl:41-140); catn_net_tensor_adjoint reads the result.
The reference itself has no derivative of contraction!, so nothing of its own
pins this; the pins are mathematical:
  * exact contraction (Dmax = -1, chi above every rank): d lnZ / d T_k equals
    the environment of T_k (the network with T_k removed, contracted by
    numpy.einsum) divided by Z, to 1e-10 of the largest entry;
  * Z is linear in each tensor, so sum_k <T_k, d lnZ / d T_k> = the number of
    tensors, at any size;
  * with chi truncation active, central differences of the engine's own lnZ
    at entries whose perturbed runs keep the same schedule.
Every test runs on the CPU emulator and, marked gpu, on the HIP library.
"""
import string

import numpy as np
import pytest

from oracle import configs as Cf
from oracle import exact as X


@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx")
    return catn, request.getfixturevalue("gpu_ctx")


EXACT = dict(Dmax=-1, chi=10_000)


def _letters(ixs):
    ids = {}
    for ix in ixs:
        for l in ix:
            if l not in ids:
                ids[l] = string.ascii_letters[len(ids)]
    return [''.join(ids[l] for l in ix) for ix in ixs]


def einsum_grads(ts, ixs):
    """(Z, [environment_k / Z]) by numpy.einsum over the whole network."""
    sub = _letters(ixs)
    Z = np.einsum(','.join(sub) + '->', *ts, optimize='greedy')
    out = []
    for k in range(len(ts)):
        rest = [t for j, t in enumerate(ts) if j != k]
        rsub = [s for j, s in enumerate(sub) if j != k]
        envk = np.einsum(','.join(rsub) + '->' + sub[k], *rest, optimize='greedy')
        out.append(envk / Z)
    return float(Z), out


def tree_network(seed, signed=False):
    """Tensors of rank 1 to 4 on a graph with two short cycles (B-D-E and
    D-E-f) and vector leaves. No two tensors share two labels: the
    reference's contraction! re-points such a parallel leg onto the eating
    node itself (src/tensor_network.jl:626-648)."""
    rng = np.random.default_rng(seed)
    net = [((3,), ["i"]), ((3, 4, 2, 2), ["i", "j", "k", "r"]), ((4,), ["j"]),
           ((2, 5, 2), ["k", "l", "s"]), ((5, 2, 2, 3), ["l", "r", "n", "p"]),
           ((2, 2), ["n", "s"]), ((3, 3), ["p", "q"]), ((3,), ["q"])]
    draw = rng.standard_normal if signed else (lambda s: rng.uniform(0.2, 1.0, s))
    return [draw(sh) for sh, _ in net], [ix for _, ix in net]


def net_grads(catn, ctx, ts, ixs, **kw):
    tn = catn.tensor_network(ts, ixs, ctx=ctx, **kw)
    lz, _, psi = catn.gradient_batch([tn])
    return float(lz[0]), float(psi[0]), [tn.tensor_gradient(k) for k in range(len(ts))]


def _check_exact(catn, ctx, ts, ixs):
    Z, want = einsum_grads(ts, ixs)
    lz, psi, got = net_grads(catn, ctx, ts, ixs, **EXACT)
    assert lz == pytest.approx(np.log(abs(Z)), abs=1e-10)
    assert psi == np.sign(Z)
    scale = max(float(np.max(np.abs(w))) for w in want)
    for k, (g, w) in enumerate(zip(got, want)):
        assert g.shape == w.shape
        assert np.max(np.abs(g - w)) <= 1e-10 * scale, f"tensor {k}"
    return got


@pytest.mark.parametrize("signed", [False, True])
def test_tree_exact(env, signed):
    catn, ctx = env
    ts, ixs = tree_network(11 + signed, signed)
    _check_exact(catn, ctx, ts, ixs)


@pytest.mark.parametrize("signed", [False, True])
def test_grid3_exact(env, signed):
    catn, ctx = env
    ts, ixs = Cf.dense_lattice(3, D=3, seed=5 + signed, signed=signed)
    got = _check_exact(catn, ctx, ts, ixs)
    # linearity in each tensor
    assert sum(float(np.sum(t * g)) for t, g in zip(ts, got)) == pytest.approx(len(ts), abs=1e-9)


@pytest.mark.parametrize("dtype,tol", [("f32x6", 2e-6), ("f32", 5e-5)])
def test_fp32_storage(env, dtype, tol):
    """FP32 site storage (plain, and six-pass split-BF16 swap products): the
    exact-mode gradient to `tol` of the largest entry, lnZ to `tol`."""
    catn, ctx = env
    ts, ixs = Cf.dense_lattice(3, D=3, seed=5, signed=True)
    Z, want = einsum_grads(ts, ixs)
    lz, psi, got = net_grads(catn, ctx, ts, ixs, dtype=dtype, **EXACT)
    assert lz == pytest.approx(np.log(abs(Z)), abs=tol)
    assert psi == np.sign(Z)
    scale = max(float(np.max(np.abs(w))) for w in want)
    assert max(float(np.max(np.abs(g - w))) for g, w in zip(got, want)) <= tol * scale


def test_einsum_environment_pin():
    """The test's own environment formula, checked against central
    differences of the einsum Z: a transposed or misplaced environment fails."""
    ts, ixs = tree_network(3, signed=True)
    sub = _letters(ixs)
    Zf = lambda tt: np.einsum(','.join(sub) + '->', *tt, optimize='greedy')
    Z, want = einsum_grads(ts, ixs)
    rng = np.random.default_rng(0)
    for k in (1, 4, 6):
        idx = tuple(int(rng.integers(0, s)) for s in ts[k].shape)
        h = 1e-6
        tp = [t.copy() for t in ts]; tp[k][idx] += h
        tm = [t.copy() for t in ts]; tm[k][idx] -= h
        fd = (np.log(abs(Zf(tp))) - np.log(abs(Zf(tm)))) / (2 * h)
        assert fd == pytest.approx(want[k][idx], rel=1e-6, abs=1e-9)


def test_grid5_linearity(env):
    """5 x 5 grid, bond 2, exact: lnZ against the row-transfer contraction
    and sum_k <T_k, d lnZ / d T_k> = 25."""
    catn, ctx = env
    ts, ixs = Cf.dense_lattice(5, D=2, seed=21, signed=True)
    lz, psi, got = net_grads(catn, ctx, ts, ixs, **EXACT)
    lz_x, ph = X.grid_lnZ(ts, 5)
    assert lz == pytest.approx(lz_x, abs=1e-9)
    assert psi == pytest.approx(ph)
    assert sum(float(np.sum(t * g)) for t, g in zip(ts, got)) == pytest.approx(25.0, abs=1e-8)


def test_truncated_finite_difference(env):
    """chi truncation active (4 x 4 grid, bond 3, chi 6; lnZ off the exact
    value by 5e-3): the gradient of the executed schedule against central
    differences of the engine's lnZ."""
    catn, ctx = env
    ts, ixs = Cf.dense_lattice(4, D=3, seed=3)
    kw = dict(Dmax=-1, chi=6)
    lz, _, got = net_grads(catn, ctx, ts, ixs, **kw)
    lz_x, _ = X.grid_lnZ(ts, 4)
    assert abs(lz - lz_x) > 1e-3          # the truncation is active
    rng = np.random.default_rng(1)
    h = 1e-5
    for k in (0, 5, 10, 15):
        idx = tuple(int(rng.integers(0, s)) for s in ts[k].shape)
        vals = []
        for sgn in (1, -1):
            tt = [t.copy() for t in ts]
            tt[k][idx] += sgn * h
            tn = catn.tensor_network(tt, ixs, ctx=ctx, **kw)
            vals.append(catn.contraction(tn)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        assert got[k][idx] == pytest.approx(fd, rel=1e-5, abs=1e-7), f"tensor {k} {idx}"


def test_errors(env):
    catn, ctx = env
    ts, ixs = tree_network(2)
    tn = catn.tensor_network(ts, ixs, ctx=ctx, **EXACT)
    with pytest.raises(catn.CatnError, match="gradient_batch first"):
        tn.tensor_gradient(0)
    catn.gradient_batch([tn])
    with pytest.raises(catn.CatnError, match="no tensor"):
        tn.tensor_gradient(len(ts))
    n = int(np.prod(ts[1].shape))
    out = np.zeros(n + 1)
    with pytest.raises(catn.CatnError, match="exceeds"):
        ctx._check(ctx.lib.catn_net_tensor_adjoint(tn.h, 1, catn._dptr(out), n + 1))
