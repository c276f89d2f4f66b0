"""tsvd's fallback (src/linalg_utils.jl:16-21): when LAPACK's divide-and-conquer
SVD fails, the reference retries with the QR-iteration driver and carries on
with the same truncation. On the device a primary SVD path that does not
converge within its sweep limit hands over to a fallback one-sided Jacobi on
the operand itself (no QR preconditioner, up to 200 sweeps).

Non-convergence does not occur on ordinary inputs, so the fallback is forced
from the host (catn_net_set_fault, bit 0: every primary SVD counts as not
converged). A run that takes the fallback at every SVD must give what the
primary paths give: the same edge sequence and Dmax-cut kept counts, lnZ to
1e-12 per spin, the same discarded weight, and, through the reverse sweep's
replay, the same gradient. Runs on the CPU emulator and, marked gpu, on the
MI355X.
"""
import numpy as np
import pytest

from oracle import configs as Cf
from oracle import exact as X


@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx")
    return catn, request.getfixturevalue("gpu_ctx")


def _case(name):
    if name == "sg16_D4_chi8":
        n, edges, w, h = Cf.spin_glass_3reg(16, seed=3)
        return n, edges, w, h, 0.9, dict(Dmax=4, chi=8)
    if name == "sg40_D8_chi16":
        n, edges, w, h = Cf.spin_glass_3reg(40, seed=1)
        return n, edges, w, h, 1.0, dict(Dmax=8, chi=16)
    L = 6
    edges = X.square_edges(L)
    r = np.random.default_rng(0)
    return L * L, edges, r.choice([-1.0, 1.0], len(edges)), r.normal(0, 0.1, L * L), 0.8, \
        dict(Dmax=-1, chi=6)


@pytest.mark.parametrize("name", ["sg16_D4_chi8", "sg40_D8_chi16", "lattice6_chi6"])
def test_fallback_gives_the_primary_results(env, name):
    catn, ctx = env
    n, edges, w, h, beta, kw = _case(name)
    out = []
    for fault in (0, 1):
        tn = catn.ising_network(n, edges, w, h, beta, ctx=ctx, **kw)
        tn.set_fault(fault)
        lz, err, psi = catn.contraction(tn)
        fb = tn.profile()[1]["_large_svd"]["fallbacks"]
        out.append((lz, err, psi, tn.edge_log(), tn.cut_kept(), fb, tn.profile()[0]))
    (l0, e0, p0, s0, c0, f0, ops), (l1, e1, p1, s1, c1, f1, _) = out
    assert f0 == 0
    nsvd = sum(ops.get(k, (0,))[0] for k in ("cano_r", "cano_l", "swap", "cut"))
    assert 0 < f1 <= nsvd     # certified full-rank steps factorise by QR, not SVD
    assert s1 == s0 and c1 == c0 and p1 == p0
    assert abs(l1 - l0) / n < 1e-12
    assert abs(e1 - e0) <= 1e-8 * max(1.0, e0)


def test_fallback_replayed_by_the_reverse_sweep(env):
    """A chi-truncated network (no Dmax cut): with the fallback forced, the
    forward and the reverse sweep's replays both take it, and the gradient
    equals the primary run's."""
    catn, ctx = env
    n, edges, w, h = Cf.spin_glass_3reg(24, seed=5)
    res = []
    for fault in (0, 1):
        tn = catn.ising_network(n, edges, w, h, 1.0, ctx=ctx, Dmax=-1, chi=8)
        tn.set_fault(fault)
        lz, _, _ = catn.gradient_batch([tn])
        dw, dh, _ = catn.ising_gradient(tn, len(edges), n)
        assert tn.bwd_profile()[1] < 1e-9       # replays reproduce the forward's factors
        res.append((lz[0], dw, dh))
    assert abs(res[1][0] - res[0][0]) / n < 1e-12
    assert np.abs(res[1][1] - res[0][1]).max() < 1e-9
    assert np.abs(res[1][2] - res[0][2]).max() < 1e-9


def test_fault_flag_checks(env):
    catn, ctx = env
    tn = catn.ising_network(3, [(1, 2), (2, 3)], [1.0, 1.0], [0.0] * 3, 0.5, ctx=ctx)
    with pytest.raises(catn.CatnError, match="unknown flag"):
        tn.set_fault(2)
    rng = np.random.default_rng(0)
    z = catn.tensor_network([rng.standard_normal((2, 2)) + 1j, rng.standard_normal((2, 2))],
                            [["a", "b"], ["b", "a"]], ctx=ctx, Dmax=-1, chi=100)
    with pytest.raises(catn.CatnError, match="real networks"):
        z.set_fault(1)
