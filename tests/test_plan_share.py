"""Plan sharing in batched contraction!.

Networks of a batch whose planner states are equal (the pinned networks of
one ising instance differ only in site values) are planned once per segment
and run the same uploaded program. Each network must still get exactly what
it gets when contracted alone: the same lnZ, error, edge sequence, Dmax-cut
kept counts and maxdim_intermediate, also when the cuts of different pins
keep different counts and their plans diverge.
"""
import numpy as np
import pytest

from oracle import configs as Cf


@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx")
    return catn, request.getfixturevalue("gpu_ctx")


def _pins(n, edges):
    out = [dict()]
    for k in range(1, n + 1):
        for v in (1, 2):
            out.append(dict(pos1=k, val1=v))
    for (a, b) in edges[:6]:
        for v1 in (1, 2):
            for v2 in (1, 2):
                out.append(dict(pos1=a, val1=v1, pos2=b, val2=v2))
    return out


@pytest.mark.parametrize("case", ["sg24_D4_chi8", "sg30_D6_chi16", "lattice5_D3"])
def test_batch_equals_networks_alone(env, case):
    catn, ctx = env
    if case == "sg24_D4_chi8":
        n, edges, w, h = Cf.spin_glass_3reg(24, seed=7)
        kw = dict(Dmax=4, chi=8)
    elif case == "sg30_D6_chi16":
        n, edges, w, h = Cf.spin_glass_3reg(30, seed=2)
        kw = dict(Dmax=6, chi=16)
    else:
        from oracle import exact as X
        edges = X.square_edges(5)
        n = 25
        r = np.random.default_rng(3)
        w, h = r.uniform(-1, 1, len(edges)), r.normal(0, 0.2, n)
        kw = dict(Dmax=3, chi=32)
    pins = _pins(n, edges)
    batch = [catn.ising_network(n, edges, w, h, 1.0, ctx=ctx, **p, **kw) for p in pins]
    lz, err, psi = catn.contraction_batch(batch)
    cuts = set()
    for k, p in enumerate(pins):
        alone = catn.ising_network(n, edges, w, h, 1.0, ctx=ctx, **p, **kw)
        la, ea, pa = catn.contraction(alone)
        assert lz[k] == la and err[k] == ea and psi[k] == pa, (k, p)
        assert batch[k].edge_log() == alone.edge_log()
        assert batch[k].maxdim_intermediate == alone.maxdim_intermediate
        ca = alone.profile()[0].get("cut", (0,))[0]
        assert batch[k].profile()[0].get("cut", (0,))[0] == ca
        cuts.add(ca)
    assert max(cuts) > 0        # the Dmax cut is exercised


def test_plans_diverge_after_a_cut(env):
    """Two instances on one graph, the second with most couplings zero (rank-1
    bonds, so its cuts keep fewer than Dmax): both start from the same shared
    plan and must part after the first cut, each as when run alone."""
    catn, ctx = env
    n, edges, w, h = Cf.spin_glass_3reg(24, seed=7)
    w0 = np.array(w, dtype=float).copy()
    w0[: 3 * len(edges) // 4] = 0.0
    kw = dict(Dmax=4, chi=8)
    specs = [(w, dict()), (w0, dict()), (w, dict(pos1=3, val1=1)), (w0, dict(pos1=3, val1=2))]
    batch = [catn.ising_network(n, edges, ww, h, 1.0, ctx=ctx, **p, **kw) for ww, p in specs]
    lz, err, _ = catn.contraction_batch(batch)
    logs = []
    for k, (ww, p) in enumerate(specs):
        alone = catn.ising_network(n, edges, ww, h, 1.0, ctx=ctx, **p, **kw)
        la, ea, _ = catn.contraction(alone)
        assert lz[k] == la and err[k] == ea, k
        assert batch[k].edge_log() == alone.edge_log()
        assert batch[k].maxdim_intermediate == alone.maxdim_intermediate
        logs.append((tuple(alone.edge_log()), alone.maxdim_intermediate))
    assert logs[0] != logs[1]   # the two instances' plans did part
