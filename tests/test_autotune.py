"""Launch shape of the interpreter (lanes and LDS budget per workgroup,
Householder panel width, global-Jacobi lanes per column pair) and the
autotuner that picks it by timing a batch.

The shape may change which ops run LDS-resident and the order of sums, never
the algorithm: every shape gives the same lnZ to 1e-10 relative, and the same
edge sequence and Dmax-cut kept counts.
"""
import numpy as np
import pytest

from oracle import configs as Cf


@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx")
    return catn, request.getfixturevalue("gpu_ctx")


def _pinned(catn, ctx, n, edges, w, h, **kw):
    nets = [catn.ising_network(n, edges, w, h, 1.0, ctx=ctx, **kw)]
    for k in range(1, n + 1):
        for v in (1, 2):
            nets.append(catn.ising_network(n, edges, w, h, 1.0, pos1=k, val1=v, ctx=ctx, **kw))
    return nets


def test_launch_shape_round_trip_and_limits(env):
    catn, ctx = env
    t0, l0, lim, q0, j0 = ctx.launch()
    assert lim >= 10176          # the automatic budget of small batches fits
    assert (q0, j0) == (16, 16)  # default tiles
    try:
        ctx.set_launch(128, 6784, 32, 8)
        assert ctx.launch() == (128, 6784, lim, 32, 8)
        for bad in ((96, 0), (768, 0), (2048, 0), (256, 100), (256, lim + 1), (256, 0, 12, 16),
                    (256, 0, 16, 24), (64, 0, 16, 128)):
            with pytest.raises(catn.CatnError):
                ctx.set_launch(*bad)
        assert ctx.launch() == (128, 6784, lim, 32, 8)   # a rejected shape changes nothing
    finally:
        ctx.set_launch(t0, l0, q0, j0)


def test_every_shape_gives_the_same_contraction(env):
    catn, ctx = env
    n, edges, w, h = Cf.spin_glass_3reg(24, seed=5)
    kw = dict(Dmax=6, chi=12)
    t0, l0, lim, q0, j0 = ctx.launch()
    ref = None
    cands = [(t, l, 16, 16) for (t, l) in catn.AUTOTUNE_SHAPES + catn.AUTOTUNE_WIDE if l <= lim]
    cands += [(256, 5056, q, j) for (q, j) in catn.AUTOTUNE_TILES]
    try:
        for threads, lds, qb, jl in cands:
            ctx.set_launch(threads, lds, qb, jl)
            nets = _pinned(catn, ctx, n, edges, w, h, **kw)
            lz, err, _ = catn.contraction_batch(nets)
            seq = nets[0].edge_log()
            if ref is None:
                ref = (lz, err, seq)
                continue
            assert np.allclose(lz, ref[0], rtol=1e-10, atol=0)
            assert np.allclose(err, ref[1], rtol=1e-6, atol=1e-14)
            assert seq == ref[2]
    finally:
        ctx.set_launch(t0, l0, q0, j0)


def test_autotune_picks_a_timed_shape(env):
    catn, ctx = env
    n, edges, w, h = Cf.spin_glass_3reg(16, seed=2)
    kw = dict(Dmax=6, chi=12)
    t0, l0, lim, q0, j0 = ctx.launch()
    try:
        res = catn.autotune(ctx, lambda: _pinned(catn, ctx, n, edges, w, h, **kw))
        key = lambda r: (r["threads"], r["lds_doubles"], r["panel"], r["jacobi_lanes"])
        timed = {key(r) for r in res["table"]}
        assert key(res) in timed
        assert ctx.launch() == (res["threads"], res["lds_doubles"], lim, res["panel"],
                                res["jacobi_lanes"])
        assert all(r["lnz_rel_dev"] <= 1e-10 for r in res["table"])
        # a batch of 129 networks also tries the wide workgroups
        assert {(r["threads"], r["lds_doubles"]) for r in res["table"]} >= set(catn.AUTOTUNE_WIDE)
        assert all(r["lds_doubles"] <= lim for r in res["table"])
        assert {(r["panel"], r["jacobi_lanes"]) for r in res["table"]} >= set(catn.AUTOTUNE_TILES)
        best = min(r["seconds"] for r in res["table"])
        assert res["seconds"] == best
    finally:
        ctx.set_launch(t0, l0, q0, j0)
