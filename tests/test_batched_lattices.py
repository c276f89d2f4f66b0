"""Batched +-J lattices of uneven size and beta, FP32 storage (config 3's shape).

One contraction_batch runs every instance (one workgroup each per launch):
open L x L square lattices, w_ij uniform +-1, h = 0, beta uniform in
[0.2, 1.2], Dmax = 32, chi = 128. Per instance, lnZ/N with FP32 storage
(split-BF16 swap products) is compared with the FP64 form of the same
engine at the same (Dmax, chi), which isolates precision, and the FP64
form with the FP64 oracle on the smaller lattices. Tolerances: f32x6
(6-pass) 1e-6 per spin, f32 (3-pass) 5e-6, FP64 1e-12.
"""
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
