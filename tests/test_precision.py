"""FP32 storage (the reference's Float32 networks, cu(tn); src/adapt.jl:6-24)
with the swap theta GEMM as split-BF16 MFMA products accumulated in FP32.

Pins:
  * lnZ/N against the FP64 oracle at the same (Dmax, chi): the 6-pass split
    (three BF16 terms per operand) within 1e-6 on spin glasses and +-J
    lattices; the 3-pass split (hi/lo) within 5e-6, the accuracy its 17-bit
    operands allow;
  * one swap! on a hand-built node: the swapped two-site tensor equals the
    FP64 swap to the split's product accuracy (2^-16 relative for 3 passes,
    FP32 rounding for 6), which also checks the matrix-core lane layout on the
    device;
  * the reverse sweep on FP32 storage: d lnZ / d(w, h) of exact contractions
    agree with the FP64 sweep to 1e-5 relative;
  * a tensor read back from FP32 storage is the FP32 rounding of the input.
"""
import numpy as np
import pytest

from oracle import catn_oracle as O
from oracle import configs as Cf
from oracle import exact as X


@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx")
    return catn, request.getfixturevalue("gpu_ctx")


def _lattice(L, seed, beta):
    e = X.square_edges(L)
    r = np.random.default_rng(seed)
    return L * L, e, r.choice([-1.0, 1.0], len(e)), np.zeros(L * L), beta


CASES = {
    "sg40_D8_chi16": (*Cf.spin_glass_3reg(40, seed=1), 1.0, dict(Dmax=8, chi=16)),
    "lattice8_b0.8": (*_lattice(8, 5, 0.8), dict(Dmax=32, chi=128)),
    "lattice10_b1.2": (*_lattice(10, 6, 1.2), dict(Dmax=32, chi=128)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_lnz_against_fp64_oracle(env, case):
    catn, ctx = env
    n, e, w, h, beta, kw = CASES[case]
    ref = O.contraction(O.ising_network(n, e, w, h, beta, **kw))[0]
    l6 = catn.contraction(catn.ising_network(n, e, w, h, beta, ctx=ctx, dtype="f32x6", **kw))[0]
    l3 = catn.contraction(catn.ising_network(n, e, w, h, beta, ctx=ctx, dtype="f32", **kw))[0]
    assert abs(l6 - ref) / n <= 1e-6, (l6, ref)
    assert abs(l3 - ref) / n <= 5e-6, (l3, ref)


@pytest.mark.parametrize("dtype,tol", [("f32", 3e-5), ("f32x6", 1e-6)])
def test_split_bf16_swap(env, dtype, tol):
    catn, ctx = env
    r = np.random.default_rng(4)
    tl = r.standard_normal((6, 5, 20))     # dd = 20: two K blocks of the 16-wide MFMA
    tr = r.standard_normal((20, 7, 9))
    out = {}
    for dt in ("f64", dtype):
        net = catn.MPSNet(ctx=ctx, chi=10_000, Dmax=-1, dtype=dt)
        net.add_node(1, [tl, tr], [2, 3], cano=1)
        net.swap(1, 1, 2)
        a, b = net.sites(1)
        out[dt] = np.einsum("iak,kjb->iajb", a, b)
    ref = np.einsum("ijk,kab->iajb", tl, tr)
    assert np.abs(out["f64"] - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(out[dtype] - ref).max() < tol * np.abs(ref).max()


def test_reverse_sweep_on_fp32_storage(env):
    catn, ctx = env
    n, e, w, h = Cf.spin_glass_3reg(12, seed=3)
    kw = dict(Dmax=-1, chi=10_000)
    l64, w64, h64 = catn.lnz_and_grad(n, e, w, h, 0.9, ctx=ctx, **kw)
    l32, w32, h32 = catn.lnz_and_grad(n, e, w, h, 0.9, ctx=ctx, dtype="f32x6", **kw)
    assert abs(l32 - l64) / n < 1e-6
    assert np.abs(w32 - w64).max() < 1e-5 * max(1.0, np.abs(w64).max())
    assert np.abs(h32 - h64).max() < 1e-5 * max(1.0, np.abs(h64).max())


def test_storage_rounds_to_fp32(env):
    catn, ctx = env
    r = np.random.default_rng(7)
    t = r.standard_normal((3, 4, 5))
    net = catn.MPSNet(ctx=ctx, dtype="f32")
    net.add_node(1, [t], [2], cano=1)
    (back,) = net.sites(1)
    assert np.array_equal(back, t.astype(np.float32).astype(np.float64))
