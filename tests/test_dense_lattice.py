"""Config 4's dense-regime network: an open L x L square-lattice tensor
network with bond dimension 4 on every edge, entries i.i.d. uniform[0, 1)
(positive variant) and N(0, 1) (signed variant), Dmax = 256, chi = 1024, in
FP64 and in FP32 storage with split-BF16 swap products (3 and 6 passes),
against the exact contraction: numpy einsum of the tensors at 4 x 4, and a
row-by-row transfer contraction (exact, 4^(L+1) entries) that this file pins
against einsum at 3 x 3 and 4 x 4. Tolerances on lnZ at 4 x 4: FP64 1e-12,
six-pass 2e-6, three-pass 1e-4 (its 2^-16 products meet the sign
cancellations of the signed variant: 2.9e-5 measured). Measured on the
MI355X (one 1024-lane workgroup): 5 x 5 in 0.03 s, 6 x 6 in 0.1 s, 8 x 8 in
7.9 s (mostly one split of min(m, n) > 256). From 10 x 10 a swap
reaches swap!'s randomized-SVD switch with a sketch of min(n, 10 chi) = 10240
columns wider than theta's rows, which the device's range finder refuses (it
needs p <= m), so config 4's 12 x 12 size is not run.
"""
import math

import numpy as np
import pytest

from oracle import configs as Cf
from oracle import exact as X


@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx")
    return catn, request.getfixturevalue("gpu_ctx")


def _grid(L, D, seed, signed):
    return Cf.dense_lattice(L, D, seed=seed, signed=signed)


def _row_exact(ts, L):
    lz, ph = X.grid_lnZ(ts, L)
    return lz, math.copysign(1.0, ph.real)


@pytest.mark.parametrize("L", [3, 4])
@pytest.mark.parametrize("signed", [False, True])
def test_row_transfer_is_the_einsum_contraction(L, signed):
    """The oracle's row transfer (oracle/exact.py grid_lnZ) against the dense
    einsum at 3 x 3 and 4 x 4, real and complex entries."""
    ts, ixs = _grid(L, 4, L + signed, signed)
    ref = _exact(ts, ixs)
    lz, sg = _row_exact(ts, L)
    assert abs(lz - math.log(abs(ref))) <= 1e-12 * abs(lz) and sg == math.copysign(1.0, ref)
    zs = [t * np.exp(0.3j * t) for t in ts]
    ids = {}
    args = []
    for t, ix in zip(zs, ixs):
        args += [t, [ids.setdefault(l, len(ids)) for l in ix]]
    zref = complex(np.einsum(*args, [], optimize="greedy"))
    lzc, ph = X.grid_lnZ(zs, L)
    assert abs(lzc - math.log(abs(zref))) <= 1e-12 * abs(lzc) and abs(ph - zref / abs(zref)) < 1e-12


def _exact(ts, ixs):
    ids = {}
    args = []
    for t, ix in zip(ts, ixs):
        args += [t, [ids.setdefault(l, len(ids)) for l in ix]]
    return float(np.einsum(*args, [], optimize="greedy"))


@pytest.mark.parametrize("signed", [False, True])
@pytest.mark.parametrize("dtype,tol", [("f64", 1e-12), ("f32x6", 2e-6), ("f32", 1e-4)])
def test_dense_lattice_4x4(env, signed, dtype, tol):
    catn, ctx = env
    ts, ixs = _grid(4, 4, 4 + signed, signed)
    ref = _exact(ts, ixs)
    tn = catn.tensor_network(ts, ixs, ctx=ctx, Dmax=256, chi=1024, dtype=dtype)
    lnz, err, psi = catn.contraction(tn)
    assert abs(lnz - math.log(abs(ref))) <= tol
    assert psi == math.copysign(1.0, ref)


def _run_grid(catn, ctx, L, dtype, signed, lanes=0):
    ts, ixs = _grid(L, 4, L + signed, signed)
    saved = ctx.launch()
    if lanes:
        ctx.set_launch(lanes, 20352, 16, 16)
    try:
        tn = catn.tensor_network(ts, ixs, ctx=ctx, Dmax=256, chi=1024, dtype=dtype,
                                 arena_elems=1 << 26, scratch_elems=1 << 26)
        lnz, err, psi = catn.contraction(tn)
    finally:
        ctx.set_launch(saved[0], saved[1], saved[3], saved[4])
    return (lnz, psi), _row_exact(ts, L)


@pytest.mark.parametrize("signed", [False, True])
@pytest.mark.parametrize("dtype,tol", [("f64", 1e-10), ("f32x6", 2e-6)])
def test_dense_lattice_6x6(env, dtype, tol, signed):
    """6 x 6 against the exact row-transfer contraction (absolute lnZ
    tolerance; lnZ is 40..60). Measured with six passes: 1.5e-7 (positive),
    2e-7 (signed)."""
    catn, ctx = env
    (lnz, psi), (ref, sg) = _run_grid(catn, ctx, 6, dtype, signed)
    assert abs(lnz - ref) <= tol and psi == sg


@pytest.mark.gpu
@pytest.mark.parametrize("signed", [False, True])
def test_dense_lattice_8x8_gpu(catn, gpu_ctx, signed):
    """8 x 8, six-pass FP32 storage, one 1024-lane workgroup (7.9 s, mostly one
    split of min(m, n) > 256), against the exact row-transfer
    contraction (lnZ 78..111). Measured: 4e-7 (positive), 1.2e-6 (signed)."""
    (lnz, psi), (ref, sg) = _run_grid(catn, gpu_ctx, 8, "f32x6", signed, lanes=1024)
    assert abs(lnz - ref) <= 5e-6 and psi == sg
