"""Config 4's dense-regime network at a size the one-workgroup ops reach: an
open 4 x 4 square-lattice tensor network with bond dimension 4 on every edge,
entries i.i.d. uniform[0, 1) (positive variant) and N(0, 1) (signed variant),
Dmax = 256, chi = 1024, in FP64 and in FP32 storage with split-BF16 swap
products (3 and 6 passes), against the exact contraction (numpy einsum of the
16 tensors). Tolerances on lnZ: FP64 1e-12, six-pass 2e-6, three-pass 1e-4
(its 2^-16 products meet the sign cancellations of the signed variant: 2.9e-5
measured). At 5 x 5 the swaps' thetas pass a few thousand rows and columns
and one network no longer finishes in minutes on one workgroup, so config 4's
12 x 12 size is not run.
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx")
    return catn, request.getfixturevalue("gpu_ctx")


def _grid(L, D, rng, signed):
    ts, ixs = [], []
    for r in range(L):
        for c in range(L):
            ix = []
            if r > 0:
                ix.append(f"v{r - 1}_{c}")
            if r < L - 1:
                ix.append(f"v{r}_{c}")
            if c > 0:
                ix.append(f"h{r}_{c - 1}")
            if c < L - 1:
                ix.append(f"h{r}_{c}")
            shape = [D] * len(ix)
            ts.append(rng.standard_normal(shape) if signed else rng.uniform(0.0, 1.0, shape))
            ixs.append(ix)
    return ts, ixs


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
    ts, ixs = _grid(4, 4, np.random.default_rng(4 + signed), signed)
    ref = _exact(ts, ixs)
    tn = catn.tensor_network(ts, ixs, ctx=ctx, Dmax=256, chi=1024, dtype=dtype)
    lnz, err, psi = catn.contraction(tn)
    assert abs(lnz - math.log(abs(ref))) <= tol
    assert psi == math.copysign(1.0, ref)
