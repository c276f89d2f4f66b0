"""Open-leg (partial) contraction and result_tensor (src/tensor_network.jl:504-523),
after test/test_openleg.jl.

The pin is the dense contraction of the same tensors (numpy einsum over the
closed labels, the semantics of test/exact.jl's exact_contract), and the
device's leg order is checked through the labels it returns, entry by entry
rather than up to a permutation as the reference's test does. Runs on the CPU
emulator and, marked gpu, on the MI355X.
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx")
    return catn, request.getfixturevalue("gpu_ctx")


def _dense(ts, ixs, out):
    """sum over the closed labels, result indexed by `out` (a list of labels)"""
    letters = {}
    for ix in ixs:
        for l in ix:
            if l not in letters:
                letters[l] = chr(ord("a") + len(letters))
    spec = ",".join("".join(letters[l] for l in ix) for ix in ixs)
    return np.einsum(spec + "->" + "".join(letters[l] for l in out), *ts)


def _open_value(catn, ctx, ts, ixs, **kw):
    tn = catn.tensor_network(ts, ixs, ctx=ctx, **kw)
    lnz, err, psi = catn.contraction(tn)
    arr, labels = catn.result_tensor(tn)
    return arr * (math.exp(lnz) * psi), labels, err


def _tree(rng, n, D, nopen):
    """random tree on n nodes (node k > 1 hangs off an earlier node), with
    nopen open legs of dimension 2 spread over random nodes"""
    labels = {k: [] for k in range(n)}
    for k in range(1, n):
        par = int(rng.integers(0, k))
        labels[par].append(f"e{par}_{k}")
        labels[k].append(f"e{par}_{k}")
    for q in range(nopen):
        k = int(rng.integers(0, n))
        ix = labels[k]
        ix.insert(int(rng.integers(0, len(ix) + 1)), f"o{q}")
    ts, ixs = [], []
    for k in range(n):
        ix = labels[k] if labels[k] else [f"s{k}"]
        ixs.append(ix)
        ts.append(rng.standard_normal([2 if l[0] in "os" else D for l in ix]))
    return ts, ixs


def test_chain_with_open_physical_legs(env):
    """test/test_openleg.jl:14-26 (Float64): T1(p1,a) T2(a,p2,b) T3(b,p3)."""
    catn, ctx = env
    rng = np.random.default_rng(11)
    ts = [rng.standard_normal((2, 4)), rng.standard_normal((4, 2, 4)), rng.standard_normal((4, 2))]
    ixs = [["p1", "a"], ["a", "p2", "b"], ["b", "p3"]]
    got, labels, err = _open_value(catn, ctx, ts, ixs, Dmax=-1, chi=10_000)
    assert sorted(labels) == ["p1", "p2", "p3"]
    ref = _dense(ts, ixs, labels)
    assert got.shape == ref.shape
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    assert err < 1e-12


@pytest.mark.parametrize("n,D,nopen,seed", [(5, 3, 3, 1), (7, 2, 4, 2), (6, 3, 2, 3)])
def test_tree_with_open_legs(env, n, D, nopen, seed):
    """Open legs spread over a random tree: the surviving node carries them all
    (in its MPS order), and the result equals the dense contraction."""
    catn, ctx = env
    ts, ixs = _tree(np.random.default_rng(seed), n, D, nopen)
    got, labels, _ = _open_value(catn, ctx, ts, ixs, Dmax=-1, chi=10_000)
    open_labels = [l for ix in ixs for l in ix if l[0] in "os"]
    assert sorted(labels) == sorted(open_labels)
    ref = _dense(ts, ixs, labels)
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_closed_network_gives_scalar_one(env):
    """test/test_openleg.jl:27-36."""
    catn, ctx = env
    rng = np.random.default_rng(5)
    A, B, Cm = rng.standard_normal((2, 3)), rng.standard_normal((3, 4)), rng.standard_normal((4, 2))
    ixs = [["a", "b"], ["b", "c"], ["c", "a"]]
    ref = float(_dense([A, B, Cm], ixs, []))
    tn = catn.tensor_network([A, B, Cm], ixs, ctx=ctx, Dmax=-1, chi=10_000)
    lnz, err, psi = catn.contraction(tn)
    assert abs(math.exp(lnz) * psi - ref) <= 1e-10 * abs(ref)
    arr, labels = catn.result_tensor(tn)
    assert arr.ndim == 0 and float(arr) == 1.0 and labels == []


def test_two_open_components_are_rejected(env):
    """src/tensor_network.jl:521: more than one surviving open-leg node."""
    catn, ctx = env
    rng = np.random.default_rng(6)
    ts = [rng.standard_normal((2, 3)), rng.standard_normal((3, 2)),
          rng.standard_normal((2, 3)), rng.standard_normal((3, 2))]
    ixs = [["p1", "a"], ["a", "p2"], ["p3", "b"], ["b", "p4"]]
    tn = catn.tensor_network(ts, ixs, ctx=ctx, Dmax=-1, chi=10_000)
    catn.contraction(tn)
    with pytest.raises(catn.CatnError):
        catn.result_tensor(tn)


def test_fp32_storage_result_tensor(env):
    """The readout of an FP32-storage network (FP32 arena, six-pass BF16
    theta): the dense contraction to FP32 accuracy."""
    catn, ctx = env
    rng = np.random.default_rng(12)
    ts = [rng.standard_normal((2, 3)), rng.standard_normal((3, 2, 3)), rng.standard_normal((3, 2))]
    ixs = [["p1", "a"], ["a", "p2", "b"], ["b", "p3"]]
    got, labels, _ = _open_value(catn, ctx, ts, ixs, Dmax=-1, chi=10_000, dtype="f32x6")
    ref = _dense(ts, ixs, labels)
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
