"""Lane-parallel execution of the engine on the CPU (test-only emulator
tests/emul/catn_emul_mt.cpp: 256 host threads per workgroup, one per lane,
checked barriers, cross-lane reductions in a fixed order).

The device code must give the same results whatever the lanes' relative
progress: every barrier reached from one call site by all lanes (a divergent
barrier desynchronises the workgroup on the GPU), no lane reading data that
another lane rewrites without a barrier in between (lagging wavefronts). The
forward lnZ and the reverse-sweep gradient of truncated networks therefore
match the one-thread build to rounding, and no divergent barrier is seen.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import configs as Cf
from oracle import exact as X

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MT_LIB = os.path.join(ROOT, "tests", "emul", "libcatn_emul_mt.so")
ST_LIB = os.path.join(ROOT, "tests", "emul", "libcatn_emul.so")


@pytest.fixture(scope="module")
def ctxs(catn):
    if not (os.path.exists(MT_LIB) and os.path.exists(ST_LIB)):
        pytest.skip("emulators not built (run __graft_entry__.build())")
    return catn.Context(lib=MT_LIB), catn.Context(lib=ST_LIB)


def _divergent():
    lib = ctypes.CDLL(MT_LIB)
    lib.catn_emu_divergent_barriers.restype = ctypes.c_long
    return lib.catn_emu_divergent_barriers()


def _run(catn, ctx, n, edges, w, h, beta, **kw):
    tn = catn.ising_network(n, edges, w, h, beta, ctx=ctx, **kw)
    lz, err, _ = catn.gradient_batch([tn])
    dw, dh, _ = catn.ising_gradient(tn, len(edges), n)
    return lz[0], err[0], dw, dh, tn.edge_log()


@pytest.mark.parametrize("case", ["sg16_D6_chi12", "lattice4_Dmax3", "sg12_select2_norm2",
                                  "sg16_f32"])
def test_lane_parallel_matches_one_thread(catn, ctxs, case):
    mt, st = ctxs
    if case == "sg16_D6_chi12":
        n, edges, w, h = Cf.spin_glass_3reg(16, seed=4)
        beta, kw = 1.0, dict(Dmax=6, chi=12)
    elif case == "lattice4_Dmax3":
        edges = X.square_edges(4)
        n = 16
        r = np.random.default_rng(3)
        w, h = r.uniform(-1, 1, len(edges)), r.normal(0, 0.2, n)
        beta, kw = 0.9, dict(Dmax=3, chi=32)
    elif case == "sg16_f32":
        n, edges, w, h = Cf.spin_glass_3reg(16, seed=4)
        beta, kw = 1.0, dict(Dmax=6, chi=12, dtype="f32x6")
    else:
        n, edges, w, h = Cf.spin_glass_3reg(12, seed=2)
        beta, kw = 1.0, dict(Dmax=4, chi=8, select=2, norm_method=2)
    before = _divergent()
    a = _run(catn, mt, n, edges, w, h, beta, **kw)
    b = _run(catn, st, n, edges, w, h, beta, **kw)
    assert _divergent() == before
    assert a[4] == b[4]                              # same edge sequence
    tol = 1e-7 if kw.get("dtype") else 1e-13   # FP32 storage: rounding of the stored sites
    assert abs(a[0] - b[0]) / n < tol
    assert abs(a[1] - b[1]) <= 1e-9 * max(1.0, abs(b[1]))
    scale = max(1.0, np.abs(b[2]).max(), np.abs(b[3]).max())
    gtol = 1e-4 if kw.get("dtype") else 1e-9
    assert np.abs(a[2] - b[2]).max() < gtol * scale
    assert np.abs(a[3] - b[3]).max() < gtol * scale


@pytest.mark.parametrize("tiles", [(32, 8), (8, 64)])
def test_tile_parameters_lane_parallel(catn, ctxs, tiles):
    """Householder panel width and global-Jacobi lanes per pair at 64 lanes
    (one wavefront, which keeps the lane emulator fast), with an
    LDS budget of 2048 doubles so that the blocked Householder and the
    global-memory Jacobi run: same lnZ (1e-12 per spin) and edge sequence as
    the one-thread build, no divergent barrier."""
    mt, st = ctxs
    L = 11
    edges = X.square_edges(L)
    n = L * L
    r = np.random.default_rng(L)
    w, h = r.choice([-1.0, 1.0], len(edges)), np.zeros(n)
    kw = dict(Dmax=16, chi=64)   # SVDs up to min(m, n) = 32, 12 Jacobi sweeps
    qb, jl = tiles
    res = []
    for ctx in (mt, st):
        ctx.set_launch(64, 2048, qb, jl)
        try:
            tn = catn.ising_network(n, edges, w, h, 0.9, ctx=ctx, **kw)
            lz, err, _ = catn.contraction_batch([tn])
            res.append((lz[0], err[0], tn.edge_log()))
        finally:
            ctx.set_launch(256, 0, 0, 0)
    before = _divergent()
    assert res[0][2] == res[1][2]
    assert abs(res[0][0] - res[1][0]) / n < 1e-12
    assert abs(res[0][1] - res[1][1]) <= 1e-9 * max(1.0, abs(res[1][1]))
    assert _divergent() == before


def test_braket_lane_parallel(catn, ctxs):
    """The double-layer engine (construction splits, descriptor reshape,
    paired-edge eat!, compress) at 256 lanes: same lnZ, error and edge
    sequence as the one-thread build on a truncated random tree."""
    from test_braket import _tree
    mt, st = ctxs
    ts, ixs = _tree(np.random.default_rng(4), 9, 2, 2)
    before = _divergent()
    res = []
    for ctx in (mt, st):
        tn = catn.braket_network(ts, ixs, ctx=ctx, chi=3)
        lz, err, psi = catn.contraction(tn)
        res.append((lz, err, tn.edge_log()))
    assert _divergent() == before
    assert res[0][2] == res[1][2]
    assert abs(res[0][0] - res[1][0]) <= 1e-12 * abs(res[1][0])
    assert abs(res[0][1] - res[1][1]) <= 1e-9 * abs(res[1][1])


def test_column_block_jacobi_lane_parallel(catn, ctxs):
    """The column-block Jacobi of a large SVD (LDS budget 5056 doubles, R
    factor 100 x 100: six blocks of 18 columns staged in LDS) at 256 lanes:
    the one-thread build's singular values and reconstruction to rounding, no
    divergent barrier."""
    mt, st = ctxs
    A = np.random.default_rng(9).standard_normal((160, 100))
    before = _divergent()
    res = []
    for ctx in (mt, st):
        ctx.set_launch(256, 5056, 0, 0)
        try:
            res.append(catn.tsvd(A, cutoff=0.0, ctx=ctx))
        finally:
            ctx.set_launch(256, 0, 0, 0)
    assert _divergent() == before
    (U0, S0, V0, _, _), (U1, S1, V1, _, _) = res
    assert np.allclose(S0, S1, rtol=1e-13, atol=1e-13 * S1[0])
    assert np.allclose(U0 @ np.diag(S0) @ V0.T, A, atol=1e-11 * S1[0])
    assert np.allclose(U0.T @ U0, np.eye(100), atol=1e-12)


def test_loopy_braket_lane_parallel(catn, ctxs):
    """The loopy double layer (doubled-bond merges on a 2x3 lattice, truncated
    at chi = 3) at 256 lanes: the one-thread build's lnZ, error and edge
    sequence, no divergent barrier."""
    from test_braket import _loopy
    mt, st = ctxs
    ts, ixs = _loopy("lattice2x3_D2")
    before = _divergent()
    res = []
    for ctx in (mt, st):
        tn = catn.braket_network(ts, ixs, ctx=ctx, chi=3, loops=True)
        lz, err, psi = catn.contraction(tn)
        res.append((lz, err, tn.edge_log()))
    assert _divergent() == before
    assert res[0][2] == res[1][2]
    assert abs(res[0][0] - res[1][0]) <= 1e-12 * abs(res[1][0])
    assert abs(res[0][1] - res[1][1]) <= 1e-9 * max(abs(res[1][1]), 1e-12)


# ----------------------------------------------------------------------------
# Device code paths on the CPU (tests/emul/libcatn_emul_dev.so): the same lane
# emulator compiled with CATN_EMU_DEV, so engine.h takes its device branches
# (FP64 and BF16 matrix-core products, wavefront shuffles, one-barrier
# reductions, reflectors read in place) with each wavefront instruction
# emulated as a rendezvous of the lanes it reads. The device's own operand and
# accumulator index arithmetic then runs on host threads, where a build under
# AddressSanitizer or ThreadSanitizer checks it.

DEV_LIB = os.path.join(ROOT, "tests", "emul", "libcatn_emul_dev.so")


@pytest.fixture(scope="module")
def devctx(catn):
    if not (os.path.exists(DEV_LIB) and os.path.exists(ST_LIB)):
        pytest.skip("emulators not built (run __graft_entry__.build())")
    return catn.Context(lib=DEV_LIB), catn.Context(lib=ST_LIB)


def _wave_ops():
    lib = ctypes.CDLL(DEV_LIB)
    out = (ctypes.c_long * 3)()
    lib.catn_emu_wave_ops(out)
    return list(out)          # FP64 matrix-core, BF16 matrix-core, shuffles (lane 0)


def _dev_divergent():
    lib = ctypes.CDLL(DEV_LIB)
    lib.catn_emu_divergent_barriers.restype = ctypes.c_long
    return lib.catn_emu_divergent_barriers()


@pytest.mark.parametrize("case", ["sg40_D8_chi16", "sg16_f32x6"])
def test_device_paths_cut_gradient(catn, devctx, case):
    """Gradient runs across Dmax cuts through the device code paths at 256
    lanes. sg40_D8_chi16 is the workload of the cut-gradient case that faulted
    the GPU: it matches the one-thread build and executes no matrix-core
    product (every product there is below the tiled-GEMM threshold), only
    wavefront shuffles. sg16_f32x6 stores FP32 and runs its swap thetas as
    6-pass split-BF16 matrix-core products."""
    dev, st = devctx
    if case == "sg40_D8_chi16":
        n, edges, w, h = Cf.spin_glass_3reg(40, seed=1)
        kw = dict(Dmax=8, chi=16)
    else:
        n, edges, w, h = Cf.spin_glass_3reg(16, seed=4)
        kw = dict(Dmax=6, chi=12, dtype="f32x6")
    _wave_ops()
    before = _dev_divergent()
    a = _run(catn, dev, n, edges, w, h, 1.0, **kw)
    ops = _wave_ops()
    b = _run(catn, st, n, edges, w, h, 1.0, **kw)
    assert _dev_divergent() == before
    assert a[4] == b[4]
    f32 = "dtype" in kw
    assert abs(a[0] - b[0]) / n < (1e-7 if f32 else 1e-13)
    scale = max(1.0, np.abs(b[2]).max(), np.abs(b[3]).max())
    gtol = 1e-4 if f32 else 1e-9
    assert np.abs(a[2] - b[2]).max() < gtol * scale
    assert np.abs(a[3] - b[3]).max() < gtol * scale
    assert ops[2] > 0                       # the reductions ran as shuffles
    if f32:
        assert ops[1] > 0                   # split-BF16 theta products
    else:
        assert ops[0] == 0 and ops[1] == 0  # no matrix-core product in this workload


def test_device_paths_matrix_cores(catn, devctx):
    """A 160 x 100 SVD at an LDS budget of 5056 doubles: blocked Householder
    with the compact-WY products, the T factors and the inverse bound on the
    FP64 matrix cores (emulated), equal to the one-thread build's tiled FMA
    products to rounding."""
    dev, st = devctx
    A = np.random.default_rng(9).standard_normal((160, 100))
    res = []
    _wave_ops()
    for ctx in (dev, st):
        ctx.set_launch(256, 5056, 0, 0)
        try:
            res.append(catn.tsvd(A, cutoff=0.0, ctx=ctx))
        finally:
            ctx.set_launch(256, 0, 0, 0)
    ops = _wave_ops()
    assert ops[0] > 1000
    (U0, S0, V0, _, _), (U1, S1, V1, _, _) = res
    assert np.allclose(S0, S1, rtol=1e-13, atol=1e-13 * S1[0])
    assert np.allclose(U0 @ np.diag(S0) @ V0.T, A, atol=1e-11 * S1[0])
    assert np.allclose(U0.T @ U0, np.eye(100), atol=1e-12)


def test_wide_workgroups_lane_parallel(catn, ctxs, devctx):
    """1024 lanes per workgroup (the wide kernel: 16 wavefronts, 16 reduction
    slots per half of the one-barrier sums): a truncated forward and reverse
    sweep on the lane emulator equal to the one-thread build with no
    divergent barrier, and a 64 x 36 SVD at the minimum LDS budget through the
    device paths (emulated shuffles and matrix cores; the Householder T
    factors split their rows over partial tiles summed in LDS) equal to it to
    rounding."""
    mt, st = ctxs
    dev, _ = devctx
    lanes = 1024
    n, edges, w, h = Cf.spin_glass_3reg(12, seed=2)
    kw = dict(Dmax=4, chi=8)
    b = _run(catn, st, n, edges, w, h, 1.0, **kw)
    A = np.random.default_rng(9).standard_normal((64, 36))
    S_ref = catn.tsvd(A, cutoff=0.0, ctx=st)[1]
    before, dbefore = _divergent(), _dev_divergent()
    mt.set_launch(lanes, 0, 0, 0)
    try:
        a = _run(catn, mt, n, edges, w, h, 1.0, **kw)
    finally:
        mt.set_launch(256, 0, 0, 0)
    assert a[4] == b[4]
    assert abs(a[0] - b[0]) / n < 1e-13
    scale = max(1.0, np.abs(b[2]).max(), np.abs(b[3]).max())
    assert np.abs(a[2] - b[2]).max() < 1e-9 * scale
    assert np.abs(a[3] - b[3]).max() < 1e-9 * scale
    _wave_ops()
    dev.set_launch(lanes, 2048, 0, 0)
    try:
        U, S, V, _, _ = catn.tsvd(A, cutoff=0.0, ctx=dev)
    finally:
        dev.set_launch(256, 0, 0, 0)
    assert _wave_ops()[0] > 0                 # FP64 matrix-core products ran
    assert np.allclose(S, S_ref, rtol=1e-13, atol=1e-13 * S_ref[0])
    assert np.allclose(U @ np.diag(S) @ V.T, A, atol=1e-11 * S_ref[0])
    assert _divergent() == before and _dev_divergent() == dbefore
