"""The lifted gradient: the exact d lnZ at exactly rank-deficient splits.

The reverse sweep differentiates the executed schedule (every split's kept
count held fixed). A frustrated cycle of +-J couplings can make a split's
theta exactly rank-deficient at w; the absolute cutoff drops that zero
singular value, a perturbation of w restores it, and the exact lnZ then has a
derivative that the schedule at w does not carry. lnz_and_grad runs each
instance at w +- eta for a fixed generic eta (lift = 1e-6 of the couplings'
scale) and averages, which is the exact derivative to O(|eta|^2).

Pins:
  * the direction generator is splitmix64: its first three outputs from state
    0 are the published values of Vigna's generator;
  * combine_lifted is exact on a quadratic, with a planted sign flip failing;
  * exact contraction of +-J spin glasses that have such splits: the lifted
    gradient equals brute-force enumeration (beta <s_i s_j>, beta <s_i>) to
    1e-9, where the executed-schedule gradient (lift = 0) is off by more than
    1e-3; at sizes past enumeration it equals the pinning route (the
    reference's own observables, src/ising.jl:263-339) to 1e-8.
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


EXACT = dict(Dmax=-1, chi=256)


def test_direction_is_splitmix64(catn):
    # splitmix64 from state 0: 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
    # 0x06C45D188009454F (the generator's published first outputs)
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    gw, gh = catn.lift_direction(2, 1, seed=0)
    got = np.concatenate([gw, gh])
    assert np.array_equal(got, np.array([(z >> 11) * 2.0 ** -52 - 1.0 for z in want]))
    gw, gh = catn.lift_direction(300, 200)
    g = np.concatenate([gw, gh])
    assert np.all(np.abs(g) <= 1.0) and np.abs(g).min() > 0 and len(np.unique(g)) == 500


def _quadratic(x, a, b, H):
    return a + b @ x + 0.5 * x @ H @ x, b + H @ x


@pytest.mark.parametrize("flip", [False, True], ids=["combine", "planted_sign_flip"])
def test_combine_is_exact_on_a_quadratic(catn, flip):
    r = np.random.default_rng(3)
    nw, nh = 5, 4
    a, b = 1.7, r.standard_normal(nw + nh)
    H = r.standard_normal((nw + nh, nw + nh))
    H = H + H.T
    ew, eh = 1e-3 * r.standard_normal(nw), 1e-3 * r.standard_normal(nh)
    eta = np.concatenate([ew, eh])
    fp, gp = _quadratic(eta, a, b, H)
    fm, gm = _quadratic(-eta, a, b, H)
    if flip:
        # a mistake planted in the combination: the curvature term added, not removed
        lz = 0.5 * (fp + fm) + 0.25 * float(eta @ (gp - gm))
    else:
        lz, dw, dh = catn.combine_lifted((fp, gp[:nw], gp[nw:]), (fm, gm[:nw], gm[nw:]), (ew, eh))
        assert np.allclose(np.concatenate([dw, dh]), b, rtol=0, atol=1e-14)
    assert (abs(lz - a) < 1e-14) != flip


def _rank_deficient_instances():
    # +-J random 3-regular spin glasses whose exact contraction at chi 256 has an
    # exactly rank-deficient split that a coupling perturbation lifts
    return [(16, 3), (16, 12), (20, 6)]


@pytest.mark.parametrize("n,seed", _rank_deficient_instances(),
                         ids=[f"sg{n}_s{s}" for n, s in _rank_deficient_instances()])
def test_lifted_gradient_is_exact_at_rank_deficient_splits(env, n, seed):
    catn, ctx = env
    n, edges, w, h = Cf.spin_glass_3reg(n, seed=seed)
    beta = 1.0
    mag, cor = X.brute_observables(n, edges, w, h, beta)
    lz_exact = X.brute_lnZ(n, edges, w, h, beta)
    l0, dw0, dh0 = catn.lnz_and_grad(n, edges, w, h, beta, ctx=ctx, lift=0, **EXACT)
    lz, dw, dh = catn.lnz_and_grad(n, edges, w, h, beta, ctx=ctx, **EXACT)
    # the executed schedule's derivative misses the exact one on the cycle's couplings
    assert np.abs(dw0 - beta * cor).max() > 1e-3
    assert np.abs(dw - beta * cor).max() < 1e-9
    assert np.abs(dh - beta * mag).max() < 1e-9
    assert abs(lz - lz_exact) < 1e-11 * max(1.0, abs(lz_exact))
    assert abs(l0 - lz_exact) < 1e-11 * max(1.0, abs(lz_exact))


@pytest.mark.parametrize("dtype,tol", [("f32x6", 2e-6), ("f32", 5e-5)])
@pytest.mark.parametrize("n,seed", [(12, 3), (16, 3), (20, 6)], ids=["sg12_s3", "sg16_s3", "sg20_s6"])
def test_lifted_gradient_on_fp32_storage(env, n, seed, dtype, tol):
    """FP32 storage: the lifted singular value must clear the storage's noise
    floor, so the lift is larger (2e-3 six-pass, 1e-2 three-pass) and a
    Richardson step over eta and 2 eta cancels its O(eta^2) bias. Measured
    errors: 2e-7..4e-7 six-pass, 8e-6..1.2e-5 three-pass, where lift 0 is off
    by 1.3e-2..7.3e-2."""
    catn, ctx = env
    n, edges, w, h = Cf.spin_glass_3reg(n, seed=seed)
    beta = 0.9
    mag, cor = X.brute_observables(n, edges, w, h, beta)
    assert catn.lift_plan(None, dtype) == (catn.LIFT[dtype], 2 * catn.LIFT[dtype])
    lz, dw, dh = catn.lnz_and_grad(n, edges, w, h, beta, ctx=ctx, dtype=dtype,
                                   Dmax=-1, chi=10_000)
    assert np.abs(dw - beta * cor).max() < tol
    assert np.abs(dh - beta * mag).max() < tol
    assert abs(lz - X.brute_lnZ(n, edges, w, h, beta)) / n < 1e-6


def test_lifted_gradient_matches_pinning_past_enumeration(env):
    """n = 40 (seed 3): the executed-schedule gradient is off by 0.18 on one
    coupling; the lifted one equals the pinning route."""
    catn, ctx = env
    n, edges, w, h = Cf.spin_glass_3reg(40, seed=3)
    _, dw0, _ = catn.lnz_and_grad(n, edges, w, h, 1.0, ctx=ctx, lift=0, **EXACT)
    _, dw, dh = catn.lnz_and_grad(n, edges, w, h, 1.0, ctx=ctx, **EXACT)
    cor = catn.correlation(n, edges, w, h, 1.0, ctx=ctx, **EXACT)
    mag = catn.magnetization(n, edges, w, h, 1.0, ctx=ctx, **EXACT)
    assert np.abs(dw0 - cor).max() > 0.1
    assert np.abs(dw - cor).max() < 1e-8
    assert np.abs(dh - mag).max() < 1e-8


@pytest.mark.gpu
def test_lifted_gradient_on_the_cut_free_benchmark_instance(catn, gpu_ctx):
    """The spin glass of the cut-free comparison (n = 90, seed 1, Dmax = -1,
    chi = 256): lifted adjoint against the 721 pinned contractions."""
    n, edges, w, h = Cf.spin_glass_3reg(90, seed=1)
    kw = dict(Dmax=-1, chi=256, arena_elems=3 << 20, scratch_elems=1 << 21)
    _, dw, dh = catn.lnz_and_grad(n, edges, w, h, 1.0, ctx=gpu_ctx, **kw)
    cor = catn.correlation(n, edges, w, h, 1.0, ctx=gpu_ctx, **kw)
    mag = catn.magnetization(n, edges, w, h, 1.0, ctx=gpu_ctx, **kw)
    assert np.abs(dw - cor).max() < 1e-8
    assert np.abs(dh - mag).max() < 1e-8


def test_lift_zero_is_one_run_and_negative_is_refused(catn, emu_ctx):
    n, edges, w, h = Cf.spin_glass_3reg(12, seed=2)
    with pytest.raises(catn.CatnError, match="lift"):
        catn.lnz_and_grad(n, edges, w, h, 1.0, ctx=emu_ctx, lift=-1.0, **EXACT)
    a = catn.lnz_and_grad(n, edges, w, h, 1.0, ctx=emu_ctx, lift=0, **EXACT)
    tn = catn.ising_network(n, edges, w, h, 1.0, ctx=emu_ctx, **EXACT)
    lz, _, _ = catn.gradient_batch([tn])
    dw, dh, _ = catn.ising_gradient(tn, len(edges), n)
    assert a[0] == lz[0] and np.array_equal(a[1], dw) and np.array_equal(a[2], dh)


@pytest.mark.gpu
def test_reverse_sweep_on_wide_workgroups(catn, gpu_ctx):
    """The reverse sweep on the 1024-lane kernel (exec_backward_wide) gives the
    256-lane kernel's gradient to rounding (n = 40 spin glass, exact
    contraction, lifted)."""
    n, edges, w, h = Cf.spin_glass_3reg(40, seed=3)
    out = {}
    saved = gpu_ctx.launch()
    try:
        for lanes in (256, 1024):
            gpu_ctx.set_launch(lanes, 20352 if lanes > 256 else 0, 16, 16)
            out[lanes] = catn.lnz_and_grad(n, edges, w, h, 1.0, ctx=gpu_ctx, **EXACT)
    finally:
        gpu_ctx.set_launch(saved[0], saved[1], saved[3], saved[4])
    (l0, w0, h0), (l1, w1, h1) = out[256], out[1024]
    assert abs(l1 - l0) <= 1e-12 * abs(l0)
    assert np.abs(w1 - w0).max() < 1e-10 and np.abs(h1 - h0).max() < 1e-10


def test_prepared_gradient_memory_gives_the_same_gradient(catn, emu_ctx):
    """catn_net_prepare_gradient allocates the tape and adjoint arena ahead of
    the reverse sweep: the gradient is the same, a prepared network still runs
    a plain contraction, and a negative tape capacity is refused."""
    n, edges, w, h = Cf.spin_glass_3reg(12, seed=2)
    a = catn.ising_network(n, edges, w, h, 1.0, ctx=emu_ctx, **EXACT)
    b = catn.ising_network(n, edges, w, h, 1.0, ctx=emu_ctx, **EXACT)
    b.prepare_gradient()
    la, _, _ = catn.gradient_batch([a])
    lb, _, _ = catn.gradient_batch([b])
    assert la[0] == lb[0]
    assert np.array_equal(catn.ising_gradient(a, len(edges), n)[0], catn.ising_gradient(b, len(edges), n)[0])
    c = catn.ising_network(n, edges, w, h, 1.0, ctx=emu_ctx, **EXACT)
    c.prepare_gradient(1000)
    assert catn.contraction(c)[0] == la[0]
    with pytest.raises(catn.CatnError, match="tape_cap"):
        catn.ising_network(n, edges, w, h, 1.0, ctx=emu_ctx, **EXACT).prepare_gradient(-1)
