"""The hand-written adjoint of contraction!'s lnZ (chosen form, differentiation).

Pins, in order of strength:
  * exact contraction (no truncation): d lnZ/d w_ij = beta <s_i s_j> and
    d lnZ/d h_i = beta <s_i> against brute-force enumeration (oracle/exact.py)
    and closed forms (open chain: beta tanh(beta J); isolated spin:
    beta tanh(beta h)), to 1e-10;
  * truncated contraction (Dmax / chi active): central finite differences of
    the engine's own lnZ, step 1e-4, at a frozen schedule (the perturbed runs
    execute the same op sequence with the same kept counts), to 1e-6
    relative;
  * the reference's own observables (magnetization / correlation by pinning,
    src/ising.jl:263-339) on the same networks: equal in exact mode to 1e-10,
    and within the truncation error otherwise.
The forward of a recorded run is the plain forward: same lnZ bits, same edge
sequence. Every test runs on the CPU emulator and, marked gpu, on the HIP
library.
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


EXACT = dict(Dmax=-1, chi=10_000)


def err_of(tn):
    return tn.err_

# Gradient runs across a Dmax cut faulted the GPU and the cause is not found
# (the CPU emulators, also lane-parallel under ASan/TSan, run them cleanly);
# the gfx950 executor refuses them before any launch, so their GPU cases are
# skipped and the emulator cases keep the math covered.
CUT_REVERSE_SKIP = ("reverse sweep across a Dmax cut is disabled on the gfx950 executor: "
                    "such runs faulted the GPU, cause not found on the CPU")


def _skip_cut_gradient_on_gpu(ctx):
    if ctx.backend.startswith("hip"):
        pytest.skip(CUT_REVERSE_SKIP)


def grad(catn, ctx, n, edges, w, h, beta, **kw):
    tn = catn.ising_network(n, edges, w, h, beta, ctx=ctx, **kw)
    lz, err, _ = catn.gradient_batch([tn])
    tn.err_ = float(err[0])
    dw, dh, db = catn.ising_gradient(tn, len(edges), n)
    return float(lz[0]), dw, dh, db, tn


# ------------------------- exact contraction ------------------------------

def test_triangle_observables(env):
    """test/test_ising.jl:47-72 with the adjoint in place of pinning."""
    catn, ctx = env
    beta, edges = 0.35, [(1, 2), (2, 3), (3, 1)]
    w, h = [0.4, 0.6, -0.3], [0.1, -0.2, 0.05]
    mag, cor = X.brute_observables(3, edges, w, h, beta)
    m = catn.magnetization(3, edges, w, h, beta, ctx=ctx, method="adjoint", **EXACT)
    c = catn.correlation(3, edges, w, h, beta, ctx=ctx, method="adjoint", **EXACT)
    assert np.allclose(m, mag, rtol=0, atol=1e-12)
    assert np.allclose(c, cor, rtol=0, atol=1e-12)
    # and the reference's own pinning route agrees
    assert np.allclose(catn.magnetization(3, edges, w, h, beta, ctx=ctx, **EXACT), m, atol=1e-12)
    assert np.allclose(catn.correlation(3, edges, w, h, beta, ctx=ctx, **EXACT), c, atol=1e-12)


def test_open_chain_closed_form(env):
    """Open chain (test/test_ising.jl:14-21): Z = 2 prod_e 2cosh(beta J_e),
    so d lnZ/d J_e = beta tanh(beta J_e) and d lnZ/d beta = sum_e J_e tanh(beta J_e)."""
    catn, ctx = env
    L, beta = 6, 0.4
    J = np.array([0.7, -0.3, 1.2, 0.5, -0.9])
    edges = [(i, i + 1) for i in range(1, L)]
    lz, dw, dh, db, _ = grad(catn, ctx, L, edges, J, np.zeros(L), beta, **EXACT)
    assert abs(lz - (math.log(2) + np.sum(np.log(2 * np.cosh(beta * J))))) < 1e-12
    assert np.allclose(dw, beta * np.tanh(beta * J), atol=1e-13)
    assert np.allclose(dh, 0.0, atol=1e-13)
    assert abs(db - np.sum(J * np.tanh(beta * J))) < 1e-12


def test_isolated_spins(env):
    """test/test_ising.jl:90-110: lnZ = sum log 2cosh(beta h), d/dh = beta tanh(beta h)."""
    catn, ctx = env
    beta, h = 0.6, np.array([0.3, 0.7])
    lz, dw, dh, _, _ = grad(catn, ctx, 2, [], np.zeros(0), h, beta, **EXACT)
    assert abs(lz - np.sum(np.log(2 * np.cosh(beta * h)))) < 1e-13
    assert np.allclose(dh, beta * np.tanh(beta * h), atol=1e-14)


def _instance(kind, seed):
    r = np.random.default_rng(seed)
    if kind.startswith("lattice"):
        L = int(kind[7:])
        edges = X.square_edges(L)
        n = L * L
    else:
        n = int(kind[2:])
        edges = Cf.random_regular_edges(n, 3, r)
    # generic couplings: no exact rank deficiency for a perturbation to lift
    return n, edges, r.uniform(-1.2, 1.2, len(edges)), r.normal(0, 0.4, n), float(r.uniform(0.3, 1.1))


@pytest.mark.parametrize("kind,seed", [("lattice3", 1), ("lattice4", 2), ("sg12", 3), ("sg16", 4)])
@pytest.mark.parametrize("opts", [dict(), dict(select=0), dict(select=2), dict(norm_method=2),
                                  dict(norm_method=0), dict(reverse=False)])
def test_exact_gradient_vs_brute_force(env, kind, seed, opts):
    catn, ctx = env
    n, edges, w, h, beta = _instance(kind, seed)
    lz, dw, dh, db, tn = grad(catn, ctx, n, edges, w, h, beta, **EXACT, **opts)
    mag, cor = X.brute_observables(n, edges, w, h, beta)
    assert abs(lz - X.brute_lnZ(n, edges, w, h, beta)) < 1e-11
    assert np.abs(dw - beta * cor).max() < 1e-10
    assert np.abs(dh - beta * mag).max() < 1e-10
    # d lnZ / d beta = <E>
    E = np.dot(w, cor) + np.dot(h, mag)
    assert abs(db - E) < 1e-9
    assert tn.bwd_profile()[1] < 1e-12      # replayed factorisations == forward factors


def test_pinned_network_gradient(env):
    """The adjoint of a pinned network (ghost-0 spin, src/ising.jl:273) is the
    gradient of that network's lnZ: conditional expectations given s_k."""
    catn, ctx = env
    n, edges, w, h, beta = _instance("lattice3", 7)
    tn = catn.ising_network(n, edges, w, h, beta, pos1=5, val1=1, ctx=ctx, **EXACT)
    catn.gradient_batch([tn])
    dw, dh, _ = catn.ising_gradient(tn, len(edges), n)
    s = X._spins(n).astype(float)
    s[s[:, 4] > 0, 4] = 0.0       # val1 = 1: spin value +1 of spin 5 becomes a ghost 0
    E = s @ h + sum(w[e] * s[:, i - 1] * s[:, j - 1] for e, (i, j) in enumerate(edges))
    p = np.exp(beta * (E - E.max()))
    p /= p.sum()
    cor = np.array([p @ (s[:, i - 1] * s[:, j - 1]) for (i, j) in edges])
    assert np.abs(dw - beta * cor).max() < 1e-10


# ------------------------ truncated contraction ---------------------------

def _frozen_fd(catn, ctx, n, edges, w, h, beta, idx_w, idx_h, eps=1e-4, **kw):
    """Central differences of the engine's lnZ; each perturbed pair is kept
    only when both runs replay the unperturbed op sequence and kept counts."""
    base = catn.ising_network(n, edges, w, h, beta, ctx=ctx, **kw)
    catn.gradient_batch([base])
    rec0, _ = base.tape()
    jobs = []
    for k in idx_w:
        for sgn in (1, -1):
            ww = w.copy()
            ww[k] += sgn * eps
            jobs.append(("w", k, ww, h))
    for k in idx_h:
        for sgn in (1, -1):
            hh = h.copy()
            hh[k] += sgn * eps
            jobs.append(("h", k, w, hh))
    nets = [catn.ising_network(n, edges, ww, hh, beta, ctx=ctx, **kw) for (_, _, ww, hh) in jobs]
    lz, _, _ = catn.gradient_batch(nets)
    out = {}
    for q in range(0, len(jobs), 2):
        kind, k = jobs[q][0], jobs[q][1]
        ra, _ = nets[q].tape()
        rb, _ = nets[q + 1].tape()
        frozen = all(len(r) == len(rec0) and (r[:, :3] == rec0[:, :3]).all() for r in (ra, rb))
        out[(kind, k)] = ((lz[q] - lz[q + 1]) / (2 * eps), frozen)
    return out


@pytest.mark.parametrize("case", ["sg40_D8_chi16", "lattice6_pmJ_chi6", "lattice5_Dmax3",
                                  "sg40_chi10", "lattice6_pmJ_chi4", "sg24_chi8"])
def test_truncated_gradient_vs_finite_differences(env, case):
    """Truncated contractions: the first three reach Dmax cuts (their GPU
    cases are skipped, see CUT_REVERSE_SKIP); the last three truncate by chi
    only (Dmax = -1, no cut), so the reverse sweep differentiates truncated
    swap and canonicalisation SVDs (the broadened SVD adjoint) on every
    backend."""
    catn, ctx = env
    if "chi" not in case or "_D" in case or case == "lattice6_pmJ_chi6":
        _skip_cut_gradient_on_gpu(ctx)
    if case == "sg40_chi10":
        n, edges, w, h = Cf.spin_glass_3reg(40, seed=1)
        beta, kw = 1.0, dict(Dmax=-1, chi=10)
    elif case == "lattice6_pmJ_chi4":
        L = 6
        edges = X.square_edges(L)
        n = L * L
        r = np.random.default_rng(0)
        w, h = r.choice([-1.0, 1.0], len(edges)), r.normal(0, 0.1, n)
        beta, kw = 0.8, dict(Dmax=-1, chi=4)
    elif case == "sg24_chi8":
        n, edges, w, h = Cf.spin_glass_3reg(24, seed=5)
        beta, kw = 1.0, dict(Dmax=-1, chi=8)
    elif case == "sg40_D8_chi16":
        n, edges, w, h = Cf.spin_glass_3reg(40, seed=1)
        beta, kw = 1.0, dict(Dmax=8, chi=16)
    elif case == "lattice6_pmJ_chi6":
        L = 6
        edges = X.square_edges(L)
        n = L * L
        r = np.random.default_rng(0)
        w, h = r.choice([-1.0, 1.0], len(edges)), r.normal(0, 0.1, n)
        beta, kw = 0.8, dict(Dmax=6, chi=6)
    else:
        L = 5
        edges = X.square_edges(L)
        n = L * L
        r = np.random.default_rng(3)
        w, h = r.uniform(-1, 1, len(edges)), r.normal(0, 0.2, n)
        beta, kw = 0.9, dict(Dmax=3, chi=32)
    lz, dw, dh, _, tn = grad(catn, ctx, n, edges, w, h, beta, **kw)
    ops = tn.profile()[0]
    if case == "lattice5_Dmax3":
        assert ops.get("cut", (0,))[0] > 0          # the Dmax cut is exercised
    if kw["Dmax"] < 0:
        assert ops.get("cut", (0,))[0] == 0
        assert tn.mem()[0] > 0 and err_of(tn) > 0     # chi truncates
    iw = list(range(0, len(edges), max(1, len(edges) // 12)))
    ih = list(range(0, n, max(1, n // 6)))
    fd = _frozen_fd(catn, ctx, n, edges, np.array(w, float), np.array(h, float), beta, iw, ih, **kw)
    checked = 0
    for (kind, k), (v, frozen) in fd.items():
        if not frozen:
            continue
        a = dw[k] if kind == "w" else dh[k]
        assert abs(a - v) <= 1e-6 * max(abs(v), 1e-2), (kind, k, a, v)
        checked += 1
    assert checked >= 0.8 * len(fd)


@pytest.mark.parametrize("kw", [dict(Dmax=6, chi=12), dict(Dmax=-1, chi=8)],
                         ids=["Dmax6_chi12", "chi8"])
def test_adjoint_close_to_pinning_when_truncated(env, kw):
    """With truncation the two routes differentiate slightly different
    approximations; they agree to the truncation level. The chi-only case
    reaches no Dmax cut."""
    catn, ctx = env
    if kw["Dmax"] > 0:
        _skip_cut_gradient_on_gpu(ctx)
    n, edges, w, h = Cf.spin_glass_3reg(24, seed=5)
    m_adj = catn.magnetization(n, edges, w, h, 1.0, ctx=ctx, method="adjoint", **kw)
    m_pin = catn.magnetization(n, edges, w, h, 1.0, ctx=ctx, **kw)
    mag, _ = X.brute_observables(n, edges, w, h, 1.0)
    assert np.abs(m_adj - m_pin).max() < 5e-3
    assert np.abs(m_adj - mag).max() < 5e-3


def test_recorded_forward_is_the_plain_forward(env):
    catn, ctx = env
    _skip_cut_gradient_on_gpu(ctx)
    n, edges, w, h = Cf.spin_glass_3reg(30, seed=2)
    kw = dict(Dmax=8, chi=24)
    a = catn.ising_network(n, edges, w, h, 1.0, ctx=ctx, **kw)
    b = catn.ising_network(n, edges, w, h, 1.0, ctx=ctx, **kw)
    la, ea, pa = catn.contraction_batch([a])
    lb, eb, pb = catn.gradient_batch([b])
    assert la[0] == lb[0] and ea[0] == eb[0] and pa[0] == pb[0]
    assert a.edge_log() == b.edge_log()


def test_batch_of_uneven_instances(env):
    """Instances of different size and beta in one forward + one reverse sweep
    give the same results as one at a time."""
    catn, ctx = env
    insts = [_instance("lattice3", 11), _instance("sg12", 12), _instance("lattice4", 13)]
    insts = [(n, e, w, h, b) for (n, e, w, h, b) in insts]
    batch = catn.lnz_and_grad_batch(insts, ctx=ctx, **EXACT)
    for (n, e, w, h, b), (lz, dw, dh) in zip(insts, batch):
        l1, w1, h1 = catn.lnz_and_grad(n, e, w, h, b, ctx=ctx, **EXACT)
        assert lz == l1
        assert np.abs(dw - w1).max() < 1e-13 and np.abs(dh - h1).max() < 1e-13


def test_gradient_needs_a_fresh_network(env):
    catn, ctx = env
    n, edges, w, h, beta = _instance("lattice3", 1)
    tn = catn.ising_network(n, edges, w, h, beta, ctx=ctx, **EXACT)
    catn.contraction(tn)
    with pytest.raises(catn.CatnError):
        catn.gradient_batch([tn])
    tn2 = catn.ising_network(n, edges, w, h, beta, ctx=ctx, **EXACT)
    with pytest.raises(catn.CatnError):
        catn.ising_gradient(tn2, len(edges), n)     # no reverse sweep yet


def test_device_refuses_gradient_across_a_cut_before_any_launch(catn, emu_ctx, monkeypatch):
    """The refusal of the gfx950 executor, emulated: a gradient run whose
    schedule reaches a Dmax cut raises before its first launch (no device op
    executed), while a cut-free run is unaffected."""
    monkeypatch.setenv("CATN_EMU_REFUSE_CUT_REVERSE", "1")
    n, edges, w, h = Cf.spin_glass_3reg(40, seed=1)
    tn = catn.ising_network(n, edges, w, h, 1.0, ctx=emu_ctx, Dmax=8, chi=16)
    with pytest.raises(catn.CatnError, match="Dmax cut"):
        catn.gradient_batch([tn])
    assert tn.work()[0] == 0.0
    n, edges, w, h, beta = _instance("lattice3", 1)
    l, dw, dh = catn.lnz_and_grad(n, edges, w, h, beta, ctx=emu_ctx, **EXACT)
    assert np.isfinite(l) and np.all(np.isfinite(dw))
    # the plain forward with cuts is not affected
    tn = catn.ising_network(*Cf.spin_glass_3reg(40, seed=1), 1.0, ctx=emu_ctx, Dmax=8, chi=16)
    lz, _, _ = catn.contraction_batch([tn])
    assert np.isfinite(lz[0])
