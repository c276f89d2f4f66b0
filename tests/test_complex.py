"""Complex element types (ComplexF64), after test/test_complex.jl.

The oracle's restatement is pinned first on complex tensors by the checks of
test/test_complex.jl (:7-64): eat! of two leaves is the non-conjugating
sum(u .* v) and differs from dot(u, v); cut_bondim(_opt)! keeps a rank-2
complex bond at Dmax = 2 and preserves the product; a loop of four rank-2
complex matrices contracts losslessly at Dmax = 2; and generic chain, star
and loop networks match the dense contraction (numpy einsum, the semantics of
test/exact.jl's exact_contract) for select 0, 1, 2. A planted mistake (a
conjugating leaf dot) fails the pin.

The device path is then checked on the CPU emulator and, marked gpu, on the
MI355X: the same reference checks; random complex graphs with chi and Dmax
truncation against the oracle (edge sequence, maxdim_intermediate, Dmax-cut
kept counts, lnZ to 1e-12, phase psi to 1e-10, discarded weight); complex
double-layer networks (braket_network on complex PEPS, the bra built from
conj(T)) against the dense <psi|psi> and the oracle; and open-leg readout of
a complex network against the dense contraction. ComplexF32 storage is
checked against the dense contraction and against ComplexF64.
"""
import math

import numpy as np
import pytest

from oracle import braket_oracle as B
from oracle import catn_oracle as O
from oracle import configs as Cf


def _c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _dense(ts, ixs, out=()):
    abc = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"
    letters = {}
    for ix in ixs:
        for l in ix:
            if l not in letters:
                letters[l] = abc[len(letters)]
    spec = ",".join("".join(letters[l] for l in ix) for ix in ixs)
    return np.einsum(spec + "->" + "".join(letters[l] for l in out), *ts)


def _generic_networks(rng):
    return [
        ([_c(rng, 3, 4), _c(rng, 4, 5), _c(rng, 5, 3)], [["a", "b"], ["b", "c"], ["c", "a"]]),
        ([_c(rng, 2, 3, 4), _c(rng, 2), _c(rng, 3), _c(rng, 4)],
         [["a", "b", "c"], ["a"], ["b"], ["c"]]),
        ([_c(rng, 2, 3), _c(rng, 3, 2), _c(rng, 2, 3), _c(rng, 3, 2)],
         [["a", "b"], ["b", "c"], ["c", "d"], ["d", "a"]]),
    ]


def _loop_rank2(rng):
    mats = [_c(rng, 4, 2) @ _c(rng, 2, 4) for _ in range(4)]
    return mats, [["a", "b"], ["b", "c"], ["c", "d"], ["d", "a"]]


def _random_graph(rng, n, D):
    """Random 3-regular graph with a complex Gaussian tensor of bond dimension
    D per node (legs in edge-list order) and one open-free closed network."""
    edges = Cf.random_regular_edges(n, 3, rng)
    ixs = [[] for _ in range(n)]
    for e, (i, j) in enumerate(edges):
        ixs[i - 1].append(f"e{e}")
        ixs[j - 1].append(f"e{e}")
    ts = [_c(rng, *([D] * len(ix))) / math.sqrt(D) for ix in ixs]
    return ts, ixs


# ----------------------------- oracle pins ----------------------------------

def test_oracle_eat_leaves_non_conjugating():
    """test/test_complex.jl:8-16."""
    rng = np.random.default_rng(1)
    u, v = _c(rng, 4), _c(rng, 4)
    ni = O.MPSNode.from_tensor(u, [2], chi=32, cutoff=1e-15, norm_method=0)
    nj = O.MPSNode.from_tensor(v, [1], chi=32, cutoff=1e-15, norm_method=0)
    lg, err, ph = O.eat(ni, nj, 1, 1)
    assert not ni.mps
    assert abs(math.exp(lg) * ph - np.sum(u * v)) < 1e-12 * abs(np.sum(u * v))
    assert abs(math.exp(lg) * ph - np.vdot(u, v)) > 1e-3


@pytest.mark.parametrize("opt", [True, False])
def test_oracle_cut_keeps_complex_low_rank_bond(opt):
    """test/test_complex.jl:18-30."""
    rng = np.random.default_rng(2)
    P = _c(rng, 3, 2) @ _c(rng, 2, 4)
    Qm = _c(rng, 4, 2) @ _c(rng, 2, 3)
    tn = O.tensor_network([P, Qm], [["a", "m"], ["m", "b"]], chi=1000, Dmax=2)
    (O.cut_bondim_opt if opt else O.cut_bondim)(tn, 1, 2)
    assert tn.tensors[1].mps[1].shape[1] == 2
    got = O.mps2raw(tn.tensors[1]) @ O.mps2raw(tn.tensors[2])
    assert np.allclose(got, P @ Qm, atol=1e-8)


@pytest.mark.parametrize("svdopt", [True, False])
def test_oracle_finite_dmax_loop(svdopt):
    """test/test_complex.jl:32-43."""
    ts, ixs = _loop_rank2(np.random.default_rng(3))
    ref = _dense(ts, ixs)
    tn = O.tensor_network(ts, ixs, Dmax=2, chi=64, select=1, compress=True, svdopt=svdopt)
    lnz, err, psi = O.contraction(tn)
    assert abs(math.exp(lnz) * psi - ref) <= 1e-6 * abs(ref)


@pytest.mark.parametrize("sel", [0, 1, 2])
def test_oracle_generic_exact(sel):
    """test/test_complex.jl:45-63."""
    for ts, ixs in _generic_networks(np.random.default_rng(4)):
        ref = _dense(ts, ixs)
        tn = O.tensor_network(ts, ixs, Dmax=-1, chi=10_000, select=sel, reverse=True,
                              compress=True, norm_method=1)
        lnz, err, psi = O.contraction(tn)
        assert abs(math.exp(lnz) * psi - ref) <= 1e-10 * abs(ref)


def test_oracle_plants_a_mistake(monkeypatch):
    """A conjugating leaf dot (dot(vi, vj) in eat! case (a)) fails the exact pin."""
    ts, ixs = _generic_networks(np.random.default_rng(4))[0]
    ref = _dense(ts, ixs)
    orig = O.eat

    def eat_conj(node, nodej, idx, idxi):
        if len(node.mps) == 1:
            node.mps[0] = node.mps[0].conj()
        return orig(node, nodej, idx, idxi)

    monkeypatch.setattr(O, "eat", eat_conj)
    tn = O.tensor_network(ts, ixs, Dmax=-1, chi=10_000, select=1)
    lnz, err, psi = O.contraction(tn)
    assert abs(math.exp(lnz) * psi - ref) > 1e-3 * abs(ref)


# ----------------------------- device path ----------------------------------

@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    if request.param == "emu":
        return catn, request.getfixturevalue("emu_ctx")
    return catn, request.getfixturevalue("gpu_ctx")


def test_device_eat_leaves_non_conjugating(env):
    """test/test_complex.jl:8-16 on the device: the leaf dot of eat! case (a)."""
    catn, ctx = env
    rng = np.random.default_rng(1)
    u, v = _c(rng, 4), _c(rng, 4)
    net = catn.MPSNet(ctx=ctx, eltype="complex", norm_method=0)
    net.add_node(1, [u.reshape(1, 4, 1)], [2])
    net.add_node(2, [v.reshape(1, 4, 1)], [1])
    lg, err, ph = net.eat(1, 2, 1, 1)
    assert isinstance(ph, complex)
    assert abs(math.exp(lg) * ph - np.sum(u * v)) < 1e-12 * abs(np.sum(u * v))
    assert abs(math.exp(lg) * ph - np.vdot(u, v)) > 1e-3


@pytest.mark.parametrize("opt", [True, False])
def test_device_cut_keeps_complex_low_rank_bond(env, opt):
    """test/test_complex.jl:18-30 on the device."""
    catn, ctx = env
    rng = np.random.default_rng(2)
    P = _c(rng, 3, 2) @ _c(rng, 2, 4)
    Qm = _c(rng, 4, 2) @ _c(rng, 2, 3)
    net = catn.MPSNet(ctx=ctx, eltype="complex", chi=1000, Dmax=2)
    net.add_tensor(1, P, [-1, 2])
    net.add_tensor(2, Qm, [1, -2])
    net.cut_bondim(1, 2, opt=opt)
    s1, s2 = net.sites(1), net.sites(2)
    assert s1[1].shape[1] == 2 and s2[0].shape[1] == 2
    A = np.einsum("xai,iby->ab", s1[0], s1[1])        # mps2raw of node 1: (a, m)
    Bm = np.einsum("xai,iby->ab", s2[0], s2[1])       # node 2: (m, b)
    assert np.allclose(A @ Bm, P @ Qm, atol=1e-8)


@pytest.mark.parametrize("svdopt", [True, False])
def test_device_finite_dmax_loop(env, svdopt):
    """test/test_complex.jl:32-43 on the device."""
    catn, ctx = env
    ts, ixs = _loop_rank2(np.random.default_rng(3))
    ref = _dense(ts, ixs)
    tn = catn.tensor_network(ts, ixs, ctx=ctx, Dmax=2, chi=64, select=1, compress=True,
                             svdopt=svdopt)
    lnz, err, psi = catn.contraction(tn)
    assert isinstance(psi, complex)
    assert abs(math.exp(lnz) * psi - ref) <= 1e-6 * abs(ref)


@pytest.mark.parametrize("sel", [0, 1, 2])
@pytest.mark.parametrize("compress", [True, False])
def test_device_generic_exact(env, sel, compress):
    """test/test_complex.jl:45-63 on the device, with and without compress."""
    catn, ctx = env
    for ts, ixs in _generic_networks(np.random.default_rng(4)):
        ref = _dense(ts, ixs)
        tn = catn.tensor_network(ts, ixs, ctx=ctx, Dmax=-1, chi=10_000, select=sel, reverse=True,
                                 compress=compress, norm_method=1)
        lnz, err, psi = catn.contraction(tn)
        assert abs(math.exp(lnz) * psi - ref) <= 1e-10 * abs(ref)


@pytest.mark.parametrize("n,D,Dmax,chi,sel,norm,seed", [
    (10, 2, -1, 10_000, 1, 1, 1),    # exact
    (12, 2, 3, 6, 1, 1, 2),          # Dmax cuts and chi truncation
    (12, 3, 4, 8, 0, 2, 3),          # max-abs normalisation
    (14, 2, 2, 4, 2, 1, 4),          # sequential selection
    (10, 3, 5, 12, 1, 0, 5),         # no normalisation
])
def test_device_random_graph_matches_oracle(env, n, D, Dmax, chi, sel, norm, seed):
    """Random complex 3-regular graphs, truncated by chi and Dmax: the device
    and the oracle take the same schedule (edge sequence, maxdim, cut kept
    counts) and give the same lnZ, phase and discarded weight."""
    catn, ctx = env
    ts, ixs = _random_graph(np.random.default_rng(seed), n, D)
    kw = dict(Dmax=Dmax, chi=chi, select=sel, norm_method=norm)
    ref = O.tensor_network(ts, ixs, **kw)
    s = O.Schedule()
    lo, eo, po = O.contraction(ref, s)
    tn = catn.tensor_network(ts, ixs, ctx=ctx, **kw)
    lnz, err, psi = catn.contraction(tn)
    assert tn.edge_log() == s.edges
    assert tn.maxdim_intermediate == ref.maxdim_intermediate
    assert tn.cut_kept() == [e[2] for e in s.events if e[0] == "cut_keep"]
    assert abs(lnz - lo) <= 1e-12 * max(1.0, abs(lo))
    assert abs(psi * tn.sign - complex(po) * complex(ref.sign)) < 1e-10
    assert abs(err - eo) <= 1e-8 * max(1.0, abs(eo))
    if Dmax < 0:
        val = _dense(ts, ixs)
        assert abs(lnz - math.log(abs(val))) < 1e-10


@pytest.mark.parametrize("name", ["chain4_cplx", "chain5_cplx", "star_cplx"])
@pytest.mark.parametrize("svdopt", [True, False])
def test_device_braket_norm_matches_exact(env, name, svdopt):
    """test/test_braket.jl:246-305 with complex tensors: <psi|psi> of complex
    chains and a star, the bra chains built from conj(T)."""
    catn, ctx = env
    ts, ixs = _braket_case(name)
    ref = B.exact_norm(ts, ixs)
    tn = catn.braket_network(ts, ixs, ctx=ctx, chi=10_000, svdopt=svdopt)
    lnz, err, psi = catn.contraction(tn)
    assert abs(math.exp(lnz) * psi - ref) <= 1e-10 * ref
    assert abs(psi - 1.0) < 1e-10


def _braket_case(name):
    """complex chains (p0 b0, b0 p1 b1, ...) and a star, as in test/test_braket.jl"""
    rng = np.random.default_rng({"chain4_cplx": 11, "chain5_cplx": 12, "star_cplx": 13}[name])
    if name == "star_cplx":
        return ([_c(rng, 2, 2, 2, 2), _c(rng, 2, 2), _c(rng, 2, 2), _c(rng, 2, 2)],
                [["a", "b", "c", "pc"], ["a", "p1"], ["b", "p2"], ["c", "p3"]])
    n, D, d = (4, 3, 2) if name == "chain4_cplx" else (5, 3, 2)
    ts, ixs = [], []
    for k in range(n):
        if k == 0:
            ts.append(_c(rng, d, D)); ixs.append([f"p{k}", f"b{k}"])
        elif k == n - 1:
            ts.append(_c(rng, D, d)); ixs.append([f"b{k - 1}", f"p{k}"])
        else:
            ts.append(_c(rng, D, d, D)); ixs.append([f"b{k - 1}", f"p{k}", f"b{k}"])
    return ts, ixs


def _lattice(rng, Lx, Ly, D, d):
    """open Lx x Ly PEPS, complex Gaussian entries, physical leg in the middle"""
    ts, ixs = [], []
    for x in range(Lx):
        for y in range(Ly):
            ix = []
            if x > 0:
                ix.append(f"v{x - 1}_{y}")
            if y > 0:
                ix.append(f"h{x}_{y - 1}")
            ix.append(f"p{x}_{y}")
            if y < Ly - 1:
                ix.append(f"h{x}_{y}")
            if x < Lx - 1:
                ix.append(f"v{x}_{y}")
            ts.append(_c(rng, *[d if l.startswith("p") else D for l in ix]) / math.sqrt(2))
            ixs.append(ix)
    return ts, ixs


@pytest.mark.parametrize("Lx,Ly,D,chi,loops", [(2, 3, 2, 10_000, True), (3, 3, 2, 10_000, True),
                                               (3, 3, 3, 8, True)])
def test_device_complex_peps_norm_matches_oracle(env, Lx, Ly, D, chi, loops):
    """A complex PEPS norm (the double layer of config 5, at lattice sizes the
    dense norm and the oracle reach): the device matches the oracle's edge
    sequence, lnZ and phase; exact runs also match the dense <psi|psi>."""
    catn, ctx = env
    ts, ixs = _lattice(np.random.default_rng(300 + Lx * Ly + D), Lx, Ly, D, 2)
    bk = B.braket_network(ts, ixs, chi=chi, loops=loops)
    lo, eo, po = B.contraction(bk)
    tn = catn.braket_network(ts, ixs, ctx=ctx, chi=chi, loops=loops)
    lnz, err, psi = catn.contraction(tn)
    assert [tuple(e) for e in tn.edge_log()] == bk.edge_log
    assert abs(lnz - lo) <= 1e-12 * max(1.0, abs(lo))
    assert abs(psi - complex(po)) < 1e-10
    if chi >= 10_000:
        assert abs(lnz - math.log(B.exact_norm(ts, ixs))) < 1e-10


def test_device_complex_open_legs(env):
    """result_tensor (src/tensor_network.jl:504-523) of a complex chain with open
    physical legs against the dense contraction."""
    catn, ctx = env
    rng = np.random.default_rng(11)
    ts = [_c(rng, 2, 4), _c(rng, 4, 2, 4), _c(rng, 4, 2)]
    ixs = [["p1", "a"], ["a", "p2", "b"], ["b", "p3"]]
    tn = catn.tensor_network(ts, ixs, ctx=ctx, Dmax=-1, chi=10_000)
    lnz, err, psi = catn.contraction(tn)
    arr, labels = catn.result_tensor(tn)
    assert arr.dtype == np.complex128 and sorted(labels) == ["p1", "p2", "p3"]
    ref = _dense(ts, ixs, labels)
    got = arr * (math.exp(lnz) * psi)
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_device_complex_refusals(env):
    """No reverse sweep, Ising networks or mixed batches for complex networks:
    each is refused with an error that names it."""
    catn, ctx = env
    rng = np.random.default_rng(7)
    ts, ixs = _generic_networks(rng)[0]
    tn = catn.tensor_network(ts, ixs, ctx=ctx, Dmax=-1, chi=10_000)
    with pytest.raises(catn.CatnError, match="real networks"):
        catn.gradient_batch([tn])
    with pytest.raises(catn.CatnError, match="real"):
        catn.ising_network(3, [(1, 2), (2, 3)], [1.0, 1.0], [0.0] * 3, 0.5, ctx=ctx,
                           eltype="complex")
    a = catn.tensor_network(ts, ixs, ctx=ctx, Dmax=-1, chi=10_000)
    b = catn.tensor_network([t.real for t in ts], ixs, ctx=ctx, Dmax=-1, chi=10_000)
    with pytest.raises(catn.CatnError, match="real and complex"):
        catn.contraction_batch([a, b])


@pytest.mark.parametrize("sel", [0, 1, 2])
def test_device_complexf32_generic(env, sel):
    """ComplexF32 storage (TensorNetwork{ComplexF32}): FP32 (re, im) pairs in
    the arena, every product in FP64; the generic networks of
    test/test_complex.jl match the dense contraction to FP32 accuracy."""
    catn, ctx = env
    for ts, ixs in _generic_networks(np.random.default_rng(4)):
        ts32 = [t.astype(np.complex64).astype(np.complex128) for t in ts]
        ref = _dense(ts32, ixs)
        tn = catn.tensor_network(ts, ixs, ctx=ctx, Dmax=-1, chi=10_000, select=sel,
                                 compress=True, dtype="f32")
        lnz, err, psi = catn.contraction(tn)
        assert isinstance(psi, complex)
        assert abs(math.exp(lnz) * psi - ref) <= 2e-6 * abs(ref)


@pytest.mark.parametrize("Lx,Ly,D,chi", [(3, 3, 2, 10_000), (3, 3, 3, 8), (3, 4, 2, 12)])
def test_device_complexf32_peps_norm(env, Lx, Ly, D, chi):
    """A complex PEPS norm in ComplexF32 storage (config 5's element type)
    against the same network in ComplexF64: |d lnZ / N| <= 1e-6 and the same
    edge sequence."""
    catn, ctx = env
    ts, ixs = _lattice(np.random.default_rng(400 + Lx * Ly + D), Lx, Ly, D, 2)
    ts = [t.astype(np.complex64) for t in ts]
    a = catn.braket_network(ts, ixs, ctx=ctx, chi=chi, loops=True)
    b = catn.braket_network(ts, ixs, ctx=ctx, chi=chi, loops=True, dtype="f32")
    la, ea, pa = catn.contraction(a)
    lb, eb, pb = catn.contraction(b)
    assert a.edge_log() == b.edge_log()
    assert abs(la - lb) / (Lx * Ly) <= 1e-6
    assert abs(pb - pa) < 1e-5
    if chi >= 10_000:
        assert abs(lb - math.log(B.exact_norm([t.astype(np.complex128) for t in ts], ixs))) < 1e-5


def _spec(ket, ket_ix):
    """einsum arguments of the single ket layer with its physical legs open
    (the amplitude tensor psi)"""
    abc = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"
    letters = {}
    for ix in ket_ix:
        for l in ix:
            letters.setdefault(l, abc[len(letters)])
    out = "".join(letters[l] for ix in ket_ix for l in ix if l.startswith("p"))
    return [",".join("".join(letters[l] for l in ix) for ix in ket_ix) + "->" + out] + list(ket)


def _peps_fused(rng, L, D, d):
    """examples/peps_norm.jl:67-110: a random complex PEPS (physical leg first,
    then N, S, W, E) and its double-layer network of E tensors,
    E[(l1 l1'), ...] = sum_p A[p, l1, ...] conj(A[p, l1', ...]), each bond of
    dimension D^2 with the ket index fastest."""
    ket, ket_ix, E, E_ix = [], [], [], []
    for r in range(1, L + 1):
        for c in range(1, L + 1):
            bonds = []
            if r > 1:
                bonds.append(f"v{r - 1}_{c}")
            if r < L:
                bonds.append(f"v{r}_{c}")
            if c > 1:
                bonds.append(f"h{r}_{c - 1}")
            if c < L:
                bonds.append(f"h{r}_{c}")
            deg = len(bonds)
            A = _c(rng, d, *([D] * deg)) / math.sqrt(2)
            ket.append(A)
            ket_ix.append([f"p{r}_{c}"] + bonds)
            Am = A.reshape((d, D ** deg), order="F")
            M2 = Am.T @ Am.conj()                                  # (alpha, beta)
            T = M2.reshape([D] * (2 * deg), order="F")             # (l1..ldeg, l1'..ldeg')
            perm = [x for k in range(deg) for x in (k, deg + k)]   # (l1, l1', l2, l2', ...)
            E.append(T.transpose(perm).reshape([D * D] * deg, order="F"))
            E_ix.append(bonds)
    return ket, ket_ix, E, E_ix


@pytest.mark.parametrize("L,D,Dmax,chi,dtype", [(3, 2, -1, 10_000, "f64"), (4, 2, 8, 16, "f64"),
                                                 (4, 3, 16, 32, "f64"), (4, 3, 128, 512, "f64"),
                                                 (4, 3, 128, 512, "f32")])
def test_device_fused_peps_norm(env, L, D, Dmax, chi, dtype):
    """Config 5's route (examples/peps_norm.jl:112-116): the fused double-layer
    network of a random complex PEPS as a complex TensorNetwork with Dmax and
    chi, at config 5's D = 3, Dmax = 128, chi = 512 on a 4 x 4 lattice (no
    bond reaches Dmax or chi there) in ComplexF64 and ComplexF32 storage.
    Untruncated runs match sum |psi|^2 of the single ket layer (the amplitude
    oracle of the example, 1e-9; 1e-6 per site for ComplexF32) and give a
    real positive value; truncated runs match the oracle's schedule, lnZ and
    phase."""
    catn, ctx = env
    ket, ket_ix, E, E_ix = _peps_fused(np.random.default_rng(500 + 10 * L + D), L, D, 2)
    kw = dict(Dmax=Dmax, chi=chi)
    tn = catn.tensor_network(E, E_ix, ctx=ctx, dtype=dtype, **kw)
    lnz, err, psi = catn.contraction(tn)
    if dtype == "f32":
        amp = np.einsum(*_spec(ket, ket_ix), optimize=True)
        assert abs(lnz - math.log(float(np.sum(np.abs(amp) ** 2)))) / (L * L) <= 1e-6
        assert abs(psi - 1.0) < 1e-6
        return
    if Dmax < 0 or Dmax >= 128:
        assert err == 0.0
        amp = np.einsum(*_spec(ket, ket_ix), optimize=True)
        ref = float(np.sum(np.abs(amp) ** 2))
        assert abs(math.exp(lnz) * psi - ref) <= 1e-9 * ref
        assert abs(psi - 1.0) < 1e-9
        return
    ref = O.tensor_network(E, E_ix, **kw)
    s = O.Schedule()
    lo, eo, po = O.contraction(ref, s)
    assert tn.edge_log() == s.edges
    assert tn.cut_kept() == [e[2] for e in s.events if e[0] == "cut_keep"]
    assert abs(lnz - lo) <= 1e-12 * max(1.0, abs(lo))
    assert abs(psi - complex(po)) < 1e-9


def _ising_factor_graph(L, beta, seed):
    """+-J Ising model on an open L x L lattice as a generic network of one
    tensor per spin: T_k[s_1, ...] = sum_sigma prod_legs F[sigma, s_leg], with
    F the Boltzmann matrix B_e = exp(beta w_e sigma s) on one end of each
    coupling and the identity on the other (real entries; its splits have
    exactly degenerate singular values)."""
    from oracle import exact as X
    r = np.random.default_rng(seed)
    edges = X.square_edges(L)
    w = r.choice([-1.0, 1.0], len(edges))
    spin = np.array([1.0, -1.0])
    legs = {k: [] for k in range(1, L * L + 1)}
    for e, (a, b) in enumerate(edges):
        legs[a].append((f"e{e}", np.exp(beta * w[e] * np.outer(spin, spin))))
        legs[b].append((f"e{e}", np.eye(2)))
    ts, ixs = [], []
    for k in range(1, L * L + 1):
        t = 0.0
        for sig in range(2):
            f = np.ones(())
            for (_, F) in legs[k]:
                f = np.multiply.outer(f, F[sig])
            t = t + f
        ts.append(np.asarray(t, dtype=float))
        ixs.append([lab for (lab, _) in legs[k]])
    return L * L, edges, w, ts, ixs


@pytest.mark.parametrize("L,beta,Dmax,chi", [(4, 1.0, -1, 10_000), (8, 0.6, 32, 128),
                                          (10, 0.6, 32, 128)])
def test_device_complex_type_on_real_data_is_the_real_engine(env, L, beta, Dmax, chi):
    """The complex engine (complex one-sided Jacobi, complex rotations) on
    real data with exactly degenerate spectra (+-J couplings): lnZ, the phase
    (+1), the edge sequence and the Dmax-cut kept counts equal the real
    engine's (QR-preconditioned real Jacobi), and the exact lnZ at 4 x 4."""
    from oracle import exact as X
    catn, ctx = env
    n, edges, w, ts, ixs = _ising_factor_graph(L, beta, 10 + L)
    kw = dict(Dmax=Dmax, chi=chi, arena_elems=1 << 24, scratch_elems=1 << 24)
    real = catn.tensor_network(ts, ixs, ctx=ctx, **kw)
    cplx = catn.tensor_network([t.astype(complex) for t in ts], ixs, ctx=ctx, **kw)
    lr, er, pr = catn.contraction(real)
    lc, ec, pc = catn.contraction(cplx)
    assert abs(lc - lr) <= 1e-12 * max(1.0, abs(lr))
    assert abs(pc - pr) <= 1e-12 and abs(ec - er) <= 1e-9 * max(1.0, er)
    assert cplx.edge_log() == real.edge_log()
    assert cplx.cut_kept() == real.cut_kept()
    if L == 4:
        assert abs(lr - X.brute_lnZ(n, edges, w, np.zeros(n), beta)) <= 1e-12 * abs(lr)
    if L == 10:   # splits of min(m, n) in 17..32 ran
        assert cplx.profile()[1].get("<=32", (0,))[0] > 0
