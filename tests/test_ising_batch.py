"""catn_ising_network_batch: ising_network built for a batch with the site
values computed on the device from the packed instance format
(src/ising.jl:68-195). Each network must equal the one catn_ising_network
builds on the host: the same initial site tensors (to the last bits of exp),
the same edge sequence, lnZ and gradient."""
import numpy as np
import pytest

from oracle import configs as Cf
from oracle import exact as X


@pytest.fixture(scope="module", params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, catn):
    return catn, request.getfixturevalue("emu_ctx" if request.param == "emu" else "gpu_ctx")


def _instances():
    r = np.random.default_rng(5)
    a = Cf.spin_glass_3reg(12, seed=3)
    L = 4
    edges = X.square_edges(L)
    b = (L * L, edges, r.uniform(-1, 1, len(edges)), r.normal(0, 0.3, L * L))
    # an isolated spin (no edge) and a leaf
    c = (4, [(1, 2), (2, 3)], np.array([0.7, -0.4]), np.array([0.2, -0.1, 0.3, 0.5]))
    return [a, b, c]


def _sites(catn, tn, n):
    """Every node's site tensors, read back before contraction."""
    view = catn.MPSNet.view(tn)    # the per-node readers on this handle
    out = {}
    for node in range(1, n + 1):
        for k, site in enumerate(view.sites(node)):
            out[(node, k + 1)] = site
    return out


@pytest.mark.parametrize("dtype", ["f64", "f32x6"])
def test_device_built_networks_equal_host_built(env, dtype):
    catn, ctx = env
    insts = _instances()
    spec = [(0, 0.9), (0, 0.9, (3, 1, 0, 0)), (1, 0.6, (2, 2, 5, 1)), (2, 1.1), (2, 0.5, (4, 2, 0, 0)),
            (1, 0.3)]
    kw = dict(Dmax=4, chi=8, dtype=dtype)
    dev = catn.ising_network_batch(insts, spec, ctx=ctx, **kw)
    host = []
    for t in spec:
        n, e, w, h = insts[t[0]]
        p = t[2] if len(t) > 2 else (0, 0, 0, 0)
        host.append(catn.ising_network(n, e, w, h, t[1], pos1=p[0], val1=p[1], pos2=p[2], val2=p[3],
                                       ctx=ctx, **kw))
    tol = 1e-6 if dtype != "f64" else 1e-14
    for t, a, b in zip(spec, dev, host):
        n = insts[t[0]][0]
        sa, sb = _sites(catn, a, n), _sites(catn, b, n)
        assert sa.keys() == sb.keys()
        for k in sa:
            assert sa[k].shape == sb[k].shape
            assert np.allclose(sa[k], sb[k], rtol=tol, atol=0), (t, k)
    la, _, _ = catn.contraction_batch(dev)
    lb, _, _ = catn.contraction_batch(host)
    for t, a, b, x, y in zip(spec, dev, host, la, lb):
        assert a.edge_log() == b.edge_log()
        assert abs(x - y) <= (1e-6 if dtype != "f64" else 1e-12) * max(1.0, abs(y)), t


def test_device_built_gradient(env):
    """The chain rule of the reverse sweep works on device-built sites."""
    catn, ctx = env
    n, e, w, h = _instances()[1]
    a = catn.ising_network_batch([(n, e, w, h)], [(0, 0.7)], ctx=ctx, Dmax=-1, chi=10_000)[0]
    b = catn.ising_network(n, e, w, h, 0.7, ctx=ctx, Dmax=-1, chi=10_000)
    la, _, _ = catn.gradient_batch([a])
    lb, _, _ = catn.gradient_batch([b])
    dwa, dha, _ = catn.ising_gradient(a, len(e), n)
    dwb, dhb, _ = catn.ising_gradient(b, len(e), n)
    assert abs(la[0] - lb[0]) < 1e-12 * abs(lb[0])
    assert np.abs(dwa - dwb).max() < 1e-12 and np.abs(dha - dhb).max() < 1e-12
    mag, cor = X.brute_observables(n, e, w, h, 0.7)
    assert np.abs(dha - 0.7 * mag).max() < 1e-10


def test_pinned_batch_matches_pinning_route(env):
    """The plain form's pinning networks built in one packed batch give the
    reference's magnetization and correlation (src/ising.jl:263-339)."""
    catn, ctx = env
    n, e, w, h = _instances()[0]
    nets = catn.ising_network_batch([(n, e, w, h)], catn.pinned_networks(n, e, w, h, 0.8), ctx=ctx,
                                    Dmax=-1, chi=10_000)
    lz, _, _ = catn.contraction_batch(nets)
    mag = np.array([np.exp(lz[2 + 2 * k] - lz[0]) - np.exp(lz[1 + 2 * k] - lz[0]) for k in range(n)])
    off = 1 + 2 * n
    cor = np.array([np.exp(lz[off + 4 * q:off + 4 * q + 4] - lz[0]) @ [1, -1, -1, 1]
                    for q in range(len(e))])
    m_ref, c_ref = X.brute_observables(n, e, w, h, 0.8)
    assert np.abs(mag - m_ref).max() < 1e-10 and np.abs(cor - c_ref).max() < 1e-10
