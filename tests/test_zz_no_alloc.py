"""catn_tsvd_fixed: tsvd on caller-owned device memory with no allocation
(the library's fixed device region holds everything else). Checked through
the C ABI against catn_tsvd on the same matrices, two calls in a row, the
second with a smaller matrix (the region keeps the first call's contents)."""
import ctypes as C

import numpy as np
import pytest


def _mats():
    r = np.random.default_rng(41)
    A1 = r.standard_normal((120, 90)) @ np.diag(np.logspace(0, -18, 90)) @ r.standard_normal((90, 90))
    A2 = r.standard_normal((40, 70))
    return [(A1, 1e-12, 0), (A2, 1e-15, 30)]


def _reference(catn, ctx, A, cutoff, maxdim):
    U, _, V, disc, S = catn.tsvd(A, cutoff=cutoff, maxdim=maxdim or None, ctx=ctx)
    return U, S, V, U.shape[1], disc


def _fixed(ctx, m, n, dA, cutoff, maxdim, dU, dS, dV):
    k = C.c_int32()
    disc = C.c_double()
    ctx._check(ctx.lib.catn_tsvd_fixed(ctx.h, m, n, dA, cutoff, maxdim, C.byref(k), dU, dS, dV,
                                       C.byref(disc)))
    return k.value, disc.value


def _check(got, ref):
    U, S, V, k, disc = ref
    Uf, Sf, Vf, kf, df = got
    assert kf == k
    assert np.array_equal(Sf, S)
    assert np.array_equal(Uf, U) and np.array_equal(Vf, V)
    assert df == disc


def test_fixed_region_tsvd_emulator(catn, emu_ctx):
    """The same entry point on the CPU emulator (host buffers stand in for
    device memory)."""
    for A, cutoff, maxdim in _mats():
        m, n = A.shape
        ref = _reference(catn, emu_ctx, A, cutoff, maxdim)
        a = np.asfortranarray(A)
        U = np.zeros((m, min(m, n)), order="F")
        V = np.zeros((n, min(m, n)), order="F")
        S = np.zeros(min(m, n))
        k, disc = _fixed(emu_ctx, m, n, a.ctypes.data, cutoff, maxdim, U.ctypes.data,
                         S.ctypes.data, V.ctypes.data)
        _check((U.reshape(-1, order="F")[:m * k].reshape((m, k), order="F"), S,
                V.reshape(-1, order="F")[:n * k].reshape((n, k), order="F"), k, disc), ref)


def test_fixed_region_refuses_too_large(catn, emu_ctx):
    with pytest.raises(catn.CatnError, match="fixed device region"):
        _fixed(emu_ctx, 2000, 2000, None, 1e-15, 0, None, None, None)


class _Hip:
    """Device buffers for the test through the HIP runtime the library itself
    is linked against (the test allocates; the entry point under test does
    not)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")   # the soname libcatn_hip.so links
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p.value

    def put(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        d = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return d

    def get(self, d, count):
        out = np.zeros(count)
        assert self.rt.hipMemcpy(out.ctypes.data, d, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


@pytest.mark.gpu
def test_fixed_region_tsvd_gpu(catn, gpu_ctx):
    """On the MI355X: equal to catn_tsvd (same kernel, same launch shape) on
    both calls, the second with a smaller matrix."""
    hip = _Hip()
    try:
        for A, cutoff, maxdim in _mats():
            m, n = A.shape
            print(f"tsvd_fixed {m}x{n} cutoff={cutoff} maxdim={maxdim}", flush=True)
            ref = _reference(catn, gpu_ctx, A, cutoff, maxdim)
            r = min(m, n)
            dA = hip.put(A.ravel(order="F"))
            dU, dV, dS = hip.alloc(8 * m * r), hip.alloc(8 * n * r), hip.alloc(8 * r)
            k, disc = _fixed(gpu_ctx, m, n, dA, cutoff, maxdim, dU, dS, dV)
            U = hip.get(dU, m * k).reshape((m, k), order="F")
            V = hip.get(dV, n * k).reshape((n, k), order="F")
            _check((U, hip.get(dS, r), V, k, disc), ref)
    finally:
        hip.free()
