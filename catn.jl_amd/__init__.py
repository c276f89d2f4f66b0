"""catn_amd — MI355X contraction engine for CATN.jl networks (Python host mirror).

Mirrors the reference's public API (xuanzhaogao/CATN.jl, src/CATN.jl:15-32)
over the C ABI of `libcatn_hip.so` (include/catn_hip.h):

  ising_network(n, edges, w, h, beta; ...)  src/ising.jl:20-227
  TensorNetwork(tensors, ixs; ...)          src/tensor_network.jl:41-140
  contraction(tn) -> (lnZ, error, psi)      contraction!  src/tensor_network.jl:556-717
  contraction_batch([tn...])                many networks per persistent launch
  free_energy(tn) -> (lnZ/n, F)             src/ising.jl:239-244
  magnetization(...), correlation(...)      src/ising.jl:263-339 (pinning; all
                                            pinned networks run as one batch), or
                                            method="adjoint": one reverse sweep
  gradient_batch([tn...])                   lnZ + hand-written adjoint per network
  ising_gradient(tn) -> (dw, dh, dbeta)     d lnZ / d(w, h, beta) after the sweep
  lnz_and_grad(n, edges, w, h, beta; ...)   one ising solve: lnZ and its gradient
                                            (lifted: exact at rank-deficient splits)
  tsvd(A; cutoff, maxdim)                   src/linalg_utils.jl:15-28
  MPSNet: hand-built MPSNode primitives      src/mps_node.jl (swap!, cano_to!, ...)

The numeric work runs only in the HIP library. There is no CPU fallback: if
the library is missing or no GPU is present, creating a Context raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcatn_hip.so")

__all__ = [
    "Context", "TensorNetwork", "MPSNet", "ising_network", "tensor_network", "contraction",
    "result_tensor",
    "contraction_batch", "free_energy", "magnetization", "correlation", "tsvd",
    "default_context", "LIB_PATH", "CatnError", "gradient_batch", "ising_gradient",
    "lnz_and_grad", "lnz_and_grad_batch", "rsvd", "lift_direction", "lifted_pair",
    "combine_lifted", "richardson", "lift_plan", "LIFT", "autotune", "AUTOTUNE_WIDE",
]


class CatnError(RuntimeError):
    pass


class _Params(C.Structure):
    _fields_ = [
        ("Dmax", C.c_int32), ("chi", C.c_int32), ("cutoff", C.c_double),
        ("norm_method", C.c_int32), ("select", C.c_int32), ("reverse", C.c_int32),
        ("svdopt", C.c_int32), ("swapopt", C.c_int32), ("compress", C.c_int32),
        ("cut_bond", C.c_int32), ("seed", C.c_int32), ("arena_elems", C.c_int64),
        ("scratch_elems", C.c_int64), ("rsvd_both", C.c_int32), ("rsvd_any", C.c_int32),
        ("precision", C.c_int32), ("loops", C.c_int32), ("braket_dmax", C.c_int32),
        ("eltype", C.c_int32),
    ]


_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)

_SIGS = {
    "catn_params_default": (None, [C.POINTER(_Params)]),
    "catn_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "catn_destroy": (C.c_int, [_P]),
    "catn_last_error": (C.c_char_p, []),
    "catn_backend": (C.c_char_p, [_P]),
    "catn_ising_network": (C.c_int, [_P, C.c_int32, C.c_int32, _I32P, _DP, _DP, C.c_double,
                                     C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                     C.POINTER(_Params), C.POINTER(_P)]),
    "catn_ising_network_batch": (C.c_int, [_P, C.c_int32, _I32P, _I32P, _I32P, _DP, _DP,
                                           C.c_int32, _I32P, _DP, _I32P, C.POINTER(_Params),
                                           C.POINTER(_P)]),
    "catn_tensor_network": (C.c_int, [_P, C.c_int32, _I32P, _I64P, _DP, _I64P,
                                      C.POINTER(_Params), C.POINTER(_P)]),
    "catn_net_create": (C.c_int, [_P, C.POINTER(_Params), C.POINTER(_P)]),
    "catn_net_add_node": (C.c_int, [_P, C.c_int32, C.c_int32, _I32P, _DP, _I32P, C.c_int32]),
    "catn_net_add_tensor_node": (C.c_int, [_P, C.c_int32, C.c_int32, _I64P, _DP, _I32P]),
    "catn_net_upload": (C.c_int, [_P]),
    "catn_net_prepare_gradient": (C.c_int, [_P, C.c_int32]),
    "catn_net_set_fault": (C.c_int, [_P, C.c_int32]),
    "catn_net_free": (C.c_int, [_P]),
    "catn_contraction": (C.c_int, [_P, _DP, _DP, _DP]),
    "catn_contraction_batch": (C.c_int, [_P, C.c_int32, C.POINTER(_P), _DP, _DP, _DP]),
    "catn_gradient_batch": (C.c_int, [_P, C.c_int32, C.POINTER(_P), C.c_int32, _DP, _DP, _DP]),
    "catn_ising_gradient": (C.c_int, [_P, _DP, _DP, _DP]),
    "catn_net_site_adjoint": (C.c_int, [_P, C.c_int32, _DP, C.c_int64]),
    "catn_net_tensor_adjoint": (C.c_int, [_P, C.c_int32, _DP, C.c_int64]),
    "catn_net_bwd_profile": (C.c_int, [_P, _I64P, _DP, _I32P]),
    "catn_net_tape": (C.c_int, [_P, _I32P, _DP, C.c_int32]),
    "catn_bwd_kernel_stats": (C.c_int, [_P, _I64P, _DP, C.c_int32]),
    "catn_net_info": (C.c_int, [_P, _DP, _I32P, _I32P, _I32P, _DP]),
    "catn_net_edge_log": (C.c_int, [_P, _I32P, _I32P]),
    "catn_net_plan_trace": (C.c_int, [_P, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P]),
    "catn_net_mem": (C.c_int, [_P, _I64P, _I64P]),
    "catn_net_profile": (C.c_int, [_P, _I64P, _I32P, _I64P, _I32P, _DP]),
    "catn_net_work": (C.c_int, [_P, _DP, _DP, _I64P]),
    "catn_net_svd_profile": (C.c_int, [_P, _I64P]),
    "catn_kernel_stats": (C.c_int, [_P, _I64P, _DP, C.c_int32]),
    "catn_host_stats": (C.c_int, [_P, _DP, _DP]),
    "catn_set_threads": (C.c_int, [_P, C.c_int32]),
    "catn_braket_network": (C.c_int, [_P, C.c_int32, _I32P, _I64P, _DP, _I64P, _P, _P]),
    "catn_set_launch": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "catn_get_launch": (C.c_int, [_P, _I32P, _I32P, _I32P, _I32P, _I32P]),
    "catn_set_batch_streams": (C.c_int, [_P, C.c_int32]),
    "catn_node_swap": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _DP]),
    "catn_node_cano_to": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "catn_node_left_canonical": (C.c_int, [_P, C.c_int32]),
    "catn_node_move": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _DP]),
    "catn_node_move2tail": (C.c_int, [_P, C.c_int32, C.c_int32, _DP]),
    "catn_node_move2head": (C.c_int, [_P, C.c_int32, C.c_int32, _DP]),
    "catn_node_reverse": (C.c_int, [_P, C.c_int32]),
    "catn_node_merge": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _DP]),
    "catn_node_compress": (C.c_int, [_P, C.c_int32, C.c_int32, _DP]),
    "catn_node_eat": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _DP, _DP, _DP]),
    "catn_cut_bondim": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _DP]),
    "catn_node_get": (C.c_int, [_P, C.c_int32, _I32P, _I32P, _I32P, _I32P, C.c_int32]),
    "catn_node_site": (C.c_int, [_P, C.c_int32, C.c_int32, _DP, C.c_int64]),
    "catn_result_tensor": (C.c_int, [_P, _DP, C.c_int64, _I32P, _I32P, _I32P, C.c_int32]),
    "catn_tsvd": (C.c_int, [_P, C.c_int32, C.c_int32, _DP, C.c_double, C.c_int32, _I32P, _DP,
                            _DP, _DP, _DP]),
    "catn_tsvd_fixed": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_void_p, C.c_double, C.c_int32,
                                  _I32P, C.c_void_p, C.c_void_p, C.c_void_p, _DP]),
    "catn_rsvd": (C.c_int, [_P, C.c_int32, C.c_int32, _DP, C.c_int32, C.c_int32, C.c_int32,
                            C.c_int32, _I32P, _DP, _DP, _DP]),
}

EXPORTED_SYMBOLS = tuple(_SIGS)


def load_library(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise CatnError(f"libcatn_hip not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _dptr(a):
    return a.ctypes.data_as(_DP)


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _elems(a, cplx: bool):
    """Column-major data of an array as the C ABI takes it: doubles, or for a
    complex network interleaved (re, im) doubles (Julia's ComplexF64 layout)."""
    a = np.asarray(a)
    if cplx:
        return np.ascontiguousarray(a.astype(np.complex128).ravel(order="F")).view(np.float64)
    if np.iscomplexobj(a):
        raise CatnError("complex tensor given to a real network (eltype='real')")
    return np.ascontiguousarray(a.astype(np.float64).ravel(order="F"))


def _from_elems(buf, cplx: bool):
    return buf.view(np.complex128) if cplx else buf


class Context:
    """One device context (one HIP stream). `lib` selects the shared library
    (default: the in-tree libcatn_hip.so)."""

    def __init__(self, device: int = 0, lib: Optional[str] = None):
        self.lib = load_library(lib or LIB_PATH)
        h = _P()
        self._check(self.lib.catn_create(C.byref(h), int(device)))
        self.h = h
        self.backend = self.lib.catn_backend(h).decode()

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.catn_last_error().decode()
            raise CatnError(f"libcatn (rc={rc}): {msg}")

    def kernel_stats(self, reset=False):
        """(launches, seconds) of the interpreter kernel on this context."""
        n, sec = C.c_int64(), C.c_double()
        self._check(self.lib.catn_kernel_stats(self.h, C.byref(n), C.byref(sec), int(bool(reset))))
        return n.value, sec.value

    def bwd_kernel_stats(self, reset=False):
        """(launches, seconds) of the reverse-sweep kernel on this context."""
        n, sec = C.c_int64(), C.c_double()
        self._check(self.lib.catn_bwd_kernel_stats(self.h, C.byref(n), C.byref(sec),
                                                   int(bool(reset))))
        return n.value, sec.value

    def host_stats(self):
        """(planning seconds, device round-trip seconds) since the last reset."""
        a, b = C.c_double(), C.c_double()
        self._check(self.lib.catn_host_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_threads(self, threads: int):
        """Workgroup size of the interpreter kernel (64, 128, 256, 512 or 1024)."""
        self._check(self.lib.catn_set_threads(self.h, int(threads)))

    def set_launch(self, threads: int = 256, lds_doubles: int = 0, panel: int = 0,
                   jacobi_lanes: int = 0):
        """Launch shape of the interpreter: lanes per workgroup, LDS budget per
        workgroup in doubles (0 = automatic by batch size), Householder panel
        width (0 = 16) and global-Jacobi lanes per column pair (0 = 16)."""
        self._check(self.lib.catn_set_launch(self.h, int(threads), int(lds_doubles), int(panel),
                                             int(jacobi_lanes)))

    def set_batch_streams(self, streams: int):
        """Concurrent groups (streams) of a forward contraction_batch of
        distinct networks: 1 to 16 (default 4, the HIP runtime's default
        hardware queues; keep it at most GPU_MAX_HW_QUEUES); 1 makes every
        launch batch-wide."""
        self._check(self.lib.catn_set_batch_streams(self.h, int(streams)))

    def launch(self):
        """(threads, lds_doubles, lds_limit, panel, jacobi_lanes) of the current
        launch shape."""
        v = [C.c_int32() for _ in range(5)]
        self._check(self.lib.catn_get_launch(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def close(self):
        if getattr(self, "h", None):
            self.lib.catn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DEFAULT: Optional[Context] = None


def default_context() -> Context:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


def _params(ctx, Dmax=32, chi=32, cutoff=1e-15, norm_method=1, select=1, reverse=True,
            svdopt=True, swapopt=True, compress=False, cut_bond=False, seed=1,
            arena_elems=0, scratch_elems=0, rsvd_both=7000, rsvd_any=20000, dtype="f64",
            loops=False, eltype="real"):
    p = _Params()
    ctx.lib.catn_params_default(C.byref(p))
    p.Dmax, p.chi, p.cutoff = int(Dmax), int(chi), float(cutoff)
    p.norm_method, p.select = int(norm_method), int(select)
    p.reverse, p.svdopt, p.swapopt = int(bool(reverse)), int(bool(svdopt)), int(bool(swapopt))
    p.compress, p.cut_bond, p.seed = int(bool(compress)), int(bool(cut_bond)), int(seed)
    p.arena_elems, p.scratch_elems = int(arena_elems), int(scratch_elems)
    p.rsvd_both, p.rsvd_any = int(rsvd_both), int(rsvd_any)
    levels = {"f64": 0, "f32": 1, "f32x6": 2}
    if dtype not in levels:
        raise CatnError(f"dtype must be 'f64' (Float64), 'f32' (Float32 storage, split-BF16 "
                        f"3-pass theta) or 'f32x6' (6-pass theta), got {dtype!r}")
    p.precision = levels[dtype]
    p.loops = int(bool(loops))
    kinds = {"real": 0, "complex": 1}
    if eltype not in kinds:
        raise CatnError(f"eltype must be 'real' or 'complex', got {eltype!r}")
    p.eltype = kinds[eltype]
    return p


class TensorNetwork:
    """Handle of a device-resident network (TensorNetwork{T,AT} of the
    reference with AT = device storage)."""

    def __init__(self, ctx: Context, handle, cplx: bool = False):
        self.ctx = ctx
        self.h = handle
        self.cplx = bool(cplx)   # complex element type (ComplexF64)
        self.lnZ = 0.0
        self.psi = 1.0

    def free(self):
        if getattr(self, "h", None):
            self.ctx.lib.catn_net_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def info(self):
        sign = np.zeros(2)
        md, ni, n = C.c_int32(), C.c_int32(), C.c_int32()
        beta = C.c_double()
        self.ctx._check(self.ctx.lib.catn_net_info(self.h, _dptr(sign), C.byref(md),
                                                   C.byref(ni), C.byref(n), C.byref(beta)))
        sg = complex(sign[0], sign[1]) if self.cplx else float(sign[0])
        return dict(sign=sg, maxdim_intermediate=md.value, num_isolated=ni.value,
                    n=n.value, beta=beta.value)

    @property
    def n(self):
        return self.info()["n"]

    @property
    def beta(self):
        return self.info()["beta"]

    @property
    def sign(self):
        return self.info()["sign"]

    @property
    def maxdim_intermediate(self):
        return self.info()["maxdim_intermediate"]

    @property
    def num_isolated(self):
        return self.info()["num_isolated"]

    def edge_log(self) -> List[Tuple[int, int]]:
        cnt = C.c_int32(0)
        self.ctx._check(self.ctx.lib.catn_net_edge_log(self.h, None, C.byref(cnt)))
        buf = np.zeros(2 * max(cnt.value, 1), dtype=np.int32)
        cap = C.c_int32(cnt.value)
        self.ctx._check(self.ctx.lib.catn_net_edge_log(self.h, buf.ctypes.data_as(_I32P), C.byref(cap)))
        return [(int(buf[2 * k]), int(buf[2 * k + 1])) for k in range(cnt.value)]

    def plan_trace(self):
        """Structural trace of contraction! (src/tensor_network.jl:574-708):
        steps: (sel_i, sel_j, i, j, reversed_i, reversed_j, idx_j_in_i,
        idx_i_in_j) per step; dups: (step, k, cross) per duplicate bond;
        cuts: the kept count of every Dmax cut."""
        n = [C.c_int32(0) for _ in range(3)]
        self.ctx._check(self.ctx.lib.catn_net_plan_trace(self.h, None, C.byref(n[0]), None,
                                                         C.byref(n[1]), None, C.byref(n[2])))
        bufs = [np.zeros(w * max(k.value, 1), dtype=np.int32) for w, k in zip((8, 3, 1), n)]
        caps = [C.c_int32(k.value) for k in n]
        self.ctx._check(self.ctx.lib.catn_net_plan_trace(
            self.h, bufs[0].ctypes.data_as(_I32P), C.byref(caps[0]), bufs[1].ctypes.data_as(_I32P),
            C.byref(caps[1]), bufs[2].ctypes.data_as(_I32P), C.byref(caps[2])))
        steps = [tuple(int(x) for x in bufs[0][8 * k:8 * k + 8]) for k in range(n[0].value)]
        dups = [tuple(int(x) for x in bufs[1][3 * k:3 * k + 3]) for k in range(n[1].value)]
        return {"steps": steps, "dups": dups, "cuts": [int(x) for x in bufs[2][:n[2].value]]}

    def cut_kept(self):
        """Kept count of every Dmax cut of the contraction, in order."""
        return self.plan_trace()["cuts"]

    OPNAMES = ["nop", "cano_r", "cano_l", "swap", "permute321", "eat_a", "eat_b", "eat_c",
               "normalize", "merge_fuse", "cut", "compress_opt", "compress", "raw_split",
               "scalar_lognorm", "tsvd"]

    def profile(self):
        """Device op profile: {opname: (count, seconds)} and per SVD size bucket."""
        t = np.zeros(16, np.int64)
        c = np.zeros(16, np.int32)
        st = np.zeros(8, np.int64)
        sc = np.zeros(8, np.int32)
        hz = C.c_double()
        self.ctx._check(self.ctx.lib.catn_net_profile(
            self.h, t.ctypes.data_as(_I64P), c.ctypes.data_as(_I32P), st.ctypes.data_as(_I64P),
            sc.ctypes.data_as(_I32P), C.byref(hz)))
        ops = {self.OPNAMES[k]: (int(c[k]), float(t[k]) / hz.value) for k in range(16) if c[k]}
        buckets = {f"<={4 << k}" if k < 7 else ">256": (int(sc[k]), float(st[k]) / hz.value)
                   for k in range(8) if sc[k]}
        sub = np.zeros(16, np.int64)
        self.ctx._check(self.ctx.lib.catn_net_svd_profile(self.h, sub.ctypes.data_as(_I64P)))
        hzv = hz.value
        buckets["_large_svd"] = {"qr_s": sub[0] / hzv, "jacobi_s": sub[1] / hzv,
                                 "finish_s": sub[2] / hzv, "sweeps": int(sub[3]),
                                 "global_sweeps": int(sub[11]), "fallbacks": int(sub[12])}
        buckets["_certified_qr"] = {"qr_s": sub[4] / hzv, "inverse_s": sub[5] / hzv,
                                    "qform_s": sub[6] / hzv}
        buckets["_householder"] = {"panel_s": sub[8] / hzv, "update_s": sub[9] / hzv,
                                   "apply_q_s": sub[10] / hzv}
        buckets["_theta_s"] = sub[7] / hzv
        return ops, buckets

    def bwd_profile(self):
        """Reverse sweep: ({opname: seconds}, replay deviation, tape length)."""
        t = np.zeros(16, np.int64)
        dev = C.c_double()
        tl = C.c_int32()
        self.ctx._check(self.ctx.lib.catn_net_bwd_profile(self.h, t.ctypes.data_as(_I64P),
                                                          C.byref(dev), C.byref(tl)))
        hz = C.c_double()
        self.ctx._check(self.ctx.lib.catn_net_profile(self.h, None, None, None, None, C.byref(hz)))
        ops = {self.OPNAMES[k]: float(t[k]) / hz.value for k in range(16) if t[k]}
        return ops, dev.value, tl.value

    def tape(self):
        """Tape records as an (n, 16) int array (see catn_net_tape) and their scalars."""
        _, _, tl = self.bwd_profile()
        rec = np.zeros((max(tl, 1), 16), np.int32)
        val = np.zeros(max(tl, 1))
        self.ctx._check(self.ctx.lib.catn_net_tape(self.h, rec.ctypes.data_as(_I32P), _dptr(val), tl))
        return rec[:tl], val[:tl]

    def site_adjoint(self, slot: int, size: int) -> np.ndarray:
        """d lnZ / d (initial tensor of `slot`), flat column-major."""
        out = np.zeros(size)
        self.ctx._check(self.ctx.lib.catn_net_site_adjoint(self.h, int(slot), _dptr(out), size))
        return out

    def tensor_gradient(self, k: int) -> np.ndarray:
        """After gradient_batch on a network from tensor_network: d lnZ / d T
        for its k-th tensor (0-based), shaped like the tensor, lnZ = log|Z|."""
        shapes = getattr(self, "tensor_shapes", None)
        if shapes is None:
            raise CatnError("tensor_gradient: the network was not built by tensor_network")
        if not 0 <= k < len(shapes):
            raise CatnError(f"tensor_gradient: no tensor {k} in this network")
        size = int(np.prod(shapes[k], dtype=np.int64))
        out = np.zeros(size)
        self.ctx._check(self.ctx.lib.catn_net_tensor_adjoint(self.h, int(k), _dptr(out), size))
        return out.reshape(shapes[k], order="F")

    def prepare_gradient(self, tape_cap: int = 0):
        """Allocate the reverse sweep's tape and adjoint arena now, so that a
        later gradient_batch (same tape_cap) allocates nothing."""
        self.ctx._check(self.ctx.lib.catn_net_prepare_gradient(self.h, int(tape_cap)))

    def set_fault(self, flags: int):
        """Test hook of tsvd's fallback: flags = 1 makes every SVD of this
        network run its fallback path (catn_net_set_fault)."""
        self.ctx._check(self.ctx.lib.catn_net_set_fault(self.h, int(flags)))

    def work(self):
        """(algorithmic flops, algorithmic bytes, scratch high-water mark)."""
        f, b, h = C.c_double(), C.c_double(), C.c_int64()
        self.ctx._check(self.ctx.lib.catn_net_work(self.h, C.byref(f), C.byref(b), C.byref(h)))
        return f.value, b.value, h.value

    def mem(self):
        u, c = C.c_int64(), C.c_int64()
        self.ctx._check(self.ctx.lib.catn_net_mem(self.h, C.byref(u), C.byref(c)))
        return u.value, c.value


def ising_network(n, edges, weights, fields, beta, *, pos1=0, val1=0, pos2=0, val2=0,
                  ctx: Optional[Context] = None, **kw) -> TensorNetwork:
    """ising_network (src/ising.jl:20-227)."""
    ctx = ctx or default_context()
    e = _i32(np.asarray(edges, dtype=np.int32).reshape(-1, 2)) if len(edges) else np.zeros((0, 2), np.int32)
    w = _f64(weights) if len(weights) else np.zeros(0)
    h = _f64(fields)
    p = _params(ctx, **kw)
    out = _P()
    ctx._check(ctx.lib.catn_ising_network(
        ctx.h, int(n), int(e.shape[0]), e.ctypes.data_as(_I32P), _dptr(w), _dptr(h),
        float(beta), int(pos1), int(val1), int(pos2), int(val2), C.byref(p), C.byref(out)))
    return TensorNetwork(ctx, out)


def ising_network_batch(instances, networks, *, ctx: Optional[Context] = None,
                        **kw) -> List[TensorNetwork]:
    """ising_network for a batch with the site values built on the device
    (catn_ising_network_batch). instances: (n, edges, weights, fields) tuples;
    networks: (instance index, beta) or (instance index, beta, (pos1, val1,
    pos2, val2)) tuples. Returns one network per entry of `networks`."""
    ctx = ctx or default_context()
    ns = _i32([int(q[0]) for q in instances])
    nes = _i32([len(q[1]) for q in instances])
    e = _i32(np.concatenate([np.asarray(q[1], np.int32).reshape(-1, 2) for q in instances
                             if len(q[1])] or [np.zeros((0, 2), np.int32)]).reshape(-1))
    w = _f64(np.concatenate([np.asarray(q[2], float).reshape(-1) for q in instances] or [[]]))
    h = _f64(np.concatenate([np.asarray(q[3], float).reshape(-1) for q in instances]))
    inst = _i32([int(t[0]) for t in networks])
    beta = _f64([float(t[1]) for t in networks])
    pins = _i32([list(t[2]) if len(t) > 2 else [0, 0, 0, 0] for t in networks]).reshape(-1)
    p = _params(ctx, **kw)
    out = (_P * len(networks))()
    ctx._check(ctx.lib.catn_ising_network_batch(
        ctx.h, len(instances), ns.ctypes.data_as(_I32P), nes.ctypes.data_as(_I32P),
        e.ctypes.data_as(_I32P), _dptr(w), _dptr(h), len(networks), inst.ctypes.data_as(_I32P),
        _dptr(beta), pins.ctypes.data_as(_I32P), C.byref(p), out))
    return [TensorNetwork(ctx, C.c_void_p(out[k])) for k in range(len(networks))]


def pinned_networks(n, edges, weights, fields, beta, *, inst=0):
    """The network list of the pinning route for one instance (src/ising.jl:
    263-339): the unpinned network, then (k, 1), (k, 2) for every spin k, then
    (a, v1, b, v2) for every edge (a, b) and v1, v2 in (1, 2)."""
    nets = [(inst, beta)]
    for k in range(1, n + 1):
        for v in (1, 2):
            nets.append((inst, beta, (k, v, 0, 0)))
    for (a, b) in edges:
        for v1 in (1, 2):
            for v2 in (1, 2):
                nets.append((inst, beta, (a, v1, b, v2)))
    return nets


def tensor_network(tensors: Sequence[np.ndarray], ixs: Sequence[Sequence], *,
                   ctx: Optional[Context] = None, **kw) -> TensorNetwork:
    """TensorNetwork(tensors, ixs; ...) (src/tensor_network.jl:41-140). Tensors
    are taken with Julia (column-major) index order: t[i1, i2, ...]."""
    ctx = ctx or default_context()
    # complex tensors make a complex network (TensorNetwork{ComplexF64}) unless
    # eltype is given
    cplx = kw.pop("eltype", None)
    cplx = any(np.iscomplexobj(t) for t in tensors) if cplx is None else cplx == "complex"
    lab_ids = {}
    labels, ndims, dims, data = [], [], [], []
    for t, ix in zip(tensors, ixs):
        t = np.asarray(t)
        if t.ndim != len(ix):
            raise CatnError("tensor rank does not match its label list")
        ndims.append(t.ndim)
        dims.extend(t.shape)
        data.append(_elems(t, cplx))
        for l in ix:
            if l not in lab_ids:
                lab_ids[l] = len(lab_ids)
            labels.append(lab_ids[l])
    nd = _i32(ndims)
    dm = np.ascontiguousarray(np.asarray(dims, dtype=np.int64))
    lb = np.ascontiguousarray(np.asarray(labels, dtype=np.int64))
    dt = _f64(np.concatenate(data) if data else np.zeros(0))
    p = _params(ctx, eltype="complex" if cplx else "real", **kw)
    out = _P()
    ctx._check(ctx.lib.catn_tensor_network(ctx.h, len(tensors), nd.ctypes.data_as(_I32P),
                                           dm.ctypes.data_as(_I64P), _dptr(dt),
                                           lb.ctypes.data_as(_I64P), C.byref(p), C.byref(out)))
    tn = TensorNetwork(ctx, out, cplx)
    tn.tensor_shapes = [np.asarray(t).shape for t in tensors]
    # open labels in order of first appearance: the device's sentinels -1, -2, ...
    count = {}
    for ix in ixs:
        for l in ix:
            count[l] = count.get(l, 0) + 1
    tn.open_labels = [l for l in lab_ids if count[l] == 1]
    return tn


TensorNetwork_ = tensor_network


def braket_network(tensors: Sequence[np.ndarray], ixs: Sequence[Sequence], *,
                   ctx: Optional[Context] = None, chi: int = 64, **kw) -> TensorNetwork:
    """braket_network(tensors, ixs; chi, cutoff, norm_method, svdopt, swapopt)
    (src/braket.jl:369-496): the double-layer network of <psi|psi>, the bra
    chains built from conj(T); complex tensors make a complex network. contraction() of the result returns (ln <psi|psi>, error, psi).
    Tensors are taken with Julia (column-major) index order. loops=True accepts
    graphs with cycles, which the reference rejects: when an eat closes a
    cycle, the doubled bonds to the shared neighbour are merged (ket and bra
    legs separately, the older bond as the fast index), and with Dmax > 0 a
    merged bond wider than Dmax is cut as cut_bondim_opt! does."""
    ctx = ctx or default_context()
    cplx = kw.pop("eltype", None)
    cplx = any(np.iscomplexobj(t) for t in tensors) if cplx is None else cplx == "complex"
    lab_ids = {}
    labels, ndims, dims, data = [], [], [], []
    for t, ix in zip(tensors, ixs):
        t = np.asarray(t)
        if t.ndim != len(ix):
            raise CatnError("tensor rank does not match its label list")
        ndims.append(t.ndim)
        dims.extend(t.shape)
        data.append(_elems(t, cplx))
        for l in ix:
            if l not in lab_ids:
                lab_ids[l] = len(lab_ids)
            labels.append(lab_ids[l])
    nd = _i32(ndims)
    dm = np.ascontiguousarray(np.asarray(dims, dtype=np.int64))
    lb = np.ascontiguousarray(np.asarray(labels, dtype=np.int64))
    dt = _f64(np.concatenate(data) if data else np.zeros(0))
    braket_dmax = kw.pop("Dmax", -1)
    p = _params(ctx, chi=chi, eltype="complex" if cplx else "real", **kw)
    p.braket_dmax = int(braket_dmax)
    out = _P()
    ctx._check(ctx.lib.catn_braket_network(ctx.h, len(tensors), nd.ctypes.data_as(_I32P),
                                           dm.ctypes.data_as(_I64P), _dptr(dt),
                                           lb.ctypes.data_as(_I64P), C.byref(p), C.byref(out)))
    return TensorNetwork(ctx, out, cplx)


def contraction(tn: TensorNetwork):
    """contraction!(tn) -> (lnZ, error, psi); psi is complex for a complex network."""
    lz, err = C.c_double(), C.c_double()
    psi = np.zeros(2)
    tn.ctx._check(tn.ctx.lib.catn_contraction(tn.h, C.byref(lz), C.byref(err), _dptr(psi)))
    ps = complex(psi[0], psi[1]) if tn.cplx else float(psi[0])
    tn.lnZ, tn.psi = lz.value, ps
    return lz.value, err.value, ps


def result_tensor(tn: TensorNetwork):
    """result_tensor(tn) (src/tensor_network.jl:504-523) after contraction(tn):
    (array, labels). array is the dense tensor of the one surviving open-leg
    node (sites fused on the device), its legs in that node's MPS order, a
    permutation of the open labels given in `labels`; the full open result is
    array * exp(lnZ) * psi. A closed network gives a 0-d array holding 1.0 and
    no labels. More than one open-leg node raises CatnError, as the reference
    does. The network is consumed."""
    kmax = 64
    dims = np.zeros(kmax, dtype=np.int32)
    labs = np.zeros(kmax, dtype=np.int32)
    rank = C.c_int32()
    # a query call (no output buffer) gives the rank, dims and labels
    tn.ctx._check(tn.ctx.lib.catn_result_tensor(tn.h, None, 0, C.byref(rank),
                                                dims.ctypes.data_as(_I32P),
                                                labs.ctypes.data_as(_I32P), kmax))
    one = np.array(1.0 + 0j) if tn.cplx else np.array(1.0)
    if rank.value == 0:
        return one, []
    ed = 2 if tn.cplx else 1
    out = np.zeros(ed * int(np.prod(dims[: rank.value].astype(np.int64))))
    tn.ctx._check(tn.ctx.lib.catn_result_tensor(tn.h, _dptr(out), out.size, C.byref(rank),
                                                dims.ctypes.data_as(_I32P),
                                                labs.ctypes.data_as(_I32P), kmax))
    out = _from_elems(out, tn.cplx)
    r = rank.value
    if r == 0:
        return one, []
    names = getattr(tn, "open_labels", None)
    labels = [names[-int(v) - 1] if names is not None and v < 0 else int(v) for v in labs[:r]]
    return out.reshape(tuple(int(x) for x in dims[:r]), order="F"), labels


def contraction_batch(tns: Sequence[TensorNetwork]):
    """contraction! of many networks, one workgroup each per launch."""
    if not tns:
        return np.zeros(0), np.zeros(0), np.zeros(0)
    ctx = tns[0].ctx
    cplx = tns[0].cplx
    arr = (_P * len(tns))(*[t.h for t in tns])
    lz, err, psi = np.zeros(len(tns)), np.zeros(len(tns)), np.zeros((2 if cplx else 1) * len(tns))
    ctx._check(ctx.lib.catn_contraction_batch(ctx.h, len(tns), arr, _dptr(lz), _dptr(err), _dptr(psi)))
    psi = _from_elems(psi, cplx)
    for k, t in enumerate(tns):
        t.lnZ, t.psi = float(lz[k]), psi[k].item()
    return lz, err, psi


def gradient_batch(tns: Sequence[TensorNetwork], tape_cap: int = 0):
    """contraction! of many networks with a tape, then the hand-written reverse
    sweep of lnZ (one workgroup per network, one launch). Returns (lnZ, error,
    psi) exactly as contraction_batch; the gradients stay with the networks."""
    if not tns:
        return np.zeros(0), np.zeros(0), np.zeros(0)
    ctx = tns[0].ctx
    arr = (_P * len(tns))(*[t.h for t in tns])
    lz, err, psi = np.zeros(len(tns)), np.zeros(len(tns)), np.zeros(len(tns))
    ctx._check(ctx.lib.catn_gradient_batch(ctx.h, len(tns), arr, int(tape_cap), _dptr(lz),
                                           _dptr(err), _dptr(psi)))
    for k, t in enumerate(tns):
        t.lnZ, t.psi = float(lz[k]), float(psi[k])
    return lz, err, psi


# launch shapes tried by autotune(): (lanes per workgroup, LDS doubles per
# workgroup); 0 = the automatic budget. 20352, 10176, 6784 and 5056 doubles
# let 1, 2, 3 and 4 workgroups share one CU's 160 KB of LDS.
AUTOTUNE_SHAPES = ((256, 0), (256, 5056), (256, 6784), (256, 10176), (256, 20352),
                   (128, 5056), (128, 10176))
# wide workgroups (one per CU): tried when the batch has at most WIDE_MAX_BATCH
# networks. They pay where single ops are large (config 3's 64 lattices:
# 26.7 s at 256 lanes, 16.5 s at 512, 12.6 s at 1024) and cost 10-40% on
# networks of small ops.
AUTOTUNE_WIDE = ((512, 20352), (1024, 20352))
WIDE_MAX_BATCH = 256    # the MI355X's CUs
# tile parameters tried at the fastest shape: (Householder panel width,
# global-Jacobi lanes per column pair)
AUTOTUNE_TILES = ((16, 16), (32, 16), (8, 16), (16, 32), (16, 8), (16, 64))


def autotune(ctx: Context, build, run=None, shapes=AUTOTUNE_SHAPES, tiles=AUTOTUNE_TILES,
             repeats=1, rtol=1e-10, wide=AUTOTUNE_WIDE):
    """Pick the interpreter's launch shape for a workload by timing it.

    build() returns a fresh batch of networks for ctx (a run consumes its
    batch); run(nets) executes one (default contraction_batch) and returns
    lnZ per network. Stage 1 times each (lanes, LDS budget) shape within the
    device's LDS limit at the default tiles, and the `wide` shapes when the
    batch has at most WIDE_MAX_BATCH networks; stage 2 times each (panel width,
    Jacobi lanes) tile pair at the fastest shape. Times are best of
    `repeats`. A candidate counts only if its lnZ agree with the first
    candidate's to rtol (shape and tiles change rounding, never the
    algorithm). The winner is set on ctx. Returns {"threads",
    "lds_doubles", "panel", "jacobi_lanes", "seconds", "table": [...]}."""
    run = run or (lambda nets: contraction_batch(nets)[0])
    limit = ctx.launch()[2]
    table, ref, size = [], [None], [0]

    def timed(threads, lds, panel, jl):
        ctx.set_launch(threads, lds, panel, jl)
        best, lz = math.inf, None
        for _ in range(max(1, int(repeats))):
            nets = build()
            size[0] = len(nets)
            t0 = time.perf_counter()
            lz = np.asarray(run(nets), dtype=float)
            best = min(best, time.perf_counter() - t0)
            for t in nets:
                t.free()
        if ref[0] is None:
            ref[0] = lz
        dev = (float(np.max(np.abs(lz - ref[0]) / np.maximum(np.abs(ref[0]), 1e-300)))
               if lz.size else 0.0)
        row = {"threads": threads, "lds_doubles": lds, "panel": panel, "jacobi_lanes": jl,
               "seconds": best, "lnz_rel_dev": dev}
        table.append(row)
        return row

    for threads, lds in shapes:
        if lds <= limit:
            timed(threads, lds, 16, 16)
    if size[0] <= WIDE_MAX_BATCH:
        for threads, lds in wide:
            if lds <= limit:
                timed(threads, lds, 16, 16)
    ok = [r for r in table if r["lnz_rel_dev"] <= rtol]
    if not ok:
        raise CatnError("autotune: no launch shape ran")
    s1 = min(ok, key=lambda r: r["seconds"])
    for panel, jl in tiles:
        if (panel, jl) != (16, 16) and jl <= s1["threads"]:
            timed(s1["threads"], s1["lds_doubles"], panel, jl)
    win = min((r for r in table if r["lnz_rel_dev"] <= rtol), key=lambda r: r["seconds"])
    ctx.set_launch(win["threads"], win["lds_doubles"], win["panel"], win["jacobi_lanes"])
    return {"threads": win["threads"], "lds_doubles": win["lds_doubles"], "panel": win["panel"],
            "jacobi_lanes": win["jacobi_lanes"], "seconds": win["seconds"], "table": table}


def ising_gradient(tn: TensorNetwork, nedges: int, n: int):
    """(d lnZ/d w, d lnZ/d h, d lnZ/d beta) of an ising_network after
    gradient_batch. Exact contraction gives beta <s_i s_j> and beta <s_i>."""
    dw, dh, db = np.zeros(nedges), np.zeros(n), C.c_double()
    tn.ctx._check(tn.ctx.lib.catn_ising_gradient(tn.h, _dptr(dw), _dptr(dh), C.byref(db)))
    return dw, dh, db.value


# Default lift per storage precision. A lifted singular value is of order
# eta times an instance-dependent factor (down to about 1e-2 on +-J spin
# glasses), and must clear the storage's noise floor (2^-22 of the largest
# for six-pass FP32, 2^-18 for three-pass) by a margin: FP64 takes one
# symmetric pair at 1e-6 (bias ~1e-12); FP32 storage takes pairs at eta and
# 2 eta combined by a Richardson step, which cancels the O(eta^2) bias of the
# larger lift.
LIFT = {"f64": 1e-6, "f32x6": 2e-3, "f32": 1e-2}


def lift_direction(nedges: int, n: int, seed: int = 0x5EED):
    """The fixed perturbation direction of the lifted gradient: one value in
    [-1, 1) per coupling, then one per field, from splitmix64 of the index
    (the same sequence in every host binding)."""
    m = (1 << 64) - 1
    out = np.empty(nedges + n)
    for i in range(nedges + n):
        z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & m
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        z ^= z >> 31
        out[i] = (z >> 11) * 2.0 ** -52 - 1.0
    return out[:nedges], out[nedges:]


def lifted_pair(weights, fields, lift: float):
    """(w + eta_w, h + eta_h), (w - eta_w, h - eta_h) and (eta_w, eta_h), with
    eta = lift * max(1, |w|_inf, |h|_inf) * lift_direction."""
    w = np.asarray(weights, float).reshape(-1)
    h = np.asarray(fields, float).reshape(-1)
    gw, gh = lift_direction(len(w), len(h))
    s = lift * max(1.0, float(np.max(np.abs(w), initial=0.0)), float(np.max(np.abs(h), initial=0.0)))
    ew, eh = s * gw, s * gh
    return (w + ew, h + eh), (w - ew, h - eh), (ew, eh)


def combine_lifted(plus, minus, eta):
    """lnZ and gradient at w from the two runs at w +- eta:
    grad = (g+ + g-)/2 and lnZ = (l+ + l-)/2 - eta.(g+ - g-)/4, both exact to
    O(|eta|^2) (g+ - g- = 2 H eta + O(|eta|^3) removes the curvature term of
    the average)."""
    (lp, dwp, dhp), (lm, dwm, dhm), (ew, eh) = plus, minus, eta
    lz = 0.5 * (lp + lm) - 0.25 * (float(ew @ (dwp - dwm)) + float(eh @ (dhp - dhm)))
    return lz, 0.5 * (dwp + dwm), 0.5 * (dhp + dhm)


def lift_plan(lift: Optional[float], dtype: str = "f64"):
    """The lifts of one gradient solve: () for lift = 0, (eta,) for FP64
    storage, (eta, 2 eta) for FP32 storage (Richardson step)."""
    if lift is None:
        lift = LIFT.get(dtype, LIFT["f64"])
    if not lift >= 0:
        raise CatnError(f"lift must be >= 0, got {lift!r}")
    if lift == 0:
        return ()
    return (lift,) if dtype == "f64" else (lift, 2.0 * lift)


def richardson(results):
    """Combine the lifted solves at eta and 2 eta (each exact to
    c eta^2 + O(eta^4)): gradient (4 g(eta) - g(2 eta)) / 3; lnZ from the
    smaller lift, whose curvature correction already leaves O(eta^4)."""
    if len(results) == 1:
        return results[0]
    (l1, w1, h1), (_, w2, h2) = results
    return l1, (4.0 * w1 - w2) / 3.0, (4.0 * h1 - h2) / 3.0


def lnz_and_grad_batch(instances, *, ctx=None, tape_cap: int = 0,
                       lift: Optional[float] = None, **kw):
    """Many ising instances [(n, edges, w, h, beta), ...] (any sizes and betas)
    in one forward + reverse sweep: [(lnZ, dlnZ/dw, dlnZ/dh), ...].

    The reverse sweep differentiates the executed schedule, with every split's
    kept count held fixed. At an exactly rank-deficient split (a singular value
    that is zero at w, such as a frustrated cycle of +-J couplings gives, and
    that the cutoff drops) a perturbation of w restores that singular value,
    and the exact lnZ has a derivative that the schedule at w does not carry.
    With a lift every instance is run at w +- eta (lifted_pair), where a
    generic eta lifts such zeros above the cutoff so that the schedule is
    smooth there; the average gradient is then the exact one to O(|eta|^2)
    (combine_lifted), and for FP32 storage a second pair at 2 eta removes that
    term (richardson). All runs share the batch. lift=None takes the default
    of the storage precision (LIFT); lift=0 gives the executed schedule's
    derivative at w from one run per instance."""
    ctx = ctx or default_context()
    lifts = lift_plan(lift, kw.get("dtype", "f64"))
    runs = []      # per instance: (n, edges, [(w, h) per network], [eta per pair])
    for (n, e, w, h, b) in instances:
        if not lifts:
            runs.append((n, e, b, [(w, h)], []))
            continue
        pts, etas = [], []
        for s in lifts:
            p, m, eta = lifted_pair(w, h, s)
            pts += [p, m]
            etas.append(eta)
        runs.append((n, e, b, pts, etas))
    nets = [ising_network(n, e, pw, ph, b, ctx=ctx, **kw)
            for (n, e, b, pts, _) in runs for (pw, ph) in pts]
    lz, _, _ = gradient_batch(nets, tape_cap)
    out, q = [], 0
    for (n, e, b, pts, etas) in runs:
        res = []
        for _ in pts:
            dw, dh, _ = ising_gradient(nets[q], len(e), n)
            res.append((float(lz[q]), dw, dh))
            q += 1
        if not etas:
            out.append(res[0])
            continue
        pairs = [combine_lifted(res[2 * k], res[2 * k + 1], etas[k]) for k in range(len(etas))]
        out.append(richardson(pairs))
    return out


def lnz_and_grad(n, edges, weights, fields, beta, *, ctx=None, **kw):
    """lnZ and (d lnZ/d w, d lnZ/d h) of one ising_network by forward
    contraction and reverse sweep (lifted runs unless lift=0; see
    lnz_and_grad_batch)."""
    (l, dw, dh), = lnz_and_grad_batch([(n, edges, weights, fields, beta)], ctx=ctx, **kw)
    return l, dw, dh


def free_energy(tn: TensorNetwork):
    """free_energy (src/ising.jl:239-244)."""
    lnZ, _, _ = contraction(tn)
    n, beta = tn.n, tn.beta
    lz = lnZ / n
    return lz, -lz / beta


def magnetization(n, edges, weights, fields, beta, *, ctx=None, method="pinning", **kw):
    """magnetization by pinning (src/ising.jl:263-290); the 1+2n networks are
    contracted as one batch. method="adjoint": <s_i> = (d lnZ/d h_i) / beta
    from one reverse sweep."""
    ctx = ctx or default_context()
    if method == "adjoint":
        _, _, dh = lnz_and_grad(n, edges, weights, fields, beta, ctx=ctx, **kw)
        return dh / beta
    if method != "pinning":
        raise CatnError(f"unknown method {method!r}")
    nets = [ising_network(n, edges, weights, fields, beta, ctx=ctx, **kw)]
    for k in range(1, n + 1):
        for v in (1, 2):
            nets.append(ising_network(n, edges, weights, fields, beta, pos1=k, val1=v, ctx=ctx, **kw))
    lz, _, _ = contraction_batch(nets)
    lz0 = lz[0]
    mag = np.zeros(n)
    for k in range(n):
        p = np.exp(lz[1 + 2 * k: 3 + 2 * k] - lz0)
        mag[k] = p[1] - p[0]
    return mag


def correlation(n, edges, weights, fields, beta, *, ctx=None, method="pinning", **kw):
    """correlation by pinning (src/ising.jl:309-339); 1+4|E| networks in one
    batch. method="adjoint": <s_i s_j> = (d lnZ/d w_ij) / beta."""
    ctx = ctx or default_context()
    if method == "adjoint":
        _, dw, _ = lnz_and_grad(n, edges, weights, fields, beta, ctx=ctx, **kw)
        return dw / beta
    if method != "pinning":
        raise CatnError(f"unknown method {method!r}")
    nets = [ising_network(n, edges, weights, fields, beta, ctx=ctx, **kw)]
    for (m, nn) in edges:
        for v1 in (1, 2):
            for v2 in (1, 2):
                nets.append(ising_network(n, edges, weights, fields, beta, pos1=m, val1=v1,
                                          pos2=nn, val2=v2, ctx=ctx, **kw))
    lz, _, _ = contraction_batch(nets)
    lz0 = lz[0]
    cor = np.zeros(len(edges))
    for e in range(len(edges)):
        p = np.exp(lz[1 + 4 * e: 5 + 4 * e] - lz0)
        cor[e] = p[0] - p[1] - p[2] + p[3]
    return cor


def tsvd(A, cutoff=1e-15, maxdim=None, *, ctx=None):
    """tsvd (src/linalg_utils.jl:15-28) -> (U, S, V, discarded); S is the kept
    part; the full spectrum is returned as the fifth value."""
    ctx = ctx or default_context()
    A = np.asfortranarray(np.asarray(A, dtype=np.float64))
    m, n = A.shape
    r = min(m, n)
    U = np.zeros((m, r), order="F")
    V = np.zeros((n, r), order="F")
    S = np.zeros(r)
    k = C.c_int32()
    disc = C.c_double()
    a = A.ravel(order="F")
    Ub = np.zeros(m * r)
    Vb = np.zeros(n * r)
    ctx._check(ctx.lib.catn_tsvd(ctx.h, m, n, _dptr(a), float(cutoff), int(maxdim or 0),
                                 C.byref(k), _dptr(Ub), _dptr(S), _dptr(Vb), C.byref(disc)))
    kk = k.value
    U = Ub[: m * kk].reshape((m, kk), order="F")
    V = Vb[: n * kk].reshape((n, kk), order="F")
    return U, S[:kk].copy(), V, disc.value, S


def rsvd(A, k, oversample=10, power=10, *, seed=1, ctx=None):
    """rsvd(A, k, oversample, power; rng) (src/linalg_utils.jl:44-58) on the
    device -> (U, S, V). The Gaussian sketch comes from a counter-based
    generator seeded with `seed`."""
    ctx = ctx or default_context()
    A = np.asfortranarray(np.asarray(A, dtype=np.float64))
    m, n = A.shape
    p = min(n, oversample * k)
    kk = C.c_int32()
    Ub, Vb, S = np.zeros(m * p), np.zeros(n * p), np.zeros(p)
    ctx._check(ctx.lib.catn_rsvd(ctx.h, m, n, _dptr(A.ravel(order="F")), int(k), int(oversample),
                                 int(power), int(seed), C.byref(kk), _dptr(Ub), _dptr(S), _dptr(Vb)))
    q = kk.value
    return Ub[: m * q].reshape((m, q), order="F"), S[:q].copy(), Vb[: n * q].reshape((n, q), order="F")


class MPSNet:
    """Hand-built MPSNode objects on the device, for the per-primitive API of
    src/mps_node.jl. Node ids and positions are 1-based as in the reference."""

    cplx = False   # eltype="complex": MPSNode{ComplexF64}

    def __init__(self, ctx: Optional[Context] = None, **kw):
        self.ctx = ctx or default_context()
        p = _params(self.ctx, **kw)
        self.cplx = p.eltype == 1
        out = _P()
        self.ctx._check(self.ctx.lib.catn_net_create(self.ctx.h, C.byref(p), C.byref(out)))
        self.net = TensorNetwork(self.ctx, out, self.cplx)
        self.h = out

    @classmethod
    def view(cls, tn: TensorNetwork) -> "MPSNet":
        """The per-node primitives and readers on an existing network."""
        v = cls.__new__(cls)
        v.ctx, v.h, v.net, v.cplx = tn.ctx, tn.h, tn, tn.cplx
        return v

    def add_node(self, id, sites: Sequence[np.ndarray], neighbor, cano=None):
        dims, data = [], []
        for s in sites:
            s = np.asarray(s)
            assert s.ndim == 3
            dims.extend(s.shape)
            data.append(_elems(s, self.cplx))
        d = _i32(dims)
        dt = _f64(np.concatenate(data) if data else np.zeros(0))
        nb = _i32(neighbor)
        self.ctx._check(self.ctx.lib.catn_net_add_node(
            self.h, int(id), len(sites), d.ctypes.data_as(_I32P), _dptr(dt),
            nb.ctypes.data_as(_I32P), int(len(sites) if cano is None else cano)))

    def add_tensor(self, id, tensor, neighbor):
        """MPSNode(tensor, neighbor; chi, cutoff) via raw2mps on the device."""
        t = np.asarray(tensor)
        dm = np.ascontiguousarray(np.asarray(t.shape, dtype=np.int64))
        dt = _elems(t, self.cplx)
        nb = _i32(neighbor)
        self.ctx._check(self.ctx.lib.catn_net_add_tensor_node(
            self.h, int(id), t.ndim, dm.ctypes.data_as(_I64P), _dptr(dt), nb.ctypes.data_as(_I32P)))

    def _err(self, fn, *args):
        e = C.c_double()
        self.ctx._check(fn(self.h, *args, C.byref(e)))
        return e.value

    def swap(self, id, i, j):
        return self._err(self.ctx.lib.catn_node_swap, int(id), int(i), int(j))

    def move(self, id, a, b):
        return self._err(self.ctx.lib.catn_node_move, int(id), int(a), int(b))

    def move2tail(self, id, idx):
        return self._err(self.ctx.lib.catn_node_move2tail, int(id), int(idx))

    def move2head(self, id, idx):
        return self._err(self.ctx.lib.catn_node_move2head, int(id), int(idx))

    def merge(self, id, j, cross=False):
        return self._err(self.ctx.lib.catn_node_merge, int(id), int(j), int(bool(cross)))

    def compress(self, id, opt=True):
        return self._err(self.ctx.lib.catn_node_compress, int(id), int(bool(opt)))

    def cut_bondim(self, i, idx_j_in_i, opt=True):
        return self._err(self.ctx.lib.catn_cut_bondim, int(i), int(idx_j_in_i), int(bool(opt)))

    def cano_to(self, id, idx):
        self.ctx._check(self.ctx.lib.catn_node_cano_to(self.h, int(id), int(idx)))

    def left_canonical(self, id):
        self.ctx._check(self.ctx.lib.catn_node_left_canonical(self.h, int(id)))

    def reverse(self, id):
        self.ctx._check(self.ctx.lib.catn_node_reverse(self.h, int(id)))

    def eat(self, i, j, idx, idxi):
        ln, err = C.c_double(), C.c_double()
        ph = np.zeros(2)
        self.ctx._check(self.ctx.lib.catn_node_eat(self.h, int(i), int(j), int(idx), int(idxi),
                                                   C.byref(ln), C.byref(err), _dptr(ph)))
        return ln.value, err.value, complex(ph[0], ph[1]) if self.cplx else float(ph[0])

    def node(self, id):
        """(order, cano, neighbor list, site dims list)"""
        o, c = C.c_int32(), C.c_int32()
        self.ctx._check(self.ctx.lib.catn_node_get(self.h, int(id), C.byref(o), C.byref(c),
                                                   None, None, 0))
        cap = max(o.value, 1)
        nb = np.zeros(cap + 8, dtype=np.int32)
        dims = np.zeros(3 * (cap + 8), dtype=np.int32)
        self.ctx._check(self.ctx.lib.catn_node_get(self.h, int(id), C.byref(o), C.byref(c),
                                                   nb.ctypes.data_as(_I32P),
                                                   dims.ctypes.data_as(_I32P), cap + 8))
        # neighbour count can differ from the order transiently (never at API boundaries)
        return o.value, c.value, nb, [tuple(dims[3 * k: 3 * k + 3]) for k in range(o.value)]

    def neighbor(self, id, count):
        return [int(x) for x in self.node(id)[2][:count]]

    def sites(self, id):
        o, c, _, dims = self.node(id)
        out = []
        for k, (dl, d, dr) in enumerate(dims):
            buf = np.zeros((2 if self.cplx else 1) * int(dl) * int(d) * int(dr))
            self.ctx._check(self.ctx.lib.catn_node_site(self.h, int(id), k + 1, _dptr(buf), buf.size))
            out.append(_from_elems(buf, self.cplx).reshape((int(dl), int(d), int(dr)), order="F"))
        return out
