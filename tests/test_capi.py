"""The C ABI library loads and exports every entry point that
include/catn_hip.h declares (no compute calls: the CPU suite has no GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "catn_hip.h")
HIP_LIB = os.path.join(ROOT, "catn.jl_amd", "lib", "libcatn_hip.so")
EMU_LIB = os.path.join(ROOT, "tests", "emul", "libcatn_emul.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(catn_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "catn_contraction" in names and "catn_tsvd" in names and len(names) >= 30


@pytest.mark.parametrize("lib", [HIP_LIB, EMU_LIB])
def test_library_exports_every_declared_symbol(lib):
    assert os.path.exists(lib), f"{lib} not built (run __graft_entry__.build())"
    so = ctypes.CDLL(lib)
    missing = [n for n in declared() if not hasattr(so, n)]
    assert not missing, missing


def test_python_binding_matches_header(catn):
    assert set(catn.EXPORTED_SYMBOLS) <= set(declared())
    assert set(declared()) <= set(catn.EXPORTED_SYMBOLS)


def test_hip_library_has_no_cpu_fallback():
    """The product library contains the gfx950 code object and no emulator."""
    data = open(HIP_LIB, "rb").read()
    assert b"gfx950" in data
    assert b"cpu-emulator" not in data


def test_params_defaults_and_element_type(catn, emu_ctx):
    """catn_params_default: the reference's TensorNetwork defaults, the double
    layer's merged-bond Dmax off (braket_dmax = -1) and real elements. An
    Ising network (real site tensors) refuses the complex element type, and
    an element type other than 0 / 1 is refused."""
    p = catn._Params()
    emu_ctx.lib.catn_params_default(ctypes.byref(p))
    assert (p.Dmax, p.chi, p.select, p.norm_method) == (32, 32, 1, 1)
    assert p.braket_dmax == -1 and p.eltype == 0 and p.loops == 0
    with pytest.raises(catn.CatnError, match="real"):
        catn.ising_network(2, [(1, 2)], [1.0], [0.0, 0.0], 0.5, ctx=emu_ctx, eltype="complex")
    p.eltype = 7
    out = ctypes.c_void_p()
    assert emu_ctx.lib.catn_net_create(emu_ctx.h, ctypes.byref(p), ctypes.byref(out)) != 0
    assert "eltype" in emu_ctx.lib.catn_last_error().decode()


def test_braket_ignores_tensor_network_dmax(catn, emu_ctx):
    """catn_braket_network reads the merged-bond Dmax from braket_dmax only: a
    call with Dmax = 2 and braket_dmax = -1 (a C caller that set only `loops`
    on top of other defaults) cuts nothing and equals the uncut run, while
    braket_dmax = 2 cuts."""
    import numpy as np
    from test_braket import _loopy
    ts, ixs = _loopy("lattice3x3_D2")
    lab = {}
    for ix in ixs:
        for l in ix:
            lab.setdefault(l, len(lab))
    nd = np.array([t.ndim for t in ts], np.int32)
    dm = np.array([d for t in ts for d in t.shape], np.int64)
    lb = np.array([lab[l] for ix in ixs for l in ix], np.int64)
    dt = np.concatenate([np.asarray(t, float).ravel(order="F") for t in ts])

    def build(Dmax, braket_dmax):
        p = catn._params(emu_ctx, Dmax=Dmax, chi=10_000, loops=True)
        p.braket_dmax = braket_dmax
        out = ctypes.c_void_p()
        I32, I64, D = (ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                       ctypes.POINTER(ctypes.c_double))
        emu_ctx._check(emu_ctx.lib.catn_braket_network(
            emu_ctx.h, len(ts), nd.ctypes.data_as(I32), dm.ctypes.data_as(I64),
            dt.ctypes.data_as(D), lb.ctypes.data_as(I64), ctypes.byref(p), ctypes.byref(out)))
        return catn.contraction(catn.TensorNetwork(emu_ctx, out))
    uncut = build(-1, -1)
    assert build(2, -1)[0] == uncut[0]
    cut = build(-1, 2)
    assert abs(cut[0] - uncut[0]) > 1e-6 and np.isfinite(cut[0])


JULIA = os.path.join(ROOT, "julia", "CATNHip.jl")
_CTYPE = {"int32_t": "Int32", "int64_t": "Int64", "double": "Float64"}


def _c_struct_fields():
    text = open(HEADER).read()
    body = re.search(r"typedef struct catn_params \{(.*?)\} catn_params;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ctype, names = decl.split(None, 1)
        out += [(n.strip(), _CTYPE[ctype]) for n in names.split(",")]
    return out


def _c_arity():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    ar = {}
    for m in re.finditer(r"\b(catn_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        args = m.group(2).strip()
        ar[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return ar


def test_julia_binding_params_layout_matches_header():
    """julia/CATNHip.jl's Params is catn_params field for field (name, type,
    order), and params() fills every field."""
    jl = open(JULIA).read()
    body = re.search(r"struct Params\n(.*?)\nend", jl, re.S).group(1)
    fields = [(m.group(1), m.group(2)) for m in re.finditer(r"^\s*(\w+)::(\w+)", body, re.M)]
    assert fields == _c_struct_fields()
    call = re.search(r"Params\((Dmax, chi.*?)\)\nend", jl, re.S).group(1)
    assert len([a for a in call.split(",") if a.strip()]) == len(fields)


def test_julia_binding_ccalls_match_header():
    """Every ccall of the binding names a declared entry point with the
    header's number of arguments."""
    jl = open(JULIA).read()
    ar = _c_arity()
    calls = re.findall(r"ccall\(\(:(catn_\w+), lib\), \w+,\s*\(([^()]*(?:\([^()]*\)[^()]*)*)\)", jl)
    assert len(calls) >= 12
    for name, types in calls:
        assert name in ar, name
        n = 0 if not types.strip() else len([t for t in re.split(r",(?![^{]*\})", types) if t.strip()])
        assert n == ar[name], (name, n, ar[name])
