"""C-ABI and host-logic tests that need no GPU.

* libnfftm.so loads and exports every entry point include/nfftm.h declares;
* argument validation happens in the library before any device call and maps
  to the reference's exception classes (src/utils.jl:46-55 checkNodes ->
  ArgumentError; unknown window / precompute -> error(...));
* the accuracy-parameter rules of AbstractNFFTs/src/misc.jl:44-81;
* the backend registry and TimingStats of the host mirror.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import nfftm
from nfftm import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(REPO, "include", "nfftm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nfftm_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 18
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.EXPORTS)
    assert L.nfftm_version().decode().startswith("nfftm")


def test_no_cpu_fallback_when_library_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIBPATH", str(tmp_path / "missing.so"))
    with pytest.raises(_lib.LibraryMissing):
        _lib.load()


def _create(k, N, prec=0, m=4, sigma=2.0, window=0, precompute=3):
    L = _lib.load()
    k = np.asarray(k, dtype=np.float64)
    D, M = k.shape
    kj = np.ascontiguousarray(k[::-1].T)
    h = ctypes.c_void_p()
    Nj = (ctypes.c_int64 * D)(*tuple(N)[::-1])
    rc = L.nfftm_plan_create(ctypes.byref(h), prec, D, Nj, M,
                             kj.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), m, sigma,
                             window, precompute, 1, None)
    return rc, L.nfftm_last_error().decode()


@pytest.mark.parametrize("k,msg", [
    ([[-0.6, 0.9], [0.5, -0.5]], "range"),   # test/constructors.jl:181-182
    ([[-0.3, 0.3], [0.3, np.nan]], "range"),  # test/constructors.jl:183-184
    ([[0.50000001, 0.0]], "range"),
])
def test_check_nodes_is_argument_error(k, msg):
    rc, err = _create(np.atleast_2d(k), (4,) * np.atleast_2d(k).shape[0])
    assert rc == _lib.ERR_ARG and msg in err


def test_unsupported_options():
    k = np.zeros((2, 3))
    assert _create(k, (4, 4), window=7)[0] == _lib.ERR_UNSUPPORTED
    assert _create(k, (4, 4), precompute=7)[0] == _lib.ERR_UNSUPPORTED
    assert _create(k, (4, 4), m=9)[0] == _lib.ERR_UNSUPPORTED
    assert _create(k, (4, 4), prec=5)[0] == _lib.ERR_UNSUPPORTED
    assert _create(k, (4, 4), sigma=0.9)[0] == _lib.ERR_ARG
    assert _create(k, (0, 4))[0] == _lib.ERR_ARG


def test_null_plan_calls_fail_cleanly():
    L = _lib.load()
    b = ctypes.c_void_p()
    assert L.nfftm_batch_create(ctypes.byref(b), 0, None, None, 0) == _lib.ERR_ARG
    plans = (ctypes.c_void_p * 2)(None, None)
    assert L.nfftm_batch_create(ctypes.byref(b), 2, plans, None, 0) == _lib.ERR_ARG
    assert L.nfftm_batch_forward(None, None, None, None) == _lib.ERR_ARG
    assert L.nfftm_plan_tune(None) == _lib.ERR_ARG
    L.nfftm_batch_destroy(None)
    for name in ("nfftm_forward", "nfftm_adjoint", "nfftm_convolve", "nfftm_convolve_transpose",
                 "nfftm_deconvolve", "nfftm_deconvolve_transpose"):
        assert getattr(L, name)(None, None, None) == _lib.ERR_ARG
    assert L.nfftm_nodes(None, None, 0) == _lib.ERR_ARG
    L.nfftm_plan_destroy(None)


def test_host_mirror_errors_map_to_reference_classes():
    with pytest.raises(nfftm.ArgumentError):
        nfftm.plan_nfft(np.zeros((1, 4)), (2, 2))  # test/constructors.jl:15
    with pytest.raises(nfftm.ArgumentError):
        nfftm.plan_nfft(np.array([[-0.6, 0.9], [0.5, -0.5]]), (2, 2))
    with pytest.raises(nfftm.NFFTError):
        nfftm.plan_nfft(np.zeros((2, 4)), (2, 2), window="hann")
    # directional plans: node rows must match len(dims) (src/precomputation.jl:19-21),
    # dims a range of consecutive axes
    with pytest.raises(nfftm.ArgumentError):
        nfftm.plan_nfft(np.zeros((2, 4)), (2, 2), dims=1)
    with pytest.raises(nfftm.ArgumentError):
        nfftm.plan_nfft(np.zeros((2, 4)), (2, 2, 2), dims=(0, 2))


def test_accuracy_params_rules():
    """AbstractNFFTs/src/misc.jl:44-81: reltol wins, m counts only with sigma,
    default reltol 1e-9 -> m = 5, sigma = 2."""
    assert nfftm.accuracy_params() == (5, 2.0, 1e-9)
    assert nfftm.accuracy_params(m=3) == (5, 2.0, 1e-9)
    assert nfftm.accuracy_params(m=3, sigma=1.5) == (3, 1.5, 1e-5)
    assert nfftm.accuracy_params(m=4, σ=2.0) == (4, 2.0, 1e-7)
    assert nfftm.accuracy_params(reltol=1e-5, m=8, sigma=3.0) == (3, 2.0, 1e-5)
    assert nfftm.reltol_to_params(1e-9) == (5, 2.0)
    assert nfftm.params_to_reltol(5, 2.0) == pytest.approx(1e-9)


def test_sortnodes_order_matches_sortslices():
    """sortslices(k, dims=2) sorts Julia columns lexicographically by Julia
    row 1, 2, ...; Julia row 1 is the last Python row."""
    k = np.array([[0.1, 0.1, -0.2, 0.1],     # Python axis 0 (Julia row 2)
                  [0.3, -0.4, 0.3, 0.3]])    # Python axis 1 (Julia row 1)
    order = np.lexsort(k)
    cols = [tuple(k[::-1, j]) for j in order]  # Julia-order tuples
    assert cols == sorted(cols)


def test_backend_registry_scoping():
    """test/constructors.jl:3-13."""
    b = nfftm.backend()
    nfftm.set_active_backend_(None)
    assert nfftm.active_backend() is None
    with nfftm.with_backend(b):
        assert isinstance(nfftm.active_backend(), nfftm.MI355XBackend)
    assert nfftm.active_backend() is None
    with pytest.raises(nfftm.NFFTError):
        nfftm.plan_nfft(np.zeros((1, 2)), (4,))
    nfftm.activate_()
    assert isinstance(nfftm.active_backend(), nfftm.MI355XBackend)


def test_timingstats_fields():
    t = nfftm.TimingStats()
    assert [f for f in t.__dataclass_fields__] == ["pre", "conv", "fft", "deconv", "conv_adjoint",
                                                   "fft_adjoint", "deconv_adjoint"]
    assert str(t).startswith("Timing: pre = 0.0000 s")


def test_flag_values_agree_across_bindings():
    """The plan flags of include/nfftm.h, the Python mirror and the Julia
    binding carry the same values."""
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(REPO, "include", "nfftm.h")).read(), flags=re.S)
    hdr = {k: int(v) for k, v in re.findall(r"\bNFFTM_FLAG_([A-Z_]+)\s*=\s*(\d+)", txt)}
    assert set(hdr) == {"BLOCKING", "TIMING", "GRAPHS", "ROCFFT", "SEPARATE_FOLD", "SPARSE"}
    for k, v in hdr.items():
        assert getattr(_lib, "FLAG_" + k) == v, k
    jl = open(os.path.join(REPO, "nfft.jl_amd", "julia", "MI355XNFFT.jl")).read()
    jlf = {k: int(v) for k, v in re.findall(r"const NFFTM_FLAG_([A-Z_]+) = Cint\((\d+)\)", jl)}
    assert jlf == hdr


def _split_top(s, sep=","):
    """Split s at separators outside (), {} and []."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == sep and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return [x.strip() for x in out]


def header_prototypes():
    """{name: (return type, [parameter types])} of include/nfftm.h, types
    normalised ('const' dropped, arrays as pointers)."""
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(REPO, "include", "nfftm.h")).read(), flags=re.S)
    protos = {}
    for ret, name, params in re.findall(r"([A-Za-z_][\w\s\*]*?)\b(nfftm_\w+)\s*\(([^)]*)\)\s*;", txt):
        ret = ret.split()[-1] if "*" not in ret else ret
        types = []
        for prm in _split_top(params):
            if prm in ("", "void"):
                continue
            arr = "[" in prm
            prm = re.sub(r"\[.*?\]", "", prm).replace("const", " ")
            toks = prm.replace("*", " * ").split()
            base = [t for t in toks[:-1]]  # the last token is the parameter name
            ty = " ".join(t for t in base if t != "*") + "*" * base.count("*") + ("*" if arr else "")
            types.append(ty.replace(" *", "*"))
        protos[name] = (" ".join(ret.replace("const", " ").split()).replace(" *", "*"), types)
    return protos


# C parameter type -> Julia ccall argument types that pass it correctly
_JL_OK = {
    "int": {"Cint"}, "int64_t": {"Int64"}, "double": {"Cdouble", "Float64"},
    "void*": {"Ptr{Cvoid}"}, "double*": {"Ptr{Float64}", "Ref{Float64}", "Ptr{Cvoid}"},
    "int64_t*": {"Ptr{Int64}", "Ref{Int64}"}, "int*": {"Ptr{Cint}", "Ref{Cint}"},
    "void**": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"}, "void* *": {"Ptr{Ptr{Cvoid}}"},
    "nfftm_plan": {"Ptr{Cvoid}"}, "nfftm_batch": {"Ptr{Cvoid}"}, "nfftm_toeplitz": {"Ptr{Cvoid}"},
    "nfftm_plan*": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"}, "nfftm_batch*": {"Ref{Ptr{Cvoid}}"},
    "nfftm_toeplitz*": {"Ref{Ptr{Cvoid}}"},
}
_JL_RET = {"int": "Cint", "void": "Cvoid", "char*": "Cstring"}


def julia_ccalls():
    """[(name, return type, [argument types])] of every ccall in the shim."""
    src = open(os.path.join(REPO, "nfft.jl_amd", "julia", "MI355XNFFT.jl")).read()
    out = []
    for mt in re.finditer(r"ccall\(\(:(\w+),\s*libnfftm\),", src):
        depth, j = 0, mt.end()
        while depth > 0 or src[j] != ")":  # up to the ccall's own closing parenthesis
            depth += src[j] in "({[" and 1 or src[j] in ")}]" and -1 or 0
            j += 1
        parts = _split_top(src[mt.end():j])
        ret, tup = parts[0], parts[1]
        assert tup.startswith("(") and tup.endswith(")"), (mt.group(1), tup)
        out.append((mt.group(1), ret, _split_top(tup[1:-1])))
    return out


def test_julia_ccalls_match_header():
    """Every ccall of the Julia binding names an exported entry point with the
    header's arity, argument types and return type."""
    protos = header_prototypes()
    assert "nfftm_convolve_typed" in protos and len(protos) == len(header_symbols())
    calls = julia_ccalls()
    assert len(calls) >= 20
    seen = set()
    for name, ret, args in calls:
        assert name in protos, name
        cret, cargs = protos[name]
        assert _JL_RET[cret] == ret, (name, ret, cret)
        assert len(args) == len(cargs), (name, args, cargs)
        for a, c in zip(args, cargs):
            assert a in _JL_OK[c], (name, a, c)
        seen.add(name)
    # the stage operators, sdc and Toeplitz are reached from Julia
    for need in ("nfftm_convolve_typed", "nfftm_convolve_transpose_typed", "nfftm_deconvolve",
                 "nfftm_deconvolve_transpose", "nfftm_sdc", "nfftm_toeplitz_create", "nfftm_toeplitz_kernel",
                 "nfftm_toeplitz_apply", "nfftm_toeplitz_destroy"):
        assert need in seen, need


def test_element_type_codes_agree():
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(REPO, "include", "nfftm.h")).read(), flags=re.S)
    hdr = {k: int(v) for k, v in re.findall(r"\bNFFTM_([RC]\d\d)\s*=\s*(\d+)", txt)}
    assert hdr == {"R32": 0, "R64": 1, "C32": 2, "C64": 3, "R16": 4, "C16": 5}
    for k, v in hdr.items():
        assert getattr(_lib, k) == v
    jl = open(os.path.join(REPO, "nfft.jl_amd", "julia", "MI355XNFFT.jl")).read()
    jmap = dict(re.findall(r"(\w+) => Cint\((\d)\)", jl[jl.index("const ELT"):jl.index("elt(::Type")]))
    names = {"Float32": "R32", "Float64": "R64", "ComplexF32": "C32", "ComplexF64": "C64", "Float16": "R16",
             "ComplexF16": "C16"}
    assert {names[k]: int(v) for k, v in jmap.items()} == hdr
