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
