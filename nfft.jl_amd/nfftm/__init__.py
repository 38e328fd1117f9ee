"""nfftm -- NFFT plans on AMD MI355X behind the AbstractNFFTs operator interface.

Python host mirror of the plan API of JuliaMath/NFFT.jl (AbstractNFFTs/src/
interface.jl, derived.jl, misc.jl): ``plan_nfft``, ``mul_`` (``mul!``),
``adjoint``, ``nodes_`` (``nodes!``), ``size_in``/``size_out``, the optional
``convolve_``/``convolve_transpose_``/``deconvolve_``/``deconvolve_transpose_``,
``nfft``/``nfft_adjoint``, ``TimingStats`` and the backend registry.  Every
transform runs in libnfftm.so (hand-written gfx950 HIP kernels + rocFFT)
through its C ABI (include/nfftm.h); there is no CPU path.

Array conventions are Python's: an image of size ``N = (N_0, ..., N_{D-1})``
is a C-order tensor of shape ``N`` and the nodes ``k`` are a ``(D, M)`` array
whose row ``d`` pairs with axis ``d``:

    fHat_j = sum_n f[n] exp(-2 pi i sum_d k[d, j] * (n_d - N_d // 2)).

A C-order array of shape N is byte-for-byte a Julia column-major array of size
reverse(N), so the C ABI (Julia order) is called with N and the rows of k
reversed.  Data tensors live on the GPU (torch tensors on ``cuda``; numpy or
CPU inputs are copied there); nodes are host arrays, as ``k::Matrix`` in the
reference (AbstractNFFTs/src/derived.jl:29-30).
"""
from __future__ import annotations

import contextlib
import contextvars
import ctypes
import dataclasses
import math
import time

import numpy as np
import torch

from . import _lib as _L

__all__ = [
    "FULL", "TENSOR", "LINEAR", "POLYNOMIAL", "ArgumentError", "DimensionMismatch", "NFFTError",
    "TimingStats", "accuracy_params", "reltol_to_params", "params_to_reltol", "MI355XBackend",
    "backend", "activate_", "active_backend", "set_active_backend_", "with_backend", "NFFTPlan",
    "AdjointPlan", "plan_nfft", "mul_", "adjoint", "size_in", "size_out", "nodes_", "convolve_",
    "convolve_transpose_", "deconvolve_", "deconvolve_transpose_", "nfft", "nfft_adjoint",
]

# PrecomputeFlags, AbstractNFFTs/src/misc.jl:13-18
FULL, TENSOR, LINEAR, POLYNOMIAL = _L.FULL, _L.TENSOR, _L.LINEAR, _L.POLYNOMIAL
_WINDOWS = {"kaiser_bessel": 0, "gauss": 1, "spline": 2, "kaiser_bessel_rev": 3, "cosh_type": 4}


class ArgumentError(ValueError):
    """Julia's ArgumentError (src/utils.jl:50, src/convolution.jl:52, 148)."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch (src/utils.jl:101-103, src/convolution.jl:26-29)."""


class NFFTError(RuntimeError):
    """Julia's error(...) (unsupported option) and device/FFT failures."""


def _check(rc):
    if rc == _L.OK:
        return
    msg = _L.load().nfftm_last_error().decode(errors="replace")
    if rc == _L.ERR_ARG:
        raise ArgumentError(msg)
    if rc == _L.ERR_DIM:
        raise DimensionMismatch(msg)
    raise NFFTError(f"[{rc}] {msg}")


# ---------------------------------------------------------------------------
# TimingStats and accuracy parameters (AbstractNFFTs/src/misc.jl:22-81)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class TimingStats:
    pre: float = 0.0
    conv: float = 0.0
    fft: float = 0.0
    deconv: float = 0.0
    conv_adjoint: float = 0.0
    fft_adjoint: float = 0.0
    deconv_adjoint: float = 0.0

    def __str__(self):
        t = self
        tot = (t.deconv + t.fft + t.conv) or 1.0
        tota = (t.deconv_adjoint + t.fft_adjoint + t.conv_adjoint) or 1.0
        return (f"Timing: pre = {t.pre:.4f} s deconv = {t.deconv:.4f} / {t.deconv_adjoint:.4f} s "
                f"fft = {t.fft:.4f} / {t.fft_adjoint:.4f} s conv = {t.conv:.4f} / "
                f"{t.conv_adjoint:.4f} s\n                       deconv = "
                f"{100 * t.deconv / tot:.4f} / {100 * t.deconv_adjoint / tota:.4f} % fft = "
                f"{100 * t.fft / tot:.4f} / {100 * t.fft_adjoint / tota:.4f} % conv = "
                f"{100 * t.conv / tot:.4f} / {100 * t.conv_adjoint / tota:.4f} %")


def reltol_to_params(reltol):
    """AbstractNFFTs/src/misc.jl:44-48."""
    w = math.ceil(math.log10(1.0 / reltol)) + 1
    return w // 2, 2.0


def params_to_reltol(m, sigma):
    """AbstractNFFTs/src/misc.jl:50-53."""
    return 10.0 ** (-(2 * m - 1))


def accuracy_params(**kw):
    """accuracyParams (AbstractNFFTs/src/misc.jl:66-81): ``reltol`` wins;
    ``m`` is honoured only together with ``sigma``; default reltol = 1e-9."""
    kw = _norm_kwargs(kw)
    if "reltol" in kw and kw["reltol"] is not None:
        reltol = kw["reltol"]
        m, sigma = reltol_to_params(reltol)
    elif kw.get("m") is not None and kw.get("sigma") is not None:
        m, sigma = kw["m"], kw["sigma"]
        reltol = params_to_reltol(m, sigma)
    else:
        reltol = 1e-9
        m, sigma = reltol_to_params(reltol)
    return int(m), float(sigma), float(reltol)


def _norm_kwargs(kw):
    kw = dict(kw)
    if "σ" in kw:
        kw["sigma"] = kw.pop("σ")
    return kw


# ---------------------------------------------------------------------------
# Backend registry (AbstractNFFTs/src/interface.jl:1-21, src/NFFT.jl:40-42)
# ---------------------------------------------------------------------------
class AbstractNFFTBackend:
    pass


class MI355XBackend(AbstractNFFTBackend):
    def __repr__(self):
        return "MI355XBackend()"


_global_backend = [None]
_scoped_backend = contextvars.ContextVar("nfft_backend", default=None)


def backend():
    return MI355XBackend()


def set_active_backend_(b):
    _global_backend[0] = b


def activate_():
    set_active_backend_(backend())


def active_backend():
    s = _scoped_backend.get()
    return s if s is not None else _global_backend[0]


@contextlib.contextmanager
def with_backend(b):
    """``with(nfft_backend => b) do ... end`` (docs/src/abstract.md)."""
    tok = _scoped_backend.set(b)
    try:
        yield
    finally:
        _scoped_backend.reset(tok)


activate_()


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def _device():
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    if not torch.cuda.is_available():
        # argument validation still runs (and fails loudly) in the library
        return ctypes.c_void_p(None)
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ctype(prec):
    return torch.complex128 if prec == _L.F64 else torch.complex64


def _to_dev(x, dtype):
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(x))
    if not isinstance(x, torch.Tensor):
        x = torch.as_tensor(x)
    if not x.is_complex():
        x = x.to(torch.float64 if dtype == torch.complex128 else torch.float32)
    x = x.to(device=_device(), dtype=dtype)
    return x.contiguous()


class _CudaArray:
    """__cuda_array_interface__ view of device memory owned by a plan."""

    def __init__(self, ptr, shape, typestr, owner):
        self._owner = owner
        self.__cuda_array_interface__ = {
            "shape": tuple(shape), "typestr": typestr, "data": (int(ptr), False),
            "version": 3, "strides": None,
        }


# ---------------------------------------------------------------------------
# The plan
# ---------------------------------------------------------------------------
class NFFTPlan:
    """MI355X NFFT plan, the counterpart of ``NFFTPlan``/``GPU_NFFTPlan``
    (src/implementation.jl:16-106, ext/NFFTGPUArraysExt/implementation.jl:3-59).

    Parameters mirror ``NFFTParams`` (src/implementation.jl:3-14): ``m``,
    ``sigma`` (or ``σ``), ``reltol``, ``window``, ``precompute`` (default
    POLYNOMIAL), ``sortNodes``, ``blocking``; ``dtype`` selects FP32
    (``torch.float32``, default when k is float32) or FP64.
    """

    def __init__(self, k, N, *, dims=None, fftflags=None, timing=None, dtype=None, **kw):
        t0 = time.perf_counter()
        kw = _norm_kwargs(kw)
        k = _as_nodes(k)
        N = (int(N),) if np.isscalar(N) else tuple(int(n) for n in N)
        D = len(N)
        dims = _dims_range(dims, D)
        if k.shape[0] != len(dims):
            raise ArgumentError(f"Nodes x have dimension {k.shape[0]} != {len(dims)}")
        m, sigma, reltol = accuracy_params(**kw)
        window = str(kw.get("window", "kaiser_bessel")).lstrip(":")  # a Julia symbol name works too
        if window not in _WINDOWS:
            raise NFFTError(f"Window {window} not yet implemented!")
        precompute = int(kw.get("precompute", POLYNOMIAL))
        blocking = bool(kw.get("blocking", True))
        sort_nodes = bool(kw.get("sortNodes", False))
        if dtype is None:
            dtype = torch.float32 if k.dtype == np.float32 else torch.float64
        prec = _L.F64 if dtype in (torch.float64, torch.complex128) else _L.F32
        if sort_nodes:  # sortslices(k, dims=2), mutating k (src/precomputation.jl:52-54)
            order = np.lexsort(k)
            k[:] = k[:, order]
        self._lib = _L.load()
        self._handle = ctypes.c_void_p()
        kj = np.ascontiguousarray(k[::-1].T, dtype=np.float64)
        Nj = (ctypes.c_int64 * D)(*N[::-1])
        flags = ((_L.FLAG_BLOCKING if blocking else 0) | (_L.FLAG_GRAPHS if kw.get("graphs", False) else 0)
                 | (_L.FLAG_ROCFFT if kw.get("rocfft", False) else 0)
                 | (0 if kw.get("fused_fold", True) else _L.FLAG_SEPARATE_FOLD)
                 | (_L.FLAG_SPARSE if kw.get("sparse_matrix", False) else 0))
        kp = kj.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        if len(dims) == D:
            _check(self._lib.nfftm_plan_create(
                ctypes.byref(self._handle), prec, D, Nj, k.shape[1], kp, m, sigma, _WINDOWS[window],
                precompute, flags, _stream()))
        else:
            # Julia axes of the C-order axes dims: D - 1 - a
            _check(self._lib.nfftm_plan_create_dims(
                ctypes.byref(self._handle), prec, D, Nj, D - 1 - dims[-1], D - 1 - dims[0], k.shape[1], kp,
                m, sigma, _WINDOWS[window], precompute, flags, _stream()))
        self.k = k
        self.N = N
        self.D = D
        self.dims = dims
        self.prec = prec
        self.dtype = _ctype(prec)
        self.params = dict(m=m, sigma=sigma, reltol=reltol, window=window, precompute=precompute,
                           sortNodes=sort_nodes, blocking=blocking)
        self._read_sizes()
        if timing is not None:
            timing.pre = time.perf_counter() - t0

    @property
    def directional(self):
        return len(self.dims) < self.D

    @property
    def Nsub(self):
        """N along dims (the size of one transformed slice)."""
        return tuple(self.N[a] for a in self.dims)

    def _read_sizes(self):
        Nt = (ctypes.c_int64 * self.D)()
        M = ctypes.c_int64()
        sg = ctypes.c_double()
        _check(self._lib.nfftm_sizes(self._handle, Nt, ctypes.byref(M), ctypes.byref(sg)))
        self.Ntilde = tuple(int(x) for x in Nt)[::-1]
        self.J = int(M.value)
        self.params["sigma"] = float(sg.value)

    # AbstractNFFTs surface ----------------------------------------------------
    @property
    def M(self):
        return self.J

    @property
    def NOut(self):
        """Output size: J in place of the transformed axes (initParams,
        src/precomputation.jl:38-48)."""
        return self.N[:self.dims[0]] + (self.J,) + self.N[self.dims[-1] + 1:]

    @property
    def Ñ(self):
        return self.Ntilde

    def size_in(self):
        return self.N

    def size_out(self):
        return self.NOut

    @property
    def H(self):
        return AdjointPlan(self)

    def __mul__(self, f):
        return self.forward(f)

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                self._lib.nfftm_plan_destroy(h)
            except Exception:
                pass
            self._handle = None

    def __repr__(self):
        r = (f"NFFTPlan (MI355X) with {self.J} sampling points for an input array of size "
             f"{self.N} and an output array of size {self.NOut}")
        return r + (f" with dims {self.dims}" if self.directional else "")

    def copy(self):
        """copy(p), src/implementation.jl:45-66."""
        c = NFFTPlan.__new__(NFFTPlan)
        c._lib = self._lib
        c._handle = ctypes.c_void_p()
        _check(self._lib.nfftm_plan_copy(self._handle, ctypes.byref(c._handle)))
        c.k, c.N, c.D, c.dims, c.prec, c.dtype = self.k, self.N, self.D, self.dims, self.prec, self.dtype
        c.params = dict(self.params)
        c._read_sizes()
        return c

    def nodes_(self, k):
        """nodes!(p, k), src/implementation.jl:108-141."""
        k = _as_nodes(k)
        if k.shape[0] != len(self.dims):
            raise ArgumentError(f"Nodes x have dimension {k.shape[0]} != {len(self.dims)}")
        if self.params["sortNodes"]:
            order = np.lexsort(k)
            k[:] = k[:, order]
        kj = np.ascontiguousarray(k[::-1].T, dtype=np.float64)
        torch.cuda.current_stream().synchronize()
        _check(self._lib.nfftm_nodes(self._handle,
                                     kj.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     k.shape[1]))
        self.k = k
        self._read_sizes()
        return self

    @property
    def tmpVec(self):
        """The plan's oversampled grid (a view of device memory it owns)."""
        ptr = ctypes.c_void_p()
        _check(self._lib.nfftm_grid(self._handle, ctypes.byref(ptr)))
        ts = "<c16" if self.prec == _L.F64 else "<c8"
        shape = tuple(self.Ntilde[a] for a in self.dims)  # one slice of a directional plan
        return torch.as_tensor(_CudaArray(ptr.value, shape, ts, self), device=_device())

    # transforms -----------------------------------------------------------------
    def _prep(self, x, shape, what):
        x = _to_dev(x, self.dtype)
        if tuple(x.shape) != tuple(shape):
            raise DimensionMismatch(f"Data is not consistent with NFFTPlan ({what} has shape "
                                    f"{tuple(x.shape)}, expected {tuple(shape)})")
        return x

    def _out(self, out, shape, what):
        if out is None:
            return torch.empty(shape, dtype=self.dtype, device=_device()), None
        if not isinstance(out, torch.Tensor) or tuple(out.shape) != tuple(shape):
            raise DimensionMismatch(f"Data is not consistent with NFFTPlan ({what})")
        if out.dtype == self.dtype and out.is_contiguous() and out.is_cuda:
            return out, None
        return torch.empty(shape, dtype=self.dtype, device=_device()), out

    def _run(self, fn, inp, outp, timing, verbose, fields):
        self._lib.nfftm_set_stream(self._handle, _stream())
        want = timing is not None or verbose
        if want:
            _check(self._lib.nfftm_set_timing(self._handle, 1))
        try:
            _check(fn(self._handle, ctypes.c_void_p(inp.data_ptr()),
                      ctypes.c_void_p(outp.data_ptr())))
        finally:
            if want:
                self._lib.nfftm_set_timing(self._handle, 0)
        if want:
            t = (ctypes.c_double * 7)()
            _check(self._lib.nfftm_timing(self._handle, t))
            vals = {name: t[i] for i, name in enumerate(
                ("pre", "conv", "fft", "deconv", "conv_adjoint", "fft_adjoint", "deconv_adjoint"))}
            if timing is not None:
                for f in fields:
                    setattr(timing, f, vals[f])
            if verbose:
                print("Timing: " + " ".join(f"{f}={vals[f]}" for f in fields))

    def forward(self, f, out=None, *, timing=None, verbose=False):
        """mul!(fHat, p, f), src/implementation.jl:155-172."""
        f = self._prep(f, self.N, "f")
        o, user = self._out(out, self.NOut, "fHat")
        self._run(self._lib.nfftm_forward, f, o, timing, verbose, ("deconv", "fft", "conv"))
        if user is not None:
            user.copy_(o)
            return user
        return o

    def adjoint(self, fhat, out=None, *, timing=None, verbose=False):
        """mul!(f, adjoint(p), fHat), src/implementation.jl:176-193."""
        fhat = self._prep(fhat, self.NOut, "fHat")
        o, user = self._out(out, self.N, "f")
        self._run(self._lib.nfftm_adjoint, fhat, o, timing, verbose,
                  ("conv_adjoint", "fft_adjoint", "deconv_adjoint"))
        if user is not None:
            user.copy_(o)
            return user
        return o

    def grad_nodes(self, f, ybar):
        """Node gradient kbar[d, j] = Re(conj(ybar_j) * d fHat_j / d k[d, j]) of
        fHat = p * f: the node cotangent that the reference's rrule leaves as
        NoTangent() (AbstractNFFTs/ext/AbstractNFFTsChainRulesCoreExt.jl:14-45).
        Returns a (D, M) float64 device tensor in the axis order of k."""
        if self.directional:  # as the library answers (NFFTM_ERR_UNSUPPORTED)
            raise NFFTError("node gradient of a directional plan")
        f = self._prep(f, self.N, "f")
        ybar = self._prep(ybar, (self.J,), "ybar")
        kb = torch.empty((self.J, self.D), dtype=torch.float64, device=_device())
        self._lib.nfftm_set_stream(self._handle, _stream())
        _check(self._lib.nfftm_grad_nodes(self._handle, ctypes.c_void_p(f.data_ptr()),
                                          ctypes.c_void_p(ybar.data_ptr()),
                                          ctypes.c_void_p(kb.data_ptr())))
        return kb.t().flip(0)

    def grad_nodes_grid(self, g, ybar):
        """The node gradient of grad_nodes from the transformed grid g of the
        forward transform (the plan's tmpVec right after ``p * f``, or a copy
        of it) instead of f: only the window-derivative gather runs
        (nfftm_grad_nodes_grid)."""
        if self.directional:
            raise NFFTError("node gradient of a directional plan")
        g = self._prep(g, self.Ntilde, "g")
        ybar = self._prep(ybar, (self.J,), "ybar")
        kb = torch.empty((self.J, self.D), dtype=torch.float64, device=_device())
        self._lib.nfftm_set_stream(self._handle, _stream())
        _check(self._lib.nfftm_grad_nodes_grid(self._handle, ctypes.c_void_p(g.data_ptr()),
                                               ctypes.c_void_p(ybar.data_ptr()),
                                               ctypes.c_void_p(kb.data_ptr())))
        return kb.t().flip(0)

    def tune(self):
        """Autotune the tile edge and workgroup size of the gridding kernels on
        this plan's nodes (nfftm_plan_tune); returns the chosen shape."""
        torch.cuda.current_stream().synchronize()
        self._lib.nfftm_set_stream(self._handle, _stream())
        _check(self._lib.nfftm_plan_tune(self._handle))
        return self.tile_shape()

    def set_tiles(self, tile, threads):
        """Set the tile shape (nfftm_plan_set_tiles): tile edge for D <= 2,
        geometry variant 0 / 1 for D = 3."""
        torch.cuda.current_stream().synchronize()
        _check(self._lib.nfftm_plan_set_tiles(self._handle, int(tile), int(threads)))
        return self.tile_shape()

    def binning(self):
        """(perm, items) of the plan's node binning (nfftm_binning): perm[s] is
        the caller index of the node at sorted position s; items is an
        (nitems, 4) array of (tile, begin, end, 0) in Julia tile order."""
        n = ctypes.c_int()
        _check(self._lib.nfftm_binning(self._handle, ctypes.byref(n), None, None))
        perm = np.empty(self.J, dtype=np.int32)
        items = np.empty((n.value, 4), dtype=np.int32)
        _check(self._lib.nfftm_binning(self._handle, ctypes.byref(n), perm.ctypes.data_as(ctypes.c_void_p),
                                       items.ctypes.data_as(ctypes.c_void_p)))
        return perm, items

    def tile_shape(self):
        """(tile edge per dimension, workgroup size) of the gridding kernels."""
        t = (ctypes.c_int * 3)()
        nt = ctypes.c_int()
        _check(self._lib.nfftm_tile_shape(self._handle, t, ctypes.byref(nt)))
        return tuple(t[: self.D])[::-1], nt.value

    # optional interface (AbstractNFFTs/src/interface.jl:217-243) ---------------
    def _stage(self, fn, inp, in_shape, out, out_shape, names):
        real_in = isinstance(inp, torch.Tensor) and not inp.is_complex() or (
            isinstance(inp, np.ndarray) and not np.iscomplexobj(inp))
        real_out = isinstance(out, torch.Tensor) and not out.is_complex()
        if real_out and not real_in:
            raise ArgumentError(f"Complex input {names[0]} requires Complex output {names[1]}")
        x = self._prep(inp, in_shape, names[0])
        if not isinstance(out, torch.Tensor) or tuple(out.shape) != tuple(out_shape):
            raise DimensionMismatch(f"size({names[1]}) does not match the plan")
        o = out if (out.dtype == self.dtype and out.is_contiguous() and out.is_cuda) else \
            torch.empty(out_shape, dtype=self.dtype, device=_device())
        self._lib.nfftm_set_stream(self._handle, _stream())
        _check(fn(self._handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(o.data_ptr())))
        if o is not out:
            out.copy_(o.real if real_out else o)
        return out

    _ELT = {torch.float32: _L.R32, torch.float64: _L.R64, torch.complex64: _L.C32,
            torch.complex128: _L.C64, torch.float16: _L.R16, torch.complex32: _L.C16}

    def _typed(self, fn, inp, in_shape, out, out_shape, names):
        """Stage operator on tensors of any real or complex type (the
        reference's RealOrComplex methods), through the typed C entry point:
        the library converts to and from the plan's complex type."""
        if isinstance(inp, np.ndarray):
            inp = torch.from_numpy(np.ascontiguousarray(inp))
        if not isinstance(inp, torch.Tensor):
            inp = torch.as_tensor(inp)
        if not isinstance(out, torch.Tensor):
            raise ArgumentError(f"{names[1]} must be a device tensor")
        if inp.dtype not in self._ELT or out.dtype not in self._ELT:
            raise ArgumentError(f"element types {inp.dtype}, {out.dtype} not supported")
        if inp.is_complex() and not out.is_complex():
            raise ArgumentError(f"Complex input {names[0]} requires Complex output {names[1]}")
        if tuple(inp.shape) != tuple(in_shape):
            raise DimensionMismatch(f"size({names[0]})={tuple(inp.shape)} does not match {tuple(in_shape)}")
        if tuple(out.shape) != tuple(out_shape):
            raise DimensionMismatch(f"size({names[1]})={tuple(out.shape)} does not match {tuple(out_shape)}")
        x = inp.to(device=_device()).contiguous()
        o = out if (out.is_contiguous() and out.is_cuda) else torch.empty_like(out, device=_device())
        self._lib.nfftm_set_stream(self._handle, _stream())
        _check(fn(self._handle, ctypes.c_void_p(x.data_ptr()), self._ELT[x.dtype], x.numel(),
                  ctypes.c_void_p(o.data_ptr()), self._ELT[o.dtype], o.numel()))
        if o is not out:
            out.copy_(o)
        return out

    def convolve_(self, g, fhat):
        """convolve!(p, g, fHat), src/convolution.jl:20-53 (g and fHat real or
        complex, any precision)."""
        if self.directional:
            return self._stage(self._lib.nfftm_convolve, g, self.Ntilde, fhat, self.NOut, ("g", "fHat"))
        return self._typed(self._lib.nfftm_convolve_typed, g, self.Ntilde, fhat, (self.J,), ("g", "fHat"))

    def convolve_transpose_(self, fhat, g):
        """convolve_transpose!(p, fHat, g), src/convolution.jl:115-149 (fHat and
        g real or complex, any precision)."""
        if self.directional:
            return self._stage(self._lib.nfftm_convolve_transpose, fhat, self.NOut, g, self.Ntilde,
                               ("fHat", "g"))
        return self._typed(self._lib.nfftm_convolve_transpose_typed, fhat, (self.J,), g, self.Ntilde,
                           ("fHat", "g"))

    def deconvolve_(self, f, g):
        """deconvolve!(p, f, g), src/deconvolution.jl:2-10 (after fill!(g, 0))."""
        return self._stage(self._lib.nfftm_deconvolve, f, self.N, g, self.Ntilde, ("f", "g"))

    def deconvolve_transpose_(self, g, f):
        """deconvolve_transpose!(p, g, f), src/deconvolution.jl:49-57."""
        return self._stage(self._lib.nfftm_deconvolve_transpose, g, self.Ntilde, f, self.N,
                           ("g", "f"))

    def fft_(self, g, direction=-1):
        """p.forwardFFT * g (direction -1) or p.backwardFFT * g (+1), in place."""
        if not (isinstance(g, torch.Tensor) and g.is_cuda and g.dtype == self.dtype
                and g.is_contiguous() and tuple(g.shape) == self.Ntilde):
            raise DimensionMismatch("g must be a contiguous device tensor of size Ñ")
        self._lib.nfftm_set_stream(self._handle, _stream())
        _check(self._lib.nfftm_fft(self._handle, ctypes.c_void_p(g.data_ptr()), int(direction)))
        return g


class BatchPlan:
    """Many transforms per launch (nfftm_batch_*): plans of one precision and
    D <= 2 with different N, M and nodes, and ``ntrans[i]`` images per plan
    that share its nodes.  Slots are (plan i, image c) in order of i then c;
    ``forward``/``adjoint`` take and return one tensor per slot and compute
    exactly what ``plans[i] * f`` / ``plans[i].H * fhat`` compute, with every
    stage of all slots in one launch and each direction replayed from a HIP
    graph.  The plans must stay alive and keep their nodes while the batch is
    used."""

    def __init__(self, plans, ntrans=None, *, autotune=True, graph=True):
        self._lib = _L.load()
        self.plans = list(plans)
        n = len(self.plans)
        if n == 0:
            raise ArgumentError("a batch needs at least one plan")
        self.ntrans = [1] * n if ntrans is None else [int(x) for x in ntrans]
        if len(self.ntrans) != n:
            raise DimensionMismatch("ntrans needs one entry per plan")
        self.slots = [(i, c) for i in range(n) for c in range(self.ntrans[i])]
        handles = (ctypes.c_void_p * n)(*[p._handle.value for p in self.plans])
        nt = (ctypes.c_int * n)(*self.ntrans)
        flags = (_L.BATCH_AUTOTUNE if autotune else 0) | (0 if graph else _L.BATCH_NO_GRAPH)
        torch.cuda.current_stream().synchronize()
        self._handle = ctypes.c_void_p()
        _check(self._lib.nfftm_batch_create(ctypes.byref(self._handle), n, handles, nt, flags))
        ns, g = ctypes.c_int(), ctypes.c_int()
        _check(self._lib.nfftm_batch_info(self._handle, ctypes.byref(ns), ctypes.byref(g)))
        self.nslots, self.graph = ns.value, bool(g.value)
        self.dtype = self.plans[0].dtype

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                self._lib.nfftm_batch_destroy(h)
            except Exception:
                pass
            self._handle = None

    def _call(self, fn, xs, in_shape, out_shape, outs):
        if len(xs) != self.nslots:
            raise DimensionMismatch(f"expected {self.nslots} inputs (one per slot), got {len(xs)}")
        ins = [self.plans[i]._prep(x, in_shape(self.plans[i]), "input")
               for x, (i, _) in zip(xs, self.slots)]
        if outs is None:
            outs = [torch.empty(out_shape(self.plans[i]), dtype=self.dtype, device=_device())
                    for (i, _) in self.slots]
        for o, (i, _) in zip(outs, self.slots):
            if (not isinstance(o, torch.Tensor) or o.dtype != self.dtype or not o.is_contiguous()
                    or tuple(o.shape) != tuple(out_shape(self.plans[i]))):
                raise DimensionMismatch("outputs must be contiguous device tensors of the plan shapes")
        pin = (ctypes.c_void_p * self.nslots)(*[x.data_ptr() for x in ins])
        pout = (ctypes.c_void_p * self.nslots)(*[o.data_ptr() for o in outs])
        _check(fn(self._handle, pin, pout, _stream()))
        return outs

    def forward(self, fs, outs=None):
        """fHat of every slot: plans[i] * fs[s] for slot s = (i, c)."""
        return self._call(self._lib.nfftm_batch_forward, fs, lambda p: p.N, lambda p: (p.J,), outs)

    def adjoint(self, fhats, outs=None):
        """f of every slot: adjoint(plans[i]) * fhats[s]."""
        return self._call(self._lib.nfftm_batch_adjoint, fhats, lambda p: (p.J,), lambda p: p.N, outs)


class AdjointPlan:
    """``adjoint(p)`` (AbstractNFFTs/src/interface.jl:118-125)."""

    def __init__(self, parent):
        self.parent = parent

    def size_in(self):
        return self.parent.size_out()

    def size_out(self):
        return self.parent.size_in()

    @property
    def H(self):
        return self.parent

    def __mul__(self, fhat):
        return self.parent.adjoint(fhat)

    def copy(self):
        return AdjointPlan(self.parent.copy())

    def __repr__(self):
        return "Adjoint " + repr(self.parent)


def _as_nodes(k):
    if isinstance(k, torch.Tensor):
        k = k.detach().cpu().numpy()
    k = np.asarray(k)
    if k.dtype not in (np.float32, np.float64):
        k = k.astype(np.float64)
    if k.ndim == 1:  # derived.jl:23-27, a vector of 1-D nodes
        k = k.reshape(1, -1)
    if k.ndim != 2:
        raise ArgumentError("nodes must be a (D, M) matrix")
    return k


# ---------------------------------------------------------------------------
# functional surface (AbstractNFFTs/src/derived.jl)
# ---------------------------------------------------------------------------
def plan_nfft(*args, **kw):
    """plan_nfft([backend,] [Q,] k, N; kwargs...) -> NFFTPlan (derived.jl:7-36)."""
    args = list(args)
    if args and isinstance(args[0], AbstractNFFTBackend):
        args.pop(0)
    elif args and args[0] is None:
        raise NFFTError("No default backend available!")
    if args and isinstance(args[0], type):  # array type Q (derived.jl:29-32)
        args.pop(0)
    if active_backend() is None:
        raise NFFTError("No default backend available!")
    k, N = args[0], args[1]
    dims = kw.pop("dims", None)
    Nt = (int(N),) if np.isscalar(N) else tuple(int(n) for n in N)
    if dims is not None and _dims_range(dims, len(Nt)) != tuple(range(len(Nt))):
        return DirectionalPlan(k, Nt, dims, **kw)
    return NFFTPlan(k, N, **kw)


def _dims_range(dims, D):
    """dims (an int, a range or a sequence of consecutive C-order axes; None
    for all) as a tuple of axes, checked as initParams checks them."""
    if dims is None:
        return tuple(range(D))
    d = np.atleast_1d(np.asarray(list(dims) if isinstance(dims, range) else dims, dtype=np.int64))
    if d.ndim != 1 or len(d) == 0 or np.any(np.diff(d) != 1) or d[0] < 0 or d[-1] >= D:
        raise ArgumentError(f"dims must be a range of consecutive axes of a {D}-dimensional array")
    return tuple(int(x) for x in d)


class DirectionalPlan(NFFTPlan):
    """NFFT along a contiguous range of axes of a larger array
    (``plan_nfft(k, N; dims)``, src/directional.jl:1-156 with initParams'
    dims handling at src/precomputation.jl:3-56).  The other axes are batch
    axes: the transform equals the ``len(dims)``-dimensional NFFT of every
    slice, and the output replaces the transformed axes by one axis of
    length J.  ``dims`` are axes of the C-order array (an int or a range of
    consecutive ints); node row d pairs with axis dims[d].

    One C-ABI plan (nfftm_plan_create_dims) transforms every slice in place
    in the caller's arrays: the kernels read and write the slices with their
    strides, so no axis permutation is made.  Node gradients, sdc, Toeplitz
    kernels, batches and the typed stage operators are not defined for it
    (NFFTM_ERR_UNSUPPORTED)."""

    def __init__(self, k, N, dims, **kw):
        N = (int(N),) if np.isscalar(N) else tuple(int(n) for n in N)
        super().__init__(k, N, dims=dims, **kw)

    def copy(self):
        c = super().copy()
        c.__class__ = DirectionalPlan
        return c


def mul_(out, p, inp, **kw):
    """mul!(out, p, inp) for a plan or its adjoint; returns out."""
    if isinstance(p, AdjointPlan):
        return p.parent.adjoint(inp, out, **kw)
    return p.forward(inp, out, **kw)


def adjoint(p):
    return p.H


def size_in(p):
    return p.size_in()


def size_out(p):
    return p.size_out()


def nodes_(p, k):
    return p.nodes_(k)


def convolve_(p, g, fhat):
    return p.convolve_(g, fhat)


def convolve_transpose_(p, fhat, g):
    return p.convolve_transpose_(fhat, g)


def deconvolve_(p, f, g):
    return p.deconvolve_(f, g)


def deconvolve_transpose_(p, g, f):
    return p.deconvolve_transpose_(g, f)


def nfft(k, f, **kw):
    """nfft(k, f) = plan_nfft(k, size(f)) * f (derived.jl:134-145)."""
    f = f if isinstance(f, torch.Tensor) else torch.as_tensor(np.asarray(f))
    return plan_nfft(k, tuple(f.shape), **kw) * f


def nfft_adjoint(k, N, fhat, **kw):
    """nfft_adjoint(k, N, fHat) = adjoint(plan_nfft(k, N)) * fHat (derived.jl:148-152)."""
    return plan_nfft(k, N, **kw).H * fhat


def version():
    return _L.load().nfftm_version().decode()
