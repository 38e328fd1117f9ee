"""NFFTTools on the MI355X plan: sampling density compensation and the
Toeplitz embedding of the normal operator.

Host mirror of NFFTTools/src/samplingDensity.jl and NFFTTools/src/Toeplitz.jl
over the C ABI (``nfftm_sdc``, ``nfftm_toeplitz_*``): the same function names,
argument meaning and errors.  Arrays are Python's (C order, the reverse of
Julia's dimension order), as in the plan mirror.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as _L
from . import (ArgumentError, DimensionMismatch, NFFTPlan, _as_nodes, _check, _device, _stream,
               plan_nfft)


def _real(p):
    return torch.float64 if p.prec == _L.F64 else torch.float32


def sdc(p: NFFTPlan, iters: int = 20, *, weights=None, weights_tmp=None, workg=None, workf=None,
        workv=None):
    """weights = sdc(p; iters=20, ...) (NFFTTools/src/samplingDensity.jl:59-86).

    Pipe & Menon density compensation: M positive real weights scaled so that
    A' diag(w) A 1 ~ 1.  ``weights`` (M real values of the plan's precision on
    the device) is the initial value and is overwritten; the work-buffer
    keywords of the reference are accepted and unused (the library holds its
    own device work arrays)."""
    del weights_tmp, workg, workf, workv
    if weights is None:
        weights = torch.ones(p.J, dtype=_real(p), device=_device())
    return sdc_(p, iters, weights, scale=True)


def sdc_(p: NFFTPlan, iters: int, weights, scale: bool = True):
    """sdc!(p, iters, weights, ...) (NFFTTools/src/samplingDensity.jl:93-155):
    with ``scale`` the 7-argument form (final least-squares scale), without
    it the 5-argument form.  Updates ``weights`` in place and returns it."""
    if not isinstance(weights, torch.Tensor) or weights.dtype != _real(p) or not weights.is_cuda \
            or not weights.is_contiguous():
        raise ArgumentError(f"weights must be a contiguous device tensor of {_real(p)}")
    if tuple(weights.shape) != (p.J,):
        raise DimensionMismatch(f"weights has shape {tuple(weights.shape)}, expected ({p.J},)")
    L = _L.load()
    L.nfftm_set_stream(p._handle, _stream())
    _check(L.nfftm_sdc(p._handle, int(iters), ctypes.c_void_p(weights.data_ptr()), int(bool(scale))))
    return weights


class Toeplitz:
    """FFTs of size 2N and a work array for the Toeplitz embedding of an image
    of size ``shape`` (D <= 3): ``kernel`` and ``apply`` are
    calculateToeplitzKernel! and convolveToeplitzKernel! with their plans
    and buffers preallocated (NFFTTools/src/Toeplitz.jl:119-125, 236-250)."""

    def __init__(self, shape, dtype=torch.complex64):
        self.shape = tuple(int(s) for s in shape)
        self.shape_os = tuple(2 * s for s in self.shape)
        self.dtype = dtype
        prec = _L.F64 if dtype == torch.complex128 else _L.F32
        self._lib = _L.load()
        self._h = ctypes.c_void_p()
        N = (ctypes.c_int64 * len(self.shape))(*self.shape[::-1])
        _check(self._lib.nfftm_toeplitz_create(ctypes.byref(self._h), prec, len(self.shape), N, _stream()))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.nfftm_toeplitz_destroy(h)
            self._h = None

    def _dev(self, x, shape, what):
        if not isinstance(x, torch.Tensor) or x.dtype != self.dtype or not x.is_cuda or not x.is_contiguous():
            raise ArgumentError(f"{what} must be a contiguous device tensor of {self.dtype}")
        if tuple(x.shape) != tuple(shape):
            raise DimensionMismatch(f"{what} has shape {tuple(x.shape)}, expected {tuple(shape)}")
        return x

    def kernel(self, p: NFFTPlan, lam):
        """lam = FFT(fftshift(adjoint(p) * ones)) for a plan over 2N."""
        lam = self._dev(lam, self.shape_os, "lambda")
        self._lib.nfftm_set_stream(p._handle, _stream())
        _check(self._lib.nfftm_toeplitz_kernel(self._h, p._handle, ctypes.c_void_p(lam.data_ptr())))
        return lam

    def apply(self, lam, y):
        """y <- corner of IFFT(FFT(pad(y)) .* lam), in place."""
        lam = self._dev(lam, self.shape_os, "lambda")
        y = self._dev(y, self.shape, "y")
        _check(self._lib.nfftm_toeplitz_apply(self._h, ctypes.c_void_p(lam.data_ptr()),
                                              ctypes.c_void_p(y.data_ptr())))
        return y


def _cplx_of(tr, dtype):
    if dtype is not None:
        return torch.complex128 if dtype in (torch.float64, torch.complex128) else torch.complex64
    return torch.complex64 if np.asarray(tr).dtype == np.float32 else torch.complex128


def calculateToeplitzKernel(shape, tr, *, m=4, sigma=2.0, window="kaiser_bessel", dtype=None, **kw):
    """lam = calculateToeplitzKernel(shape, tr; m=4, σ=2.0, window)
    (NFFTTools/src/Toeplitz.jl:86-93): the kernel of size 2 .* shape, of the
    complex type of the trajectory's precision (float32 nodes -> complex64)."""
    if "σ" in kw:
        sigma = kw.pop("σ")
    tr = _as_nodes(tr)
    shape = tuple(int(s) for s in shape)
    if tr.shape[0] != len(shape):
        raise ArgumentError(f"trajectory has {tr.shape[0]} rows for a {len(shape)}-D image")
    cd = _cplx_of(tr, dtype)
    p = plan_nfft(tr, tuple(2 * s for s in shape), m=m, sigma=sigma, window=window,
                  dtype=torch.float64 if cd == torch.complex128 else torch.float32, **kw)
    t = Toeplitz(shape, cd)
    lam = torch.empty(t.shape_os, dtype=cd, device=_device())
    return t.kernel(p, lam)


def calculateToeplitzKernel_(f, p: NFFTPlan, tr, toeplitz: Toeplitz | None = None):
    """calculateToeplitzKernel!(f, p, tr, fftplan) (NFFTTools/src/Toeplitz.jl:119-125):
    nodes!(p, tr), then the kernel into f (size of p, twice the image)."""
    p.nodes_(tr)
    t = toeplitz or Toeplitz(tuple(s // 2 for s in f.shape), f.dtype)
    return t.kernel(p, f)


def convolveToeplitzKernel_(y, lam, toeplitz: Toeplitz | None = None):
    """convolveToeplitzKernel!(y, lam[, fftplan, ifftplan, xOS1, xOS2])
    (NFFTTools/src/Toeplitz.jl:236-250): y <- A'A y through the kernel, in
    place.  Pass a Toeplitz object to reuse its FFT plans and work array."""
    t = toeplitz or Toeplitz(tuple(y.shape), y.dtype)
    return t.apply(lam, y)
