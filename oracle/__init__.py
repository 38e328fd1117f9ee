"""CPU oracle for the NFFT hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker / reported baseline.  The
shipped library never calls it.

``liboracle.so`` (built from ``nfft_oracle.cpp`` by ``oracle/Makefile``) is an
FP64 restatement of NFFT.jl's CPU plan:

* window, LUT, polynomial fit, deconvolution factors -- src/windowFunctions.jl,
  src/precomputation.jl
* deconvolve!/deconvolve_transpose! -- src/deconvolution.jl:2-92
* convolve!/convolve_transpose! (blocked, non-blocked, FULL) -- src/convolution.jl
* direct NDFT -- src/direct.jl:99-151

The oversampled FFT stage (src/implementation.jl:87-88, 161, 182: in-place,
unnormalised, sign -1 forward / +1 backward) is done here with numpy's
pocketfft (or scipy.fft when ``workers`` > 1).

Array conventions are Python's (the same as the ``nfftm`` host package): an
image of size ``N = (N_0, ..., N_{D-1})`` is a C-order numpy array of shape
``N`` and nodes ``k`` are a ``(D, M)`` array whose row d pairs with axis d.
A C-order array of shape N is, byte for byte, a Julia column-major array of
size reverse(N), so the C restatement is called with N and the rows of k
reversed (Julia's dimension 1 is the last Python axis).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

FULL, TENSOR, LINEAR, POLYNOMIAL = 1, 2, 3, 4

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        L = ctypes.CDLL(_LIBPATH)
        P = ctypes.c_void_p
        D = ctypes.c_double
        I = ctypes.c_int
        LL = ctypes.c_longlong
        dp = ctypes.POINTER(ctypes.c_double)
        lp = ctypes.POINTER(ctypes.c_longlong)
        L.orc_bessel_i0.restype = D
        L.orc_bessel_i0.argtypes = [D]
        L.orc_window.restype = D
        L.orc_window.argtypes = [D, D, I, D]
        L.orc_window_hat.restype = D
        L.orc_window_hat.argtypes = [D, D, I, D]
        L.orc_lut_size.restype = LL
        L.orc_lut_size.argtypes = [I]
        L.orc_lut.argtypes = [I, D, LL, dp]
        L.orc_poly.argtypes = [I, D, dp]
        L.orc_deconv_factors.argtypes = [LL, LL, I, D, I, dp]
        L.orc_num_threads.restype = I
        L.orc_plan_create.restype = P
        L.orc_plan_create.argtypes = [I, lp, LL, dp, I, D, I, I, I, I]
        L.orc_window_w.restype = D
        L.orc_window_w.argtypes = [I, D, D, I, D]
        L.orc_window_hat_w.restype = D
        L.orc_window_hat_w.argtypes = [I, D, D, I, D]
        L.orc_bessel_j0.restype = D
        L.orc_bessel_j0.argtypes = [D]
        L.orc_plan_destroy.argtypes = [P]
        L.orc_plan_info.argtypes = [P, lp, lp, dp]
        for name in ("orc_deconvolve", "orc_deconvolve_transpose", "orc_convolve",
                     "orc_convolve_transpose"):
            getattr(L, name).argtypes = [P, dp, dp]
        L.orc_ndft.argtypes = [I, lp, LL, dp, dp, dp, lp, LL]
        L.orc_ndft_adjoint.argtypes = [I, lp, LL, dp, dp, dp, lp, LL]
        L.orc_window_du.restype = D
        L.orc_window_du.argtypes = [D, I, D]
        L.orc_grad_gather.argtypes = [P, dp, dp, dp]
        L.orc_ndft_grad.argtypes = [I, lp, LL, dp, dp, dp, dp, lp, LL]
        L.orc_ndft_adjoint_box.argtypes = [I, lp, LL, dp, dp, dp, lp, lp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _lp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))


def _julia_nodes(k):
    """(D, M) Python-order nodes -> Julia D x M column-major buffer (rows
    reversed), D, M."""
    k = np.asarray(k, dtype=np.float64)
    if k.ndim == 1:
        k = k.reshape(1, -1)
    return np.ascontiguousarray(k[::-1].T), k.shape[0], k.shape[1]


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


# ----------------------------------------------------------------------------
# scalar helpers
# ----------------------------------------------------------------------------
def bessel_i0(x):
    return lib().orc_bessel_i0(float(x))


def window(k, Nt, m, sigma):
    """Kaiser-Bessel psi(k; Nt, m, sigma), src/windowFunctions.jl:21-34."""
    return lib().orc_window(float(k), float(Nt), int(m), float(sigma))


def window_hat(n, Nt, m, sigma):
    """Kaiser-Bessel psi_hat(n; Nt, m, sigma), src/windowFunctions.jl:36-39."""
    return lib().orc_window_hat(float(n), float(Nt), int(m), float(sigma))


def lut_size(m):
    return int(lib().orc_lut_size(int(m)))


def lut(m, sigma, K=None):
    """precomputeLinInterp, src/precomputation.jl:291-300 (K+2 entries)."""
    K = lut_size(m) if K is None else K
    out = np.empty(K + 2)
    lib().orc_lut(int(m), float(sigma), int(K), _dp(out))
    return out


def poly(m, sigma):
    """precomputePolyInterp, src/precomputation.jl:302-320 -> (2m+1, 2m)."""
    out = np.empty((2 * m) * (2 * m + 1))
    lib().orc_poly(int(m), float(sigma), _dp(out))
    return out.reshape(2 * m, 2 * m + 1).T.copy()


def deconv_factors(N, Nt, m, sigma, cheb=True):
    """precomputeWindowHatInvLUT for one dim, src/precomputation.jl:347-358."""
    out = np.empty(int(N))
    lib().orc_deconv_factors(int(N), int(Nt), int(m), float(sigma), int(bool(cheb)), _dp(out))
    return out


def grid_size(N, sigma):
    """Oversampled grid size, src/precomputation.jl:25-27."""
    return tuple(int(np.ceil(sigma * n)) // 2 * 2 for n in N)


# ----------------------------------------------------------------------------
# NDFT (src/direct.jl:99-151)
# ----------------------------------------------------------------------------
def ndft(k, f, sel=None):
    """fhat_j = sum_n f_n exp(-2 pi i n.k_j) for j in ``sel`` (default all)."""
    kj, D, M = _julia_nodes(k)
    f = _c128(f)
    N = np.array(f.shape[::-1], dtype=np.int64)  # Julia order
    assert len(N) == D
    if sel is None:
        out = np.empty(M, dtype=np.complex128)
        lib().orc_ndft(D, _lp(N), M, _dp(kj), _dp(f.view(np.float64)),
                       _dp(out.view(np.float64)), None, 0)
    else:
        sel = np.ascontiguousarray(sel, dtype=np.int64)
        out = np.empty(len(sel), dtype=np.complex128)
        lib().orc_ndft(D, _lp(N), M, _dp(kj), _dp(f.view(np.float64)),
                       _dp(out.view(np.float64)), _lp(sel), len(sel))
    return out


def ndft_adjoint(k, N, fhat, sel=None):
    """f_n = sum_j fhat_j exp(+2 pi i n.k_j); ``sel`` lists flat (C-order)
    indices of the image, default all (returned with shape N)."""
    kj, D, M = _julia_nodes(k)
    fhat = _c128(fhat)
    Na = np.array(tuple(N)[::-1], dtype=np.int64)
    if sel is None:
        out = np.empty(int(np.prod(N)), dtype=np.complex128)
        lib().orc_ndft_adjoint(D, _lp(Na), M, _dp(kj), _dp(fhat.view(np.float64)),
                               _dp(out.view(np.float64)), None, 0)
        return out.reshape(tuple(N))
    sel = np.ascontiguousarray(sel, dtype=np.int64)
    out = np.empty(len(sel), dtype=np.complex128)
    lib().orc_ndft_adjoint(D, _lp(Na), M, _dp(kj), _dp(fhat.view(np.float64)),
                           _dp(out.view(np.float64)), _lp(sel), len(sel))
    return out


def ndft_adjoint_box(k, N, fhat, idx):
    """Adjoint NDFT on the product set idx[0] x ... x idx[D-1] of image
    indices (Python axis order, one index array per axis); returns an array
    of shape (len(idx[0]), ..., len(idx[D-1])).  Same sums as
    ``ndft_adjoint``, with the phase factorised per node."""
    kj, D, M = _julia_nodes(k)
    assert len(idx) == D == len(N)
    fhat = _c128(fhat)
    Na = np.array(tuple(N)[::-1], dtype=np.int64)
    lists = [np.asarray(i, dtype=np.int64) for i in idx[::-1]]  # Julia order
    flat = np.ascontiguousarray(np.concatenate(lists))
    cnt = np.array([len(i) for i in lists], dtype=np.int64)
    out = np.empty(int(np.prod(cnt)), dtype=np.complex128)
    lib().orc_ndft_adjoint_box(D, _lp(Na), M, _dp(kj), _dp(fhat.view(np.float64)),
                               _dp(out.view(np.float64)), _lp(flat), _lp(cnt))
    return out.reshape(tuple(len(i) for i in idx))


def ndft_grad(k, f, ybar, sel=None):
    """Closed-form node gradient of the NDFT: kbar[d, q] =
    Re(conj(ybar_j) * sum_n f_n (-2 pi i n_d) exp(-2 pi i n.k_j)), j = sel[q]
    (default all nodes).  Python axis order: returns (D, len(sel))."""
    kj, D, M = _julia_nodes(k)
    f = _c128(f)
    ybar = _c128(ybar)
    N = np.array(f.shape[::-1], dtype=np.int64)
    nout = M if sel is None else len(sel)
    out = np.empty((nout, D))
    if sel is None:
        lib().orc_ndft_grad(D, _lp(N), M, _dp(kj), _dp(f.view(np.float64)),
                            _dp(ybar.view(np.float64)), _dp(out), None, 0)
    else:
        sel = np.ascontiguousarray(sel, dtype=np.int64)
        lib().orc_ndft_grad(D, _lp(N), M, _dp(kj), _dp(f.view(np.float64)),
                            _dp(ybar.view(np.float64)), _dp(out), _lp(sel), len(sel))
    return np.ascontiguousarray(out.T[::-1])


def window_du(u, m, sigma):
    """d psi / du of the Kaiser-Bessel window, u in grid units."""
    return lib().orc_window_du(float(u), int(m), float(sigma))


# ----------------------------------------------------------------------------
# The plan (src/implementation.jl)
# ----------------------------------------------------------------------------
# window ids of src/windowFunctions.jl:4-19 (getWindow)
WINDOWS = {"kaiser_bessel": 0, "gauss": 1, "spline": 2, "kaiser_bessel_rev": 3, "cosh_type": 4}


def window_w(window, k, Nt, m, sigma):
    """psi(k) of a named window (src/windowFunctions.jl:21-134)."""
    return lib().orc_window_w(WINDOWS[window], float(k), float(Nt), int(m), float(sigma))


def window_hat_w(window, n, Nt, m, sigma):
    """psi_hat(n) of a named window (src/windowFunctions.jl:21-134)."""
    return lib().orc_window_hat_w(WINDOWS[window], float(n), float(Nt), int(m), float(sigma))


def bessel_j0(x):
    return lib().orc_bessel_j0(float(x))


class OraclePlan:
    """FP64 CPU restatement of ``NFFTPlan`` (src/implementation.jl:16-193)."""

    def __init__(self, k, N, m=4, sigma=2.0, precompute=LINEAR, blocking=True, cheb=True,
                 workers=1, window="kaiser_bessel"):
        kj, D, M = _julia_nodes(k)
        self.D, self.M, self.N = D, M, tuple(int(n) for n in N)
        assert len(self.N) == D
        Na = np.array(self.N[::-1], dtype=np.int64)
        self._p = lib().orc_plan_create(D, _lp(Na), M, _dp(kj), int(m), float(sigma),
                                        int(precompute), int(bool(blocking)), int(bool(cheb)),
                                        WINDOWS[window])
        if not self._p:
            raise ValueError("nodes must lie in [-1/2, 1/2]")
        Nt = np.zeros(D, dtype=np.int64)
        ls = np.zeros(1, dtype=np.int64)
        sg = np.zeros(1)
        lib().orc_plan_info(self._p, _lp(Nt), _lp(ls), _dp(sg))
        self.Nt = tuple(int(x) for x in Nt[::-1])
        self.LUTSize = int(ls[0])
        self.sigma = float(sg[0])
        self.m = m
        self.workers = workers

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            lib().orc_plan_destroy(p)
            self._p = None

    # stages -----------------------------------------------------------------
    def deconvolve(self, f):
        f = _c128(f)
        g = np.empty(self.Nt, dtype=np.complex128)
        lib().orc_deconvolve(self._p, _dp(f.view(np.float64)), _dp(g.view(np.float64)))
        return g

    def deconvolve_transpose(self, g):
        g = _c128(g)
        f = np.empty(self.N, dtype=np.complex128)
        lib().orc_deconvolve_transpose(self._p, _dp(g.view(np.float64)), _dp(f.view(np.float64)))
        return f

    def convolve(self, g):
        g = _c128(g)
        fhat = np.empty(self.M, dtype=np.complex128)
        lib().orc_convolve(self._p, _dp(g.view(np.float64)), _dp(fhat.view(np.float64)))
        return fhat

    def convolve_transpose(self, fhat):
        fhat = _c128(fhat)
        g = np.empty(self.Nt, dtype=np.complex128)
        lib().orc_convolve_transpose(self._p, _dp(fhat.view(np.float64)),
                                     _dp(g.view(np.float64)))
        return g

    def fft(self, g):
        """forwardFFT: in-place unnormalised, sign -1 (src/implementation.jl:87)."""
        if self.workers > 1:
            import scipy.fft
            return scipy.fft.fftn(g, workers=self.workers, overwrite_x=True)
        return np.fft.fftn(g)

    def bfft(self, g):
        """backwardFFT: unnormalised, sign +1 (src/implementation.jl:88)."""
        if self.workers > 1:
            import scipy.fft
            return scipy.fft.ifftn(g, norm="forward", workers=self.workers, overwrite_x=True)
        return np.fft.ifftn(g, norm="forward")

    # mul! -------------------------------------------------------------------
    def forward(self, f):
        """mul!(fHat, p, f), src/implementation.jl:155-172."""
        g = self.deconvolve(np.asarray(f).reshape(self.N))
        g = self.fft(g)
        return self.convolve(g)

    def adjoint(self, fhat):
        """mul!(f, adjoint(p), fHat), src/implementation.jl:176-193."""
        g = self.convolve_transpose(fhat)
        g = self.bfft(g)
        return self.deconvolve_transpose(g)

    def grad_nodes(self, f, ybar):
        """Node gradient kbar[d, j] = Re(conj(ybar_j) d fhat_j / d k_{d,j}) of
        fhat = forward(f), by differentiating the window of the interpolation
        sum.  Python axis order: returns (D, M)."""
        g = self.fft(self.deconvolve(np.asarray(f).reshape(self.N)))
        g = _c128(g)
        ybar = _c128(ybar)
        out = np.empty((self.M, self.D))
        lib().orc_grad_gather(self._p, _dp(g.view(np.float64)), _dp(ybar.view(np.float64)),
                              _dp(out))
        return np.ascontiguousarray(out.T[::-1])


def sdc(plan, iters=20, weights=None, scale=True):
    """Sampling density compensation over the oracle's stages (FP64): sdc! of
    NFFTTools/src/samplingDensity.jl:93-118 (Pipe & Menon iterations: spread
    the weights, divide the grid by its first-iteration maximum, interpolate,
    divide) and, with ``scale``, :133-155 (least-squares scale c with
    A' diag(w) A 1 ~ 1)."""
    w = np.ones(plan.M) if weights is None else np.array(weights, dtype=np.float64)
    sf = None
    for i in range(iters):
        g = plan.convolve_transpose(w.astype(np.complex128)).real
        if i == 0:
            sf = g.max()
        g = g / sf
        wt = plan.convolve(g.astype(np.complex128)).real / sf
        if np.any(wt <= 0):
            raise ValueError("non-positive weights")
        w = w / wt
    if scale:
        v = plan.adjoint(plan.forward(np.ones(plan.N, dtype=np.complex128)) * w)
        w = w * (v.sum().real / np.sum(np.abs(v) ** 2))
    return w


def toeplitz_kernel_explicit(shape, tr):
    """calculateToeplitzKernel_explicit of NFFTTools/src/Toeplitz.jl:133-156
    (FP64, Python axis order): lambda = FFT(fftshift(E)) over the 2N grid,
    E[n] = sum_j exp(2 pi i sum_d k_dj (n_d - N_d)), n_d in [0, 2 N_d) -- the
    reference's getMatrixElement with 1-based idx = n + 1.  Formed as a sum
    of separable products per node (chunked), then numpy's FFT."""
    tr = np.asarray(tr, dtype=np.float64)
    L = [2 * int(s) for s in shape]
    D = len(L)
    E = np.zeros(L, dtype=np.complex128)
    for j0 in range(0, tr.shape[1], 256):
        k = tr[:, j0:j0 + 256]
        facs = [np.exp(2j * np.pi * np.outer(k[d], np.arange(L[d]) - L[d] // 2)) for d in range(D)]
        if D == 1:
            E += facs[0].sum(axis=0)
        elif D == 2:
            E += facs[0].T @ facs[1]
        else:
            E += np.einsum("ja,jb,jc->abc", facs[0], facs[1], facs[2], optimize=True)
    return np.fft.fftn(np.fft.fftshift(E))


def relerr(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
