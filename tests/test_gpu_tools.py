"""NFFTTools on the GPU plan (nfftm_sdc, nfftm_toeplitz_* through the C ABI).

Bars:
* sdc: the reference's known answer (NFFT.jl test/samplingDensity.jl:12-27,
  weights 1/prod(N) on a unit-spaced grid, Julia's default Float32 rtol
  3.45e-4) for every precompute mode and both precisions; against the FP64
  oracle restatement on random nodes rel-l2 <= 1e-12 (FP64) and <= 1e-5
  (FP32); caller-provided weight buffers give the same values.
* Toeplitz: the kernel against the exact sum (calculateToeplitzKernel_explicit
  restated in the oracle) with the reference's tolerances (NFFT.jl
  test/testToeplitz.jl: rtol 1e-6 in FP64, 1e-5 in FP32), for square, odd,
  rectangular and 3-D sizes; the product through the kernel against
  nfft_adjoint(nfft(x)) at rtol 1e-5 (FP32) as there.
"""
import numpy as np
import pytest
import torch

import nfftm
import nfftm.tools as T
import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    if isinstance(b, torch.Tensor):
        b = b.detach().cpu().numpy()
    return O.relerr(a, b)


def unit_grid(N):
    fun = lambda n: np.arange(n) / n - 0.5
    return np.array([[x, y] for x in fun(N[0]) for y in fun(N[1])]).T


@pytest.mark.parametrize("pre", [nfftm.LINEAR, nfftm.FULL, nfftm.TENSOR, nfftm.POLYNOMIAL],
                         ids=["linear", "full", "tensor", "poly"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_sdc_unit_grid(pre, dtype):
    print(f"sdc unit grid pre={pre} {dtype}", flush=True)
    N = (9, 8)
    p = nfftm.plan_nfft(unit_grid(N), N, m=5, sigma=2.0, precompute=pre, dtype=dtype)
    w = T.sdc(p, iters=10)
    assert w.dtype == dtype and w.shape == (72,)
    wn = w.cpu().numpy()
    assert np.all(wn > 0)
    np.testing.assert_allclose(wn, 1 / 72, rtol=3.45e-4)
    # caller-provided buffer: same values, in place
    w2 = torch.ones(72, dtype=dtype, device="cuda")
    out = T.sdc(p, iters=10, weights=w2)
    assert out.data_ptr() == w2.data_ptr()
    assert torch.equal(w2, w)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_sdc_vs_oracle(dtype):
    print(f"sdc oracle {dtype}", flush=True)
    rng = np.random.default_rng(11)
    k = rng.random((2, 3000)) - 0.5
    k[0] *= 0.8  # uneven density
    N = (32, 24)
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=dtype)
    orc = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=O.LINEAR)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    w = T.sdc(p, iters=8)
    assert rel(w, O.sdc(orc, iters=8)) < tol
    w0 = torch.ones(p.J, dtype=dtype, device="cuda")
    T.sdc_(p, 8, w0, scale=False)
    assert rel(w0, O.sdc(orc, iters=8, scale=False)) < tol


def test_sdc_3d_and_errors():
    print("sdc 3d", flush=True)
    rng = np.random.default_rng(12)
    k = rng.random((3, 4000)) - 0.5
    N = (16, 16, 16)
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float64)
    orc = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=O.LINEAR)
    assert rel(T.sdc(p, iters=5), O.sdc(orc, iters=5)) < 1e-12
    with pytest.raises(nfftm.DimensionMismatch):
        T.sdc_(p, 2, torch.ones(10, dtype=torch.float64, device="cuda"))
    with pytest.raises(nfftm.ArgumentError):
        T.sdc_(p, 2, torch.ones(p.J, dtype=torch.float32, device="cuda"))


CASES = [((32, 32), torch.complex128, 1000, 1e-6), ((33, 33), torch.complex128, 1000, 1e-6),
         ((64, 64), torch.complex128, 1000, 1e-6), ((32, 32), torch.complex64, 10000, 1e-5),
         ((32, 33), torch.complex64, 1000, 1e-5), ((16, 16, 16), torch.complex64, 1000, 1e-5),
         ((12, 10, 8), torch.complex128, 1000, 1e-6)]


@pytest.mark.parametrize("shape,cd,M,rtol", CASES, ids=[f"{'x'.join(map(str, c[0]))}-{str(c[1])[6:]}" for c in CASES])
def test_toeplitz_kernel_vs_explicit(shape, cd, M, rtol):
    print(f"toeplitz kernel {shape} {cd}", flush=True)
    rng = np.random.default_rng([len(shape), sum(shape)])
    tr = rng.random((len(shape), M)) - 0.5
    if cd == torch.complex64:
        tr = tr.astype(np.float32)
    lam = T.calculateToeplitzKernel(shape, tr, m=4, sigma=2)
    assert lam.dtype == cd and tuple(lam.shape) == tuple(2 * s for s in shape)
    assert rel(lam, O.toeplitz_kernel_explicit(shape, tr.astype(np.float64))) < rtol


def test_toeplitz_preallocated_and_product():
    """calculateToeplitzKernel! with a preallocated 2N plan, then
    convolveToeplitzKernel! against nfft_adjoint(nfft(x))."""
    print("toeplitz product", flush=True)
    rng = np.random.default_rng(21)
    Nx = 32
    tr = (rng.random((2, 10000)) - 0.5).astype(np.float32)
    lam0 = T.calculateToeplitzKernel((Nx, Nx), tr, m=4, sigma=2)
    p2 = nfftm.plan_nfft(rng.random((2, 10000)) - 0.5, (2 * Nx, 2 * Nx), m=4, sigma=2.0,
                         dtype=torch.float32)
    t = T.Toeplitz((Nx, Nx), torch.complex64)
    lam = torch.empty((2 * Nx, 2 * Nx), dtype=torch.complex64, device="cuda")
    T.calculateToeplitzKernel_(lam, p2, tr, toeplitz=t)
    assert rel(lam, lam0) < 1e-6
    x = torch.as_tensor(rng.standard_normal((Nx, Nx)) + 1j * rng.standard_normal((Nx, Nx)),
                        dtype=torch.complex64, device="cuda")
    p = nfftm.plan_nfft(tr, (Nx, Nx), m=4, sigma=2.0, dtype=torch.float32)
    xN = p.H * (p * x)
    y = x.clone()
    T.convolveToeplitzKernel_(y, lam, toeplitz=t)
    torch.cuda.synchronize()
    assert rel(y, xN) < 1e-5
    # the same product twice in a row on the same object (work array reuse)
    y2 = x.clone()
    T.convolveToeplitzKernel_(y2, lam, toeplitz=t)
    assert torch.equal(y2, y)
    with pytest.raises(nfftm.DimensionMismatch):
        t.kernel(p, lam)  # plan over N, not 2N
