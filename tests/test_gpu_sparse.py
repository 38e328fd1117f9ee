"""Sparse-matrix convolution (NFFTM_FLAG_SPARSE, Python sparse_matrix=True):
the reference GPU backend's form of precompute = FULL
(ext/NFFTGPUArraysExt/implementation.jl:3-136, B built by
src/precomputation.jl:81-120) as a comparison baseline.  The plan stores the
W^D window weights and grid indices of every node on the device and convolves
by a gather product (forward) and an atomic scatter (adjoint).

Bars: against the default plan of the same precompute mode (the same
weights, another summation order) FP64 rel-l2 <= 1e-13, FP32 <= 2e-6;
against the FP64 oracle <= 1e-12; nodes! rebuilds the matrix; the FP32
adjoint differs bitwise from the tiled path's (atomic float sums against
fixed point), which shows that the flag selects the matrix path.
"""
import numpy as np
import pytest
import torch

import nfftm
import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    if isinstance(b, torch.Tensor):
        b = b.detach().cpu().numpy()
    return O.relerr(a, b)


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


CASES = [((64,), nfftm.FULL, torch.float64), ((24, 20), nfftm.LINEAR, torch.float64),
         ((16, 12, 10), nfftm.FULL, torch.float64), ((16, 16, 16), nfftm.LINEAR, torch.float32),
         ((32, 32), nfftm.TENSOR, torch.float32), ((48,), nfftm.POLYNOMIAL, torch.float32)]
ORC = {nfftm.FULL: O.FULL, nfftm.LINEAR: O.LINEAR, nfftm.TENSOR: O.TENSOR, nfftm.POLYNOMIAL: O.POLYNOMIAL}


@pytest.mark.parametrize("N,pre,dtype", CASES,
                         ids=[f"{'x'.join(map(str, c[0]))}-{c[1]}-{str(c[2])[6:]}" for c in CASES])
def test_sparse_matrix_matches_default(N, pre, dtype):
    print(f"sparse N={N} pre={pre} {dtype}", flush=True)
    rng = np.random.default_rng([len(N), sum(N), pre])
    M = 2000
    D = len(N)
    k = rng.random((D, M)) - 0.5
    kw = dict(m=4, sigma=2.0, precompute=pre, dtype=dtype)
    p = nfftm.plan_nfft(k, N, sparse_matrix=True, **kw)
    q = nfftm.plan_nfft(k, N, **kw)
    cd = torch.complex128 if dtype == torch.float64 else torch.complex64
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    for rnd in range(2):
        f = torch.as_tensor(crandn(rng, N), dtype=cd, device="cuda")
        fh = torch.as_tensor(crandn(rng, M), dtype=cd, device="cuda")
        y, yq = p * f, q * f
        x, xq = p.H * fh, q.H * fh
        torch.cuda.synchronize()
        assert rel(y, yq) < tol, rel(y, yq)
        assert rel(x, xq) < tol, rel(x, xq)
        if dtype == torch.float32 and D == 3:
            assert not torch.equal(x, xq)
        if dtype == torch.float64:
            orc = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=ORC[pre])
            assert rel(y, orc.forward(f.cpu().numpy())) < 1e-12
            assert rel(x, orc.adjoint(fh.cpu().numpy())) < 1e-12
        # nodes! rebuilds the matrix
        k = rng.random((D, M)) - 0.5
        p.nodes_(k)
        q.nodes_(k)


def test_sparse_matrix_standalone_convolutions():
    """convolve! / convolve_transpose! on a caller grid through the matrix."""
    print("sparse convolve", flush=True)
    rng = np.random.default_rng(41)
    N, M = (20, 16), 1500
    k = rng.random((2, M)) - 0.5
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.FULL, dtype=torch.float64, sparse_matrix=True)
    q = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.FULL, dtype=torch.float64)
    g = torch.as_tensor(crandn(rng, p.Ntilde), dtype=torch.complex128, device="cuda")
    fh = torch.as_tensor(crandn(rng, M), dtype=torch.complex128, device="cuda")
    y = torch.empty(M, dtype=torch.complex128, device="cuda")
    yq = torch.empty_like(y)
    p.convolve_(g, y)
    q.convolve_(g, yq)
    gt = torch.empty(tuple(p.Ntilde), dtype=torch.complex128, device="cuda")
    gq = torch.empty_like(gt)
    p.convolve_transpose_(fh, gt)
    q.convolve_transpose_(fh, gq)
    torch.cuda.synchronize()
    assert rel(y, yq) < 1e-13
    assert rel(gt, gq) < 1e-13
