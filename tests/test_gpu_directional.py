"""Directional plans (plan_nfft(k, N; dims), src/directional.jl) on the GPU.

As the reference's test (NFFT.jl test/accuracy.jl:83-165): the NFFT along
``dims`` equals the ``len(dims)``-dimensional NFFT of every slice; here
against the GPU plan over N[dims] slice by slice (FP64 rel-l2 <= 1e-13: the
batch launches compute the same values) and against the CPU oracle slice by
slice (FP64 <= 1e-12).  Axes are those of the C-order array.
"""
import itertools

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


def slices(shape, dims):
    """index tuples of the slices: ':' on dims, every index elsewhere"""
    other = [range(n) if a not in dims else [slice(None)] for a, n in enumerate(shape)]
    return list(itertools.product(*other))


CASES = [((8, 12), (0,)), ((8, 12), (1,)), ((10, 8, 6), (0,)), ((10, 8, 6), (1,)), ((10, 8, 6), (2,)),
         ((10, 8, 12), (0, 1)), ((10, 8, 12), (1, 2)), ((4, 8, 6, 10), (1, 2, 3))]


@pytest.mark.parametrize("N,dims", CASES, ids=[f"{'x'.join(map(str, c[0]))}-d{''.join(map(str, c[1]))}" for c in CASES])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_directional_equals_slices(N, dims, dtype):
    print(f"directional N={N} dims={dims} {dtype}", flush=True)
    rng = np.random.default_rng([len(N), sum(N), len(dims)])
    J = int(np.prod(N))
    k = rng.random((len(dims), J)) - 0.5
    cd = torch.complex128 if dtype == torch.float64 else torch.complex64
    p_dir = nfftm.plan_nfft(k, N, dims=dims, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=dtype)
    Nd = tuple(N[a] for a in dims)
    assert p_dir.size_out() == N[:dims[0]] + (J,) + N[dims[-1] + 1:]
    f = torch.as_tensor(rng.standard_normal(N) + 1j * rng.standard_normal(N), dtype=cd, device="cuda")
    fhat = p_dir * f
    g = p_dir.H * fhat
    p = nfftm.plan_nfft(k, Nd, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=dtype)
    orc = O.OraclePlan(k, Nd, m=4, sigma=2.0, precompute=O.LINEAR) if dtype == torch.float64 else None
    ref_fhat = torch.empty_like(fhat)
    ref_g = torch.empty_like(g)
    ofh, og = np.empty(fhat.shape, complex), np.empty(g.shape, complex)
    fn = f.cpu().numpy()
    fhn = fhat.cpu().numpy()
    for idx in slices(N, dims):
        oidx = idx[:dims[0]] + (slice(None),) + idx[dims[-1] + 1:]
        ref_fhat[oidx] = p * f[idx]
        ref_g[idx] = p.H * fhat[oidx]
        if orc is not None:
            ofh[oidx] = orc.forward(fn[idx])
            og[idx] = orc.adjoint(fhn[oidx])
    tol = 1e-13 if dtype == torch.float64 else 1e-6
    assert rel(fhat, ref_fhat) < tol
    assert rel(g, ref_g) < tol
    if orc is not None:
        assert rel(fhat, ofh) < 1e-12
        assert rel(g, og) < 1e-12
