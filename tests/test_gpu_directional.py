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


def check_directional(N, dims, dtype, pre=nfftm.LINEAR, opre=O.LINEAR, J=None):
    rng = np.random.default_rng([len(N), sum(N), len(dims)])
    J = int(np.prod(N)) if J is None else J
    k = rng.random((len(dims), J)) - 0.5
    cd = torch.complex128 if dtype == torch.float64 else torch.complex64
    p_dir = nfftm.plan_nfft(k, N, dims=dims, m=4, sigma=2.0, precompute=pre, dtype=dtype)
    Nd = tuple(N[a] for a in dims)
    assert p_dir.size_out() == N[:dims[0]] + (J,) + N[dims[-1] + 1:]
    f = torch.as_tensor(rng.standard_normal(N) + 1j * rng.standard_normal(N), dtype=cd, device="cuda")
    fhat = p_dir * f
    g = p_dir.H * fhat
    p = nfftm.plan_nfft(k, Nd, m=4, sigma=2.0, precompute=pre, dtype=dtype)
    orc = O.OraclePlan(k, Nd, m=4, sigma=2.0, precompute=opre) if dtype == torch.float64 else None
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
    return p


@pytest.mark.parametrize("N,dims", CASES, ids=[f"{'x'.join(map(str, c[0]))}-d{''.join(map(str, c[1]))}" for c in CASES])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_directional_equals_slices(N, dims, dtype):
    print(f"directional N={N} dims={dims} {dtype}", flush=True)
    check_directional(N, dims, dtype)


# 1-D directions long enough for the tiled kernels: with the fitted-polynomial
# window they evaluate the window in the kernel and store their results in
# sorted order, gathered back along the strided (axis 0) or contiguous axis.
POLY_CASES = [((512, 6), (0,)), ((6, 512), (1,)), ((3, 400, 4), (1,))]


@pytest.mark.parametrize("N,dims", POLY_CASES, ids=[f"{'x'.join(map(str, c[0]))}-d{''.join(map(str, c[1]))}"
                                                    for c in POLY_CASES])
@pytest.mark.parametrize("pre", ["POLYNOMIAL", "TENSOR"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_directional_polynomial_1d(N, dims, pre, dtype):
    print(f"directional {pre} N={N} dims={dims} {dtype}", flush=True)
    p = check_directional(N, dims, dtype, getattr(nfftm, pre), getattr(O, pre), J=3000)
    assert p.tile_shape()[0][0] < p.Ntilde[0]  # the tiled path ran


STAGE_CASES = [((6, 16, 10), (1,)), ((12, 10, 8), (0, 1)), ((8, 10, 12), (1, 2))]


@pytest.mark.parametrize("N,dims", STAGE_CASES, ids=[f"{'x'.join(map(str, c[0]))}-d{''.join(map(str, c[1]))}"
                                                     for c in STAGE_CASES])
def test_directional_stage_operators(N, dims):
    """deconvolve!, convolve!, convolve_transpose!, deconvolve_transpose! and the
    FFT of a directional plan on the whole arrays (src/directional.jl:1-156)
    equal the slice-wise stage operators of the plan over N[dims] (FP64,
    rel-l2 <= 1e-13)."""
    print(f"directional stages N={N} dims={dims}", flush=True)
    rng = np.random.default_rng([7, len(N), len(dims)])
    J = 300
    k = rng.random((len(dims), J)) - 0.5
    cd = torch.complex128
    p_dir = nfftm.plan_nfft(k, N, dims=dims, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float64)
    Nd = tuple(N[a] for a in dims)
    p = nfftm.plan_nfft(k, Nd, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float64)
    G = p_dir.Ntilde
    assert G == tuple(p.Ntilde[dims.index(a)] if a in dims else N[a] for a in range(len(N)))
    f = torch.as_tensor(rng.standard_normal(N) + 1j * rng.standard_normal(N), dtype=cd, device="cuda")
    fh = torch.as_tensor(rng.standard_normal(p_dir.NOut) + 1j * rng.standard_normal(p_dir.NOut), dtype=cd,
                         device="cuda")
    g = torch.zeros(G, dtype=cd, device="cuda")
    p_dir.deconvolve_(f, g)
    fh_d = torch.zeros(p_dir.NOut, dtype=cd, device="cuda")
    p_dir.convolve_(g, fh_d)
    g2 = torch.zeros(G, dtype=cd, device="cuda")
    p_dir.convolve_transpose_(fh, g2)
    f2 = torch.zeros(N, dtype=cd, device="cuda")
    p_dir.deconvolve_transpose_(g2, f2)
    g3 = g2.clone()
    p_dir.fft_(g3, -1)
    r_g, r_fh = torch.zeros_like(g), torch.zeros_like(fh_d)
    r_g2, r_f2, r_g3 = torch.zeros_like(g2), torch.zeros_like(f2), torch.zeros_like(g3)
    for idx in slices(N, dims):
        oidx = idx[:dims[0]] + (slice(None),) + idx[dims[-1] + 1:]
        gs = torch.zeros(p.Ntilde, dtype=cd, device="cuda")
        p.deconvolve_(f[idx].contiguous(), gs)
        r_g[idx] = gs
        o = torch.zeros(J, dtype=cd, device="cuda")
        p.convolve_(g[idx].contiguous(), o)
        r_fh[oidx] = o
        gs2 = torch.zeros(p.Ntilde, dtype=cd, device="cuda")
        p.convolve_transpose_(fh[oidx].contiguous(), gs2)
        r_g2[idx] = gs2
        fs = torch.zeros(Nd, dtype=cd, device="cuda")
        p.deconvolve_transpose_(g2[idx].contiguous(), fs)
        r_f2[idx] = fs
        gs3 = g2[idx].clone()  # a copy: the FFT is in place (g2[idx] may already be contiguous)
        p.fft_(gs3, -1)
        r_g3[idx] = gs3
    torch.cuda.synchronize()
    errs = {name: rel(a, b) for name, a, b in (("deconvolve", g, r_g), ("convolve", fh_d, r_fh),
                                                ("convolve_transpose", g2, r_g2),
                                                ("deconvolve_transpose", f2, r_f2), ("fft", g3, r_g3))}
    print(errs, flush=True)
    assert all(e < 1e-13 for e in errs.values()), errs


def test_directional_nodes_copy_graphs_errors():
    """nodes! on a directional plan equals a fresh plan, copy() transforms the
    same, graph replay (three replays with other work between) equals eager
    launches, and the operations it does not define raise."""
    N, dims = (6, 32, 20), (1, 2)
    rng = np.random.default_rng(41)
    k1 = rng.random((2, 900)) - 0.5
    k2 = rng.random((2, 700)) - 0.5
    cd = torch.complex64
    f = torch.as_tensor(rng.standard_normal(N) + 1j * rng.standard_normal(N), dtype=cd, device="cuda")
    p = nfftm.plan_nfft(k1, N, dims=dims, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32)
    p.nodes_(k2)
    q = nfftm.plan_nfft(k2, N, dims=dims, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32)
    y = p * f
    assert y.shape == (6, 700) and rel(y, q * f) < 1e-6
    c = p.copy()
    assert isinstance(c, nfftm.DirectionalPlan) and rel(c * f, y) < 1e-6
    x = p.H * y
    pg = nfftm.plan_nfft(k2, N, dims=dims, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32,
                         graphs=True)
    yg = torch.empty_like(y)
    xg = torch.empty_like(x)
    for _ in range(3):
        pg.forward(f, yg)
        pg.adjoint(yg, xg)
        torch.cuda.synchronize()
        _ = (f * 2).sum().item()
        assert rel(yg, y) < 1e-6 and rel(xg, x) < 1e-6
    with pytest.raises(nfftm.NFFTError):
        p.grad_nodes(f, y)
    with pytest.raises(nfftm.NFFTError):
        nfftm.BatchPlan([p], [1])
