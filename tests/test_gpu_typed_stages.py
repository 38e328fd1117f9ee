"""convolve! and convolve_transpose! on real and complex buffers of half,
single and double precision, through the typed C entry points
(nfftm_convolve_typed / nfftm_convolve_transpose_typed).

Modelled on NFFT.jl test/convolve.jl:31-146: a 2-D plan with N = (9, 8)
(Julia order), J = 101 nodes, m = 5, sigma = 2, every precompute mode, FP32
and FP64 plans; the ground truth is the FP64 plan on FP64 buffers (itself
checked against the oracle to 1e-12); every (input type, output type) pair
must agree with it to the tolerance of the worst precision involved
(Julia's isapprox default rtol = sqrt(eps(T)): 3.1e-2 for Float16, 3.5e-4
for Float32, 1.5e-8 for Float64), as the reference's `f/fmax ≈ T.(fff/fmax)`.
Complex input with real output raises ArgumentError, wrong sizes
DimensionMismatch (test/convolve.jl:42-54).
"""
import numpy as np
import pytest
import torch

import nfftm
import oracle as O

pytestmark = pytest.mark.gpu

N = (8, 9)  # Python order of the reference's (9, 8)
J = 101
RTOL = {16: 2 ** -5.5, 32: 2 ** -11.5, 64: 2 ** -26}  # sqrt(eps) per precision (bits)
REAL = {16: torch.float16, 32: torch.float32, 64: torch.float64}
CPX = {16: torch.complex32, 32: torch.complex64, 64: torch.complex128}


def close(a, b, rtol, scale=1.0):
    """isapprox(a / scale, b / scale; rtol), in FP64 on the host."""
    a = a.detach().to("cpu").to(torch.complex128).numpy() / scale
    b = b.detach().to("cpu").to(torch.complex128).numpy() / scale
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


def nodes():
    return np.random.default_rng([77, 1]).random((2, J)) - 0.5


def test_typed_errors():
    print("typed stage operators: error contract", flush=True)
    p = nfftm.plan_nfft(nodes().astype(np.float32), N, m=5, sigma=2.0, precompute=nfftm.LINEAR)
    Nt = p.Ntilde
    dev = "cuda"
    with pytest.raises(nfftm.ArgumentError):
        p.convolve_(torch.ones(Nt, dtype=torch.complex64, device=dev), torch.zeros(J, device=dev))
    with pytest.raises(nfftm.ArgumentError):
        p.convolve_transpose_(torch.ones(J, dtype=torch.complex64, device=dev), torch.zeros(Nt, device=dev))
    with pytest.raises(nfftm.DimensionMismatch):
        p.convolve_(torch.ones(Nt, dtype=torch.complex64, device=dev), torch.zeros(2, dtype=torch.complex64, device=dev))
    with pytest.raises(nfftm.DimensionMismatch):
        p.convolve_(torch.ones(N, dtype=torch.complex64, device=dev), torch.zeros(J, dtype=torch.complex64, device=dev))
    with pytest.raises(nfftm.DimensionMismatch):
        p.convolve_transpose_(torch.ones(2, dtype=torch.complex64, device=dev), torch.zeros(Nt, dtype=torch.complex64, device=dev))
    with pytest.raises(nfftm.DimensionMismatch):
        p.convolve_transpose_(torch.ones(J, dtype=torch.complex64, device=dev), torch.zeros(N, dtype=torch.complex64, device=dev))
    # the library itself checks lengths (any binding): a short output buffer
    L = p._lib
    g = torch.ones(Nt, dtype=torch.complex64, device=dev)
    f = torch.zeros(J - 1, dtype=torch.complex64, device=dev)
    rc = L.nfftm_convolve_typed(p._handle, g.data_ptr(), nfftm._L.C32, g.numel(), f.data_ptr(), nfftm._L.C32,
                                f.numel())
    assert rc == nfftm._L.ERR_DIM
    rc = L.nfftm_convolve_typed(p._handle, g.data_ptr(), nfftm._L.C32, g.numel(), f.data_ptr(), nfftm._L.R32,
                                J)
    assert rc == nfftm._L.ERR_ARG


@pytest.mark.parametrize("pre", [nfftm.LINEAR, nfftm.FULL, nfftm.TENSOR, nfftm.POLYNOMIAL],
                         ids=["LINEAR", "FULL", "TENSOR", "POLYNOMIAL"])
@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_typed_convolve(pre, cplx):
    print(f"typed convolve! pre={pre} complex={cplx}", flush=True)
    k = nodes()
    p32 = nfftm.plan_nfft(k.astype(np.float32), N, m=5, sigma=2.0, precompute=pre)
    p64 = nfftm.plan_nfft(k, N, m=5, sigma=2.0, precompute=pre)
    assert p64.Ntilde == (16, 18)
    T = CPX if cplx else REAL
    rng = np.random.default_rng([77, 2, pre, int(cplx)])
    g64 = torch.as_tensor(rng.random(p64.Ntilde) + (1j * rng.random(p64.Ntilde) if cplx else 0),
                          dtype=T[64], device="cuda")
    fff = torch.empty(J, dtype=T[64], device="cuda")
    assert p64.convolve_(g64, fff) is fff
    orc = O.OraclePlan(k, N, m=5, sigma=2.0, precompute=pre)
    assert O.relerr(fff.cpu().numpy(), orc.convolve(g64.cpu().numpy())) < 1e-12
    fmax = float(fff.abs().max())
    for tg in (16, 32, 64):
        for tf in (32, 64):
            for p, tp in ((p64, 64), (p32, 32)):
                f = torch.empty(J, dtype=T[tf], device="cuda")
                assert p.convolve_(g64.to(T[tg]), f) is f
                assert close(f, fff, RTOL[min(tg, tf, tp)], fmax), (tg, tf, tp)
    for p in (p64, p32):  # Float16 output with the input scaled (test/convolve.jl:88-93)
        f16 = torch.empty(J, dtype=T[16], device="cuda")
        p.convolve_(g64 / fmax, f16)
        assert close(f16, fff / fmax, RTOL[16])


@pytest.mark.parametrize("pre", [nfftm.LINEAR, nfftm.FULL, nfftm.TENSOR, nfftm.POLYNOMIAL],
                         ids=["LINEAR", "FULL", "TENSOR", "POLYNOMIAL"])
@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_typed_convolve_transpose(pre, cplx):
    print(f"typed convolve_transpose! pre={pre} complex={cplx}", flush=True)
    k = nodes()
    p32 = nfftm.plan_nfft(k.astype(np.float32), N, m=5, sigma=2.0, precompute=pre)
    p64 = nfftm.plan_nfft(k, N, m=5, sigma=2.0, precompute=pre)
    T = CPX if cplx else REAL
    rng = np.random.default_rng([77, 3, pre, int(cplx)])
    f64 = torch.as_tensor(rng.random(J) + (1j * rng.random(J) if cplx else 0), dtype=T[64], device="cuda")
    ggg = torch.empty(p64.Ntilde, dtype=T[64], device="cuda")
    assert p64.convolve_transpose_(f64, ggg) is ggg
    orc = O.OraclePlan(k, N, m=5, sigma=2.0, precompute=pre)
    assert O.relerr(ggg.cpu().numpy(), orc.convolve_transpose(f64.cpu().numpy())) < 1e-12
    gmax = float(ggg.abs().max())
    for tg in (32, 64):
        for tf in (16, 32, 64):
            for p, tp in ((p64, 64), (p32, 32)):
                g = torch.empty(p64.Ntilde, dtype=T[tg], device="cuda")
                assert p.convolve_transpose_(f64.to(T[tf]), g) is g
                assert close(g, ggg, RTOL[min(tg, tf, tp)], gmax), (tg, tf, tp)
    for p in (p64, p32):
        g16 = torch.empty(p64.Ntilde, dtype=T[16], device="cuda")
        p.convolve_transpose_(f64 / gmax, g16)
        assert close(g16, ggg / gmax, RTOL[16])
