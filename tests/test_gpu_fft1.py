"""Three-pass FFT stage of 1-D plans on large grids (k_fft1.hip) against the
whole-grid rocFFT stage and the CPU oracle.

1-D grids of Nt = n1 Lm^2 points (Lm = 256, 128 or 64; n1 = 2^k or 5 2^k)
run their FFT stage as three passes of line FFTs with the deconvolution
fused into the first forward and the last adjoint pass (the default);
``rocfft=True`` selects rocFFT over the whole grid with separate
deconvolution passes.  Both must give the same transforms: FP64 rel-l2
<= 1e-13 between the two paths and <= 1e-12 against the oracle; FP32
<= 2e-6 between the paths.  The cases cover n1 = 5 M for M = 16 .. 256
(radix-5 combine), n1 = 2^k, each Lm, odd and even N (the fftshift split
Na = N / 2 of src/deconvolution.jl:22-45), and the contention-stress grid
of N = 2^24 at sigma = 1.25.
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


# (N, Nt): n1 x Lm^2 in the comment
CASES = [
    (1 << 20, 1310720),        # 320 x 64^2 (5 x 64)
    (1 << 19, 1 << 20),        # 64 x 128^2
    (3 * (1 << 18) + 1, 5242880),  # 80 x 256^2 (5 x 16), odd N
    (1 << 21, 2621440),        # 160 x 128^2 (5 x 32)
    (1 << 22, 5 * (1 << 21)),  # 160 x 256^2 (5 x 32)
    (1 << 23, 1 << 24),        # 256 x 256^2
    (1 << 24, 20971520),       # 320 x 256^2 (5 x 64)
    (3000000, 1 << 25),        # 512 x 256^2
]
F32_ONLY = [
    (1 << 24, 1 << 26),        # 1024 x 256^2
    (1 << 25, 5 * (1 << 24)),  # 1280 x 256^2 (5 x 256)
]


def make(k, N, Nt, dtype, rocfft, precompute=nfftm.LINEAR):
    p = nfftm.plan_nfft(k, (N,), m=4, sigma=Nt / N, precompute=precompute, dtype=dtype, rocfft=rocfft)
    assert p.Ntilde == (Nt,), (p.Ntilde, Nt)
    return p


def run_case(N, Nt, dtype, M=20000, oracle=False):
    print(f"fft1 N={N} Nt={Nt} dtype={dtype}", flush=True)
    rng = np.random.default_rng([N % 1000003, Nt % 1000003, dtype == torch.float64])
    k = rng.random((1, M)) - 0.5
    p = make(k, N, Nt, dtype, False)
    q = make(k, N, Nt, dtype, True)
    cd = torch.complex128 if dtype == torch.float64 else torch.complex64
    f = torch.as_tensor(crandn(rng, (N,)), dtype=cd, device="cuda")
    fh = torch.as_tensor(crandn(rng, M), dtype=cd, device="cuda")
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    y, yq = p * f, q * f
    x, xq = p.H * fh, q.H * fh
    torch.cuda.synchronize()
    assert rel(y, yq) < tol
    assert rel(x, xq) < tol
    if oracle:
        orc = O.OraclePlan(k, (N,), m=4, sigma=Nt / N, precompute=O.LINEAR)
        assert orc.Nt == (Nt,)
        assert rel(y, orc.forward(f.cpu().numpy())) < 1e-12
        assert rel(x, orc.adjoint(fh.cpu().numpy())) < 1e-12


@pytest.mark.parametrize("N,Nt", CASES, ids=[f"N{c[0]}_Nt{c[1]}" for c in CASES])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_fft1_matches_rocfft(N, Nt, dtype):
    run_case(N, Nt, dtype, oracle=dtype == torch.float64 and Nt <= (1 << 21))


@pytest.mark.parametrize("N,Nt", F32_ONLY, ids=[f"N{c[0]}_Nt{c[1]}" for c in F32_ONLY])
def test_fft1_matches_rocfft_large_f32(N, Nt):
    run_case(N, Nt, torch.float32, M=5000)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_fft1_grad_timing_and_repeat(dtype):
    """Node gradient (forward FFT stage), TimingStats (the fused first pass
    counts as deconvolution, the other passes as the FFT) and repeated
    transforms with other GPU work in between, TENSOR precompute."""
    rng = np.random.default_rng(91)
    N, Nt, M = 1 << 20, 1310720, 30000
    k = rng.random((1, M)) - 0.5
    p = make(k, N, Nt, dtype, False, nfftm.TENSOR)
    q = make(k, N, Nt, dtype, True, nfftm.TENSOR)
    cd = torch.complex128 if dtype == torch.float64 else torch.complex64
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    t = nfftm.TimingStats()
    for it in range(3):
        f = torch.as_tensor(crandn(rng, (N,)), dtype=cd, device="cuda")
        fh = torch.as_tensor(crandn(rng, M), dtype=cd, device="cuda")
        g1 = p.grad_nodes(f, fh)
        g2 = q.grad_nodes(f, fh)
        x = p.adjoint(fh, timing=t)
        _ = torch.randn(1 << 18, device="cuda").sum()
        y = p.forward(f, timing=t)
        torch.cuda.synchronize()
        assert rel(g1, g2) < tol, it
        assert rel(x, q.H * fh) < tol, it
        assert rel(y, q * f) < tol, it
    for key in ("conv", "fft", "deconv", "conv_adjoint", "fft_adjoint", "deconv_adjoint"):
        assert getattr(t, key) > 0, (key, t)


def test_fft1_not_used_off_the_factorisations():
    """Grids outside the factorisations keep the rocFFT stage and still match
    the oracle (Nt = 3 x 2^18 has no n1 x Lm^2 split with a supported n1)."""
    rng = np.random.default_rng(3)
    N, Nt, M = 1 << 18, 786432, 5000
    k = rng.random((1, M)) - 0.5
    p = make(k, N, Nt, torch.float64, False)
    f = torch.as_tensor(crandn(rng, (N,)), dtype=torch.complex128, device="cuda")
    orc = O.OraclePlan(k, (N,), m=4, sigma=Nt / N, precompute=O.LINEAR)
    assert rel(p * f, orc.forward(f.cpu().numpy())) < 1e-12


FOLD_CASES = [(nfftm.TENSOR, torch.float32), (nfftm.TENSOR, torch.float64),
              (nfftm.LINEAR, torch.float32), (nfftm.LINEAR, torch.float64)]


@pytest.mark.parametrize("pre,dtype", FOLD_CASES, ids=["tensor-f32", "tensor-f64", "linear-f32", "linear-f64"])
def test_fft1_fused_fold_clustered(pre, dtype):
    """The adjoint's fold fused into the first FFT pass against the separate
    fold (``fused_fold=False``) and the rocFFT path, on nodes clustered at
    the ends of the period (Beta(1/2, 1/2), as the contention-stress
    workload) so that end tiles hold many work items and most points near
    them take the general fold; FP64 <= 1e-13 between the paths and
    <= 1e-12 against the oracle, FP32 <= 2e-6 between the paths."""
    print(f"fft1 fused fold {pre} {dtype}", flush=True)
    rng = np.random.default_rng([11, pre, dtype == torch.float64])
    N, Nt, M = 1 << 20, 1310720, 400000
    k = (rng.beta(0.5, 0.5, (1, M)) - 0.5)
    kw = dict(m=4, sigma=Nt / N, precompute=pre, dtype=dtype)
    p = nfftm.plan_nfft(k, (N,), **kw)
    q = nfftm.plan_nfft(k, (N,), fused_fold=False, **kw)
    r = nfftm.plan_nfft(k, (N,), rocfft=True, **kw)
    assert p.Ntilde == (Nt,)
    cd = torch.complex128 if dtype == torch.float64 else torch.complex64
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    for it in range(2):
        fh = torch.as_tensor(crandn(rng, M), dtype=cd, device="cuda")
        x, xq, xr = p.H * fh, q.H * fh, r.H * fh
        torch.cuda.synchronize()
        assert rel(x, xq) < tol, it
        assert rel(x, xr) < tol, it
    if dtype == torch.float64 and pre == nfftm.LINEAR:
        orc = O.OraclePlan(k, (N,), m=4, sigma=Nt / N, precompute=O.LINEAR)
        assert rel(x, orc.adjoint(fh.cpu().numpy())) < 1e-12


@pytest.mark.parametrize("pre", [nfftm.LINEAR, nfftm.TENSOR], ids=["linear", "tensor"])
def test_fft1_empty_nodes(pre):
    """M = 0 on a three-pass grid: the forward is empty and the adjoint, with
    its fold inside the first FFT pass (tiles without work items), is zero;
    nodes! then brings nodes back and the plan matches the rocFFT path."""
    print(f"fft1 empty nodes {pre}", flush=True)
    N, Nt = 1 << 20, 1310720
    p = nfftm.plan_nfft(np.zeros((1, 0)), (N,), m=4, sigma=Nt / N, precompute=pre, dtype=torch.float32)
    assert p.Ntilde == (Nt,)
    y = p * torch.ones(N, dtype=torch.complex64, device="cuda")
    assert y.shape == (0,)
    x = p.H * torch.zeros(0, dtype=torch.complex64, device="cuda")
    assert float(x.abs().max()) == 0.0
    rng = np.random.default_rng(5)
    k = rng.random((1, 3000)) - 0.5
    p.nodes_(k)
    q = make(k, N, Nt, torch.float32, True, pre)
    fh = torch.as_tensor(crandn(rng, 3000), dtype=torch.complex64, device="cuda")
    assert rel(p.H * fh, q.H * fh) < 2e-6
