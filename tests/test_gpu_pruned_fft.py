"""Pruned FFT passes (k_fft.hip) against the whole-grid rocFFT stage.

On power-of-two grids a plan runs its FFT stage as pruned line passes with
the deconvolution fused (the default); ``rocfft=True`` selects rocFFT over
the whole oversampled grid with separate deconvolution passes.  Both must
give the same transforms: FP64 rel-l2 <= 1e-13 between the two paths and
<= 1e-12 against the CPU oracle; FP32 <= 2e-6 between the paths.  The cases
cover D = 1, 2, 3, odd and even N (the fftshift split Na = N / 2 of
src/deconvolution.jl:22-45), grids with different sizes per dimension, and
line lengths from 16 to 1024 points.
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


# (N in C order, sigma); every oversampled size is a power of two
CASES = [
    ((100,), 2.56),            # Nt = 256, N even
    ((61,), 256 / 61),         # Nt = 256, N odd
    ((300,), 1024 / 300),      # Nt = 1024
    ((32, 64), 2.0),           # Nt = (64, 128)
    ((15, 30), 64 / 30),       # Nt = (32, 64), odd N
    ((8, 16, 32), 2.0),        # Nt = (16, 32, 64)
    ((16, 8, 31), 2.05),       # Nt = (32, 16, 64), odd N
    ((64, 64, 64), 2.0),       # Nt = 128^3
]


def make(k, N, sigma, dtype, rocfft):
    return nfftm.plan_nfft(k, N, m=4, sigma=sigma, precompute=nfftm.LINEAR, dtype=dtype, rocfft=rocfft)


@pytest.mark.parametrize("N,sigma", CASES, ids=[f"{'x'.join(map(str, c[0]))}" for c in CASES])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_pruned_matches_rocfft(N, sigma, dtype):
    print(f"pruned N={N} sigma={sigma:.4f} dtype={dtype}", flush=True)
    rng = np.random.default_rng([len(N), sum(N), dtype == torch.float64])
    M = 3000 if np.prod(N) < 1e5 else 40000
    k = rng.random((len(N), M)) - 0.5
    p = make(k, N, sigma, dtype, False)
    q = make(k, N, sigma, dtype, True)
    assert all(n & (n - 1) == 0 for n in p.Ntilde), p.Ntilde
    cd = torch.complex128 if dtype == torch.float64 else torch.complex64
    f = torch.as_tensor(crandn(rng, N), dtype=cd, device="cuda")
    fh = torch.as_tensor(crandn(rng, M), dtype=cd, device="cuda")
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    y, yq = p * f, q * f
    x, xq = p.H * fh, q.H * fh
    torch.cuda.synchronize()
    assert rel(y, yq) < tol
    assert rel(x, xq) < tol
    if dtype == torch.float64 and np.prod(N) <= 4096:
        orc = O.OraclePlan(k, N, m=4, sigma=sigma, precompute=O.LINEAR)
        assert rel(y, orc.forward(f.cpu().numpy())) < 1e-12
        assert rel(x, orc.adjoint(fh.cpu().numpy())) < 1e-12


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_pruned_grad_and_repeat(dtype):
    """Node gradient (forward FFT stage) on both paths, and repeated
    transforms on one plan with other work in between (the pruned passes
    leave grid regions unwritten that no later pass reads)."""
    rng = np.random.default_rng(77)
    N, M = (16, 32, 16), 5000
    k = rng.random((3, M)) - 0.5
    p = make(k, N, 2.0, dtype, False)
    q = make(k, N, 2.0, dtype, True)
    cd = torch.complex128 if dtype == torch.float64 else torch.complex64
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    for it in range(3):
        f = torch.as_tensor(crandn(rng, N), dtype=cd, device="cuda")
        fh = torch.as_tensor(crandn(rng, M), dtype=cd, device="cuda")
        g1 = p.grad_nodes(f, fh)
        g2 = q.grad_nodes(f, fh)
        x = p.H * fh
        _ = torch.randn(1 << 18, device="cuda").sum()
        y = p * f
        torch.cuda.synchronize()
        assert rel(g1, g2) < tol, it
        assert rel(x, q.H * fh) < tol, it
        assert rel(y, q * f) < tol, it


def test_pruned_timing_stages():
    """TimingStats on the pruned path: the fused dimension-1 pass counts as
    deconvolution, the other passes as the FFT."""
    rng = np.random.default_rng(5)
    N, M = (32, 32, 32), 20000
    k = rng.random((3, M)) - 0.5
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32)
    q = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32, rocfft=True)
    f = torch.as_tensor(crandn(rng, N), dtype=torch.complex64, device="cuda")
    fh = torch.as_tensor(crandn(rng, M), dtype=torch.complex64, device="cuda")
    t = nfftm.TimingStats()
    y = p.forward(f, timing=t)
    x = p.adjoint(fh, timing=t)
    torch.cuda.synchronize()
    for key in ("conv", "fft", "deconv", "conv_adjoint", "fft_adjoint", "deconv_adjoint"):
        assert getattr(t, key) > 0, (key, t)
    assert rel(y, q * f) < 2e-6
    assert rel(x, q.H * fh) < 2e-6


def test_pruned_graph_replay():
    """HIP-graph replay (NFFTM_FLAG_GRAPHS) of transforms on the pruned FFT
    path: every replay equals the direct launches, with other device work in
    between."""
    rng = np.random.default_rng(17)
    N, M = (32, 32, 16), 30000
    k = rng.random((3, M)) - 0.5
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32, graphs=True)
    q = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32)
    f = torch.as_tensor(crandn(rng, N), dtype=torch.complex64, device="cuda")
    fh = torch.as_tensor(crandn(rng, M), dtype=torch.complex64, device="cuda")
    y = torch.empty(M, dtype=torch.complex64, device="cuda")
    x = torch.empty(N, dtype=torch.complex64, device="cuda")
    y0, x0 = q * f, q.H * fh
    for it in range(4):
        f.mul_(1.0)  # same values, other device work in between
        p.forward(f, y)
        p.adjoint(fh, x)
        _ = torch.randn(1 << 16, device="cuda").sum()
        torch.cuda.synchronize()
        assert rel(y, y0) < 1e-6, it
        assert rel(x, x0) < 1e-6, it



# 3-D adjoint with the fold fused into the first pruned pass (z lines of 256
# points; FP32 16 x 16 x 8 tiles, FP64 8 x 8 x 8): N in C order, z first
FUSED = [((128, 16, 16), torch.float32), ((128, 32, 16), torch.float32),
         ((128, 16, 16), torch.float64), ((128, 16, 32), torch.float64)]


@pytest.mark.parametrize("N,dtype", FUSED, ids=[f"{'x'.join(map(str, c[0]))}-{str(c[1])[6:]}" for c in FUSED])
def test_fused_fold_first_pass(N, dtype):
    """The fused adjoint equals the fold + separate pass (bit for bit in FP32,
    whose spreading sums in fixed point; FP64 spreading adds floats in LDS in
    arrival order, so two runs agree to rounding: <= 1e-14), never writes the
    folded grid (z positions outside S_z keep what the grid held), and
    matches the oracle (FP64 <= 1e-12, FP32 <= 1e-5 against the FP64
    oracle)."""
    print(f"fused fold N={N} {dtype}", flush=True)
    rng = np.random.default_rng([sum(N), dtype == torch.float64])
    M = 20000
    k = rng.random((3, M)) - 0.5
    kw = dict(m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=dtype)
    p = nfftm.plan_nfft(k, N, **kw)
    q = nfftm.plan_nfft(k, N, fused_fold=False, **kw)
    cd = torch.complex128 if dtype == torch.float64 else torch.complex64
    fh = torch.as_tensor(crandn(rng, M), dtype=cd, device="cuda")
    gap = slice(N[0] // 2, 2 * N[0] - N[0] // 2)  # z positions outside S_z
    for plan in (p, q):
        plan.tmpVec.fill_(7.0)
    same = torch.equal if dtype == torch.float32 else (lambda a, b: rel(a, b) < 1e-14)
    x, xq = p.H * fh, q.H * fh
    torch.cuda.synchronize()
    assert same(x, xq), rel(x, xq)
    assert bool((p.tmpVec[gap] == 7.0).all())
    assert not bool((q.tmpVec[gap] == 7.0).all())
    # a second adjoint on new values through the same plan
    fh2 = torch.as_tensor(crandn(rng, M), dtype=cd, device="cuda")
    assert same(p.H * fh2, q.H * fh2)
    orc = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=O.LINEAR)
    assert rel(x, orc.adjoint(fh.cpu().numpy())) < (1e-12 if dtype == torch.float64 else 1e-5)


def test_fused_fold_falls_back_on_crowded_tiles():
    """A tile with more nodes than one work item holds keeps the separate fold
    (the folded grid is written) and still matches the oracle."""
    print("fused fold fallback", flush=True)
    rng = np.random.default_rng(33)
    N = (128, 16, 16)
    k = rng.random((3, 12000)) - 0.5
    k[:, :3000] = 0.01 * rng.random((3, 3000))  # 3000 nodes in one tile
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float64)
    fh = torch.as_tensor(crandn(rng, k.shape[1]), dtype=torch.complex128, device="cuda")
    p.tmpVec.fill_(7.0)
    x = p.H * fh
    torch.cuda.synchronize()
    assert not bool((p.tmpVec[N[0] // 2:2 * N[0] - N[0] // 2] == 7.0).all())
    orc = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=O.LINEAR)
    assert rel(x, orc.adjoint(fh.cpu().numpy())) < 1e-12


def test_fused_fold_graph_replay():
    """HIP-graph replay of adjoints that take the fused fold: every replay
    equals the direct launches (FP32 fixed-point spreading: bit for bit)."""
    print("fused fold graph replay", flush=True)
    rng = np.random.default_rng(37)
    N, M = (128, 16, 16), 20000
    k = rng.random((3, M)) - 0.5
    kw = dict(m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32)
    p = nfftm.plan_nfft(k, N, graphs=True, **kw)
    q = nfftm.plan_nfft(k, N, **kw)
    fh = torch.as_tensor(crandn(rng, M), dtype=torch.complex64, device="cuda")
    x = torch.empty(N, dtype=torch.complex64, device="cuda")
    x0 = q.H * fh
    for it in range(4):
        p.adjoint(fh, x)
        _ = torch.randn(1 << 16, device="cuda").sum()
        torch.cuda.synchronize()
        assert torch.equal(x, x0), it
    q.tmpVec.fill_(7.0)
    _ = q.H * fh
    torch.cuda.synchronize()
    assert bool((q.tmpVec[N[0] // 2:2 * N[0] - N[0] // 2] == 7.0).all())  # the fused path ran
