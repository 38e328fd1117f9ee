"""GPU parity: libnfftm.so (through the nfftm host package and its C ABI)
against the CPU oracle (FP64 restatement of NFFT.jl) and the direct NDFT.

Tolerances (written here): FP64 GPU vs oracle rel-l2 <= 1e-12;
FP32 GPU vs FP64 NDFT rel-l2 <= 1e-5; the reference's own accuracy contract
(test/accuracy.jl:46-47, test/gpu.jl:22): Kaiser-Bessel, m = 5, sigma = 2,
rel-l2 vs NDFT < 1e-7 in FP64.
"""
import numpy as np
import pytest
import torch

import nfftm
import oracle as O

pytestmark = pytest.mark.gpu

F64_VS_ORACLE = 1e-12
F32_VS_NDFT = 1e-5
KB_M5_VS_NDFT = 1e-7


def rel(a, b):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return O.relerr(a, b)


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def banner(msg):
    print(msg, flush=True)


# test/accuracy.jl:109-112: N sets and (precompute, blocking) combinations
ACC_N = [(255,), (31, 33), (11, 12, 14), (6, 5, 6, 6)]
ACC_MODES = [(nfftm.LINEAR, True), (nfftm.LINEAR, False), (nfftm.FULL, False),
             (nfftm.TENSOR, True), (nfftm.POLYNOMIAL, True), (nfftm.POLYNOMIAL, False)]


@pytest.mark.parametrize("N", ACC_N, ids=lambda n: "x".join(map(str, n)))
@pytest.mark.parametrize("pre,blocking", ACC_MODES, ids=lambda v: str(v))
def test_accuracy_contract(N, pre, blocking):
    banner(f"accuracy N={N} pre={pre} blocking={blocking}")
    rng = np.random.default_rng([7, len(N), pre, int(blocking)])
    D = len(N)
    J = int(np.prod(N))
    k = rng.random((D, J)) - 0.5
    fhat = rng.random(J) + 1j * rng.random(J)
    f_ndft = O.ndft_adjoint(k, N, fhat)
    g_ndft = O.ndft(k, f_ndft)
    orc = O.OraclePlan(k, N, 5, 2.0, pre, blocking)
    f_orc = orc.adjoint(fhat)
    g_orc = orc.forward(f_ndft)

    p64 = nfftm.plan_nfft(k, N, m=5, sigma=2.0, precompute=pre, blocking=blocking,
                          dtype=torch.float64)
    f64 = p64.H * fhat
    g64 = p64 * f_ndft
    assert rel(f64, f_ndft) < KB_M5_VS_NDFT
    assert rel(g64, g_ndft) < KB_M5_VS_NDFT
    assert rel(f64, f_orc) < F64_VS_ORACLE
    assert rel(g64, g_orc) < F64_VS_ORACLE

    p32 = nfftm.plan_nfft(k, N, m=5, sigma=2.0, precompute=pre, blocking=blocking,
                          dtype=torch.float32)
    assert rel(p32.H * fhat, f_ndft) < F32_VS_NDFT
    assert rel(p32 * f_ndft, g_ndft) < F32_VS_NDFT


def _c1():
    # 2-D 64x64, M = 4096 uniform nodes, m = 4, sigma = 2, LINEAR
    rng = np.random.default_rng([123, 1])
    N, M = (64, 64), 4096
    k = rng.random((2, M)) - 0.5
    return N, k, crandn(rng, N), crandn(rng, M)


@pytest.mark.parametrize("pre", [nfftm.LINEAR, nfftm.TENSOR, nfftm.POLYNOMIAL, nfftm.FULL])
def test_c1_2d_vs_oracle_and_ndft(pre):
    banner(f"C1 2-D 64x64 M=4096 m=4 pre={pre}")
    N, k, f, fh = _c1()
    orc = O.OraclePlan(k, N, 4, 2.0, pre, pre != nfftm.FULL)
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=pre, dtype=torch.float64)
    y, x = p * f, p.H * fh
    assert rel(y, orc.forward(f)) < F64_VS_ORACLE
    assert rel(x, orc.adjoint(fh)) < F64_VS_ORACLE
    assert rel(y, O.ndft(k, f)) < 2e-7
    assert rel(x, O.ndft_adjoint(k, N, fh)) < 2e-7
    p32 = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=pre, dtype=torch.float32)
    assert rel(p32 * f, O.ndft(k, f)) < F32_VS_NDFT
    assert rel(p32.H * fh, O.ndft_adjoint(k, N, fh)) < F32_VS_NDFT


@pytest.mark.parametrize("N,M,m", [((64, 64), 4096, 4), ((20, 18, 16), 3000, 4), ((300,), 700, 6),
                                   ((16, 16, 16), 5000, 5), ((40, 36), 2000, 3)])
def test_stages_vs_oracle(N, M, m):
    """deconvolve!, convolve!, convolve_transpose!, deconvolve_transpose! each
    against the oracle (FP64)."""
    banner(f"stages N={N} M={M} m={m}")
    rng = np.random.default_rng([11, M])
    D = len(N)
    k = rng.random((D, M)) - 0.5
    f = crandn(rng, N)
    fh = crandn(rng, M)
    orc = O.OraclePlan(k, N, m, 2.0, O.LINEAR, True)
    p = nfftm.plan_nfft(k, N, m=m, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float64)
    dev = torch.device("cuda")
    g = torch.empty(p.Ntilde, dtype=torch.complex128, device=dev)
    p.deconvolve_(torch.as_tensor(f), g)
    assert rel(g, orc.deconvolve(f)) < F64_VS_ORACLE
    G = crandn(rng, p.Ntilde)
    out = torch.empty(M, dtype=torch.complex128, device=dev)
    p.convolve_(torch.as_tensor(G), out)
    assert rel(out, orc.convolve(G)) < F64_VS_ORACLE
    p.convolve_transpose_(torch.as_tensor(fh), g)
    assert rel(g, orc.convolve_transpose(fh)) < F64_VS_ORACLE
    fo = torch.empty(N, dtype=torch.complex128, device=dev)
    p.deconvolve_transpose_(torch.as_tensor(G), fo)
    assert rel(fo, orc.deconvolve_transpose(G)) < F64_VS_ORACLE


def test_fft_stage_vs_numpy():
    banner("fft stage")
    rng = np.random.default_rng(5)
    for N in [(64, 64), (10, 12, 14), (255,), (6, 5, 6, 6)]:
        D = len(N)
        p = nfftm.plan_nfft(rng.random((D, 10)) - 0.5, N, m=2, sigma=2.0, dtype=torch.float64)
        G = crandn(rng, p.Ntilde)
        g = torch.as_tensor(G).cuda()
        p.fft_(g, -1)
        assert rel(g, np.fft.fftn(G)) < 1e-13
        p.fft_(g, +1)
        assert rel(g, G * np.prod(p.Ntilde)) < 1e-13


def test_issue_106_lut_edge():
    """test/issues.jl:5-15: LINEAR LUT bounds at node 0.008333333f0, N=(240,)."""
    banner("issue 106")
    k = np.array([[np.float32(0.008333333)]], dtype=np.float32).astype(np.float64)
    k = np.concatenate([k, np.zeros((1, 1))], axis=1)
    for blocking in (False, True):
        p = nfftm.plan_nfft(k, (240,), precompute=nfftm.LINEAR, blocking=blocking,
                            dtype=torch.float64)
        out = p.H * np.ones(2, dtype=np.complex128)
        ref = O.ndft_adjoint(k, (240,), np.ones(2))
        assert rel(out, ref) < KB_M5_VS_NDFT


def test_empty_nodes():
    banner("M = 0")
    p = nfftm.plan_nfft(np.zeros((2, 0)), (16, 16), m=4, sigma=2.0, dtype=torch.float64)
    assert p.size_out() == (0,)
    y = p * np.ones((16, 16), dtype=np.complex128)
    assert y.shape == (0,)
    x = p.H * np.zeros(0, dtype=np.complex128)
    assert float(x.abs().max()) == 0.0


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 1e-5)])
def test_adjointness(dtype, tol):
    """<A f, y> = <f, A^H y> for the approximate operator."""
    banner(f"adjointness {dtype}")
    rng = np.random.default_rng(3)
    for N, M in [((32, 30, 28), 6000), ((100, 120), 9000), ((5000,), 4000)]:
        D = len(N)
        k = rng.random((D, M)) - 0.5
        p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=dtype)
        f = crandn(rng, N)
        y = crandn(rng, M)
        Af = (p * f).cpu().numpy().astype(np.complex128)
        Ahy = (p.H * y).cpu().numpy().astype(np.complex128)
        lhs = np.vdot(y, Af)
        rhs = np.vdot(Ahy, f)
        assert abs(lhs - rhs) / abs(lhs) < tol


def test_error_contract():
    """test/convolve.jl:43-54 and test/constructors.jl:15, 181-184."""
    banner("errors")
    rng = np.random.default_rng(9)
    N, J = (9, 8), 101
    k = rng.random((2, J)) - 0.5
    p = nfftm.plan_nfft(k, N, m=5, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32)
    Nt = p.Ntilde
    dev = torch.device("cuda")
    with pytest.raises(nfftm.ArgumentError):
        p.convolve_(torch.ones(Nt, dtype=torch.complex64), torch.zeros(J, device=dev))
    with pytest.raises(nfftm.ArgumentError):
        p.convolve_transpose_(torch.ones(J, dtype=torch.complex64), torch.zeros(Nt, device=dev))
    with pytest.raises(nfftm.DimensionMismatch):
        p.convolve_(torch.ones(Nt, dtype=torch.complex64), torch.zeros(2, dtype=torch.complex64,
                                                                       device=dev))
    with pytest.raises(nfftm.DimensionMismatch):
        p.convolve_(torch.ones(N, dtype=torch.complex64), torch.zeros(J, dtype=torch.complex64,
                                                                      device=dev))
    with pytest.raises(nfftm.DimensionMismatch):
        p.convolve_transpose_(torch.ones(2, dtype=torch.complex64),
                              torch.zeros(Nt, dtype=torch.complex64, device=dev))
    with pytest.raises(nfftm.DimensionMismatch):
        p.convolve_transpose_(torch.ones(J, dtype=torch.complex64),
                              torch.zeros(N, dtype=torch.complex64, device=dev))
    with pytest.raises(nfftm.DimensionMismatch):
        p * np.ones((8, 9), dtype=np.complex64)
    with pytest.raises(nfftm.ArgumentError):
        nfftm.plan_nfft(np.zeros((1, 4)), (2, 2))
    with pytest.raises(nfftm.ArgumentError):
        nfftm.plan_nfft(np.array([[-0.6, 0.9], [0.5, -0.5]]), (2, 2))
    with pytest.raises(nfftm.ArgumentError):
        nfftm.plan_nfft(np.array([[-0.3, 0.3], [0.3, np.nan]]), (2, 2))


def test_nodes_and_copy():
    """test/constructors.jl:42-70 (nodes! equals a fresh plan) and copy."""
    banner("nodes!/copy")
    rng = np.random.default_rng(4)
    Nx = 32
    trj1 = rng.random((2, 1000)) - 0.5
    trj2 = rng.random((2, 1300)) - 0.5
    p1 = nfftm.plan_nfft(trj1, (Nx, Nx), dtype=torch.float64)
    p2 = nfftm.plan_nfft(trj2, (Nx, Nx), dtype=torch.float64)
    p2.nodes_(trj1)
    assert p2.size_out() == p1.size_out()
    f = crandn(rng, (Nx, Nx))
    assert rel(p2 * f, (p1 * f).cpu().numpy()) < 1e-15
    pc = p1.copy()
    assert pc.Ntilde == p1.Ntilde and pc.size_out() == p1.size_out()
    assert rel(pc * f, (p1 * f).cpu().numpy()) < 1e-15
    y = crandn(rng, 1000)
    assert rel(pc.H * y, (p1.H * y).cpu().numpy()) < 1e-15


def test_grid_wraps_small():
    """Ntilde smaller than the stencil (the stencil wraps the whole grid)."""
    banner("tiny grids")
    rng = np.random.default_rng(12)
    for N in [(2, 2), (3,), (2, 3, 2)]:
        D = len(N)
        k = rng.random((D, 7)) - 0.5
        orc = O.OraclePlan(k, N, 5, 2.0, O.LINEAR, True)
        p = nfftm.plan_nfft(k, N, m=5, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float64)
        f = crandn(rng, N)
        y = crandn(rng, 7)
        assert rel(p * f, orc.forward(f)) < F64_VS_ORACLE
        assert rel(p.H * y, orc.adjoint(y)) < F64_VS_ORACLE


@pytest.mark.parametrize("N", [(16, 16, 16), (24, 24, 24), (24, 16, 32)], ids=lambda n: "x".join(map(str, n)))
@pytest.mark.parametrize("m", [2, 4, 6])
def test_3d_whole_tile_grids(N, m):
    """Grids that are whole multiples of the 3-D tile take the block-layout
    scratch and its fold; adjoint and forward against the oracle."""
    banner(f"3-D whole tiles N={N} m={m}")
    rng = np.random.default_rng([43, m, N[2]])
    k = rng.random((3, 3000)) - 0.5
    fh = crandn(rng, 3000)
    f = crandn(rng, N)
    orc = O.OraclePlan(k, N, m, 2.0, O.LINEAR, True)
    x_ref, y_ref = orc.adjoint(fh), orc.forward(f)
    p64 = nfftm.plan_nfft(k, N, m=m, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float64)
    assert rel(p64.H * fh, x_ref) < F64_VS_ORACLE
    assert rel(p64 * f, y_ref) < F64_VS_ORACLE
    p32 = nfftm.plan_nfft(k, N, m=m, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32)
    tol = 1e-5 if m >= 4 else 1e-3  # m = 2 is a coarse window; FP32 only adds rounding
    assert rel(p32.H * fh, x_ref) < tol
    assert rel(p32 * f, y_ref) < tol


@pytest.mark.parametrize("variant,threads", [(0, 512), (0, 1024), (1, 512), (1, 1024)])
@pytest.mark.parametrize("N", [(32, 32, 32), (20, 24, 36)], ids=lambda n: "x".join(map(str, n)))
def test_3d_tile_variants(variant, threads, N):
    """Both 3-D tile geometries (cubic, and flat within half the LDS) and
    both workgroup sizes, on a whole-tile grid (block fold) and an irregular
    one (piece fold), against the oracle."""
    banner(f"3-D variant={variant} threads={threads} N={N}")
    rng = np.random.default_rng([47, variant, threads, N[0]])
    k = rng.random((3, 4000)) - 0.5
    fh, f = crandn(rng, 4000), crandn(rng, N)
    orc = O.OraclePlan(k, N, 4, 2.0, O.LINEAR, True)
    x_ref, y_ref = orc.adjoint(fh), orc.forward(f)
    for dt, tol in ((torch.float64, F64_VS_ORACLE), (torch.float32, 1e-5)):
        p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=dt)
        try:
            shape = p.set_tiles(variant, threads)
        except nfftm.NFFTError:
            continue  # this variant does not fit the grid
        assert shape[1] == threads
        assert rel(p.H * fh, x_ref) < tol
        assert rel(p * f, y_ref) < tol


@pytest.mark.parametrize("N", [(40, 36), (24, 20, 16)], ids=lambda n: "x".join(map(str, n)))
def test_graph_replay_single_plan(N):
    """A (input, output) pair seen twice is replayed from a captured graph:
    every call -- fresh data in the same buffers, unrelated device work in
    between, a nodes! in the middle -- matches the oracle."""
    banner(f"graph replay N={N}")
    rng = np.random.default_rng([53, len(N)])
    D, M = len(N), 3000
    k = rng.random((D, M)) - 0.5
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float64, graphs=True)
    f = torch.empty(N, dtype=torch.complex128, device="cuda")
    fh = torch.empty(M, dtype=torch.complex128, device="cuda")
    y = torch.empty(M, dtype=torch.complex128, device="cuda")
    x = torch.empty(N, dtype=torch.complex128, device="cuda")
    for call in range(5):
        if call == 3:
            k = rng.random((D, M)) - 0.5
            p.nodes_(k)
        fn, fhn = crandn(rng, N), crandn(rng, M)
        f.copy_(torch.as_tensor(fn))
        fh.copy_(torch.as_tensor(fhn))
        p.forward(f, y)
        p.adjoint(fh, x)
        _ = torch.randn(1 << 18, device="cuda").sum()
        torch.cuda.synchronize()
        orc = O.OraclePlan(k, N, 4, 2.0, O.LINEAR, True)
        assert rel(y, orc.forward(fn)) < F64_VS_ORACLE, call
        assert rel(x, orc.adjoint(fhn)) < F64_VS_ORACLE, call
