"""GPU node gradient (nfftm_grad_nodes and nfftm_grad_nodes_grid through the
C ABI) and the differentiation rules built on it.

The reference's rrules return NoTangent() for the nodes
(AbstractNFFTs/ext/AbstractNFFTsChainRulesCoreExt.jl:14-45); the library adds
kbar[d, j] = Re(conj(ybar_j) d fhat_j / d k[d, j]).  Bars, written here:
* FP64 GPU vs the oracle's gradient (same formula, CPU): rel-l2 <= 1e-12;
* FP32 GPU vs the oracle: rel-l2 <= 1e-5;
* FP64 GPU vs FP64 central differences (h = 1e-6) of the loss
  ||A(k) f - y||^2 evaluated by the GPU plan, on 256 nodes of the 2-D spiral
  workload (256^2 image, 65,536 nodes): rel-l2 <= 1e-6.
"""
import numpy as np
import pytest
import torch

import nfftm
import nfftm.autograd as AD
import oracle as O

pytestmark = pytest.mark.gpu

H = 1e-6


def rel(a, b):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    if isinstance(b, torch.Tensor):
        b = b.detach().cpu().numpy()
    return O.relerr(a, b)


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def banner(msg):
    print(msg, flush=True)


def spiral_nodes(interleaves=8, samples=8192, turns=16):
    """Archimedean spiral: r(t) = t / 2, theta(t) = 2 pi turns t + 2 pi q /
    interleaves, t = i / samples; radius <= 1/2."""
    t = np.arange(samples) / samples
    ks = []
    for q in range(interleaves):
        th = 2 * np.pi * turns * t + 2 * np.pi * q / interleaves
        r = 0.5 * t
        ks.append(np.stack([r * np.cos(th), r * np.sin(th)]))
    return np.concatenate(ks, axis=1)


GRAD_N = [(70,), (24, 20), (10, 12, 8), (5, 6, 4, 6)]
GRAD_MODES = [(nfftm.POLYNOMIAL, True), (nfftm.LINEAR, True), (nfftm.TENSOR, True),
              (nfftm.FULL, False), (nfftm.POLYNOMIAL, False)]


@pytest.mark.parametrize("N", GRAD_N, ids=lambda n: "x".join(map(str, n)))
@pytest.mark.parametrize("pre,blocking", GRAD_MODES, ids=lambda v: str(v))
def test_grad_vs_oracle(N, pre, blocking):
    banner(f"grad N={N} pre={pre} blocking={blocking}")
    rng = np.random.default_rng([17, len(N), pre, int(blocking)])
    D, M = len(N), 300
    k = rng.random((D, M)) - 0.5
    f = crandn(rng, N)
    ybar = crandn(rng, M)
    ref = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=pre, blocking=blocking).grad_nodes(f, ybar)
    p64 = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=pre, blocking=blocking,
                          dtype=torch.float64)
    assert rel(p64.grad_nodes(f, ybar), ref) < 1e-12
    p32 = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=pre, blocking=blocking,
                          dtype=torch.float32)
    assert rel(p32.grad_nodes(f, ybar), ref) < 1e-5


def test_grad_m6_matches_ndft_closed_form():
    banner("grad m=6 vs closed form")
    rng = np.random.default_rng(19)
    N, M = (32, 28), 500
    k = rng.random((2, M)) - 0.5
    f, ybar = crandn(rng, N), crandn(rng, M)
    p = nfftm.plan_nfft(k, N, m=6, sigma=2.0, dtype=torch.float64)
    assert rel(p.grad_nodes(f, ybar), O.ndft_grad(k, f, ybar)) < 1e-8


def _spiral_problem():
    rng = np.random.default_rng([123, 4])
    k = spiral_nodes()
    N = (256, 256)
    f = crandn(rng, N).astype(np.complex64).astype(np.complex128)
    delta = rng.normal(0.0, 1e-3, size=k.shape)
    ky = np.clip(k + delta, -0.5, 0.5)  # perturbed trajectory, kept in range
    return rng, k, ky, N, f


def test_spiral_grad_vs_finite_differences():
    banner("spiral 256^2, M=65536: grad vs FD")
    rng, k, ky, N, f = _spiral_problem()
    kw = dict(m=4, sigma=2.0, precompute=nfftm.POLYNOMIAL, dtype=torch.float64)
    y = nfftm.plan_nfft(ky, N, **kw) * f
    p = nfftm.plan_nfft(k.copy(), N, **kw)
    r = p * f - y
    g = p.grad_nodes(f, 2 * r).cpu().numpy()
    sel = rng.choice(k.shape[1], 256, replace=False)
    fd = np.zeros((2, sel.size))
    for d in range(2):
        rows = []
        for s in (+1, -1):
            kk = k.copy()
            kk[d, sel] += s * H
            rows.append((torch.abs(p.nodes_(kk) * f - y) ** 2).cpu().numpy()[sel])
        fd[d] = (rows[0] - rows[1]) / (2 * H)
    err = rel(g[:, sel], fd)
    print(f"spiral FP64 grad vs FD: {err:.3e}", flush=True)
    assert err < 1e-6
    # the CPU oracle's gradient on the same nodes
    orc = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=O.POLYNOMIAL)
    assert rel(g, orc.grad_nodes(f, 2 * r.cpu().numpy())) < 1e-12
    # FP32 plan: error of its gradient against the FP64 one
    p.nodes_(k.copy())
    p32 = nfftm.plan_nfft(k.copy(), N, m=4, sigma=2.0, precompute=nfftm.POLYNOMIAL,
                          dtype=torch.float32)
    e32 = rel(p32.grad_nodes(f, 2 * r), g)
    print(f"spiral FP32 grad vs FP64 grad: {e32:.3e}", flush=True)
    assert e32 < 1e-5


def test_autograd_function_and_rrule():
    banner("autograd")
    rng = np.random.default_rng(23)
    N, M = (40, 36), 2000
    k = torch.tensor(rng.random((2, M)) - 0.5, dtype=torch.float64, requires_grad=True)
    x = torch.tensor(crandn(rng, N), dtype=torch.complex128, device="cuda", requires_grad=True)
    y = torch.tensor(crandn(rng, M), dtype=torch.complex128, device="cuda")
    yh = AD.nfft(k, x, m=4, sigma=2.0)
    loss = (torch.abs(yh - y) ** 2).sum()
    loss.backward()
    kn = k.detach().numpy()
    r = (yh - y).detach()
    _, pullback = AD.rrule_nfft(kn, x.detach(), m=4, sigma=2.0, dtype=torch.float64)
    kbar, xbar = pullback(2 * r)
    assert k.grad.device.type == "cpu" and k.grad.dtype == torch.float64
    assert rel(k.grad, kbar) < 1e-14
    assert rel(x.grad, xbar) < 1e-14
    # against the exact (direct-sum) operator: the m = 4 window's accuracy,
    # ~1e-7 for the transform and ~1e-6 for its node derivative
    xn = x.detach().cpu().numpy()
    assert rel(x.grad, O.ndft_adjoint(kn, N, 2 * r.cpu().numpy())) < 1e-6
    assert rel(k.grad, O.ndft_grad(kn, xn, 2 * r.cpu().numpy())) < 1e-5


GRID_CASES = [((64, 48), nfftm.POLYNOMIAL), ((300,), nfftm.LINEAR), ((12, 10, 16), nfftm.TENSOR),
              ((30, 20), nfftm.LINEAR)]


@pytest.mark.parametrize("N,pre", GRID_CASES, ids=lambda v: str(v))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_grad_from_forward_grid(N, pre, dtype):
    """nfftm_grad_nodes_grid on the grid a forward transform leaves (the
    plan's own grid, and a copy of it after other transforms) equals
    nfftm_grad_nodes recomputing that grid, bit for bit: the same gather on
    the same grid.  Power-of-two and other grids (pruned passes and rocFFT)."""
    banner(f"grad from grid N={N} pre={pre} {dtype}")
    rng = np.random.default_rng([len(N), sum(N), pre])
    M = 700
    k = rng.random((len(N), M)) - 0.5
    cd = torch.complex128 if dtype == torch.float64 else torch.complex64
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=pre, dtype=dtype)
    f = torch.as_tensor(crandn(rng, N), dtype=cd, device="cuda")
    ybar = torch.as_tensor(crandn(rng, M), dtype=cd, device="cuda")
    ref = p.grad_nodes(f, ybar)
    _ = p * f
    g = p.tmpVec.clone()
    assert torch.equal(p.grad_nodes_grid(p.tmpVec, ybar), ref)
    _ = p.H * ybar  # the plan's grid now holds other values
    assert torch.equal(p.grad_nodes_grid(g, ybar), ref)
    with pytest.raises(nfftm.DimensionMismatch):
        p.grad_nodes_grid(g.reshape(-1)[:-1], ybar)


def test_grad_edge_cases():
    banner("grad edge cases")
    rng = np.random.default_rng(29)
    N = (16, 16)
    k = rng.random((2, 50)) - 0.5
    k[:, 0] = [-0.5, 0.5]   # the closed ends of the node range
    f = crandn(rng, N)
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, dtype=torch.float64)
    assert torch.all(p.grad_nodes(f, np.zeros(50, complex)) == 0)
    ref = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=O.POLYNOMIAL).grad_nodes(f, np.ones(50))
    assert rel(p.grad_nodes(f, np.ones(50, complex)), ref) < 1e-12
    p0 = nfftm.plan_nfft(np.zeros((2, 0)), N, m=4, sigma=2.0, dtype=torch.float64)
    assert p0.grad_nodes(f, np.zeros(0, complex)).shape == (2, 0)
    with pytest.raises(nfftm.DimensionMismatch):
        p.grad_nodes(f, np.zeros(49, complex))
