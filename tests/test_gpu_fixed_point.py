"""Stress of the FP32 3-D spreading (and the FP32 2-D spreading at sigma >=
1.9), which accumulates each work item's padded tile in packed 32+32-bit
fixed point with a per-item scale 2^(30 - e), 2^e > sum_j |fHat_j|_inf *
max(w)^D.

Its rounding error grows with the nodes per work item (up to 1024) and with
the dynamic range of fHat, so the cases here are the hard ones:
* one tile crowded with 3000 nodes (three full work items) plus uniform
  background nodes, fHat magnitudes log-uniform over 1e-2 .. 1e4 with a few
  1e-30 entries;
* Gaussian clusters on a 128^3 grid that fill many tiles beyond 1024 nodes;
* data of magnitude 1e-30 (the scale exponent clamp; the result must scale
  linearly) and non-finite data (the output must not be finite garbage).
Bars: FP32 vs FP64 plan rel-l2 <= 1e-5 (the FP32 parity bar); FP64 vs
the oracle <= 1e-12.
"""
import numpy as np
import pytest
import torch

import nfftm
import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else b
    return O.relerr(a, b)


def wide_range(rng, M, lo, hi):
    mag = 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), M)
    return (mag * np.exp(2j * np.pi * rng.random(M))).astype(np.complex64)


def plans(k, N):
    kw = dict(m=4, sigma=2.0, precompute=nfftm.LINEAR)
    return (nfftm.plan_nfft(k, N, dtype=torch.float32, **kw), nfftm.plan_nfft(k, N, dtype=torch.float64, **kw))


def test_crowded_tile_wide_range():
    print("FP32 3-D fixed point: 3000 nodes in one tile, fHat over 1e-2..1e4", flush=True)
    rng = np.random.default_rng([91, 1])
    N = (32, 32, 32)  # grid 64^3; FP32 tile 8 (z, Python axis 0) x 16 x 16
    box = np.array([8 / 64, 16 / 64, 16 / 64])[:, None]
    crowd = rng.random((3, 3000)) * box * 0.999  # inside the tile at the origin
    k = np.concatenate([crowd, rng.random((3, 2000)) - 0.5], axis=1)
    fh = wide_range(rng, k.shape[1], 1e-2, 1e4)
    fh[rng.choice(k.shape[1], 10, replace=False)] = 1e-30
    p32, p64 = plans(k, N)
    x32, x64 = p32.H * fh, p64.H * fh
    e = rel(x32, x64)
    print(f"adjoint FP32 vs FP64 {e:.3e}", flush=True)
    assert e < 1e-5
    orc = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=O.LINEAR)
    assert rel(x64, orc.adjoint(fh)) < 1e-12
    # the spreading stage alone (no FFT rounding): grid of FP32 vs FP64
    g32 = torch.empty(p32.Ntilde, dtype=torch.complex64, device="cuda")
    g64 = torch.empty(p64.Ntilde, dtype=torch.complex128, device="cuda")
    p32.convolve_transpose_(torch.as_tensor(fh), g32)
    p64.convolve_transpose_(torch.as_tensor(fh), g64)
    eg = rel(g32, g64)
    print(f"convolve_transpose! FP32 vs FP64 {eg:.3e}", flush=True)
    assert eg < 1e-5


def test_clusters_beyond_item_size_128cubed():
    print("FP32 3-D fixed point: Gaussian clusters on 128^3 (many tiles > 1024 nodes)", flush=True)
    rng = np.random.default_rng([91, 2])
    N = (64, 64, 64)
    centers = rng.random((3, 40)) - 0.5
    pts = [c[:, None] + 0.01 * rng.standard_normal((3, 6000)) for c in centers.T]
    k = np.concatenate(pts + [rng.random((3, 100000)) - 0.5], axis=1)
    k = ((k + 0.5) % 1.0) - 0.5
    fh = wide_range(rng, k.shape[1], 1e-1, 1e3)
    p32, p64 = plans(k, N)
    e_adj = rel(p32.H * fh, p64.H * fh)
    f = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    e_fwd = rel(p32 * f, p64 * f)
    print(f"FP32 vs FP64: adjoint {e_adj:.3e} forward {e_fwd:.3e}", flush=True)
    assert e_adj < 1e-5 and e_fwd < 1e-5


def test_tiny_and_nonfinite_data():
    print("FP32 3-D fixed point: data of magnitude 1e-30, and Inf", flush=True)
    rng = np.random.default_rng([91, 3])
    N = (32, 32, 32)
    k = rng.random((3, 20000)) - 0.5
    fh = (rng.standard_normal(k.shape[1]) + 1j * rng.standard_normal(k.shape[1])).astype(np.complex64)
    p32, p64 = plans(k, N)
    x1 = (p32.H * fh).cpu().numpy().astype(np.complex128)
    tiny = (fh.astype(np.complex128) * 1e-30).astype(np.complex64)
    xt = (p32.H * tiny).cpu().numpy().astype(np.complex128)
    e = O.relerr(xt * 1e30, x1)
    print(f"adjoint(1e-30 fHat) * 1e30 vs adjoint(fHat): {e:.3e}", flush=True)
    assert e < 1e-5
    assert rel(p64.H * tiny, xt) < 1e-5
    bad = fh.copy()
    bad[123] = np.inf
    xb = (p32.H * bad).cpu().numpy()
    assert not np.all(np.isfinite(xb))  # no finite garbage
    assert np.all(np.isfinite((p32.H * fh).cpu().numpy()))  # the next call is clean again


def test_crowded_tile_wide_range_large_grid():
    """A crowded tile next to quiet ones and data over 1e-1 .. 1e3 on a
    128 x 128 x 256 grid (8192 work items): against the FP64 plan
    <= 1e-5, and the fold fused into the first FFT pass against the separate
    fold bit for bit (both sum the same fixed-point slots in the same order)."""
    print("FP32 3-D fixed point: 128x128x256 grid, crowded tile, wide range", flush=True)
    rng = np.random.default_rng([91, 4])
    N = (128, 64, 64)  # C order: z first
    box = np.array([8 / 256, 16 / 128, 16 / 128])[:, None]
    crowd = 0.2 + rng.random((3, 2500)) * box * 0.999
    k = np.concatenate([crowd, rng.random((3, 150000)) - 0.5], axis=1)
    fh = wide_range(rng, k.shape[1], 1e-1, 1e3)
    p32, p64 = plans(k, N)
    q32 = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32, fused_fold=False)
    x32, x64, xq = p32.H * fh, p64.H * fh, q32.H * fh
    e64 = rel(x32, x64)
    print(f"adjoint FP32 vs FP64 {e64:.3e}", flush=True)
    assert e64 < 1e-5 and torch.equal(x32, xq)
    # a second call on other data through the same plan
    fh2 = wide_range(rng, k.shape[1], 1e-2, 1e2)
    assert rel(p32.H * fh2, p64.H * fh2) < 1e-5


# FP32 2-D spreading (sigma >= 1.9) accumulates in the same packed fixed point
# with a per-item bound over at most 2048 nodes (twice the mean tile load).

def test_2d_crowded_tiles_wide_range():
    print("FP32 2-D fixed point: crowded tiles, fHat over 1e-2..1e4", flush=True)
    rng = np.random.default_rng([91, 5])
    N = (256, 256)
    crowd = 0.02 * rng.random((2, 20000))  # the tiles at the origin, many full items
    k = np.concatenate([crowd, rng.random((2, 40000)) - 0.5], axis=1)
    fh = wide_range(rng, k.shape[1], 1e-2, 1e4)
    fh[rng.choice(k.shape[1], 10, replace=False)] = 1e-30
    p32, p64 = plans(k, N)
    e = rel(p32.H * fh, p64.H * fh)
    g32 = torch.empty(p32.Ntilde, dtype=torch.complex64, device="cuda")
    g64 = torch.empty(p64.Ntilde, dtype=torch.complex128, device="cuda")
    p32.convolve_transpose_(torch.as_tensor(fh), g32)
    p64.convolve_transpose_(torch.as_tensor(fh), g64)
    eg = rel(g32, g64)
    print(f"adjoint FP32 vs FP64 {e:.3e}, convolve_transpose! {eg:.3e}", flush=True)
    assert e < 1e-5 and eg < 1e-5


def test_2d_tiny_and_nonfinite_data():
    print("FP32 2-D fixed point: 1e-30 data, and Inf", flush=True)
    rng = np.random.default_rng([91, 6])
    N = (128, 96)
    k = rng.random((2, 30000)) - 0.5
    fh = (rng.standard_normal(k.shape[1]) + 1j * rng.standard_normal(k.shape[1])).astype(np.complex64)
    p32, p64 = plans(k, N)
    x1 = (p32.H * fh).cpu().numpy().astype(np.complex128)
    tiny = (fh.astype(np.complex128) * 1e-30).astype(np.complex64)
    xt = (p32.H * tiny).cpu().numpy().astype(np.complex128)
    assert O.relerr(xt * 1e30, x1) < 1e-5
    assert rel(p64.H * fh, x1) < 1e-5
    bad = fh.copy()
    bad[77] = np.inf
    assert not np.all(np.isfinite((p32.H * bad).cpu().numpy()))
    assert np.all(np.isfinite((p32.H * fh).cpu().numpy()))
