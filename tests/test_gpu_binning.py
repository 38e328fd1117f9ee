"""Node binning on the device (nodes!, src/implementation.jl:108-141; the
tile binning of src/precomputation.jl:459-520).

The device result (nfftm_binning: sorted permutation and work items) must
equal, entry for entry, a numpy restatement: shift into [0, 1) as
shiftNodes! (src/utils.jl:32-44), tile of floor(k * Ntilde) per dimension,
a stable sort by tile (so a tile's nodes keep the caller's order), and work
items of at most CH nodes per tile (3-D: 1024; D <= 2: twice the mean
nodes per tile, a power of two in 256 .. 2048 / 4096).  Cases: 1-D, 2-D and 3-D in both
precisions, nodes on the range ends (-1/2, +1/2, -0.0, a negative node that
shifts to exactly 1), tiles crowded beyond one work item, empty tiles,
nodes! with a different M, and a plan on the node-parallel path.  The
binning is deterministic: two plans on the same nodes give bit-identical
FP32 3-D adjoints even where tiles hold several work items.
"""
import numpy as np
import pytest
import torch

import nfftm

pytestmark = pytest.mark.gpu

EPS = 2.220446049250313e-16


def item_size(D, M, Nt, T):
    """Most nodes of one work item: 3-D, 1024; D <= 2, twice the mean nodes
    per tile, a power of two in 256 .. 2048 (2-D) / 4096 (1-D)."""
    if D == 3:
        return 1024
    ntile = int(np.prod([-(-Nt[d] // T[d]) for d in range(D)]))
    mean, cap, ch = -(-M // ntile), {1: 4096, 2: 2048}[D], 256
    while ch < 2 * mean and ch < cap:
        ch *= 2
    return ch


def expected(k, Nt, T):
    """Python axis order in, Julia tile numbering out (dimension 1 fastest)."""
    D, M = k.shape
    kj, Ntj, Tj = k[::-1], Nt[::-1], T[::-1]
    CH = item_size(D, M, Ntj, Tj)
    key = np.zeros(M, dtype=np.int64)
    stride = 1
    for d in range(D):
        x = kj[d].astype(np.float64).copy()
        x[x < 0] += 1.0
        x[x == 1.0] -= EPS
        c = (x * Ntj[d]).astype(np.int64)
        c[c >= Ntj[d]] -= Ntj[d]
        key += (c // Tj[d]) * stride
        stride *= -(-Ntj[d] // Tj[d])
    perm = np.argsort(key, kind="stable").astype(np.int32)
    counts = np.bincount(key, minlength=stride)
    items, off = [], 0
    for t, c in enumerate(counts):
        for b in range(off, off + c, CH):
            items.append((t, b, min(off + c, b + CH), 0))
        off += c
    return perm, np.array(items, dtype=np.int32).reshape(-1, 4)


def edge_nodes(D, rng):
    e = np.array([-0.5, 0.5, -0.0, 0.0, -1e-17, 0.5 - EPS, -0.5 + EPS, 1e-300])
    return np.stack([rng.permutation(e) for _ in range(D)])


def check(p, k):
    perm, items = p.binning()
    T, _ = p.tile_shape()
    ep, ei = expected(k, p.Ntilde, T)
    assert np.array_equal(perm, ep)
    assert np.array_equal(items, ei)
    return items


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("case", ["1d", "2d", "3d"])
def test_device_binning_equals_restatement(dtype, case):
    print(f"binning {case} {dtype}", flush=True)
    rng = np.random.default_rng([17, len(case), dtype == torch.float64])
    N = {"1d": (4096,), "2d": (96, 128), "3d": (32, 48, 32)}[case]
    D = len(N)
    crowd = 0.02 * rng.random((D, 12000)) - 0.01  # one or two crowded tiles near the origin
    k = np.concatenate([rng.random((D, 30000)) - 0.5, crowd, edge_nodes(D, rng)], axis=1)
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=dtype)
    items = check(p, k)
    assert len(items) > 0 and (np.diff(items[:, 0]) >= 0).all()
    T, _ = p.tile_shape()
    ch = item_size(D, k.shape[1], p.Ntilde[::-1], T[::-1])
    assert (items[:, 2] - items[:, 1]).max() == ch  # crowded: a full item
    k2 = rng.random((D, 777)) - 0.5  # nodes! with another M
    p.nodes_(k2)
    check(p, k2)


def test_node_parallel_plan_is_identity():
    print("binning: node-parallel plan", flush=True)
    rng = np.random.default_rng(3)
    k = rng.random((2, 500)) - 0.5
    p = nfftm.plan_nfft(k, (8, 8), m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float64)
    perm, items = p.binning()
    assert len(items) == 0 and np.array_equal(perm, np.arange(500))


def test_binning_deterministic_fp32_3d():
    print("binning determinism: FP32 3-D adjoint on crowded tiles", flush=True)
    rng = np.random.default_rng(4)
    k = np.concatenate([rng.random((3, 50000)) - 0.5, 0.03 * rng.random((3, 20000))], axis=1)
    fh = (rng.standard_normal(k.shape[1]) + 1j * rng.standard_normal(k.shape[1])).astype(np.complex64)
    outs = []
    for _ in range(2):
        p = nfftm.plan_nfft(k, (64, 64, 64), m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32)
        outs.append((p.H * fh).cpu())
        del p
    assert torch.equal(outs[0], outs[1])
