"""Calls of one plan (or one batch) on alternating torch streams.

A plan owns its grid and scratch; a batch owns its grids, scratch and call
table.  When consecutive calls come on different streams, the library makes
the new stream wait for the work already enqueued on the previous one
(nfftm_set_stream / nfftm_batch_*), so results equal those of the same calls
on one stream: bit-identical, since the kernels are deterministic (FP32 3-D
spreading accumulates in integers; 2-D in FP64 with a fixed item order per
cell only up to LDS atomic order, hence the 1e-6 bar for the batch).
"""
import numpy as np
import pytest
import torch

import nfftm

pytestmark = pytest.mark.gpu


def crandn(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def test_plan_on_alternating_streams():
    print("plan calls alternating between two streams", flush=True)
    rng = np.random.default_rng([55, 1])
    N, M = (64, 64, 64), 1 << 19
    k = rng.random((3, M)) - 0.5
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32)
    fs = [torch.as_tensor(crandn(rng, N), device="cuda") for _ in range(6)]
    fhs = [torch.as_tensor(crandn(rng, M), device="cuda") for _ in range(6)]
    yr = [p.forward(f).clone() for f in fs]
    xr = [p.adjoint(fh).clone() for fh in fhs]
    ys = [torch.empty(M, dtype=torch.complex64, device="cuda") for _ in range(6)]
    xs = [torch.empty(N, dtype=torch.complex64, device="cuda") for _ in range(6)]
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for i in range(6):
        with torch.cuda.stream(s1 if i % 2 == 0 else s2):
            p.forward(fs[i], ys[i])
            p.adjoint(fhs[i], xs[i])
    torch.cuda.synchronize()
    for i in range(6):
        assert torch.equal(ys[i], yr[i]), i
        assert torch.equal(xs[i], xr[i]), i


def test_batch_on_alternating_streams():
    print("batch calls alternating between two streams", flush=True)
    rng = np.random.default_rng([55, 2])
    plans = [nfftm.plan_nfft(rng.random((2, M)) - 0.5, N, m=4, sigma=2.0, precompute=nfftm.LINEAR,
                             dtype=torch.float32)
             for N, M in [((256, 256), 200000), ((128, 128), 50000), ((512, 512), 300000)]]
    b = nfftm.BatchPlan(plans, [2, 1, 1], autotune=False)
    calls = []
    for _ in range(4):
        fs = [torch.as_tensor(crandn(rng, plans[i].N), device="cuda") for i, _ in b.slots]
        fh = [torch.as_tensor(crandn(rng, plans[i].J), device="cuda") for i, _ in b.slots]
        calls.append((fs, fh))
    ref = [([y.clone() for y in b.forward(fs)], [x.clone() for x in b.adjoint(fh)]) for fs, fh in calls]
    outs = [([torch.empty(plans[i].J, dtype=torch.complex64, device="cuda") for i, _ in b.slots],
             [torch.empty(plans[i].N, dtype=torch.complex64, device="cuda") for i, _ in b.slots])
            for _ in calls]
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for c, (fs, fh) in enumerate(calls):
        with torch.cuda.stream(s1 if c % 2 == 0 else s2):
            b.forward(fs, outs[c][0])
            b.adjoint(fh, outs[c][1])
    torch.cuda.synchronize()
    for c in range(len(calls)):
        for s in range(b.nslots):
            assert rel(outs[c][0][s], ref[c][0][s]) < 1e-6, (c, s)
            assert rel(outs[c][1][s], ref[c][1][s]) < 1e-6, (c, s)
