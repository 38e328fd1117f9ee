"""GPU batches (nfftm_batch_* through the C ABI) and the autotuner.

A batch must compute exactly what its plans compute one by one: every slot
is compared with the single-plan transform (FP64: rel-l2 <= 1e-13, FP32:
<= 1e-6; the batched rocFFT and the order of LDS additions may differ in the
last bits) and, in FP64, with the CPU oracle (<= 1e-12).  Graph replays are
checked call after call against direct launches, with other device work in
between.
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


def radial_nodes(spokes, samples):
    """Golden-angle radial trajectory: spoke s at angle s * pi / phi, radius
    -1/2 + i / samples."""
    phi = (1 + 5 ** 0.5) / 2
    th = np.arange(spokes) * np.pi / phi
    r = -0.5 + np.arange(samples) / samples
    return np.stack([np.outer(np.cos(th), r).ravel(), np.outer(np.sin(th), r).ravel()])


def mixed_plans(rng, dtype, n2d=6):
    plans, ks = [], []
    for q in range(n2d):
        N = [(16, 20), (32, 32), (24, 40), (64, 48)][q % 4]
        M = int(rng.integers(200, 3000))
        k = rng.random((2, M)) - 0.5
        plans.append(nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=dtype))
        ks.append((k, N))
    k = rng.random((1, 900)) - 0.5  # a 1-D plan in the same batch
    plans.append(nfftm.plan_nfft(k, (300,), m=4, sigma=2.0, precompute=nfftm.POLYNOMIAL, dtype=dtype))
    ks.append((k, (300,)))
    k = radial_nodes(24, 64)  # shared nodes, three images
    plans.append(nfftm.plan_nfft(k, (32, 32), m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=dtype))
    ks.append((k, (32, 32)))
    ntrans = [1] * (len(plans) - 1) + [3]
    return plans, ks, ntrans


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("graph", [True, False], ids=["graph", "direct"])
def test_batch_matches_single_plans(dtype, graph):
    print(f"batch dtype={dtype} graph={graph}", flush=True)
    rng = np.random.default_rng([31, int(graph), dtype == torch.float64])
    plans, ks, ntrans = mixed_plans(rng, dtype)
    b = nfftm.BatchPlan(plans, ntrans, autotune=False, graph=graph)
    assert b.nslots == sum(ntrans)
    if graph:
        assert b.graph
    cd = torch.complex128 if dtype == torch.float64 else torch.complex64
    tol = 1e-13 if dtype == torch.float64 else 1e-6
    for call in range(3):  # replays with fresh data, other device work in between
        fs = [torch.as_tensor(crandn(rng, plans[i].N), dtype=cd, device="cuda") for i, _ in b.slots]
        fh = [torch.as_tensor(crandn(rng, plans[i].J), dtype=cd, device="cuda") for i, _ in b.slots]
        yb = b.forward(fs)
        xb = b.adjoint(fh)
        _ = torch.randn(1 << 20, device="cuda").sum()  # unrelated eager work
        torch.cuda.synchronize()
        for s, (i, c) in enumerate(b.slots):
            assert rel(yb[s], plans[i] * fs[s]) < tol, (call, s)
            assert rel(xb[s], plans[i].H * fh[s]) < tol, (call, s)
        if dtype == torch.float64 and call == 0:
            for s, (i, c) in enumerate(b.slots):
                k, N = ks[i]
                orc = O.OraclePlan(k, N, m=4, sigma=2.0,
                                   precompute=O.POLYNOMIAL if len(N) == 1 else O.LINEAR)
                assert rel(yb[s], orc.forward(fs[s].cpu().numpy())) < 1e-12
                assert rel(xb[s], orc.adjoint(fh[s].cpu().numpy())) < 1e-12


def test_autotune_keeps_results():
    print("autotune", flush=True)
    rng = np.random.default_rng(37)
    for N, M in [((128, 96), 20000), ((2048,), 30000)]:
        k = rng.random((len(N), M)) - 0.5
        f = crandn(rng, N)
        fh = crandn(rng, M)
        p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float64)
        y0, x0 = (p * f).cpu().numpy(), (p.H * fh).cpu().numpy()
        shape = p.tune()
        print(N, "tuned", shape, flush=True)
        assert shape[1] in (256, 512, 1024)
        assert rel(p * f, y0) < 1e-13
        assert rel(p.H * fh, x0) < 1e-13
    p3 = nfftm.plan_nfft(rng.random((3, 5000)) - 0.5, (32, 32, 32), m=4, sigma=2.0,
                         dtype=torch.float32)
    f3 = crandn(rng, (32, 32, 32))
    y0 = (p3 * f3).cpu().numpy()
    assert p3.tune()[1] in (512, 1024)
    assert rel(p3 * f3, y0) < 1e-6


def test_batch_invalidated_by_nodes_and_errors():
    print("batch errors", flush=True)
    rng = np.random.default_rng(41)
    p1 = nfftm.plan_nfft(rng.random((2, 500)) - 0.5, (32, 32), dtype=torch.float64)
    p2 = nfftm.plan_nfft(rng.random((2, 700)) - 0.5, (16, 24), dtype=torch.float64)
    b = nfftm.BatchPlan([p1, p2], autotune=False)
    f = [torch.zeros((32, 32), dtype=torch.complex128, device="cuda"),
         torch.zeros((16, 24), dtype=torch.complex128, device="cuda")]
    b.forward(f)
    p2.nodes_(rng.random((2, 700)) - 0.5)
    with pytest.raises(nfftm.ArgumentError):
        b.forward(f)
    with pytest.raises(nfftm.DimensionMismatch):
        nfftm.BatchPlan([p1], autotune=False).forward(f)
    p3 = nfftm.plan_nfft(rng.random((3, 100)) - 0.5, (16, 16, 16), dtype=torch.float64)
    with pytest.raises(nfftm.NFFTError):
        nfftm.BatchPlan([p1, p3], autotune=False)
    p4 = nfftm.plan_nfft(rng.random((2, 100)) - 0.5, (16, 16), dtype=torch.float32)
    with pytest.raises(nfftm.ArgumentError):
        nfftm.BatchPlan([p1, p4], autotune=False)


def test_batch_radial_coils_autotuned():
    """32 images on one golden-angle radial trajectory (402 spokes x 512
    samples, 256^2 image) plus a few uneven plans, autotuned, FP32."""
    print("radial coils", flush=True)
    rng = np.random.default_rng([123, 3])
    kr = radial_nodes(402, 512)
    pr = nfftm.plan_nfft(kr, (256, 256), m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32)
    others = []
    for N in [(64, 64), (128, 128), (256, 256), (512, 512)]:
        M = int(10 ** rng.uniform(4, 5))
        others.append(nfftm.plan_nfft(rng.random((2, M)) - 0.5, N, m=4, sigma=2.0,
                                      precompute=nfftm.LINEAR, dtype=torch.float32))
    plans = [pr] + others
    b = nfftm.BatchPlan(plans, [32, 1, 1, 1, 1], autotune=True)
    fs = [torch.as_tensor(crandn(rng, plans[i].N), dtype=torch.complex64, device="cuda") for i, _ in b.slots]
    fh = [torch.as_tensor(crandn(rng, plans[i].J), dtype=torch.complex64, device="cuda") for i, _ in b.slots]
    yb, xb = b.forward(fs), b.adjoint(fh)
    torch.cuda.synchronize()
    for s in (0, 7, 31, 32, 33, 34, 35):
        i, _ = b.slots[s]
        assert rel(yb[s], plans[i] * fs[s]) < 1e-6
        assert rel(xb[s], plans[i].H * fh[s]) < 1e-6
    # one coil against the FP64 direct NDFT on sampled outputs
    sel = rng.choice(kr.shape[1], 256, replace=False)
    ref = O.ndft(kr, fs[5].cpu().numpy().astype(np.complex128), sel)
    assert rel(yb[5].cpu().numpy()[sel], ref) < 1e-5
