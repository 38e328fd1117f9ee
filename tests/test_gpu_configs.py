"""GPU parity at full workload sizes.

* 3-D headline workload: N = 128^3, M = 2^21 nodes uniform in
  [-1/2, 1/2)^3, m = 4, sigma = 2 (grid 256^3), LINEAR.  FP64 GPU plan vs
  the CPU oracle plan on the full output (<= 1e-12); FP32 vs FP64 plan
  (<= 1e-5); both vs the FP64 NDFT on 1024 sampled nodes (forward) and a
  16 x 8 x 8 box of 1024 image points (adjoint); the adjoint identity.
  NDFT bars: FP32 <= 1e-5; FP64 <= 5e-7 (the window's own error at m = 4,
  sigma = 2 is ~1e-7).
* Batch workload: 64 uneven 2-D plans (N in {64, 128, 256, 512}^2, M
  log-uniform in [1e4, 1e6], uniform nodes) plus 32 images sharing a 256^2
  golden-angle radial trajectory (402 spokes x 512 samples), m = 4,
  sigma = 2, LINEAR, FP32, one autotuned batch.  Every slot against its
  single-plan transform (<= 1e-6) and every plan against the FP64 NDFT on
  sampled outputs (<= 1e-5).
* 1-D contention workload: N = 2^24, M = 2^24 nodes Beta(1/2, 1/2) - 1/2
  (clustered at the edges of the node range, so a few tiles hold most
  nodes), m = 6, sigma = 1.25 (Ntilde = 20,971,520 = 2^22 * 5), TENSOR.
  FP32 vs FP64 GPU plan (full rel-l2), both vs the FP64 NDFT on sampled
  outputs, FP64 GPU vs the CPU oracle plan (full), and the adjoint identity
  <A f, y> = <f, A^H y>.  Bars: FP64 vs oracle <= 1e-12; FP32 vs NDFT
  <= 1e-5; FP64 vs NDFT <= 1e-7 (the window accuracy at m = 6,
  sigma = 1.25 is ~4e-8).
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


def test_1d_edge_clustered_2p24():
    print("1-D N=2^24 M=2^24 Beta(1/2,1/2) m=6 sigma=1.25 TENSOR", flush=True)
    rng = np.random.default_rng([123, 5])
    n = 1 << 24
    k = (rng.beta(0.5, 0.5, size=n) - 0.5).reshape(1, n)
    f = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    fh = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    kw = dict(m=6, sigma=1.25, precompute=nfftm.TENSOR)
    p32 = nfftm.plan_nfft(k, (n,), dtype=torch.float32, **kw)
    assert p32.Ntilde == (20971520,)
    y32, x32 = p32 * f, p32.H * fh
    p64 = nfftm.plan_nfft(k, (n,), dtype=torch.float64, **kw)
    y64, x64 = p64 * f, p64.H * fh
    e_fwd, e_adj = rel(y32, y64), rel(x32, x64)
    print(f"FP32 vs FP64: forward {e_fwd:.3e} adjoint {e_adj:.3e}", flush=True)
    assert e_fwd < 1e-5 and e_adj < 1e-5
    sel = rng.choice(n, 48, replace=False)
    ndf = O.ndft(k, f.astype(np.complex128), sel)
    nda = O.ndft_adjoint(k, (n,), fh.astype(np.complex128), sel)
    y64n, x64n = y64.cpu().numpy(), x64.cpu().numpy()
    print(f"FP64 vs NDFT: forward {rel(y64n[sel], ndf):.3e} adjoint {rel(x64n[sel], nda):.3e}",
          flush=True)
    assert rel(y64n[sel], ndf) < 1e-7 and rel(x64n[sel], nda) < 1e-7
    assert rel(y32.cpu().numpy()[sel], ndf) < 1e-5 and rel(x32.cpu().numpy()[sel], nda) < 1e-5
    # adjoint identity on the FP64 plan
    lhs = np.vdot(fh.astype(np.complex128), y64n)
    rhs = np.vdot(x64n, f.astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-11
    del p32, p64, y32, x32
    torch.cuda.empty_cache()
    orc = O.OraclePlan(k, (n,), m=6, sigma=1.25, precompute=O.TENSOR, workers=O.lib().orc_num_threads())
    assert rel(y64n, orc.forward(f)) < 1e-12
    assert rel(x64n, orc.adjoint(fh)) < 1e-12


def radial_nodes(spokes, samples):
    """Golden-angle radial trajectory: spoke s at angle s * pi / phi, radius
    -1/2 + i / samples."""
    phi = (1 + 5 ** 0.5) / 2
    th = np.arange(spokes) * np.pi / phi
    r = -0.5 + np.arange(samples) / samples
    return np.stack([np.outer(np.cos(th), r).ravel(), np.outer(np.sin(th), r).ravel()])


def test_3d_128cubed_M2p21():
    print("3-D N=128^3 M=2^21 uniform m=4 sigma=2 LINEAR", flush=True)
    rng = np.random.default_rng([123, 2])
    N, M = (128, 128, 128), 1 << 21
    k = rng.random((3, M)) - 0.5
    f = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    fh = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    kw = dict(m=4, sigma=2.0, precompute=nfftm.LINEAR)
    p64 = nfftm.plan_nfft(k, N, dtype=torch.float64, **kw)
    assert p64.Ntilde == (256, 256, 256)
    y64 = (p64 * f).cpu().numpy()
    x64 = (p64.H * fh).cpu().numpy()
    del p64
    torch.cuda.empty_cache()
    p32 = nfftm.plan_nfft(k, N, dtype=torch.float32, **kw)
    y32 = (p32 * f).cpu().numpy()
    x32 = (p32.H * fh).cpu().numpy()
    del p32
    torch.cuda.empty_cache()
    e_fwd, e_adj = rel(y32, y64), rel(x32, x64)
    print(f"FP32 vs FP64 plan: forward {e_fwd:.3e} adjoint {e_adj:.3e}", flush=True)
    assert e_fwd < 1e-5 and e_adj < 1e-5
    orc = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=O.LINEAR, workers=O.lib().orc_num_threads())
    e_of, e_oa = rel(y64, orc.forward(f)), rel(x64, orc.adjoint(fh))
    print(f"FP64 vs oracle plan: forward {e_of:.3e} adjoint {e_oa:.3e}", flush=True)
    assert e_of < 1e-12 and e_oa < 1e-12
    del orc
    sel = np.sort(rng.choice(M, 1024, replace=False))
    ndf = O.ndft(k, f.astype(np.complex128), sel)
    box = [np.sort(rng.choice(n, c, replace=False)) for n, c in zip(N, (16, 8, 8))]
    nda = O.ndft_adjoint_box(k, N, fh.astype(np.complex128), box)
    errs = (rel(y64[sel], ndf), rel(x64[np.ix_(*box)], nda), rel(y32[sel], ndf), rel(x32[np.ix_(*box)], nda))
    print("vs NDFT (1024 nodes / 1024 image points): FP64 fwd %.3e adj %.3e, FP32 fwd %.3e adj %.3e" % errs,
          flush=True)
    assert errs[0] < 5e-7 and errs[1] < 5e-7
    assert errs[2] < 1e-5 and errs[3] < 1e-5
    # adjoint identity <A f, fh> = <f, A^H fh> (FP64 plan)
    lhs = np.vdot(fh.astype(np.complex128), y64)
    rhs = np.vdot(x64, f.astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-11


def test_batch_64_uneven_plans_32_radial_coils():
    print("batch: 64 uneven 2-D plans + 32 radial coils, FP32, autotuned", flush=True)
    rng = np.random.default_rng([123, 3])
    specs = []
    for _ in range(64):
        n = int(rng.choice([64, 128, 256, 512]))
        M = int(round(10 ** rng.uniform(4, 6)))
        specs.append(((n, n), rng.random((2, M)) - 0.5, 1))
    specs.append(((256, 256), radial_nodes(402, 512), 32))
    plans = [nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float32)
             for N, k, _ in specs]
    b = nfftm.BatchPlan(plans, [n for _, _, n in specs], autotune=True)
    assert b.nslots == 96
    fs = [(rng.standard_normal(plans[i].N) + 1j * rng.standard_normal(plans[i].N)).astype(np.complex64)
          for i, _ in b.slots]
    fh = [(rng.standard_normal(plans[i].J) + 1j * rng.standard_normal(plans[i].J)).astype(np.complex64)
          for i, _ in b.slots]
    fs_d = [torch.as_tensor(x, device="cuda") for x in fs]
    fh_d = [torch.as_tensor(x, device="cuda") for x in fh]
    for call in range(2):  # the second call replays the captured graphs
        yb = [y.cpu().numpy() for y in b.forward(fs_d)]
        xb = [x.cpu().numpy() for x in b.adjoint(fh_d)]
        worst = 0.0
        for s, (i, _) in enumerate(b.slots):
            e = max(rel(yb[s], plans[i] * fs_d[s]), rel(xb[s], plans[i].H * fh_d[s]))
            worst = max(worst, e)
            assert e < 1e-6, (call, s)
        print(f"call {call}: batch vs single plans, worst slot {worst:.3e}", flush=True)
    worst = [0.0, 0.0]
    for s, (i, c) in enumerate(b.slots):
        if c not in (0, 31):  # every plan once, two of the radial images
            continue
        N, k, _ = specs[i]
        sel = np.sort(rng.choice(k.shape[1], min(64, k.shape[1]), replace=False))
        ref = O.ndft(k, fs[s].astype(np.complex128), sel)
        box = [np.sort(rng.choice(n, 4, replace=False)) for n in N]
        refa = O.ndft_adjoint_box(k, N, fh[s].astype(np.complex128), box)
        e_f, e_a = rel(yb[s][sel], ref), rel(xb[s][np.ix_(*box)], refa)
        worst = [max(worst[0], e_f), max(worst[1], e_a)]
        assert e_f < 1e-5 and e_a < 1e-5, (s, e_f, e_a)
    print(f"FP32 batch vs FP64 NDFT: forward {worst[0]:.3e} adjoint {worst[1]:.3e}", flush=True)
