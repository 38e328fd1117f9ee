"""GPU parity at full workload sizes that the oracle cannot brute-force.

* 1-D contention workload: N = 2^24, M = 2^24 nodes Beta(1/2, 1/2) - 1/2
  (clustered at the edges of the node range, so a few tiles hold most
  nodes), m = 6, sigma = 1.25 (Ntilde = 20,971,520 = 2^22 * 5), TENSOR.
  FP32 vs FP64 GPU plan (full rel-l2), both vs the FP64 NDFT on sampled
  outputs, FP64 GPU vs the CPU oracle plan (full), and the adjoint identity
  <A f, y> = <f, A^H y>.
Bars written here: FP64 vs oracle <= 1e-12; FP32 vs NDFT <= 1e-5; FP64 vs
NDFT <= 1e-7 (the window accuracy at m = 6, sigma = 1.25 is ~4e-8).
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
