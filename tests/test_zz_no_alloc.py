"""nfftm_sdc_noalloc (the sdc entry point without device allocation) against
nfftm_sdc through the C ABI.

Two calls in a row on one GPU share the library's fixed workspace, the
second with fewer nodes, a smaller image and the other scale setting than the
first, each checked against nfftm_sdc on the same plan and weights: FP64
rel-l2 <= 1e-13 and FP32 <= 1e-6 (the same kernels in the same order; the
gridding sums in LDS differ only in the order of their atomic additions).  A
size beyond the workspace fails with NFFTM_ERR_ARG before any device work.
"""
import ctypes

import numpy as np
import pytest
import torch

import nfftm
from nfftm import _lib

pytestmark = pytest.mark.gpu


def _call(fn, p, iters, w, scale):
    L = _lib.load()
    L.nfftm_set_stream(p._handle, nfftm._stream())
    rc = getattr(L, fn)(p._handle, int(iters), ctypes.c_void_p(w.data_ptr()), int(scale))
    torch.cuda.synchronize()
    return rc, L.nfftm_last_error().decode()


def _rel(a, b):
    a = a.double().cpu().numpy()
    b = b.double().cpu().numpy()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _pair(k, N, dtype, iters, scale, w0=None):
    p = nfftm.plan_nfft(k, N, m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=dtype)
    if w0 is None:
        w0 = torch.ones(p.J, dtype=dtype, device="cuda")
    w_ref = w0.clone()
    w_new = w0.clone()
    rc, msg = _call("nfftm_sdc", p, iters, w_ref, scale)
    assert rc == 0, msg
    rc, msg = _call("nfftm_sdc_noalloc", p, iters, w_new, scale)
    assert rc == 0, msg
    return w_ref, w_new


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_sdc_noalloc_two_calls(dtype):
    tol = 1e-13 if dtype == torch.float64 else 1e-6
    rng = np.random.default_rng(23)
    # first call: 2-D, 6000 nodes, with the final scale
    k1 = rng.random((2, 6000)) - 0.5
    k1[1] *= 0.7
    print(f"sdc_noalloc {dtype} call 1: N=(48, 40) M=6000 scale=1", flush=True)
    a, b = _pair(k1, (48, 40), dtype, 6, True)
    assert torch.all(b > 0)
    e1 = _rel(b, a)
    # second call: fewer nodes, smaller image, no scale, non-unit start weights
    k2 = rng.random((2, 1500)) - 0.5
    w0 = torch.as_tensor(0.5 + rng.random(1500), dtype=dtype, device="cuda")
    print(f"sdc_noalloc {dtype} call 2: N=(20, 24) M=1500 scale=0", flush=True)
    c, d = _pair(k2, (20, 24), dtype, 4, False, w0)
    e2 = _rel(d, c)
    print(f"sdc_noalloc {dtype}: rel-l2 {e1:.2e}, {e2:.2e}", flush=True)
    assert e1 <= tol and e2 <= tol


def test_sdc_noalloc_3d_fixed_point():
    # the 3-D FP32 tile kernels (fixed-point LDS sums: order independent)
    rng = np.random.default_rng(29)
    k = rng.random((3, 4000)) - 0.5
    print("sdc_noalloc 3-D f32: N=(16, 16, 16) M=4000", flush=True)
    a, b = _pair(k, (16, 16, 16), torch.float32, 5, True)
    assert _rel(b, a) <= 1e-6


def test_sdc_noalloc_size_limit():
    # FP64: 2 M complex values of 16 bytes exceed the 64 MiB workspace
    rng = np.random.default_rng(31)
    M = 2_200_000
    k = rng.random((1, M)) - 0.5
    p = nfftm.plan_nfft(k, (256,), m=4, sigma=2.0, precompute=nfftm.LINEAR, dtype=torch.float64)
    w = torch.ones(M, dtype=torch.float64, device="cuda")
    print("sdc_noalloc size limit: 1-D M=2.2e6 f64", flush=True)
    rc, msg = _call("nfftm_sdc_noalloc", p, 2, w, False)
    assert rc == _lib.ERR_ARG and "workspace" in msg, (rc, msg)
    assert torch.all(w == 1)  # nothing ran
    # the allocating entry point still takes it
    rc, msg = _call("nfftm_sdc", p, 1, w, False)
    assert rc == 0, msg
