"""All windows of src/windowFunctions.jl on the GPU plan.

Per window and precompute mode: FP64 against the oracle (rel-l2 <= 1e-12,
2-D and 3-D), FP64 against the NDFT within the reference's per-window
tolerances (NFFT.jl test/accuracy.jl:40-72, m = 5, sigma = 2), FP32 against
the NDFT within max(that tolerance, 1e-5).  The node gradient of a
non-Kaiser-Bessel window needs a fitted-polynomial mode and raises
otherwise.
"""
import numpy as np
import pytest
import torch

import nfftm
import oracle as O

pytestmark = pytest.mark.gpu

WINDOWS = ["kaiser_bessel", "gauss", "spline", "kaiser_bessel_rev", "cosh_type"]
EPS = {"kaiser_bessel": 1e-7, "cosh_type": 1e-7, "gauss": 1e-3, "kaiser_bessel_rev": 1e-6, "spline": 1e-4}
PRE = [(nfftm.LINEAR, O.LINEAR), (nfftm.FULL, O.FULL), (nfftm.TENSOR, O.TENSOR),
       (nfftm.POLYNOMIAL, O.POLYNOMIAL)]


def rel(a, b):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return O.relerr(a, b)


@pytest.mark.parametrize("pre", PRE, ids=["linear", "full", "tensor", "poly"])
@pytest.mark.parametrize("window", WINDOWS)
def test_window_parity(window, pre):
    print(f"window {window} pre {pre[0]}", flush=True)
    rng = np.random.default_rng([WINDOWS.index(window), pre[0]])
    for N in [(31, 33), (11, 12, 14)]:
        J = int(np.prod(N))
        k = rng.random((len(N), J)) - 0.5
        f = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        fh = rng.standard_normal(J) + 1j * rng.standard_normal(J)
        orc = O.OraclePlan(k, N, m=5, sigma=2.0, precompute=pre[1], window=window)
        p64 = nfftm.plan_nfft(k, N, m=5, sigma=2.0, precompute=pre[0], window=window, dtype=torch.float64)
        p32 = nfftm.plan_nfft(k, N, m=5, sigma=2.0, precompute=pre[0], window=window, dtype=torch.float32)
        y64, x64 = p64 * f, p64.H * fh
        y32, x32 = p32 * f, p32.H * fh
        torch.cuda.synchronize()
        assert rel(y64, orc.forward(f)) < 1e-12, N
        assert rel(x64, orc.adjoint(fh)) < 1e-12, N
        yn, xn = O.ndft(k, f), O.ndft_adjoint(k, N, fh)
        assert rel(y64, yn) < EPS[window], N
        assert rel(x64, xn) < EPS[window], N
        assert rel(y32, yn) < max(EPS[window], 1e-5), N
        assert rel(x32, xn) < max(EPS[window], 1e-5), N


@pytest.mark.parametrize("pre", [(nfftm.POLYNOMIAL, O.POLYNOMIAL), (nfftm.TENSOR, O.TENSOR)],
                         ids=["poly", "tensor"])
@pytest.mark.parametrize("window", [w for w in WINDOWS if w != "kaiser_bessel"])
def test_window_gradient_vs_oracle(window, pre):
    """Node gradient of a non-Kaiser-Bessel window (the derivative of the
    fitted polynomial) against the oracle's gradient of the same operator:
    FP64 rel-l2 <= 1e-12, FP32 <= 1e-5."""
    print(f"window gradient {window} pre {pre[0]}", flush=True)
    rng = np.random.default_rng([5, WINDOWS.index(window), pre[0]])
    N = (16, 20)
    k = rng.random((2, 300)) - 0.5
    f = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    y = rng.standard_normal(300) + 1j * rng.standard_normal(300)
    ref = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=pre[1], window=window).grad_nodes(f, y)
    for dtype, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        q = nfftm.plan_nfft(k, N, m=4, sigma=2.0, window=window, precompute=pre[0], dtype=dtype)
        g = q.grad_nodes(f, y)
        assert rel(g, ref) < tol, (dtype, rel(g, ref))


def test_window_gradient_modes():
    rng = np.random.default_rng(5)
    k = rng.random((2, 300)) - 0.5
    f = torch.as_tensor(rng.standard_normal((16, 16)) + 0j, dtype=torch.complex64, device="cuda")
    y = torch.as_tensor(rng.standard_normal(300) + 0j, dtype=torch.complex64, device="cuda")
    p = nfftm.plan_nfft(k, (16, 16), m=4, sigma=2.0, window="gauss", precompute=nfftm.LINEAR,
                        dtype=torch.float32)
    with pytest.raises(nfftm.NFFTError):
        p.grad_nodes(f, y)
    with pytest.raises(nfftm.NFFTError):
        nfftm.plan_nfft(k, (16, 16), window="hann")
