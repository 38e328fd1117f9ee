"""CPU pins of the oracle's NFFTTools restatements (oracle.sdc,
oracle.toeplitz_kernel_explicit).

* sdc: the reference's known answer (NFFT.jl test/samplingDensity.jl:12-27): a
  9 x 8 grid of unit-spaced nodes gives weights 1/prod(N) for every
  precompute mode (there with Julia's default rtol for Float32, 3.45e-4; the
  FP64 oracle meets 1e-6), and the closed form of the final scale: after it,
  Re(sum v) = sum |v|^2 for v = A' diag(w) A 1.
* toeplitz_kernel_explicit: one node at the origin gives the constant E = 1,
  so lambda = prod(2N) at frequency 0 and zero elsewhere; on random nodes E
  equals the oracle's C NDFT adjoint of ones over the 2N grid, transformed
  (an independent code path for the same sum).
"""
import numpy as np
import pytest

import oracle as O


def unit_grid(N):
    fun = lambda n: np.arange(n) / n - 0.5
    return np.array([[x, y] for x in fun(N[0]) for y in fun(N[1])]).T


@pytest.mark.parametrize("pre", [O.LINEAR, O.FULL, O.TENSOR, O.POLYNOMIAL])
def test_sdc_unit_grid_known_answer(pre):
    N = (9, 8)
    p = O.OraclePlan(unit_grid(N), N, m=5, sigma=2.0, precompute=pre)
    w = O.sdc(p, iters=10)
    assert w.shape == (72,) and np.all(w > 0)
    np.testing.assert_allclose(w, 1 / 72, rtol=1e-6)


def test_sdc_scale_closed_form():
    rng = np.random.default_rng(3)
    k = rng.random((2, 400)) - 0.5
    p = O.OraclePlan(k, (16, 12), m=4, sigma=2.0, precompute=O.LINEAR)
    w = O.sdc(p, iters=5)
    v = p.adjoint(p.forward(np.ones((16, 12), complex)) * w)
    assert abs(v.sum().real - np.sum(np.abs(v) ** 2)) < 1e-10 * np.sum(np.abs(v) ** 2)
    # without the scale the weights differ from the scaled ones by one factor
    w0 = O.sdc(p, iters=5, scale=False)
    c = w / w0
    assert np.ptp(c) < 1e-12 * abs(c[0])


def test_toeplitz_explicit_single_node():
    for shape in [(8,), (6, 5), (4, 3, 2)]:
        lam = O.toeplitz_kernel_explicit(shape, np.zeros((len(shape), 1)))
        ref = np.zeros([2 * s for s in shape], complex)
        ref.flat[0] = np.prod([2 * s for s in shape])
        assert np.abs(lam - ref).max() < 1e-9


@pytest.mark.parametrize("shape", [(10,), (8, 6), (5, 4, 3)])
def test_toeplitz_explicit_vs_ndft_adjoint(shape):
    rng = np.random.default_rng(len(shape))
    tr = rng.random((len(shape), 200)) - 0.5
    L = tuple(2 * s for s in shape)
    E = O.ndft_adjoint(tr, L, np.ones(200))
    ref = np.fft.fftn(np.fft.fftshift(E))
    assert O.relerr(O.toeplitz_kernel_explicit(shape, tr), ref) < 1e-12
