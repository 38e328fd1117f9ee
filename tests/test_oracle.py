"""Pins the CPU oracle (oracle/, an FP64 restatement of NFFT.jl) before it is
trusted as the checker of the GPU path.  CPU only.

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong sign or index, transposed operand) fails it:
* window / transform / NDFT against closed forms evaluated with mpmath
  (tests/golden/golden.json, written by tests/golden/make_golden.py);
* the NDFT against numpy's FFT on equispaced nodes (where the NDFT is a DFT);
* the full oracle NFFT against the NDFT within the reference's own accuracy
  contract (test/accuracy.jl:41-73: Kaiser-Bessel, m = 5, sigma = 2,
  rel-l2 < 1e-7) for every precompute mode and odd / 4-D sizes;
* adjointness of convolve!/convolve_transpose!, blocked == non-blocked,
  TENSOR == POLYNOMIAL, the LUT and polynomial tables against independent
  evaluations, the Chebyshev deconvolution against the exact 1/psi_hat.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_window_vs_mpmath():
    for g in GOLD["window"]:
        v = O.window(g["k"], g["Nt"], g["m"], g["sigma"])
        assert v == pytest.approx(g["value"], rel=1e-13, abs=1e-300), g


def test_window_hat_vs_mpmath():
    for g in GOLD["window_hat"]:
        v = O.window_hat(g["n"], g["Nt"], g["m"], g["sigma"])
        assert v == pytest.approx(g["value"], rel=1e-13), g


def test_ndft_vs_mpmath():
    for g in GOLD["ndft"]:
        N = tuple(g["N"])
        k = np.array(g["k"])
        f = np.array([complex(*z) for z in g["f"]]).reshape(N)
        fh = np.array([complex(*z) for z in g["fhat"]])
        fw = np.array([complex(*z) for z in g["forward"]])
        ad = np.array([complex(*z) for z in g["adjoint"]]).reshape(N)
        assert O.relerr(O.ndft(k, f), fw) < 1e-14
        assert O.relerr(O.ndft_adjoint(k, N, fh), ad) < 1e-14


@pytest.mark.parametrize("N", [(9,), (6, 7), (5, 4, 6)])
def test_ndft_adjoint_box_equals_pointwise(N):
    """The factorised product-set adjoint NDFT equals the pointwise adjoint
    NDFT (itself pinned by the mpmath fixtures above) on every point of the
    set, including repeated and unsorted indices."""
    rng = np.random.default_rng(7)
    D = len(N)
    M = 300
    k = rng.random((D, M)) - 0.5
    fh = crandn(rng, M)
    idx = [rng.choice(n, size=min(n, 3 + d), replace=False) for d, n in enumerate(N)]
    idx[0] = np.concatenate([idx[0], idx[0][:1]])
    got = O.ndft_adjoint_box(k, N, fh, idx)
    assert got.shape == tuple(len(i) for i in idx)
    mesh = np.meshgrid(*idx, indexing="ij")
    flat = np.ravel_multi_index([g.ravel() for g in mesh], N)
    ref = O.ndft_adjoint(k, N, fh, flat).reshape(got.shape)
    assert O.relerr(got, ref) < 1e-13
    full = O.ndft_adjoint(k, N, fh)
    assert O.relerr(got, full[np.ix_(*idx)]) < 1e-13


@pytest.mark.parametrize("N", [(16,), (7,), (6, 5), (4, 3, 5)])
def test_ndft_equispaced_is_shifted_dft(N):
    """On the grid k = i/N - 1/2 (per axis) the NDFT is an FFT with shifts."""
    rng = np.random.default_rng(1)
    D = len(N)
    grids = np.meshgrid(*[np.arange(n) / n - 0.5 for n in N], indexing="ij")
    k = np.stack([g.ravel() for g in grids])
    f = crandn(rng, N)
    got = O.ndft(k, f).reshape(N)
    # sum_n f_n exp(-2 pi i (n - N//2)(j/N - 1/2)), n,j = 0..N-1 per axis
    ref = f.copy()
    for d in range(D):
        n = N[d]
        idx = np.arange(n)
        E = np.exp(-2j * np.pi * np.outer(idx / n - 0.5, idx - n // 2))
        ref = np.moveaxis(np.tensordot(E, np.moveaxis(ref, d, 0), axes=(1, 0)), 0, d)
    assert O.relerr(got, ref) < 1e-13
    # and the same through numpy's FFT: with h = N//2,
    # F_j = exp(2 pi i h (j/N - 1/2)) * FFT(f_n (-1)^n)_j
    shifted = f * np.prod(np.meshgrid(*[np.exp(1j * np.pi * np.arange(n)) for n in N],
                                      indexing="ij"), axis=0)
    phase = np.prod(np.meshgrid(*[np.exp(2j * np.pi * (n // 2) * (np.arange(n) / n - 0.5))
                                  for n in N], indexing="ij"), axis=0)
    assert O.relerr(got, np.fft.fftn(shifted) * phase) < 1e-12


def test_lut_against_window():
    """precomputeLinInterp: LUT[i] = psi(i m / K) on a unit grid."""
    for m in (2, 4, 5):
        K = O.lut_size(m)
        lut = O.lut(m, 2.0, K)
        assert len(lut) == K + 2
        for i in [0, 1, K // 3, K // 2, K - 1, K, K + 1]:
            ref = O.window(i * m / K, 1.0, m, 2.0)
            assert lut[i] == pytest.approx(ref, rel=1e-15, abs=1e-300)
    for g in GOLD["window"]:
        if g["Nt"] == 1 and g["m"] == 4 and g["sigma"] == 2.0 and 0 <= g["k"] <= 4:
            K = O.lut_size(4)
            x = g["k"] * K / 4
            if abs(x - round(x)) < 1e-12:
                assert O.lut(4, 2.0, K)[int(round(x))] == pytest.approx(g["value"], rel=1e-13)


@pytest.mark.parametrize("m", [2, 4, 5, 6])
def test_polynomial_fit_independent(m):
    """precomputePolyInterp against an independent least-squares fit (numpy
    SVD lstsq) on the same samples, compared in value space."""
    sigma = 2.0
    P = O.poly(m, sigma)
    deg, W = 2 * m + 1, 2 * m
    assert P.shape == (deg, W)
    ns = 2 * deg
    xs = np.linspace(-0.5, 0.5, ns)
    V = np.vander(xs, deg, increasing=True)
    xt = np.linspace(-0.5, 0.5, 301)
    Vt = np.vander(xt, deg, increasing=True)
    scale = O.window(0.0, 1.0, m, sigma)
    for l in range(W):
        y = np.array([O.window(m - (l + 0.5) + x, 1.0, m, sigma) for x in xs])
        c, *_ = np.linalg.lstsq(V, y, rcond=None)
        assert np.max(np.abs(Vt @ P[:, l] - Vt @ c)) < 1e-9 * scale
        truth = np.array([O.window(m - (l + 0.5) + x, 1.0, m, sigma) for x in xt])
        assert np.max(np.abs(Vt @ P[:, l] - truth)) < 1e-3 * scale


@pytest.mark.parametrize("N,Nt,m,sigma,tol", [(64, 128, 4, 2.0, 5e-13), (31, 62, 5, 2.0, 5e-13),
                                              (255, 510, 5, 2.0, 5e-13), (5, 10, 5, 2.0, 5e-13),
                                              (128, 256, 8, 2.0, 5e-13),
                                              (2 ** 24, 20971520, 6, 1.25, 5e-11)])
def test_deconv_chebyshev_vs_exact(N, Nt, m, sigma, tol):
    """The reference's deconvolution factors are 1/cheb(j), cheb a 30-node
    Chebyshev interpolant of psi_hat (src/precomputation.jl:347-358).  It
    equals 1/psi_hat to ~1e-13 at sigma = 2; at m = 6, sigma = 1.25 (the
    steeper transform of the 1-D workload) it is off by up to ~2.3e-11, so
    exact factors would NOT be parity-equivalent there and the library builds
    the same interpolant."""
    sl = slice(None, None, max(1, N // 4096))
    cheb = O.deconv_factors(N, Nt, m, sigma, cheb=True)[sl]
    ex = O.deconv_factors(N, Nt, m, sigma, cheb=False)[sl]
    assert np.max(np.abs(cheb / ex - 1)) < tol
    for g in GOLD["window_hat"]:
        if g["Nt"] == Nt and g["m"] == m and g["sigma"] == sigma and N <= 100000:
            i = g["n"] + N // 2
            if 0 <= i < N:
                assert 1.0 / ex[i] == pytest.approx(g["value"], rel=1e-13)


def test_chebyshev_interpolant_short_range():
    """Over a range of 2..4 integers the 30-node interpolant of the smooth
    psi_hat is exact to rounding."""
    for N in (2, 3, 4):
        cheb = O.deconv_factors(N, 2 * N, 2, 2.0, cheb=True)
        ex = O.deconv_factors(N, 2 * N, 2, 2.0, cheb=False)
        assert np.max(np.abs(cheb / ex - 1)) < 1e-14


ACC_N = [(255,), (31, 33), (11, 12, 14), (6, 5, 6, 6)]
ACC_MODES = [(O.LINEAR, True), (O.LINEAR, False), (O.FULL, False), (O.TENSOR, True),
             (O.POLYNOMIAL, True), (O.POLYNOMIAL, False)]


@pytest.mark.parametrize("N", ACC_N, ids=lambda n: "x".join(map(str, n)))
@pytest.mark.parametrize("pre,blocking", ACC_MODES, ids=str)
def test_oracle_accuracy_contract(N, pre, blocking):
    """test/accuracy.jl:109-139 restated on the oracle: rel-l2 < 1e-7."""
    rng = np.random.default_rng([2, len(N), pre, int(blocking)])
    D, J = len(N), int(np.prod(N))
    k = rng.random((D, J)) - 0.5
    fhat = rng.random(J) + 1j * rng.random(J)
    f = O.ndft_adjoint(k, N, fhat)
    g = O.ndft(k, f)
    p = O.OraclePlan(k, N, 5, 2.0, pre, blocking)
    assert O.relerr(p.adjoint(fhat), f) < 1e-7
    assert O.relerr(p.forward(f), g) < 1e-7


def test_oracle_detects_planted_sign_error():
    """A sign error in the exponent (nodes negated) must fail the same check."""
    rng = np.random.default_rng(3)
    N, J = (20, 18), 360
    k = rng.random((2, J)) - 0.5
    f = crandn(rng, N)
    p_wrong = O.OraclePlan(-k * (np.abs(k) < 0.5) + 0.0, N, 5, 2.0, O.LINEAR, True)
    assert O.relerr(p_wrong.forward(f), O.ndft(k, f)) > 0.5
    p_swap = O.OraclePlan(k[::-1], N, 5, 2.0, O.LINEAR, True)  # transposed axes
    assert O.relerr(p_swap.forward(f), O.ndft(k, f)) > 0.5


@pytest.mark.parametrize("N,M,m", [((64, 64), 4096, 4), ((20, 18, 16), 3000, 4), ((300,), 700, 6)])
def test_oracle_adjointness_and_blocking(N, M, m):
    rng = np.random.default_rng([4, M])
    D = len(N)
    k = rng.random((D, M)) - 0.5
    pb = O.OraclePlan(k, N, m, 2.0, O.LINEAR, True)
    pn = O.OraclePlan(k, N, m, 2.0, O.LINEAR, False)
    g = crandn(rng, pb.Nt)
    y = crandn(rng, M)
    Bg = pb.convolve(g)
    Bty = pb.convolve_transpose(y)
    assert abs(np.vdot(y, Bg) - np.vdot(Bty, g)) / abs(np.vdot(y, Bg)) < 1e-13
    assert O.relerr(Bg, pn.convolve(g)) < 1e-13
    assert O.relerr(Bty, pn.convolve_transpose(y)) < 1e-13
    pt = O.OraclePlan(k, N, m, 2.0, O.TENSOR, True)
    pp = O.OraclePlan(k, N, m, 2.0, O.POLYNOMIAL, True)
    assert O.relerr(pt.convolve(g), pp.convolve(g)) < 1e-14


def test_oracle_issue_106():
    """test/issues.jl:5-15 on the oracle: LINEAR LUT edge at 0.008333333f0."""
    k = np.array([[float(np.float32(0.008333333)), 0.0]])
    for blocking in (True, False):
        p = O.OraclePlan(k, (240,), 5, 2.0, O.LINEAR, blocking)
        out = p.adjoint(np.ones(2))
        assert O.relerr(out, O.ndft_adjoint(k, (240,), np.ones(2))) < 1e-7


def test_grid_size_rule():
    """Ntilde = (ceil(sigma N) / 2) * 2 and sigma <- Ntilde/N along the first
    (Julia) dimension (src/precomputation.jl:25-29)."""
    rng = np.random.default_rng(0)
    for N, sigma in [((31, 33), 2.0), ((10,), 1.25), ((7, 9), 1.5), ((2 ** 24,), 1.25)]:
        k = rng.random((len(N), 3)) - 0.5
        p = O.OraclePlan(k, N, 2, sigma, O.LINEAR, True) if np.prod(N) < 1e6 else None
        Nt = O.grid_size(N, sigma)
        assert all(x % 2 == 0 for x in Nt)
        if p is not None:
            assert p.Nt == Nt
            assert p.sigma == Nt[-1] / N[-1]
