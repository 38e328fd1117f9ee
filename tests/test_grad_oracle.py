"""Pins the oracle's node gradient (CPU only).

The reference has no node gradient: its rrules return NoTangent() for the
nodes (AbstractNFFTs/ext/AbstractNFFTsChainRulesCoreExt.jl:14-45).  The
gradient of fhat = A(k) f that the library adds is pinned here without it:

* the analytic window derivative against mpmath's derivative of the closed
  form (tests/golden/golden.json, written by tests/golden/make_golden.py);
* the closed-form NDFT gradient sum_n f_n (-2 pi i n_d) e^{-2 pi i n.k} against
  central differences of the direct-sum loss;
* the plan gradient (window derivative in the interpolation sum) against
  central differences of the oracle plan's own forward transform, and against
  the closed form, for every precompute mode.

Loss: L(k) = ||A(k) f - y||^2, whose node gradient is the contraction with
ybar = 2 (A(k) f - y).
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
H = 1e-6  # finite-difference step on the node coordinates


def test_window_derivative_matches_mpmath():
    for g in GOLD["window_du"]:
        got = O.window_du(g["u"], g["m"], g["sigma"])
        assert got == pytest.approx(g["value"], rel=1e-12, abs=1e-9 * abs(O.window(0, 1, g["m"], g["sigma"]))), g


def _data(rng, N, M):
    D = len(N)
    k = rng.random((D, M)) - 0.5
    f = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    y = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    return k, f, y


def _fd(loss_rows, k, D):
    """Central differences of the per-node loss rows |A(k) f - y|_j^2: row j
    depends only on node j, so one evaluation per (dim, sign) suffices."""
    out = np.zeros_like(k)
    for d in range(D):
        kp, km = k.copy(), k.copy()
        kp[d] += H
        km[d] -= H
        out[d] = (loss_rows(kp) - loss_rows(km)) / (2 * H)
    return out


@pytest.mark.parametrize("N", [(17,), (12, 10), (6, 5, 4)])
def test_ndft_gradient_closed_form_vs_fd(N):
    rng = np.random.default_rng([7, len(N)])
    M = 25
    k, f, y = _data(rng, N, M)
    k = np.clip(k, -0.5 + 2 * H, 0.5 - 2 * H)
    r = O.ndft(k, f) - y
    g = O.ndft_grad(k, f, 2 * r)
    fd = _fd(lambda kk: np.abs(O.ndft(kk, f) - y) ** 2, k, len(N))
    assert O.relerr(g, fd) < 1e-8
    # the check discriminates: a sign error or swapped axes fail it
    assert O.relerr(-g, fd) > 1.0
    if len(N) > 1:
        assert O.relerr(g[::-1], fd) > 0.1


@pytest.mark.parametrize("pc,tol", [(O.FULL, 1e-8), (O.POLYNOMIAL, 1e-8), (O.TENSOR, 1e-8),
                                    (O.LINEAR, 1e-5)])
@pytest.mark.parametrize("N,m", [((40,), 4), ((12, 14), 4), ((12, 10), 6), ((6, 8, 6), 3)])
def test_plan_gradient_vs_fd_of_plan(pc, tol, N, m):
    """The gradient is that of the operator the plan applies: FD of the
    plan's forward agrees to FD accuracy.  (LINEAR: the table interpolant
    has a piecewise-constant slope; the analytic derivative of the sampled
    window differs from its finite differences by the table error / h; at
    m = 3 the table has 128 samples per grid cell, the FD step stays inside
    one segment and sees its slope, 1e-3 off the window's derivative.)"""
    if pc == O.LINEAR and m < 4:
        tol = 5e-3
    rng = np.random.default_rng([11, len(N), m])
    k, f, y = _data(rng, N, 30)
    k = np.clip(k, -0.5 + 2 * H, 0.5 - 2 * H)

    def rows(kk):
        return np.abs(O.OraclePlan(kk, N, m=m, sigma=2.0, precompute=pc).forward(f) - y) ** 2

    P = O.OraclePlan(k, N, m=m, sigma=2.0, precompute=pc)
    g = P.grad_nodes(f, 2 * (P.forward(f) - y))
    assert O.relerr(g, _fd(rows, k, len(N))) < tol


@pytest.mark.parametrize("N", [(64,), (20, 18), (8, 10, 6)])
def test_plan_gradient_vs_closed_form(N):
    """At m = 6, sigma = 2 the window-derivative gradient equals the exact
    NDFT gradient to the transform's own accuracy."""
    rng = np.random.default_rng([13, len(N)])
    k, f, y = _data(rng, N, 40)
    ybar = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    ref = O.ndft_grad(k, f, ybar)
    for pc in (O.FULL, O.POLYNOMIAL, O.LINEAR):
        g = O.OraclePlan(k, N, m=6, sigma=2.0, precompute=pc).grad_nodes(f, ybar)
        assert O.relerr(g, ref) < 1e-8, pc


def test_gradient_is_linear_in_cotangent_and_zero_for_zero():
    rng = np.random.default_rng(3)
    N = (10, 12)
    k, f, y = _data(rng, N, 20)
    P = O.OraclePlan(k, N, m=4, sigma=2.0, precompute=O.POLYNOMIAL)
    a = rng.standard_normal(20) + 1j * rng.standard_normal(20)
    b = rng.standard_normal(20) + 1j * rng.standard_normal(20)
    ga, gb, gab = P.grad_nodes(f, a), P.grad_nodes(f, b), P.grad_nodes(f, 2 * a - 3 * b)
    assert O.relerr(gab, 2 * ga - 3 * gb) < 1e-13
    assert np.all(P.grad_nodes(f, np.zeros(20)) == 0)
