"""CPU pins of the oracle's window pairs (src/windowFunctions.jl:21-134).

* psi_hat is the Fourier transform of psi: Nt * int psi(k) cos(2 pi n k) dk by
  adaptive quadrature matches window_hat for every window.  The Gaussian's
  psi_hat is that of the untruncated Gaussian, so it differs by the cut tail
  exp(-pi m) (~1.5e-7 at m = 5); the Kaiser-Bessel window jumps to b/pi at
  its support edge, which limits the quadrature to ~1e-10.
* J0 against scipy.special.j0.
* The reference's accuracy test (NFFT.jl test/accuracy.jl:40-72: m = 5,
  sigma = 2, NFFT against NDFT within 1e-7 kaiser_bessel, 1e-7 cosh_type,
  1e-3 gauss, 1e-6 kaiser_bessel_rev, 1e-4 spline) for every precompute
  mode, on the oracle's FP64 plan.
"""
import numpy as np
import pytest
import scipy.integrate as si
import scipy.special as sp

import oracle as O

TOL_FT = {"kaiser_bessel": 1e-9, "gauss": 1e-6, "spline": 1e-10, "kaiser_bessel_rev": 1e-10,
          "cosh_type": 1e-10}
EPS = {"kaiser_bessel": 1e-7, "cosh_type": 1e-7, "gauss": 1e-3, "kaiser_bessel_rev": 1e-6, "spline": 1e-4}


@pytest.mark.parametrize("window", list(O.WINDOWS))
def test_window_hat_is_fourier_transform(window):
    m, sig, Nt = 5, 2.0, 64
    for n in (0, 3, 10, 17):
        ft = si.quad(lambda k: O.window_w(window, k, Nt, m, sig) * np.cos(2 * np.pi * n * k),
                     -m / Nt, m / Nt, limit=400, epsabs=0, epsrel=1e-13)[0] * Nt
        hat = O.window_hat_w(window, n, Nt, m, sig)
        assert abs(ft - hat) <= TOL_FT[window] * abs(hat), (n, ft, hat)


def test_bessel_j0():
    for x in (0.0, 0.5, 3.7, 7.3, 12.9, -4.2):
        assert abs(O.bessel_j0(x) - sp.j0(x)) < 1e-15


@pytest.mark.parametrize("pre", [O.LINEAR, O.FULL, O.TENSOR, O.POLYNOMIAL])
@pytest.mark.parametrize("window", list(O.WINDOWS))
def test_oracle_accuracy_per_window(window, pre):
    rng = np.random.default_rng([O.WINDOWS[window], pre])
    N = (31, 33)
    J = int(np.prod(N))
    k = rng.random((2, J)) - 0.5
    p = O.OraclePlan(k, N, m=5, sigma=2.0, precompute=pre, window=window)
    fhat = rng.random(J) + 1j * rng.random(J)
    f = O.ndft_adjoint(k, N, fhat)
    assert O.relerr(p.adjoint(fhat), f) < EPS[window]
    g = O.ndft(k, f)
    assert O.relerr(p.forward(f), g) < EPS[window]
