"""Writes tests/golden/golden.json: closed-form values the oracle is pinned to.

Everything here is computed from the mathematics with mpmath at 40 digits, not
from the reference (which is Julia and not runnable here):

* Kaiser-Bessel window psi(k; Nt, m, sigma) = sinh(b sqrt(m^2 - (Nt k)^2)) /
  (pi sqrt(m^2 - (Nt k)^2)), b = pi (2 - 1/sigma)  (src/windowFunctions.jl:21-34)
* its transform psi_hat(n) = I0(m sqrt(b^2 - (2 pi n / Nt)^2))  (:36-39)
* tiny direct NDFTs f_hat_j = sum_n f_n exp(-2 pi i n.k_j) and adjoints
  (src/direct.jl:99-151), n_d = i_d - floor(N_d / 2), Python axis order.

Run: python tests/golden/make_golden.py
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))


def psi(k, Nt, m, sigma):
    k, Nt, m, sigma = mp.mpf(k), mp.mpf(Nt), mp.mpf(m), mp.mpf(sigma)
    b = mp.pi * (2 - 1 / sigma)
    if abs(k) < m / Nt:
        arg = mp.sqrt(m ** 2 - (Nt * k) ** 2)
        return mp.sinh(b * arg) / (mp.pi * arg)
    if abs(k) > m / Nt:
        return mp.mpf(0)
    return b / mp.pi


def psi_hat(n, Nt, m, sigma):
    n, Nt, m, sigma = mp.mpf(n), mp.mpf(Nt), mp.mpf(m), mp.mpf(sigma)
    b = mp.pi * (2 - 1 / sigma)
    return mp.besseli(0, m * mp.sqrt(b ** 2 - (2 * mp.pi * n / Nt) ** 2))


def psi_du(u, m, sigma):
    """d psi / du in grid units (psi(u / Nt) with Nt = 1), by mpmath's
    numerical differentiation at 40 digits."""
    return mp.diff(lambda x: psi(x, 1, m, sigma), mp.mpf(u))


def ndft(k, f):
    """k: (D, M) floats, f: nested array of shape N (C order, axis d <-> k[d])."""
    f = np.asarray(f, dtype=complex)
    N = f.shape
    D, M = len(k), len(k[0])
    out = []
    for j in range(M):
        s = mp.mpc(0)
        for idx in np.ndindex(*N):
            arg = mp.mpf(0)
            for d in range(D):
                arg += mp.mpf(k[d][j]) * (idx[d] - N[d] // 2)
            s += mp.mpc(f[idx].real, f[idx].imag) * mp.expjpi(-2 * arg)
        out.append(s)
    return out


def ndft_adjoint(k, N, fhat):
    D, M = len(k), len(k[0])
    out = []
    for idx in np.ndindex(*N):
        s = mp.mpc(0)
        for j in range(M):
            arg = mp.mpf(0)
            for d in range(D):
                arg += mp.mpf(k[d][j]) * (idx[d] - N[d] // 2)
            s += mp.mpc(fhat[j].real, fhat[j].imag) * mp.expjpi(2 * arg)
        out.append(s)
    return out


def cplx(z):
    return [float(mp.re(z)), float(mp.im(z))]


def main():
    out = {"window": [], "window_hat": [], "ndft": [], "window_du": []}
    for (m, sigma, Nt) in [(4, 2.0, 1), (5, 2.0, 1), (6, 1.25, 1), (4, 2.0, 256), (8, 2.0, 1)]:
        for u in [0.0, 0.3, 1.0, 1.5, 2.999, 3.5, 3.999999, float(m) - 1e-9, float(m), float(m) + 0.25]:
            k = u / Nt
            out["window"].append({"k": k, "Nt": Nt, "m": m, "sigma": sigma,
                                  "value": float(psi(k, Nt, m, sigma))})
    for (m, sigma, Nt) in [(4, 2.0, 256), (5, 2.0, 62), (6, 1.25, 20971520), (8, 2.0, 1024)]:
        for n in [0, 1, 7, Nt // 8, Nt // 5, Nt // 4 - 1, -(Nt // 4)]:
            out["window_hat"].append({"n": n, "Nt": Nt, "m": m, "sigma": sigma,
                                      "value": float(psi_hat(n, Nt, m, sigma))})
    for (m, sigma) in [(4, 2.0), (6, 1.25), (2, 2.0), (8, 2.0)]:
        b = float(mp.pi * (2 - 1 / mp.mpf(sigma)))
        # includes points either side of b s = 1/2 (series / closed form switch)
        s_half = 0.5 / b
        u_sw = float(mp.sqrt(m ** 2 - mp.mpf(s_half) ** 2))
        for u in [0.0, 0.25, -1.3, 2.5, m - 0.5, -(m - 0.01), u_sw - 1e-6, u_sw + 1e-6, m - 1e-4]:
            out["window_du"].append({"u": u, "m": m, "sigma": sigma,
                                     "value": float(psi_du(u, m, sigma))})
    rng = np.random.default_rng(20261202)
    for N, M in [((5,), 3), ((3, 4), 4), ((2, 3, 2), 3)]:
        D = len(N)
        k = (rng.random((D, M)) - 0.5).round(6)
        k[0, 0] = -0.5  # the closed end of the node range
        f = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).round(6)
        fh = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).round(6)
        out["ndft"].append({
            "N": list(N), "k": k.tolist(),
            "f": [[float(z.real), float(z.imag)] for z in f.ravel()],
            "fhat": [[float(z.real), float(z.imag)] for z in fh],
            "forward": [cplx(z) for z in ndft(k.tolist(), f)],
            "adjoint": [cplx(z) for z in ndft_adjoint(k.tolist(), N, fh)],
        })
    with open(os.path.join(HERE, "golden.json"), "w") as fp:
        json.dump(out, fp, indent=1)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
