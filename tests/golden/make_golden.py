"""Writes the golden fixtures tests/golden/golden_T{N}.npz with mpmath at 40 digits.

Independent of the oracle and of the HIP library: the Gauss nodes come from
mpmath root finding on P_jm, the associated Legendre functions from
mpmath.legenp (its Condon-Shortley phase removed), the latitude derivative from
mpmath.diff of the same functions, and the transforms from direct summation.
The reference ships no golden vectors for this path (it has no tests), so these
fixtures, together with closed forms and the reference's own quadrature
identities, are what pins the oracle.

Run:  python tests/golden/make_golden.py   (about 8 minutes on one core)
"""
from __future__ import annotations

import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "dennou-ccm_amd", "py"))
from shtamd import synth  # noqa: E402

mp.mp.dps = 40

CASES = [  # (N, im, jm, nfield)
    (5, 16, 8, 3),
    (10, 32, 16, 3),
]


def gauss_mp(jm):
    nodes = []
    for k in range(jm):
        x0 = mp.cos(mp.pi * (4 * k + 3) / (4 * jm + 2))
        f = lambda t: mp.legendre(jm, t)  # noqa: E731
        df = lambda t: jm * (mp.legendre(jm - 1, t) - t * mp.legendre(jm, t)) / (1 - t * t)  # noqa: E731
        x = mp.findroot(f, x0, solver="newton", df=df)
        dp = mp.diff(f, x)
        w = 2 / ((1 - x * x) * dp * dp)
        nodes.append((x, w))
    nodes.sort(key=lambda t: t[0])
    mu = [t[0] for t in nodes]
    w = [t[1] for t in nodes]
    assert all(mu[i + 1] - mu[i] > mp.mpf("1e-6") for i in range(jm - 1)), "duplicate roots"
    assert abs(sum(w) - 2) < mp.mpf("1e-30")
    return mu, w


def ptilde(n, m, x):
    norm = mp.sqrt((2 * n + 1) * mp.factorial(n - m) / mp.factorial(n + m))
    return norm * (-1) ** m * mp.legenp(n, m, x)


def dptilde_dlat(n, m, lat):
    return mp.diff(lambda ph: ptilde(n, m, mp.sin(ph)), lat)


def main():
    for N, im, jm, nf in CASES:
        mu, w = gauss_mp(jm)
        lat = [mp.asin(x) for x in mu]
        cl = [mp.cos(p) for p in lat]
        ns, ms = synth.coef_index(N)
        K = ns.size
        P = np.zeros((K, jm))
        H = np.zeros((K, jm))  # cos(lat) dPt/dlat = (1-mu^2) dPt/dmu
        Pm = [[None] * jm for _ in range(K)]
        Dm = [[None] * jm for _ in range(K)]
        for k in range(K):
            n, m = int(ns[k]), int(ms[k])
            for j in range(jm):
                Pm[k][j] = ptilde(n, m, mu[j])
                Dm[k][j] = dptilde_dlat(n, m, lat[j])
                P[k, j] = float(Pm[k][j])
                H[k, j] = float(cl[j] * Dm[k][j])
        spec = synth.spectra(N, nf, cfg=9, member=N)
        grid = np.zeros((nf, jm, im))
        glon = np.zeros((nf, jm, im))
        glat = np.zeros((nf, jm, im))
        for f in range(nf):
            for j in range(jm):
                # Fourier coefficients of ring j for value, d/dlon and d/dlat
                Fv, Fl = [mp.mpc(0)] * (N + 1), [mp.mpc(0)] * (N + 1)
                for k in range(K):
                    m = int(ms[k])
                    s = mp.mpc(spec[k, f].real, spec[k, f].imag if m > 0 else 0.0)
                    Fv[m] += s * Pm[k][j]
                    Fl[m] += s * Dm[k][j]
                for i in range(im):
                    v = mp.re(Fv[0])
                    vl = mp.mpf(0)
                    vt = mp.re(Fl[0])
                    for m in range(1, N + 1):
                        e = mp.expjpi(2 * m * mp.mpf(i) / im)
                        v += 2 * mp.re(Fv[m] * e)
                        vl += 2 * mp.re(1j * m * Fv[m] * e)
                        vt += 2 * mp.re(Fl[m] * e)
                    grid[f, j, i] = float(v)
                    glon[f, j, i] = float(vl / cl[j])
                    glat[f, j, i] = float(vt)
        # forward transform of a non-band-limited grid (uniform noise in [-1,1])
        rng = np.random.default_rng(1000 + N)
        noisy = rng.uniform(-1.0, 1.0, size=(nf, jm, im))
        fwd = np.zeros((K, nf), dtype=np.complex128)
        for f in range(nf):
            Fr = [[mp.mpc(0)] * (N + 1) for _ in range(jm)]
            for j in range(jm):
                for m in range(N + 1):
                    acc = mp.mpc(0)
                    for i in range(im):
                        acc += mp.mpf(noisy[f, j, i]) * mp.expjpi(-2 * m * mp.mpf(i) / im)
                    Fr[j][m] = acc / im
            for k in range(K):
                m = int(ms[k])
                acc = mp.mpc(0)
                for j in range(jm):
                    acc += w[j] * Pm[k][j] * Fr[j][m]
                acc /= 2
                fwd[k, f] = complex(float(mp.re(acc)), float(mp.im(acc)) if m > 0 else 0.0)
        out = os.path.join(HERE, f"golden_T{N}.npz")
        np.savez_compressed(
            out, nmax=N, im=im, jm=jm, mu=np.array([float(x) for x in mu]),
            w=np.array([float(x) for x in w]), coslat=np.array([float(x) for x in cl]),
            P=P, H=H, spec=spec, grid=grid, gradlon=glon, gradlat=glat, noisy=noisy, fwd=fwd)
        print("wrote", out)


if __name__ == "__main__":
    main()
