"""Seeded synthetic inputs for the transform workloads.

Counter-based: every value is a pure function of (seed, config, member,
variable, level, n or j, m or i, part), so any subset of a workload can be
regenerated without generating the rest, and a given field is the same no matter
which batch it is drawn in.  splitmix64 supplies the bits, Box-Muller the normal
deviates (FP64).

Spectra: s_n^m ~ CN(0, n^-3) for n >= 1 (real and imaginary parts each with
variance n^-3 / 2), Im s_n^0 = 0, s_0^0 = 0.  Grid temperature: 200 + 100 cos^2(lat)
K plus N(0, 2^2) K noise per point.
"""
from __future__ import annotations

import numpy as np

MASTER_SEED = 20261202
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _uniform(counter: np.ndarray, seed: int) -> np.ndarray:
    """Uniform deviates in (0, 1) from 64-bit counters."""
    key = _splitmix64(np.uint64(seed))
    bits = _splitmix64(np.asarray(counter, dtype=np.uint64) ^ key)
    return ((bits >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def _counter(cfg, member, var, level, a, b, part):
    c = np.uint64(cfg) << np.uint64(56)
    c |= np.uint64(member) << np.uint64(44)
    c |= np.uint64(var) << np.uint64(40)
    c |= np.asarray(level, dtype=np.uint64) << np.uint64(28)
    c |= np.asarray(a, dtype=np.uint64) << np.uint64(16)
    c |= np.asarray(b, dtype=np.uint64) << np.uint64(4)
    c |= np.uint64(part)
    return c


def _normal(cfg, member, var, level, a, b, part, seed):
    u1 = _uniform(_counter(cfg, member, var, level, a, b, 2 * part), seed)
    u2 = _uniform(_counter(cfg, member, var, level, a, b, 2 * part + 1), seed)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def coef_index(nmax: int):
    """(n, m) arrays in canonical order k = off(m) + n - m."""
    ns, ms = [], []
    for m in range(nmax + 1):
        for n in range(m, nmax + 1):
            ns.append(n)
            ms.append(m)
    return np.array(ns), np.array(ms)


def spectra(nmax: int, levels: int, *, cfg: int = 0, member: int = 0, var: int = 0,
            seed: int = MASTER_SEED) -> np.ndarray:
    """Canonical complex spectra, shape (K, levels)."""
    n, m = coef_index(nmax)
    lev = np.arange(levels)
    N2, L2 = np.meshgrid(n, lev, indexing="ij")
    M2 = np.broadcast_to(m[:, None], N2.shape)
    re = _normal(cfg, member, var, L2, N2, M2, 0, seed)
    im = _normal(cfg, member, var, L2, N2, M2, 1, seed)
    with np.errstate(divide="ignore"):
        amp = np.where(N2 >= 1, np.sqrt(0.5 * np.power(np.maximum(N2, 1).astype(np.float64), -3.0)), 0.0)
    s = amp * (re + 1j * im)
    s[M2 == 0] = s[M2 == 0].real
    return np.ascontiguousarray(s)


def state_spectra(nmax: int, nlev: int, *, cfg: int = 0, member: int = 0,
                  seed: int = MASTER_SEED) -> np.ndarray:
    """Vorticity, divergence, temperature on nlev levels plus ln p_s: 3*nlev+1
    spectral fields (K, 3*nlev+1), variables 0..3."""
    parts = [spectra(nmax, nlev, cfg=cfg, member=member, var=v, seed=seed) for v in range(3)]
    parts.append(spectra(nmax, 1, cfg=cfg, member=member, var=3, seed=seed))
    return np.ascontiguousarray(np.concatenate(parts, axis=1))


def temperature_grid(coslat: np.ndarray, im: int, levels: int, *, cfg: int = 1, member: int = 0,
                     seed: int = MASTER_SEED) -> np.ndarray:
    """(levels, jm, im) grid temperature: 200 + 100 cos^2(lat) + N(0, 2^2)."""
    jm = coslat.size
    L3, J3, I3 = np.meshgrid(np.arange(levels), np.arange(jm), np.arange(im), indexing="ij")
    noise = _normal(cfg, member, 2, L3, J3, I3, 7, seed)
    return 200.0 + 100.0 * coslat[None, :, None] ** 2 + 2.0 * noise
