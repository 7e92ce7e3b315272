"""ctypes view of the CPU FP64 oracle (oracle/sht_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_sht.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_gauss.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_legendre_column.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_dlegendre_column.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        for nm in ("orc_dft_fwd", "orc_dft_inv", "orc_fft_fwd", "orc_fft_inv"):
            getattr(L, nm).argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_plan_create.restype = ctypes.c_void_p
        L.orc_plan_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_plan_destroy.argtypes = [ctypes.c_void_p]
        L.orc_plan_ncoef.argtypes = [ctypes.c_void_p]
        L.orc_spec_to_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_grid_to_spec.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.orc_spml_pack.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_spml_unpack.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_avr_lonlat.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.orc_grid_to_div.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
        L.orc_spec_lapla.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int]
        _ip = ctypes.POINTER(ctypes.c_int)
        L.orc_remap.argtypes = [_ip, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def gauss(jm: int):
    mu, w, cl = np.empty(jm), np.empty(jm), np.empty(jm)
    rc = lib().orc_gauss(jm, _p(mu), _p(w), _p(cl))
    if rc != 0:
        raise ValueError("jm must be even and >= 2")
    return mu, w, cl


def legendre_column(nmax: int, m: int, mu, coslat=None):
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    cl = np.sqrt(1.0 - mu * mu) if coslat is None else np.ascontiguousarray(coslat, dtype=np.float64)
    out = np.empty((nmax - m + 1, mu.size))
    lib().orc_legendre_column(nmax, m, mu.size, _p(mu), _p(cl), _p(out))
    return out


def dlegendre_column(nmax: int, m: int, mu, coslat=None):
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    cl = np.sqrt(1.0 - mu * mu) if coslat is None else np.ascontiguousarray(coslat, dtype=np.float64)
    out = np.empty((nmax - m + 1, mu.size))
    lib().orc_dlegendre_column(nmax, m, mu.size, _p(mu), _p(cl), _p(out))
    return out


def ring_fwd(x, mmax: int, direct: bool = False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    F = np.empty(2 * (mmax + 1))
    (lib().orc_dft_fwd if direct else lib().orc_fft_fwd)(x.size, mmax, _p(x), _p(F))
    return F[0::2] + 1j * F[1::2]


def ring_inv(F, im: int, direct: bool = False):
    F = np.asarray(F, dtype=np.complex128)
    Fr = np.ascontiguousarray(np.stack([F.real, F.imag], axis=-1).reshape(-1))
    x = np.empty(im)
    (lib().orc_dft_inv if direct else lib().orc_fft_inv)(im, F.size - 1, _p(Fr), _p(x))
    return x


class Plan:
    """CPU transform plan; spectra use the canonical complex layout
    spec[k][f] (k = off(m) + n - m), returned as complex128 arrays (K, nf)."""

    VALUE, GRADLON, GRADLAT = 0, 1, 2

    def __init__(self, nmax: int, im: int, jm: int):
        self.nmax, self.im, self.jm = nmax, im, jm
        self._h = lib().orc_plan_create(nmax, im, jm)
        if not self._h:
            raise ValueError(f"bad plan T{nmax} {im}x{jm}")
        self.K = (nmax + 1) * (nmax + 2) // 2

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and lib is not None:  # module globals may be gone at interpreter exit
            lib().orc_plan_destroy(h)
            self._h = None

    def spec_to_grid(self, spec, kind: int = 0, nthreads: int = 0):
        spec = np.asarray(spec, dtype=np.complex128)
        nf = spec.shape[1]
        sr = np.ascontiguousarray(spec).view(np.float64).reshape(-1)
        sr = np.ascontiguousarray(sr)
        g = np.empty((nf, self.jm, self.im))
        lib().orc_spec_to_grid(self._h, nf, _p(sr), _p(g), kind, nthreads)
        return g

    def grid_to_spec(self, grid, nthreads: int = 0):
        grid = np.ascontiguousarray(grid, dtype=np.float64)
        nf = grid.shape[0]
        out = np.empty((self.K, nf), dtype=np.complex128)
        lib().orc_grid_to_spec(self._h, nf, _p(grid), _p(out.view(np.float64).reshape(-1)), nthreads)
        return out

    def grid_to_div(self, a=None, b=None, nthreads: int = 0):
        """Spectra of (1/cos) dA/dlon + (1/cos) d(B cos)/dlat; A or B may be None."""
        arrs = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (a, b)]
        nf = next(x for x in arrs if x is not None).shape[0]
        out = np.empty((self.K, nf), dtype=np.complex128)
        ptrs = [None if x is None else _p(x) for x in arrs]
        lib().orc_grid_to_div(self._h, nf, ptrs[0], ptrs[1], _p(out.view(np.float64).reshape(-1)), nthreads)
        return out

    def avr_lonlat(self, grid):
        grid = np.ascontiguousarray(grid, dtype=np.float64)
        out = np.empty(grid.shape[0])
        lib().orc_avr_lonlat(self._h, grid.shape[0], _p(grid), _p(out))
        return out


def spec_lapla(spec, nmax: int, power: int):
    """(-n(n+1))^power applied to canonical spectra (negative powers: n = 0 -> 0)."""
    spec = np.ascontiguousarray(np.asarray(spec, dtype=np.complex128))
    out = np.empty_like(spec)
    lib().orc_spec_lapla(nmax, spec.shape[1], _p(spec.view(np.float64).reshape(-1)),
                         _p(out.view(np.float64).reshape(-1)), power)
    return out


def remap(send_index, recv_index, coef, nrecv: int, send):
    """The coupler remap loop; send [ndata][nsend] -> recv [ndata][nrecv]."""
    si = np.ascontiguousarray(send_index, dtype=np.int32)
    ri = np.ascontiguousarray(recv_index, dtype=np.int32)
    cf = np.ascontiguousarray(coef, dtype=np.float64)
    send = np.ascontiguousarray(send, dtype=np.float64)
    ndata, nsend = send.shape
    out = np.empty((ndata, nrecv))
    ip = ctypes.POINTER(ctypes.c_int)
    lib().orc_remap(si.ctypes.data_as(ip), ri.ctypes.data_as(ip), _p(cf), len(cf), nsend, nrecv, ndata,
                    _p(send), _p(out))
    return out


def spml_pack(spec, nmax: int):
    spec = np.ascontiguousarray(np.asarray(spec, dtype=np.complex128))
    nf = spec.shape[1]
    wa = np.empty((nf, (nmax + 1) ** 2))
    lib().orc_spml_pack(nmax, nf, _p(spec.view(np.float64).reshape(-1)), _p(wa))
    return wa


def spml_unpack(wa, nmax: int):
    wa = np.ascontiguousarray(wa, dtype=np.float64)
    nf = wa.shape[0]
    spec = np.empty(((nmax + 1) * (nmax + 2) // 2, nf), dtype=np.complex128)
    lib().orc_spml_unpack(nmax, nf, _p(wa), _p(spec.view(np.float64).reshape(-1)))
    return spec
