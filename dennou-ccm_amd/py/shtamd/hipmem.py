"""Host <-> device copies into raw device addresses (ensemble member views,
request buffers) through the HIP runtime's hipMemcpy / hipMemcpy2D.
Plumbing for the Python side; the transforms run in libsht_hip.so."""
from __future__ import annotations

import ctypes

import numpy as np

_hip = None
_H2D, _D2H = 1, 2


def _rt():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", "libamdhip64.so.6"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("HIP runtime (libamdhip64.so) not found")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy2D.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
    return _hip


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"HIP copy failed with error {rc}")


def to_device(dst: int, a: np.ndarray):
    a = np.ascontiguousarray(a)
    _check(_rt().hipMemcpy(dst, a.ctypes.data, a.nbytes, _H2D))


def from_device(src: int, shape, dtype=np.float64) -> np.ndarray:
    out = np.empty(shape, dtype=dtype)
    _check(_rt().hipMemcpy(out.ctypes.data, src, out.nbytes, _D2H))
    return out


def to_device_2d(dst: int, dst_pitch_bytes: int, a: np.ndarray):
    """Copy a 2-D host array (rows x cols, C order) into rows of pitch dst_pitch_bytes."""
    a = np.ascontiguousarray(a)
    rows = a.shape[0]
    width = a.nbytes // rows
    _check(_rt().hipMemcpy2D(dst, dst_pitch_bytes, a.ctypes.data, width, width, rows, _H2D))


def from_device_2d(src: int, src_pitch_bytes: int, shape, dtype=np.float64) -> np.ndarray:
    out = np.empty(shape, dtype=dtype)
    rows = shape[0]
    width = out.nbytes // rows
    _check(_rt().hipMemcpy2D(out.ctypes.data, width, src, src_pitch_bytes, width, rows, _D2H))
    return out


def put_member_spectra(ptr: int, ld: int, spec_c: np.ndarray):
    """Write complex spectra (K, nf) into a member view with row stride ld doubles."""
    real = np.ascontiguousarray(np.stack([spec_c.real, spec_c.imag], axis=-1).reshape(spec_c.shape[0], -1))
    to_device_2d(ptr, ld * 8, real)


def get_member_spectra(ptr: int, ld: int, K: int, nf: int) -> np.ndarray:
    r = from_device_2d(ptr, ld * 8, (K, 2 * nf))
    return r[:, 0::2] + 1j * r[:, 1::2]
