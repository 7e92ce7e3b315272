"""Python binding of libsht_hip.so (include/sht.h) through ctypes.

The library is the product; this module only loads it and passes pointers.
Device arrays are torch tensors (device memory and streams are torch's
plumbing); the transform work is the library's HIP kernels.  There is no
fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(os.path.dirname(_HERE))  # dennou-ccm_amd/
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libsht_hip.so")
HEADER = os.path.join(os.path.dirname(PKG_ROOT), "include", "sht.h")

SHT_OK = 0
SHT_ERR_INVALID = -1
SHT_ERR_NOMEM = -2
SHT_ERR_DEVICE = -3
SHT_ERR_UNSUPPORTED = -4
SHT_ERR_BUSY = -5

PREC_FP64, PREC_SPLIT_BF16, PREC_SCALED_FP16, PREC_SPLIT_BF16X6 = 0, 1, 2, 3
DIR_INVERSE, DIR_FORWARD, DIR_ROUNDTRIP = 0, 1, 2
SHAPE_AUTO, SHAPE_KSPLIT, SHAPE_WAVE, SHAPE_TILE, SHAPE_TILE_SPLIT, SHAPE_STAGED = -1, 0, 1, 2, 3, 4
SHAPES = ("ksplit", "wave", "tile", "tile_split", "staged")  # the FP64 plans' shapes
NSHAPE = len(SHAPES)
# inverse of mixed-precision plans (64 <= im <= 512, nmax < 192): Legendre
# step and ring FFTs in one persistent launch
SHAPE_FUSED = 5
SHAPE_NAMES = SHAPES + ("fused",)
FLAG_GRAD = 1
TUNE_SHAPES_ONLY = 1

_lib = None


class ShtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class _Desc(ctypes.Structure):
    _fields_ = [("nmax", ctypes.c_int), ("im", ctypes.c_int), ("jm", ctypes.c_int),
                ("kmax", ctypes.c_int), ("precision", ctypes.c_int), ("flags", ctypes.c_int)]


class _MemberDesc(ctypes.Structure):
    _fields_ = [("nmax", ctypes.c_int), ("im", ctypes.c_int), ("jm", ctypes.c_int), ("nfield", ctypes.c_int)]


class _ServerDesc(ctypes.Structure):
    _fields_ = [("ngrid", ctypes.c_int), ("nmax", ctypes.c_int * 8), ("im", ctypes.c_int * 8),
                ("jm", ctypes.c_int * 8), ("slots", ctypes.c_int), ("max_fields", ctypes.c_int),
                ("precision", ctypes.c_int), ("mem_budget", ctypes.c_size_t)]


class _RestartDesc(ctypes.Structure):
    _fields_ = [("im", ctypes.c_int), ("jm", ctypes.c_int), ("km", ctypes.c_int), ("nq", ctypes.c_int),
                ("time", ctypes.c_double), ("lon", ctypes.c_void_p), ("lat", ctypes.c_void_p),
                ("sig", ctypes.c_void_p)]


class _Request(ctypes.Structure):
    _fields_ = [("nmax", ctypes.c_int), ("nfield", ctypes.c_int), ("direction", ctypes.c_int),
                ("inp", ctypes.c_void_p), ("out", ctypes.c_void_p)]


_vp = ctypes.c_void_p
_ci = ctypes.c_int
_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); mirrors include/sht.h
SIGNATURES = {
    "sht_plan_create": (_ci, [ctypes.POINTER(_Desc), ctypes.POINTER(_vp)]),
    "sht_plan_destroy": (None, [_vp]),
    "sht_last_error": (ctypes.c_char_p, []),
    "sht_abi_version": (_ci, []),
    "sht_plan_grid": (_ci, [_vp, _dp, _dp, _dp, _dp]),
    "sht_l_nm": (_ci, [_ci, _ci]),
    "sht_nm_l": (_ci, [_ci, ctypes.POINTER(_ci), ctypes.POINTER(_ci)]),
    "sht_spec_to_grid_dev": (_ci, [_vp, _vp, _vp, _ci, _vp]),
    "sht_grid_to_spec_dev": (_ci, [_vp, _vp, _vp, _ci, _vp]),
    "sht_spec_to_grid_dev_noalloc": (_ci, [_vp, _vp, _vp, _ci, _vp]),
    "sht_spec_to_grad_dev": (_ci, [_vp, _vp, _vp, _vp, _vp, _ci, _vp]),
    "sht_grid_to_div_dev": (_ci, [_vp, _vp, _vp, _vp, _ci, _vp]),
    "sht_spec_lapla_dev": (_ci, [_vp, _vp, _vp, _ci, _ci, _vp]),
    "sht_spml_to_canon_dev": (_ci, [_vp, _vp, _vp, _ci, _vp]),
    "sht_canon_to_spml_dev": (_ci, [_vp, _vp, _vp, _ci, _vp]),
    "sht_avr_lonlat_dev": (_ci, [_vp, _vp, _vp, _ci, _vp]),
    "sht_xya_wa": (_ci, [_vp, _dp, _dp, _ci]),
    "sht_wa_xya": (_ci, [_vp, _dp, _dp, _ci]),
    "sht_xya_gradlon_wa": (_ci, [_vp, _dp, _dp, _ci]),
    "sht_xya_gradlat_wa": (_ci, [_vp, _dp, _dp, _ci]),
    "sht_wa_divlon_xya": (_ci, [_vp, _dp, _dp, _ci]),
    "sht_wa_divlat_xya": (_ci, [_vp, _dp, _dp, _ci]),
    "sht_wa_lapla_wa": (_ci, [_vp, _dp, _dp, _ci]),
    "sht_wa_laplainv_wa": (_ci, [_vp, _dp, _dp, _ci]),
    "sht_batch_create": (_ci, [ctypes.POINTER(_MemberDesc), _ci, _ci, ctypes.POINTER(_vp)]),
    "sht_batch_destroy": (None, [_vp]),
    "sht_batch_member": (_ci, [_vp, _ci, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                               ctypes.POINTER(_ci)]),
    "sht_batch_run": (_ci, [_vp, _ci, _vp]),
    "sht_batch_run_eager": (_ci, [_vp, _ci, _vp]),
    "sht_batch_autotune": (_ci, [_vp, _ci, ctypes.c_char_p]),
    "sht_server_create": (_ci, [ctypes.POINTER(_ServerDesc), ctypes.POINTER(_vp)]),
    "sht_server_destroy": (None, [_vp]),
    "sht_server_submit": (_ci, [_vp, ctypes.POINTER(_Request), ctypes.POINTER(ctypes.c_longlong)]),
    "sht_server_poll": (_ci, [_vp, ctypes.POINTER(ctypes.c_longlong), _dp, _ci, ctypes.POINTER(_ci)]),
    "sht_server_drain": (_ci, [_vp]),
    "sht_server_stats": (_ci, [_vp] + [ctypes.POINTER(ctypes.c_longlong)] * 4),
    "sht_server_failures": (_ci, [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.c_char_p, _ci]),
    "sht_plan_set_timing": (_ci, [_vp, _ci]),
    "sht_plan_stage_times": (_ci, [_vp, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "sht_remap_create": (_ci, [ctypes.POINTER(_ci), ctypes.POINTER(_ci), _dp, _ci, _ci, _ci,
                               ctypes.POINTER(_vp)]),
    "sht_remap_destroy": (None, [_vp]),
    "sht_remap_apply_dev": (_ci, [_vp, _vp, _vp, _ci, _vp]),
    "sht_plan_set_shapes": (_ci, [_vp, _ci, _ci, _ci]),
    "sht_plan_get_shapes": (_ci, [_vp, _ci, ctypes.POINTER(_ci), ctypes.POINTER(_ci)]),
    "sht_plan_fused_failures": (_ci, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "sht_plan_autotune": (_ci, [_vp, _ci, _ci, ctypes.c_char_p, _dp]),
    "sht_plan_autotune_flags": (_ci, [_vp, _ci, _ci, ctypes.c_char_p, _ci, _dp]),
    "sht_batch_autotune_flags": (_ci, [_vp, _ci, ctypes.c_char_p, _ci]),
    "sht_plan_set_classes": (_ci, [_vp, _ci, _ci, _ci, _ci]),
    "sht_plan_get_classes": (_ci, [_vp, _ci, _ci, ctypes.POINTER(_ci), ctypes.POINTER(_ci)]),
    "sht_zonal_create": (_ci, [_ci, _ci, _ci, ctypes.POINTER(_vp)]),
    "sht_zonal_destroy": (None, [_vp]),
    "sht_zonal_grid": (_ci, [_vp, _dp, _dp]),
    "sht_l_y_dev": (_ci, [_vp, _vp, _vp, _ci, _vp]),
    "sht_y_l_dev": (_ci, [_vp, _vp, _vp, _ci, _vp]),
    "sht_l_y": (_ci, [_vp, _dp, _dp, _ci]),
    "sht_y_l": (_ci, [_vp, _dp, _dp, _ci]),
    "sht_lat_cell_edges": (_ci, [_dp, _ci, _dp]),
    "sht_restart_write": (_ci, [ctypes.c_char_p, ctypes.POINTER(_RestartDesc), ctypes.POINTER(_vp)]),
    "sht_restart_read": (_ci, [ctypes.c_char_p, ctypes.POINTER(_RestartDesc), _dp, ctypes.POINTER(_vp)]),
    "sht_exchange_grid": (_ci, [_dp, _dp, _ci, _dp, _dp, _ci, _dp, _dp, ctypes.POINTER(_ci)]),
}

STAGES = ("leg_inv", "fft_inv", "fft_fwd", "leg_fwd")


def _init_torch_device():
    """PyTorch ships its own HIP runtime beside the system one this library
    links.  When the library's runtime opens the GPU first, PyTorch's finds no
    GPU afterwards ("No HIP GPUs are available"); initialised first, both work
    (the binding hands PyTorch's device pointers to the library).  So when
    PyTorch is importable and sees a GPU, it initialises before the library
    loads."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def lib():
    """Load libsht_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        _init_torch_device()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != SHT_OK:
        raise ShtError(rc, lib().sht_last_error().decode())
    return rc


def ncoef(nmax: int) -> int:
    return (nmax + 1) * (nmax + 2) // 2


def _ptr(t) -> int | None:
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()


def _dev(t, need: int, what: str) -> int | None:
    """Device address of a float64 CUDA tensor holding at least `need`
    elements; None passes through (optional outputs)."""
    if t is None:
        return None
    if not getattr(t, "is_cuda", False):
        raise ValueError(f"{what}: expected a CUDA tensor")
    if str(t.dtype) != "torch.float64":
        raise ValueError(f"{what}: expected float64, got {t.dtype}")
    if t.numel() < need:
        raise ValueError(f"{what}: {t.numel()} elements, need at least {need}")
    return _ptr(t)


def _stream_handle(stream) -> int | None:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)


def _host(a) -> "ctypes._Pointer":
    return a.ctypes.data_as(_dp)


class Plan:
    """One transform plan (SPML wa_Initial).  Device calls take torch tensors:
    spectra as float64 (K, nfield, 2) or (K*nfield*2,) in the canonical
    layout, grids as float64 (nfield, jm, im)."""

    def __init__(self, nmax: int, im: int, jm: int, kmax: int, precision: int = PREC_FP64,
                 flags: int = 0):
        d = _Desc(nmax, im, jm, kmax, precision, flags)
        h = _vp()
        check(lib().sht_plan_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.nmax, self.im, self.jm, self.kmax = nmax, im, jm, kmax
        self.precision, self.flags = precision, flags
        self.K = ncoef(nmax)

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().sht_plan_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid(self):
        lon, lat = np.empty(self.im), np.empty(self.jm)
        lw, yw = np.empty(self.im), np.empty(self.jm)
        check(lib().sht_plan_grid(self._h, _host(lon), _host(lat), _host(lw), _host(yw)))
        return lon, lat, lw, yw

    # ---- device (canonical spectra, stream ordered)
    def _ns(self, nfield):
        return self.K * 2 * nfield

    def _ng(self, nfield):
        return nfield * self.jm * self.im

    def spec_to_grid(self, spec, grid, nfield: int, stream=None):
        check(lib().sht_spec_to_grid_dev(self._h, _dev(spec, self._ns(nfield), "spec"),
                                         _dev(grid, self._ng(nfield), "grid"), nfield, _stream_handle(stream)))

    def spec_to_grid_noalloc(self, spec, grid, nfield: int, stream=None):
        """spec_to_grid through the library's static scratch (allocates nothing)."""
        check(lib().sht_spec_to_grid_dev_noalloc(self._h, _dev(spec, self._ns(nfield), "spec"),
                                                 _dev(grid, self._ng(nfield), "grid"), nfield,
                                                 _stream_handle(stream)))

    def grid_to_spec(self, grid, spec, nfield: int, stream=None):
        check(lib().sht_grid_to_spec_dev(self._h, _dev(grid, self._ng(nfield), "grid"),
                                         _dev(spec, self._ns(nfield), "spec"), nfield, _stream_handle(stream)))

    def spec_to_grad(self, spec, grid, gradlon, gradlat, nfield: int, stream=None):
        ng = self._ng(nfield)
        check(lib().sht_spec_to_grad_dev(self._h, _dev(spec, self._ns(nfield), "spec"), _dev(grid, ng, "grid"),
                                         _dev(gradlon, ng, "gradlon"), _dev(gradlat, ng, "gradlat"),
                                         nfield, _stream_handle(stream)))

    def grid_to_div(self, a, b, spec, nfield: int, stream=None):
        """Spectra of (1/cos) dA/dlon + (1/cos) d(B cos)/dlat; a or b may be None."""
        ng = self._ng(nfield)
        check(lib().sht_grid_to_div_dev(self._h, _dev(a, ng, "a"), _dev(b, ng, "b"),
                                        _dev(spec, self._ns(nfield), "spec"), nfield, _stream_handle(stream)))

    def spec_lapla(self, spec_in, spec_out, power: int, nfield: int, stream=None):
        ns = self._ns(nfield)
        check(lib().sht_spec_lapla_dev(self._h, _dev(spec_in, ns, "spec_in"), _dev(spec_out, ns, "spec_out"),
                                       power, nfield, _stream_handle(stream)))

    def spml_to_canon(self, wa, spec, nfield: int, stream=None):
        check(lib().sht_spml_to_canon_dev(self._h, _dev(wa, nfield * (self.nmax + 1) ** 2, "wa"),
                                          _dev(spec, self._ns(nfield), "spec"), nfield, _stream_handle(stream)))

    def canon_to_spml(self, spec, wa, nfield: int, stream=None):
        check(lib().sht_canon_to_spml_dev(self._h, _dev(spec, self._ns(nfield), "spec"),
                                          _dev(wa, nfield * (self.nmax + 1) ** 2, "wa"), nfield,
                                          _stream_handle(stream)))

    def avr_lonlat(self, grid, avg, nfield: int, stream=None):
        check(lib().sht_avr_lonlat_dev(self._h, _dev(grid, self._ng(nfield), "grid"), _dev(avg, nfield, "avg"),
                                       nfield, _stream_handle(stream)))

    # ---- stage timers
    def set_timing(self, enable: bool):
        check(lib().sht_plan_set_timing(self._h, 1 if enable else 0))

    def stage_times(self):
        """{stage: (total_ms, launches)} since the last call (waits for the GPU)."""
        ms = (ctypes.c_double * 4)()
        n = (ctypes.c_longlong * 4)()
        check(lib().sht_plan_stage_times(self._h, ms, n))
        return {STAGES[i]: (ms[i], n[i]) for i in range(4)}

    # ---- Legendre kernel shapes
    def set_shapes(self, nfield: int, inv: int, fwd: int):
        """Fix the shapes for calls with nfield fields (0: the plan default)."""
        check(lib().sht_plan_set_shapes(self._h, nfield, inv, fwd))

    def get_shapes(self, nfield: int):
        a, b = ctypes.c_int(), ctypes.c_int()
        check(lib().sht_plan_get_shapes(self._h, nfield, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def fused_failures(self) -> int:
        """Waits of fused-shape calls on this plan that gave up (see sht.h)."""
        n = ctypes.c_longlong(0)
        check(lib().sht_plan_fused_failures(self._h, ctypes.byref(n)))
        return n.value

    def set_classes(self, nfield: int, direction: int, m4: int, m2: int):
        """Wavenumber classes of the split shape for calls with nfield fields."""
        check(lib().sht_plan_set_classes(self._h, nfield, direction, m4, m2))

    def get_classes(self, nfield: int, direction: int):
        a, b = ctypes.c_int(), ctypes.c_int()
        check(lib().sht_plan_get_classes(self._h, nfield, direction, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune(self, nfield: int, reps: int = 20, cache_path: str | None = None, shapes_only: bool = False):
        """Time every shape at nfield fields and keep the fastest.  Returns
        ((inv_shape, fwd_shape), (inv_us, fwd_us)); the times are 0 on a cache hit."""
        us = (ctypes.c_double * 2)()
        path = cache_path.encode() if cache_path else None
        check(lib().sht_plan_autotune_flags(self._h, nfield, reps, path, TUNE_SHAPES_ONLY if shapes_only else 0, us))
        return self.get_shapes(nfield), (us[0], us[1])

    # ---- host drop-in (SPML packing), numpy arrays; out= takes a preallocated
    # result array (as a model passes its own arrays), else one is allocated
    def _host_call(self, fn, inp, in_tail, out_tail, out=None):
        inp = np.ascontiguousarray(inp, dtype=np.float64)
        if inp.ndim != 1 + len(in_tail) or inp.shape[1:] != tuple(in_tail):
            raise ValueError(f"expected an array of shape (nlev, {', '.join(map(str, in_tail))}), "
                             f"got {inp.shape}")
        shape = (inp.shape[0], *out_tail)
        if out is None:
            out = np.empty(shape)
        elif (not isinstance(out, np.ndarray) or out.dtype != np.float64 or out.shape != shape
              or not out.flags.c_contiguous or not out.flags.writeable):
            raise ValueError(f"out must be a writeable C-contiguous float64 array of shape {shape}")
        check(fn(self._h, _host(inp), _host(out), inp.shape[0]))
        return out

    def _wa(self):
        return ((self.nmax + 1) ** 2,)

    def _xy(self):
        return (self.jm, self.im)

    def xya_wa(self, wa, out=None):
        return self._host_call(lib().sht_xya_wa, wa, self._wa(), self._xy(), out)

    def wa_xya(self, xya, out=None):
        return self._host_call(lib().sht_wa_xya, xya, self._xy(), self._wa(), out)

    def xya_gradlon_wa(self, wa, out=None):
        return self._host_call(lib().sht_xya_gradlon_wa, wa, self._wa(), self._xy(), out)

    def xya_gradlat_wa(self, wa, out=None):
        return self._host_call(lib().sht_xya_gradlat_wa, wa, self._wa(), self._xy(), out)

    def wa_divlon_xya(self, xya, out=None):
        return self._host_call(lib().sht_wa_divlon_xya, xya, self._xy(), self._wa(), out)

    def wa_divlat_xya(self, xya, out=None):
        return self._host_call(lib().sht_wa_divlat_xya, xya, self._xy(), self._wa(), out)

    def wa_lapla_wa(self, wa, out=None):
        return self._host_call(lib().sht_wa_lapla_wa, wa, self._wa(), self._wa(), out)

    def wa_laplainv_wa(self, wa, out=None):
        return self._host_call(lib().sht_wa_laplainv_wa, wa, self._wa(), self._wa(), out)


class Zonal:
    """Zonal (m = 0) Legendre transform (SPML l_module, l_Initial(nm, jm)).
    Grids y (nfield, jm) S->N, Legendre coefficients l (nfield, nmax + 1);
    l_y: l_n = (1/2) sum_j w_j Pt_n(mu_j) y_j, y_l its inverse, Pt_n = sqrt(2n+1) P_n."""

    def __init__(self, nmax: int, jm: int, kmax: int = 1):
        h = _vp()
        check(lib().sht_zonal_create(nmax, jm, kmax, ctypes.byref(h)))
        self._h = h
        self.nmax, self.jm, self.kmax = nmax, jm, kmax

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().sht_zonal_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid(self):
        """(y_Lat, y_Lat_Weight)"""
        lat, w = np.empty(self.jm), np.empty(self.jm)
        check(lib().sht_zonal_grid(self._h, _host(lat), _host(w)))
        return lat, w

    def l_y_dev(self, y, l, nfield: int, stream=None):
        check(lib().sht_l_y_dev(self._h, _dev(y, nfield * self.jm, "y"), _dev(l, nfield * (self.nmax + 1), "l"),
                                nfield, _stream_handle(stream)))

    def y_l_dev(self, l, y, nfield: int, stream=None):
        check(lib().sht_y_l_dev(self._h, _dev(l, nfield * (self.nmax + 1), "l"), _dev(y, nfield * self.jm, "y"),
                                nfield, _stream_handle(stream)))

    def _host(self, fn, inp, n_in, n_out):
        inp = np.ascontiguousarray(inp, dtype=np.float64)
        one = inp.ndim == 1
        if one:
            inp = inp[None]
        if inp.ndim != 2 or inp.shape[1] != n_in:
            raise ValueError(f"expected shape ({n_in},) or (nlev, {n_in}), got {inp.shape}")
        out = np.empty((inp.shape[0], n_out))
        check(fn(self._h, _host(inp), _host(out), inp.shape[0]))
        return out[0] if one else out

    def l_y(self, y):
        return self._host(lib().sht_l_y, y, self.jm, self.nmax + 1)

    def y_l(self, l):
        return self._host(lib().sht_y_l, l, self.nmax + 1, self.jm)


def lat_cell_edges(lat_weight):
    """Cell edges of a latitude set (gmapgen_main.f90:364-368): jm + 1 values."""
    w = np.ascontiguousarray(lat_weight, dtype=np.float64)
    e = np.empty(w.size + 1)
    check(lib().sht_lat_cell_edges(_host(w), w.size, _host(e)))
    return e


def exchange_grid(lat_a, w_a, lat_b, w_b):
    """gmapgen's surface exchange grid of latitude sets A and B: (lat, weight)."""
    la, wa = (np.ascontiguousarray(x, dtype=np.float64) for x in (lat_a, w_a))
    lb, wb = (np.ascontiguousarray(x, dtype=np.float64) for x in (lat_b, w_b))
    cap = wa.size + wb.size - 1
    lat, w, n = np.empty(cap), np.empty(cap), _ci()
    check(lib().sht_exchange_grid(_host(la), _host(wa), wa.size, _host(lb), _host(wb), wb.size,
                                  _host(lat), _host(w), ctypes.byref(n)))
    return lat[:n.value].copy(), w[:n.value].copy()


RESTART_FIELDS = ("UB", "VB", "TempB", "QMixB", "PsB", "UN", "VN", "TempN", "QMixN", "PsN")


def _restart_shape(name, im, jm, km, nq):
    if name.startswith("Ps"):
        return (jm, im)
    if name.startswith("QMix"):
        return (nq, km, jm, im)
    return (km, jm, im)


def restart_write(path, time, lon, lat, sig, fields, nq: int = 1):
    """Writes the restart state (RestartFileOutput): fields maps each of
    RESTART_FIELDS to a float64 array, C order (lon fastest)."""
    lon, lat, sig = (np.ascontiguousarray(x, dtype=np.float64) for x in (lon, lat, sig))
    im, jm, km = lon.size, lat.size, sig.size
    arrs = []
    for nm in RESTART_FIELDS:
        a = np.ascontiguousarray(fields[nm], dtype=np.float64)
        if a.shape != _restart_shape(nm, im, jm, km, nq):
            raise ValueError(f"{nm}: shape {a.shape}, expected {_restart_shape(nm, im, jm, km, nq)}")
        arrs.append(a)
    d = _RestartDesc(im, jm, km, nq, time, lon.ctypes.data, lat.ctypes.data, sig.ctypes.data)
    ptrs = (_vp * 10)(*[a.ctypes.data for a in arrs])
    check(lib().sht_restart_write(str(path).encode(), ctypes.byref(d), ptrs))


def restart_read(path, im: int, jm: int, km: int, nq: int = 1):
    """Reads the restart state (RestartFileGet): (time, {name: array})."""
    out = {nm: np.empty(_restart_shape(nm, im, jm, km, nq)) for nm in RESTART_FIELDS}
    d = _RestartDesc(im, jm, km, nq, 0.0, None, None, None)
    ptrs = (_vp * 10)(*[out[nm].ctypes.data for nm in RESTART_FIELDS])
    t = ctypes.c_double()
    check(lib().sht_restart_read(str(path).encode(), ctypes.byref(d), ctypes.byref(t), ptrs))
    return t.value, out


class Batch:
    """Ensemble batch (sht_batch_*): members of uneven truncation and level
    count, transformed by one graph launch.  members: list of
    (nmax, im, jm, nfield)."""

    def __init__(self, members, precision: int = PREC_FP64):
        arr = (_MemberDesc * len(members))(*[_MemberDesc(*m) for m in members])
        h = _vp()
        check(lib().sht_batch_create(arr, len(members), precision, ctypes.byref(h)))
        self._h = h
        self.members = [tuple(m) for m in members]

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().sht_batch_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def member_ptrs(self, j: int):
        """(spec_in, grid, spec_out, ld) device addresses of member j."""
        a, b, c, ld = _vp(), _vp(), _vp(), _ci()
        check(lib().sht_batch_member(self._h, j, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                     ctypes.byref(ld)))
        return a.value, b.value, c.value, ld.value

    def run(self, direction: int = DIR_ROUNDTRIP, stream=None):
        check(lib().sht_batch_run(self._h, direction, _stream_handle(stream)))

    def autotune(self, reps: int = 10, cache_path: str | None = None, shapes_only: bool = False):
        """Tune every group's Legendre shapes (and, for FP64 groups unless
        shapes_only, the split shape's wavenumber classes) at its field count;
        recapture on the next run."""
        check(lib().sht_batch_autotune_flags(self._h, reps, cache_path.encode() if cache_path else None,
                                             TUNE_SHAPES_ONLY if shapes_only else 0))

    def run_eager(self, direction: int = DIR_ROUNDTRIP, stream=None):
        check(lib().sht_batch_run_eager(self._h, direction, _stream_handle(stream)))


class Remap:
    """Coupler remap (sht_remap_*): recv[d, r] = sum over triplets with recv
    index r of send[d, s] * coef, in triplet order; 0-based indices."""

    def __init__(self, send_index, recv_index, coef, nsend: int, nrecv: int):
        si = np.ascontiguousarray(send_index, dtype=np.int32)
        ri = np.ascontiguousarray(recv_index, dtype=np.int32)
        cf = np.ascontiguousarray(coef, dtype=np.float64)
        h = ctypes.c_void_p()
        check(lib().sht_remap_create(si.ctypes.data_as(ctypes.POINTER(_ci)), ri.ctypes.data_as(ctypes.POINTER(_ci)),
                                     _host(cf), len(cf), nsend, nrecv, ctypes.byref(h)))
        self._h = h
        self.nsend, self.nrecv = nsend, nrecv

    def apply(self, send, recv, ndata: int, stream=None):
        check(lib().sht_remap_apply_dev(self._h, _ptr(send), _ptr(recv), ndata, _stream_handle(stream)))

    def close(self):
        if self._h:
            lib().sht_remap_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request server (sht_server_*).  grids: list of (nmax, im, jm)."""

    def __init__(self, grids, slots: int = 2, max_fields: int = 128, precision: int = PREC_FP64,
                 mem_budget: int = 8 << 30):
        d = _ServerDesc()
        d.ngrid = len(grids)
        for i, (n, im, jm) in enumerate(grids):
            d.nmax[i], d.im[i], d.jm[i] = n, im, jm
        d.slots, d.max_fields, d.precision, d.mem_budget = slots, max_fields, precision, mem_budget
        h = _vp()
        check(lib().sht_server_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().sht_server_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, nmax: int, nfield: int, direction: int, inp: int, out: int) -> int:
        """Returns the ticket; raises ShtError (code SHT_ERR_BUSY when not admitted)."""
        r = _Request(nmax, nfield, direction, inp, out)
        t = ctypes.c_longlong()
        check(lib().sht_server_submit(self._h, ctypes.byref(r), ctypes.byref(t)))
        return t.value

    def poll(self, max_n: int = 4096):
        t = (ctypes.c_longlong * max_n)()
        lat = (ctypes.c_double * max_n)()
        n = _ci()
        check(lib().sht_server_poll(self._h, t, lat, max_n, ctypes.byref(n)))
        return [(t[i], lat[i]) for i in range(n.value)]

    def drain(self):
        check(lib().sht_server_drain(self._h))

    def stats(self):
        v = [ctypes.c_longlong() for _ in range(4)]
        check(lib().sht_server_stats(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("requests", "fields", "batches", "refused"), (x.value for x in v)))

    def failures(self):
        """(failed requests, reason of the last failure)"""
        n = ctypes.c_longlong()
        buf = ctypes.create_string_buffer(512)
        check(lib().sht_server_failures(self._h, ctypes.byref(n), buf, 512))
        return n.value, buf.value.decode()


def spec_to_real(spec_c: np.ndarray) -> np.ndarray:
    """complex (K, nf) -> float64 (K, nf, 2) canonical device layout."""
    return np.ascontiguousarray(np.stack([spec_c.real, spec_c.imag], axis=-1))


def real_to_spec(a: np.ndarray) -> np.ndarray:
    a = np.asarray(a)
    return a[..., 0] + 1j * a[..., 1]
