"""CPU-side checks of the C ABI: the library loads, exports every entry point
include/sht.h declares, and rejects bad descriptors before touching a GPU."""
import ctypes
import re

import pytest

import shtamd


def header_functions():
    text = open(shtamd.HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sht_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert set(header_functions()) == set(shtamd.SIGNATURES)


def test_library_exports_every_declared_symbol():
    L = shtamd.lib()
    for name in header_functions():
        assert hasattr(L, name), name
    assert L.sht_abi_version() == 1


@pytest.mark.parametrize("desc,code", [
    ((21, 64, 31, 4, 0, 0), shtamd.SHT_ERR_INVALID),      # odd jm
    ((21, 64, 20, 4, 0, 0), shtamd.SHT_ERR_INVALID),      # jm <= nmax
    ((21, 40, 32, 4, 0, 0), shtamd.SHT_ERR_INVALID),      # im < 2N+2
    ((21, 96, 32, 4, 0, 0), shtamd.SHT_ERR_UNSUPPORTED),  # im not a power of two
    ((21, 64, 32, 0, 0, 0), shtamd.SHT_ERR_INVALID),      # kmax
    ((21, 64, 32, 4, 7, 0), shtamd.SHT_ERR_INVALID),      # precision
    ((21, 64, 32, 4, 0, 8), shtamd.SHT_ERR_INVALID),      # flags
    ((-1, 64, 32, 4, 0, 0), shtamd.SHT_ERR_INVALID),      # nmax
])
def test_descriptor_validation_without_gpu(desc, code):
    d = shtamd._Desc(*desc)
    h = ctypes.c_void_p()
    rc = shtamd.lib().sht_plan_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == code
    assert not h.value
    assert shtamd.lib().sht_last_error().decode().startswith("sht_plan_create")


def test_spectral_index_helpers():
    L = shtamd.lib()
    assert L.sht_l_nm(0, 0) == 1
    assert L.sht_l_nm(1, -1) == 2 and L.sht_l_nm(1, 0) == 3 and L.sht_l_nm(1, 1) == 4
    n, m = ctypes.c_int(), ctypes.c_int()
    for nn in range(0, 30):
        for mm in range(-nn, nn + 1):
            l = L.sht_l_nm(nn, mm)
            assert L.sht_nm_l(l, ctypes.byref(n), ctypes.byref(m)) == 0
            assert (n.value, m.value) == (nn, mm)
    assert L.sht_l_nm(2, 3) == shtamd.SHT_ERR_INVALID
    # extremes: the largest n whose (n + 1)^2 fits an int, and the largest l
    assert L.sht_l_nm(46339, 46339) == 46340 ** 2
    assert L.sht_l_nm(46340, 0) == shtamd.SHT_ERR_INVALID
    for l in (46340 ** 2, 46340 ** 2 + 1, 2 ** 31 - 1):
        assert L.sht_nm_l(l, ctypes.byref(n), ctypes.byref(m)) == 0
        assert n.value * (n.value + 1) + m.value + 1 == l and abs(m.value) <= n.value


def test_null_plan_calls_fail_cleanly():
    L = shtamd.lib()
    assert L.sht_spec_to_grid_dev(None, None, None, 1, None) == shtamd.SHT_ERR_INVALID
    assert L.sht_xya_wa(None, None, None, 1) == shtamd.SHT_ERR_INVALID
    L.sht_plan_destroy(None)


def test_fortran_module_compiles_against_the_abi():
    """The amdflang build of wa_module and its drop-in program links against
    libsht_hip.so (run on the GPU in test_gpu_fortran.py)."""
    import os
    import shutil

    if not shutil.which("amdflang") and not os.path.exists("/opt/rocm/bin/amdflang"):
        pytest.skip("amdflang not available")
    libdir = os.path.dirname(shtamd.LIB_PATH)
    for mod in ("wa_module", "l_module"):
        assert os.path.exists(os.path.join(libdir, f"{mod}.mod"))
        assert os.path.exists(os.path.join(libdir, f"test_{mod}"))


def test_zonal_descriptor_validation_without_gpu():
    h = ctypes.c_void_p()
    L = shtamd.lib()
    for args in ((10, 31, 1), (10, 0, 1), (33, 32, 1), (-1, 32, 1), (10, 32, 0)):
        assert L.sht_zonal_create(*args, ctypes.byref(h)) == shtamd.SHT_ERR_INVALID, args
        assert not h.value
    assert L.sht_l_y(None, None, None, 1) == shtamd.SHT_ERR_INVALID
    L.sht_zonal_destroy(None)


# ---- gmapgen's cell edges and surface exchange grid (host utilities of the ABI)
import numpy as np  # noqa: E402

import oracle  # noqa: E402


@pytest.mark.parametrize("jm", [2, 32, 64, 128, 1024])
def test_lat_cell_edges_close_and_contain_nodes(jm):
    """tool/gmapgen/gmapgen_main.f90:364-368 (and EBM_main.f90:200-204): the
    edges start at -pi/2, close at +pi/2 because the weights sum to 2, each
    Gauss node lies inside its cell, and each cell's area sin(e_j) - sin(e_j-1)
    is the node's weight."""
    mu, w, cl = oracle.gauss(jm)
    e = shtamd.lat_cell_edges(w)
    assert e.shape == (jm + 1,) and e[0] == -np.pi / 2 and e[-1] == np.pi / 2
    lat = np.arctan2(mu, cl)
    assert np.all(lat > e[:-1]) and np.all(lat < e[1:])
    # the last interior edge agrees with the closing edge's sin to rounding
    assert abs(np.sin(e[-2]) + w[-1] - 1.0) < 1e-12
    assert np.abs(np.diff(np.sin(e))[:-1] - w[:-1]).max() < 1e-13
    # the oracle's form of the same edges (cumulative weights in mu)
    assert np.abs(np.sin(e[1:-1]) - (-1.0 + np.cumsum(w)[:-1])).max() < 1e-13


@pytest.mark.parametrize("ja,jb", [(32, 64), (64, 32), (32, 48), (128, 64), (16, 16)])
def test_exchange_grid(ja, jb):
    """gmapgen_main.f90:336-387: the merged grid's bands are the union of both
    grids' cells (sorted, zero-area bands dropped), weights sum to 2, each
    latitude is its band's midpoint; equal latitude counts give grid A."""
    mua, wa, cla = oracle.gauss(ja)
    mub, wb, clb = oracle.gauss(jb)
    la, lb = np.arctan2(mua, cla), np.arctan2(mub, clb)
    lat, w = shtamd.exchange_grid(la, wa, lb, wb)
    if ja == jb:
        assert np.array_equal(lat, la) and np.array_equal(w, wa)
        return
    ea, eb = shtamd.lat_cell_edges(wa), shtamd.lat_cell_edges(wb)
    merged = np.unique(np.round(np.concatenate([ea, eb]), 12))
    assert lat.size == merged.size - 1
    assert abs(w.sum() - 2.0) < 1e-12 and np.all(w > 1e-12)
    assert np.all(np.diff(lat) > 0)
    assert np.abs(lat - 0.5 * (merged[:-1] + merged[1:])).max() < 1e-11
    assert np.abs(w - np.diff(np.sin(merged))).max() < 1e-11
    # every band lies inside one cell of A and one of B
    for x in lat:
        assert np.searchsorted(ea, x) >= 1 and np.searchsorted(eb, x) >= 1
