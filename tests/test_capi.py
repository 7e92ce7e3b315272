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
    assert os.path.exists(os.path.join(libdir, "wa_module.mod"))
    assert os.path.exists(os.path.join(libdir, "test_wa_module"))
