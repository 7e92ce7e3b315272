"""The Fortran drop-in modules (ISO_C_BINDING over the C ABI), compiled with
amdflang by the build, run as programs on the GPU: wa_module (SPML names and
array shapes, closed-form harmonics and gradients, round trip, global mean,
divergence, Laplacian) and l_module at the energy-balance model's
l_initial(1024, 1024) (calc_Hn_exact, round trip)."""
import os
import subprocess

import pytest

import shtamd

pytestmark = pytest.mark.gpu
LIBDIR = os.path.dirname(shtamd.LIB_PATH)


@pytest.mark.parametrize("module", ["wa_module", "l_module"])
def test_fortran_dropin(module):
    exe = os.path.join(LIBDIR, f"test_{module}")
    if not os.path.exists(exe):
        pytest.skip("Fortran drop-in not built (no amdflang)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "PASS" in r.stdout
