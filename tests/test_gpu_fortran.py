"""The Fortran wa_module drop-in (ISO_C_BINDING over the C ABI), compiled with
amdflang by the build, run as a program on the GPU: SPML names and array
shapes, closed-form harmonics and gradients, round trip, global mean."""
import os
import subprocess

import pytest

import shtamd

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(shtamd.LIB_PATH), "test_wa_module")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran drop-in not built (no amdflang)")
def test_fortran_wa_module_dropin():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "PASS" in r.stdout
