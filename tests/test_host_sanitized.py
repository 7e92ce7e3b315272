"""The host-only logic of libsht_hip.so (descriptor validation, SPML index
helpers, the shape-cache parser, gmapgen's exchange grid, the request
server's admission and batching queue, the restart file writer and reader)
built alone with AddressSanitizer and
UndefinedBehaviorSanitizer and driven by tests/host/host_check.cpp on the
CPU, several threads included; and the same under ThreadSanitizer for the
threaded parts (the admission queue, the parallel staging copies).  No GPU:
these sources use no HIP."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dennou-ccm_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_host_logic_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "host_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-pthread",
           f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "tests", "host", "host_check.cpp"), os.path.join(CSRC, "sht_host.cpp"),
           os.path.join(CSRC, "sht_restart.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "PASS" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_host_logic_under_tsan(tmp_path):
    exe = str(tmp_path / "host_check_tsan")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread",
           f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "tests", "host", "host_check.cpp"), os.path.join(CSRC, "sht_host.cpp"),
           os.path.join(CSRC, "sht_restart.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "PASS" in r.stdout and "WARNING: ThreadSanitizer" not in r.stderr
