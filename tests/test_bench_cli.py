"""bench.py on the CPU: the single-GPU refusal and the committed profile
summaries the bench line reads its measured HBM traffic and matrix-core
utilisation from.  No GPU."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_profile_summaries_parse():
    """Every kernel the roofline and the MFMA fields name has a row in the
    committed summaries, with traffic near its algorithmic bytes."""
    alg = {"fft_fwd_kernel<256,float": 33.97e6, "fft_fwd_kernel<256,double": 42.47e6,
           "leg_inv_mixed_kernel": 17.05e6}
    for form, prefix in (("bf16x6", "fft_fwd_kernel<256,float"), ("fp64", "fft_fwd_kernel<256,double"),
                         ("bf16x6", "leg_inv_mixed_kernel")):
        t = bench.traffic_from_profile(form, prefix)
        assert isinstance(t, int) and 0.9 * alg[prefix] < t < 1.2 * alg[prefix], (form, prefix, t)
    for form in ("bf16x6", "fp64"):
        m = bench.mfma_from_profile(form)
        assert m and set(m) == {"leg_inv", "leg_fwd"}, (form, m)
        assert all(0.0 < v["util"] < 1.0 for v in m.values()), (form, m)
