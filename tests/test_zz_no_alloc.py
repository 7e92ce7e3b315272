"""sht_spec_to_grid_dev_noalloc, the allocation-free spectral -> grid entry
point (static device scratch), through the C ABI: bitwise equal to
sht_spec_to_grid_dev on two calls in a row, the second with fewer fields (and
a smaller grid), in FP64 and split-BF16 x6 plans; sizes beyond the static
scratch are refused before anything runs."""
import numpy as np
import pytest

import shtamd
from shtamd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _run(plan, spec, nf, noalloc):
    s = torch.from_numpy(shtamd.spec_to_real(spec)).cuda()
    g = torch.full((nf, plan.jm, plan.im), np.nan, dtype=torch.float64, device="cuda")
    (plan.spec_to_grid_noalloc if noalloc else plan.spec_to_grid)(s, g, nf)
    torch.cuda.synchronize()
    return g.cpu().numpy()


@pytest.mark.parametrize("prec", [shtamd.PREC_FP64, shtamd.PREC_SPLIT_BF16X6])
def test_noalloc_matches_regular_entry_point(prec):
    calls = [(85, 256, 128, 97), (42, 128, 64, 13)]  # second call: smaller
    for N, im, jm, nf in calls:
        print(f"noalloc prec={prec} T{N} nf={nf}", flush=True)
        p = shtamd.Plan(N, im, jm, nf, precision=prec)
        spec = synth.spectra(N, nf, cfg=31)
        a = _run(p, spec, nf, noalloc=True)
        b = _run(p, spec, nf, noalloc=False)
        assert np.array_equal(a, b)
        p.close()
    # and again on one plan, large then small field count
    p = shtamd.Plan(85, 256, 128, 97, precision=prec)
    for nf in (97, 5):
        spec = synth.spectra(85, nf, cfg=32)
        assert np.array_equal(_run(p, spec, nf, True), _run(p, spec, nf, False))
    p.close()


def test_noalloc_refuses_oversize():
    # T170 on 512 x 256 in FP64: 128 pairs * 171 m * 4 values * 8 B per field = 700 KB;
    # 64 fields need 44.8 MB > 32 MiB
    p = shtamd.Plan(170, 512, 256, 64)
    s = torch.zeros((p.K, 64, 2), dtype=torch.float64, device="cuda")
    g = torch.zeros((64, 256, 512), dtype=torch.float64, device="cuda")
    with pytest.raises(shtamd.ShtError) as e:
        p.spec_to_grid_noalloc(s, g, 64)
    assert e.value.code == shtamd.SHT_ERR_INVALID and "static scratch" in str(e.value)
    # a size that fits still runs and equals the regular path
    spec = synth.spectra(170, 40, cfg=33)
    assert np.array_equal(_run(p, spec, 40, True), _run(p, spec, 40, False))
    p.close()
