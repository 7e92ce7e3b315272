"""Mixed-precision Legendre forms against the FP64 oracle (through the C ABI).

Tolerances (relative L2 against the oracle on the same seeded inputs):
  SPLIT_BF16    (hi/lo bf16, 3 products):        5e-5   (split residual 2^-18 per operand)
  SCALED_FP16   (hi/lo fp16 scaled, 3 products): 1e-6
  SPLIT_BF16X6  (hi/mid/lo bf16, 6 products):    1e-6
The FFT, tables and cross-slice sums are FP64; the Fourier intermediate is FP32.
"""
import sys

import numpy as np
import pytest

import oracle
import shtamd
from shtamd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = {shtamd.PREC_SPLIT_BF16: 5e-5, shtamd.PREC_SCALED_FP16: 1e-6, shtamd.PREC_SPLIT_BF16X6: 1e-6}
NAME = {shtamd.PREC_SPLIT_BF16: "bf16x3", shtamd.PREC_SCALED_FP16: "fp16x3", shtamd.PREC_SPLIT_BF16X6: "bf16x6"}


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def run(plan, spec_c, nf):
    s = torch.from_numpy(shtamd.spec_to_real(spec_c)).cuda()
    g = torch.empty((nf, plan.jm, plan.im), dtype=torch.float64, device="cuda")
    s2 = torch.empty_like(s)
    plan.spec_to_grid(s, g, nf)
    plan.grid_to_spec(g, s2, nf)
    torch.cuda.synchronize()
    return g.cpu().numpy(), shtamd.real_to_spec(s2.cpu().numpy())


def forward(plan, grid, nf):
    g = torch.from_numpy(np.ascontiguousarray(grid)).cuda()
    s = torch.empty((plan.K, nf, 2), dtype=torch.float64, device="cuda")
    plan.grid_to_spec(g, s, nf)
    torch.cuda.synchronize()
    return shtamd.real_to_spec(s.cpu().numpy())


@pytest.mark.parametrize("prec", sorted(TOL))
@pytest.mark.parametrize("N,im,jm,nlev", [(21, 64, 32, 16), (42, 128, 64, 26), (85, 256, 128, 32), (170, 512, 256, 5)])
def test_mixed_vs_fp64_oracle(prec, N, im, jm, nlev):
    nf = 3 * nlev + 1
    spec = synth.state_spectra(N, nlev, cfg=2)
    ref = oracle.Plan(N, im, jm)
    g_ref = ref.spec_to_grid(spec)
    p = shtamd.Plan(N, im, jm, nf, precision=prec)
    g, s_rt = run(p, spec, nf)
    e_inv = rel(g, g_ref)
    e_fwd = rel(forward(p, g_ref, nf), ref.grid_to_spec(g_ref))
    e_rt = rel(s_rt, spec)
    print(f"{NAME[prec]} T{N} nf={nf}: inverse {e_inv:.2e} forward {e_fwd:.2e} round trip {e_rt:.2e}", flush=True)
    sys.stdout.flush()
    assert e_inv < TOL[prec] and e_fwd < TOL[prec] and e_rt < 2 * TOL[prec]
    p.close()


@pytest.mark.parametrize("prec", sorted(TOL))
def test_mixed_noisy_grid_forward(prec):
    N, im, jm, nlev = 42, 128, 64, 26
    ref = oracle.Plan(N, im, jm)
    mu, w, cl = oracle.gauss(jm)
    g = synth.temperature_grid(cl, im, nlev)
    p = shtamd.Plan(N, im, jm, nlev, precision=prec)
    e = rel(forward(p, g, nlev), ref.grid_to_spec(g))
    print(f"{NAME[prec]} noisy T42 forward {e:.2e}", flush=True)
    assert e < TOL[prec]
    p.close()


@pytest.mark.parametrize("prec", sorted(TOL))
def test_mixed_plan_gradients_stay_fp64(prec):
    N, im, jm, nf = 42, 128, 64, 12
    spec = synth.spectra(N, nf, cfg=7)
    ref = oracle.Plan(N, im, jm)
    p = shtamd.Plan(N, im, jm, nf, precision=prec, flags=shtamd.FLAG_GRAD)
    s = torch.from_numpy(shtamd.spec_to_real(spec)).cuda()
    outs = [torch.empty((nf, jm, im), dtype=torch.float64, device="cuda") for _ in range(2)]
    p.spec_to_grad(s, None, outs[0], outs[1], nf)
    torch.cuda.synchronize()
    assert rel(outs[0].cpu().numpy(), ref.spec_to_grid(spec, kind=1)) < 1e-12
    assert rel(outs[1].cpu().numpy(), ref.spec_to_grid(spec, kind=2)) < 1e-12
    p.close()


@pytest.mark.parametrize("prec", sorted(TOL))
def test_mixed_deterministic_and_ragged(prec):
    N, im, jm = 85, 256, 128
    ref = oracle.Plan(N, im, jm)
    for nf in (1, 3, 17):
        spec = synth.spectra(N, nf, cfg=11)
        p = shtamd.Plan(N, im, jm, 20, precision=prec)
        g1, s1 = run(p, spec, nf)
        g2, s2 = run(p, spec, nf)
        assert np.array_equal(g1, g2) and np.array_equal(s1, s2)
        assert rel(g1, ref.spec_to_grid(spec)) < TOL[prec]
        p.close()
