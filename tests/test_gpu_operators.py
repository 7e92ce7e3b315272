"""GPU parity of the spectral operators beside the transform, through the C ABI:
the flux divergence (SPML wa_DivLon_xya / wa_DivLat_xya and their sum) and
the Laplacian powers (wa_Lapla_wa, wa_LaplaInv_wa, hyperdiffusion), against
the CPU oracle at the FP64 bar (relative L2 <= 1e-12), plus the identities
div(grad f) = Laplacian f and div(rotational flow) = 0 computed on the GPU
alone.
"""
import sys

import numpy as np
import pytest

import oracle
import shtamd
from shtamd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def gpu_div(plan, a, b, nf):
    s = torch.empty((plan.K, nf, 2), dtype=torch.float64, device="cuda")
    plan.grid_to_div(None if a is None else dev(a), None if b is None else dev(b), s, nf)
    torch.cuda.synchronize()
    return shtamd.real_to_spec(s.cpu().numpy())


def gpu_grads(plan, spec, nf):
    s = dev(shtamd.spec_to_real(spec))
    gl = torch.empty((nf, plan.jm, plan.im), dtype=torch.float64, device="cuda")
    gt = torch.empty_like(gl)
    plan.spec_to_grad(s, None, gl, gt, nf)
    torch.cuda.synchronize()
    return gl.cpu().numpy(), gt.cpu().numpy()


CASES = [(21, 64, 32, 49, shtamd.PREC_FP64), (42, 128, 64, 79, shtamd.PREC_FP64),
         (85, 256, 128, 97, shtamd.PREC_FP64), (42, 128, 64, 20, shtamd.PREC_SPLIT_BF16X6),
         (7, 16, 10, 5, shtamd.PREC_FP64),
         # FFT blocks that cover half or a quarter of an 8-field group
         (42, 512, 64, 11, shtamd.PREC_FP64), (42, 1024, 64, 5, shtamd.PREC_FP64)]


@pytest.mark.parametrize("N,im,jm,nf,prec", CASES)
def test_divergence_vs_oracle(N, im, jm, nf, prec):
    """Arbitrary (not band-limited) grid fields: the GPU and the oracle compute
    the same quadrature, so they must agree to FP64 rounding in every precision
    form (the divergence always runs in FP64)."""
    say("div", N, im, jm, nf, prec)
    rng = np.random.default_rng(N + nf)
    a = rng.standard_normal((nf, jm, im))
    b = rng.standard_normal((nf, jm, im))
    ref = oracle.Plan(N, im, jm)
    p = shtamd.Plan(N, im, jm, nf, precision=prec, flags=shtamd.FLAG_GRAD)
    for aa, bb in ((a, b), (a, None), (None, b)):
        got = gpu_div(p, aa, bb, nf)
        want = ref.grid_to_div(aa, bb)
        assert rel(got, want) < TOL, (aa is None, bb is None)
    p.close()


@pytest.mark.parametrize("N,im,jm,nf", [(42, 128, 64, 16), (85, 256, 128, 12)])
def test_divergence_identities_on_gpu(N, im, jm, nf):
    """GPU only: gradients then divergence give the Laplacian; a rotational flow
    has zero divergence and the Laplacian of its stream function as vorticity."""
    say("identities", N, nf)
    p = shtamd.Plan(N, im, jm, nf, flags=shtamd.FLAG_GRAD)
    f = synth.spectra(N, nf, cfg=11)
    gl, gt = gpu_grads(p, f, nf)
    lap = oracle.spec_lapla(f, N, 1)
    assert rel(gpu_div(p, gl, gt, nf), lap) < 1e-12
    u, v = -gt, gl
    assert np.abs(gpu_div(p, u, v, nf)).max() < 1e-12 * np.abs(lap).max()
    assert rel(gpu_div(p, v, -u, nf), lap) < 1e-12
    p.close()


@pytest.mark.parametrize("power", [-4, -2, -1, 0, 1, 2, 4])
def test_laplacian_powers_vs_oracle(power):
    N, im, jm, nf = 85, 256, 128, 9
    say("lapla", power)
    p = shtamd.Plan(N, im, jm, nf)
    spec = synth.spectra(N, nf, cfg=12)
    spec[0] = 1.0
    s = dev(shtamd.spec_to_real(spec))
    out = torch.empty_like(s)
    p.spec_lapla(s, out, power, nf)
    p.spec_lapla(s, s, power, nf)  # in place
    torch.cuda.synchronize()
    want = oracle.spec_lapla(spec, N, power)
    got = shtamd.real_to_spec(out.cpu().numpy())
    assert rel(got, want) < 1e-15
    assert torch.equal(out, s)
    p.close()


def test_host_dropins_spml_packing():
    """wa_DivLon_xya, wa_DivLat_xya, wa_Lapla_wa, wa_LaplaInv_wa on host arrays."""
    N, im, jm, nlev = 21, 64, 32, 16
    say("host operators")
    p = shtamd.Plan(N, im, jm, nlev, flags=shtamd.FLAG_GRAD)
    ref = oracle.Plan(N, im, jm)
    rng = np.random.default_rng(5)
    g = rng.standard_normal((nlev, jm, im))
    assert rel(p.wa_divlon_xya(g), oracle.spml_pack(ref.grid_to_div(g, None), N)) < TOL
    assert rel(p.wa_divlat_xya(g), oracle.spml_pack(ref.grid_to_div(None, g), N)) < TOL
    spec = synth.spectra(N, nlev, cfg=13)
    wa = oracle.spml_pack(spec, N)
    assert rel(p.wa_lapla_wa(wa), oracle.spml_pack(oracle.spec_lapla(spec, N, 1), N)) < 1e-15
    assert rel(p.wa_laplainv_wa(wa), oracle.spml_pack(oracle.spec_lapla(spec, N, -1), N)) < 1e-15
    p.close()


def test_operator_argument_checks():
    p = shtamd.Plan(21, 64, 32, 4)  # no SHT_FLAG_GRAD: the lat term has no table
    g = torch.zeros((4, 32, 64), dtype=torch.float64, device="cuda")
    s = torch.empty((p.K, 4, 2), dtype=torch.float64, device="cuda")
    with pytest.raises(shtamd.ShtError):
        p.grid_to_div(None, g, s, 4)
    with pytest.raises(shtamd.ShtError):
        p.grid_to_div(None, None, s, 4)
    with pytest.raises(shtamd.ShtError):
        p.spec_lapla(s, s, 5, 4)
    p.grid_to_div(g, None, s, 4)  # the lon term alone needs no derivative table
    torch.cuda.synchronize()
    assert torch.count_nonzero(s) == 0
    p.close()
