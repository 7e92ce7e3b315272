"""GPU parity: the HIP library through its C ABI against the CPU FP64 oracle.

Bar (FP64 form): relative L2 <= 1e-12 against the oracle on the same seeded
inputs; the spectral -> grid -> spectral round trip within 1e-12.  The
mpmath golden fixtures are checked directly as well.
"""
import os
import sys

import numpy as np
import pytest

import oracle
import shtamd
from shtamd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-12


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_inverse(plan, spec_c, nf):
    s = dev(shtamd.spec_to_real(spec_c))
    g = torch.empty((nf, plan.jm, plan.im), dtype=torch.float64, device="cuda")
    plan.spec_to_grid(s, g, nf)
    torch.cuda.synchronize()
    return g.cpu().numpy()


def gpu_forward(plan, grid, nf):
    g = dev(grid)
    s = torch.empty((plan.K, nf, 2), dtype=torch.float64, device="cuda")
    plan.grid_to_spec(g, s, nf)
    torch.cuda.synchronize()
    return shtamd.real_to_spec(s.cpu().numpy())


CONFIGS = [  # (N, im, jm, nlev) -> 3*nlev+1 fields
    (21, 64, 32, 16),
    (42, 128, 64, 26),
    (85, 256, 128, 32),
    (170, 512, 256, 4),
]


def test_plan_grid_matches_oracle():
    for (N, im, jm, _) in CONFIGS:
        say("grid", N, im, jm)
        p = shtamd.Plan(N, im, jm, 1)
        lon, lat, lw, yw = p.grid()
        mu, w, cl = oracle.gauss(jm)
        assert np.abs(np.sin(lat) - mu).max() < 2e-15
        # both sides are within ~1e-14 relative of mpmath (GPU and host libm
        # differ in the last bits of sin/cos)
        assert np.all(np.abs(yw - w) < 2e-14 * w)
        assert np.abs(np.cos(lat) - cl).max() < 1e-14
        assert abs(lw.sum() - 2 * np.pi) < 1e-13
        p.close()


@pytest.mark.parametrize("N,im,jm,nlev", CONFIGS)
def test_fp64_inverse_forward_vs_oracle(N, im, jm, nlev):
    nf = 3 * nlev + 1
    say("fp64", N, im, jm, nf)
    spec = synth.state_spectra(N, nlev, cfg=1)
    ref = oracle.Plan(N, im, jm)
    p = shtamd.Plan(N, im, jm, nf)
    g_gpu = gpu_inverse(p, spec, nf)
    g_ref = ref.spec_to_grid(spec)
    assert rel(g_gpu, g_ref) < TOL
    s_gpu = gpu_forward(p, g_ref, nf)
    s_ref = ref.grid_to_spec(g_ref)
    assert rel(s_gpu, s_ref) < TOL
    # round trip on the GPU alone
    assert rel(gpu_forward(p, g_gpu, nf), spec) < TOL
    assert np.all(s_gpu[: N + 1].imag == 0.0)  # m = 0 imaginary parts are exactly zero
    p.close()


def test_noisy_grid_forward_and_idempotence():
    N, im, jm, nlev = 42, 128, 64, 26
    say("noisy T42", nlev)
    ref = oracle.Plan(N, im, jm)
    mu, w, cl = oracle.gauss(jm)
    g = synth.temperature_grid(cl, im, nlev)
    p = shtamd.Plan(N, im, jm, nlev)
    s_gpu = gpu_forward(p, g, nlev)
    assert rel(s_gpu, ref.grid_to_spec(g)) < TOL
    s2 = gpu_forward(p, gpu_inverse(p, s_gpu, nlev), nlev)
    assert rel(s2, s_gpu) < TOL
    p.close()


@pytest.mark.parametrize("N,im,jm,nlev", [(42, 128, 64, 26), (21, 64, 32, 16), (85, 256, 128, 8)])
def test_gradients_vs_oracle(N, im, jm, nlev):
    nf = 3 * nlev + 1
    say("grad", N, nf)
    spec = synth.state_spectra(N, nlev, cfg=1)
    ref = oracle.Plan(N, im, jm)
    p = shtamd.Plan(N, im, jm, nf, flags=shtamd.FLAG_GRAD)
    s = dev(shtamd.spec_to_real(spec))
    outs = [torch.empty((nf, jm, im), dtype=torch.float64, device="cuda") for _ in range(3)]
    p.spec_to_grad(s, outs[0], outs[1], outs[2], nf)
    torch.cuda.synchronize()
    for kind, o in enumerate(outs):
        assert rel(o.cpu().numpy(), ref.spec_to_grid(spec, kind=kind)) < TOL, kind
    # each output alone gives the same numbers
    only = torch.empty((nf, jm, im), dtype=torch.float64, device="cuda")
    p.spec_to_grad(s, None, None, only, nf)
    torch.cuda.synchronize()
    assert torch.equal(only, outs[2])
    p.close()


@pytest.mark.parametrize("case", ["golden_T5.npz", "golden_T10.npz"])
def test_mpmath_golden_on_gpu(case):
    g = np.load(os.path.join(GOLD, case))
    N, im, jm = int(g["nmax"]), int(g["im"]), int(g["jm"])
    nf = g["spec"].shape[1]
    say("golden", case)
    p = shtamd.Plan(N, im, jm, nf, flags=shtamd.FLAG_GRAD)
    s = dev(shtamd.spec_to_real(g["spec"]))
    outs = [torch.empty((nf, jm, im), dtype=torch.float64, device="cuda") for _ in range(3)]
    p.spec_to_grad(s, *outs, nf)
    torch.cuda.synchronize()
    assert rel(outs[0].cpu().numpy(), g["grid"]) < 1e-14
    assert rel(outs[1].cpu().numpy(), g["gradlon"]) < 1e-13
    assert rel(outs[2].cpu().numpy(), g["gradlat"]) < 1e-13
    assert rel(gpu_forward(p, g["noisy"], nf), g["fwd"]) < 1e-14
    p.close()


@pytest.mark.parametrize("nf", [1, 2, 7, 8, 9, 33])
def test_ragged_field_counts(nf):
    """Partial 16-column tiles and partial FFT field groups."""
    N, im, jm = 42, 128, 64
    say("ragged", nf)
    spec = synth.spectra(N, nf, cfg=3)
    ref = oracle.Plan(N, im, jm)
    p = shtamd.Plan(N, im, jm, 40)
    g = gpu_inverse(p, spec, nf)
    assert rel(g, ref.spec_to_grid(spec)) < TOL
    assert rel(gpu_forward(p, g, nf), spec) < TOL
    p.close()


@pytest.mark.parametrize("N,im,jm", [(0, 16, 2), (1, 16, 4), (5, 16, 8), (7, 16, 10), (31, 64, 48), (63, 128, 96)])
def test_small_and_unpadded_grids(N, im, jm):
    """Tiny truncations and jm/2 not a multiple of 16 (padded table rows)."""
    nf = 5
    say("small", N, im, jm)
    spec = synth.spectra(N, nf, cfg=4)
    spec[0] = np.arange(1.0, nf + 1.0)  # nonzero global means (the synthetic s_0^0 is 0)
    ref = oracle.Plan(N, im, jm)
    p = shtamd.Plan(N, im, jm, nf, flags=shtamd.FLAG_GRAD)
    g = gpu_inverse(p, spec, nf)
    assert rel(g, ref.spec_to_grid(spec)) < TOL
    assert rel(gpu_forward(p, g, nf), ref.grid_to_spec(g)) < TOL
    p.close()


def test_host_dropin_spml_packing():
    """wa_module host entry points (SPML packing) against the oracle."""
    N, im, jm, nlev = 21, 64, 32, 16
    say("host dropin")
    spec = synth.spectra(N, nlev, cfg=6)
    wa = oracle.spml_pack(spec, N)
    ref = oracle.Plan(N, im, jm)
    p = shtamd.Plan(N, im, jm, nlev, flags=shtamd.FLAG_GRAD)
    xya = p.xya_wa(wa)
    assert rel(xya, ref.spec_to_grid(spec)) < TOL
    assert rel(p.xya_gradlon_wa(wa), ref.spec_to_grid(spec, kind=1)) < TOL
    assert rel(p.xya_gradlat_wa(wa), ref.spec_to_grid(spec, kind=2)) < TOL
    assert rel(p.wa_xya(xya), wa) < TOL
    # into the caller's own arrays (out=), bitwise the same results
    xya2, wa2 = np.full_like(xya, 7.0), np.full_like(wa, 7.0)
    assert p.xya_wa(wa, out=xya2) is xya2 and np.array_equal(xya2, xya)
    assert p.wa_xya(xya, out=wa2) is wa2 and np.array_equal(wa2, p.wa_xya(xya))
    # device layout conversions
    wa_d = dev(wa)
    sc = torch.empty((p.K, nlev, 2), dtype=torch.float64, device="cuda")
    p.spml_to_canon(wa_d, sc, nlev)
    back = torch.empty_like(wa_d)
    p.canon_to_spml(sc, back, nlev)
    torch.cuda.synchronize()
    # the sqrt2 scalings round: one ulp either way
    assert np.abs(shtamd.real_to_spec(sc.cpu().numpy()) - spec).max() < 4e-16 * np.abs(spec).max()
    assert np.abs(back.cpu().numpy() - wa).max() < 4e-16 * np.abs(wa).max()
    p.close()


def test_avr_lonlat_and_parseval():
    N, im, jm, nf = 85, 256, 128, 10
    say("avr")
    spec = synth.spectra(N, nf, cfg=8)
    spec[0] = np.arange(nf) + 1.0
    ref = oracle.Plan(N, im, jm)
    p = shtamd.Plan(N, im, jm, nf)
    g = dev(gpu_inverse(p, spec, nf))
    avg = torch.empty(nf, dtype=torch.float64, device="cuda")
    p.avr_lonlat(g, avg, nf)
    torch.cuda.synchronize()
    assert np.abs(avg.cpu().numpy() - ref.avr_lonlat(g.cpu().numpy())).max() < 1e-14
    assert np.abs(avg.cpu().numpy() - spec[0].real).max() < 1e-13
    p.avr_lonlat((g * g).contiguous(), avg, nf)
    torch.cuda.synchronize()
    ns, ms = synth.coef_index(N)
    par = (np.abs(spec) ** 2 * np.where(ms == 0, 1, 2)[:, None]).sum(0)
    assert np.abs(avg.cpu().numpy() - par).max() < 1e-12 * par.max()
    p.close()


def test_full_size_t170_properties():
    """T170, 145 fields (L48 state): round trip and Parseval (size-independent)."""
    N, im, jm, nlev = 170, 512, 256, 48
    nf = 3 * nlev + 1
    say("T170 full", nf)
    spec = synth.state_spectra(N, nlev, cfg=3)
    p = shtamd.Plan(N, im, jm, nf)
    g = gpu_inverse(p, spec, nf)
    assert rel(gpu_forward(p, g, nf), spec) < TOL
    ref = oracle.Plan(N, im, jm)
    # spot-check four fields against the oracle
    sel = [0, 47, 96, 144]
    assert rel(g[sel], ref.spec_to_grid(np.ascontiguousarray(spec[:, sel]))) < TOL
    p.close()


def test_error_codes():
    p = shtamd.Plan(21, 64, 32, 4)
    s = torch.zeros((p.K, 5, 2), dtype=torch.float64, device="cuda")
    g = torch.zeros((5, 32, 64), dtype=torch.float64, device="cuda")
    with pytest.raises(shtamd.ShtError) as e:
        p.spec_to_grid(s, g, 5)  # nfield > kmax
    assert e.value.code == shtamd.SHT_ERR_INVALID
    with pytest.raises(shtamd.ShtError) as e:
        p.spec_to_grad(s, None, None, g, 2)  # GradLat without SHT_FLAG_GRAD
    assert e.value.code == shtamd.SHT_ERR_INVALID
    # the binding refuses wrong dtypes, short buffers and mis-shaped host arrays
    # before any pointer reaches the library
    with pytest.raises(ValueError):
        p.spec_to_grid(s.float(), g, 2)
    with pytest.raises(ValueError):
        p.spec_to_grid(s, g[:1], 2)
    with pytest.raises(ValueError):
        p.grid_to_spec(g.cpu(), s, 2)
    with pytest.raises(ValueError):
        p.wa_lapla_wa(np.zeros(22 * 22))  # 1-D: no level axis
    with pytest.raises(ValueError):
        p.xya_wa(np.zeros((2, 22 * 22 + 1)))
    wa = np.zeros((2, 22 * 22))
    for bad in (np.zeros((2, 32, 63)), np.zeros((2, 32, 64), dtype=np.float32),
                np.zeros((2, 64, 32)).transpose(0, 2, 1), np.zeros((3, 32, 64))):
        with pytest.raises(ValueError):
            p.xya_wa(wa, out=bad)  # wrong shape, dtype, layout or level count
    p.close()


@pytest.mark.parametrize("im", [1024, 2048])
@pytest.mark.parametrize("prec,tol", [(shtamd.PREC_FP64, TOL), (shtamd.PREC_SPLIT_BF16X6, 1e-6)])
def test_long_rings(im, prec, tol):
    """Ring lengths 1024 and 2048 (three FFT passes, G = 2 and 1 fields per
    FFT block, a quarter and an eighth of the Fourier intermediate's 8-field
    groups; with 7 fields the one group is partly filled and the last
    blocks hold no field).  Gradients too (the derivative table and the
    GradLon/GradLat FFT modes)."""
    N, jm, nf = 42, 64, 7
    say("long rings", im, prec)
    spec = synth.spectra(N, nf, cfg=16)
    spec[0] = np.arange(1.0, nf + 1.0)
    ref = oracle.Plan(N, im, jm)
    p = shtamd.Plan(N, im, jm, nf, precision=prec, flags=shtamd.FLAG_GRAD)
    g = gpu_inverse(p, spec, nf)
    assert rel(g, ref.spec_to_grid(spec)) < tol
    assert rel(gpu_forward(p, g, nf), ref.grid_to_spec(g)) < 2 * tol
    s = dev(shtamd.spec_to_real(spec))
    outs = [torch.empty((nf, jm, im), dtype=torch.float64, device="cuda") for _ in range(3)]
    p.spec_to_grad(s, *outs, nf)
    torch.cuda.synchronize()
    for kind, o in enumerate(outs):
        assert rel(o.cpu().numpy(), ref.spec_to_grid(spec, kind=kind)) < tol
    p.close()


@pytest.mark.parametrize("im", [512, 1024, 2048])
@pytest.mark.parametrize("prec,tol", [(shtamd.PREC_FP64, TOL), (shtamd.PREC_SPLIT_BF16X6, 1e-6)])
def test_partial_field_groups_on_long_rings(im, prec, tol):
    """On grids of 512 longitudes and more an FFT block holds 4, 2 or 1 of
    the 8 fields of a Fourier-intermediate group; field counts that leave
    the last group partly filled (blocks with no field at all, blocks that
    share a group) must give each field the same result as the oracle, in
    both directions, and a field's result must not depend on the batch."""
    N, jm = 42, 64
    ref = oracle.Plan(N, im, jm)
    spec_all = synth.spectra(N, 17, cfg=21)
    p = shtamd.Plan(N, im, jm, 17, precision=prec)
    first = None
    for nf in (1, 3, 5, 9, 17):
        say("partial groups", im, prec, nf)
        spec = spec_all[:, :nf]
        g = gpu_inverse(p, spec, nf)
        assert rel(g, ref.spec_to_grid(spec)) < tol, nf
        s = gpu_forward(p, g, nf)
        assert rel(s, ref.grid_to_spec(g)) < 2 * tol, nf
        if first is None:
            first = (g[0], s[:, 0])
        else:  # field 0 alone and in every batch: the same numbers
            assert np.array_equal(g[0], first[0]) and np.array_equal(s[:, 0], first[1]), nf
    p.close()


def test_library_first_then_torch():
    """A fresh process that creates a plan before any PyTorch GPU work, then
    uses PyTorch on the GPU: the binding initialises PyTorch's device before
    it loads the library (PyTorch ships its own HIP runtime)."""
    import subprocess
    import sys as _sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys; sys.path[:0] = [%r, %r]\n"
            "import numpy as np, shtamd\n"
            "p = shtamd.Plan(21, 64, 32, 4)\n"
            "g = p.xya_wa(np.zeros((4, 22 * 22)))\n"
            "import torch\n"
            "t = torch.ones(8, device='cuda')\n"
            "print('ok', float(t.sum()), float(np.abs(g).max()))\n") % (root, os.path.join(root, "dennou-ccm_amd", "py"))
    say("library first, then torch")
    r = subprocess.run([_sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and "ok 8.0 0.0" in r.stdout, (r.stdout, r.stderr[-2000:])
