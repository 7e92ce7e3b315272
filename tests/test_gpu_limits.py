"""The largest plans the library accepts (nmax 255, im 512, jm up to 512).

These sizes reach kernel instances no smaller plan does: the FP64 forward
with 8 row tiles of 16 degrees, the inverse with 16 tiles of 16 latitude
pairs, the staged kernels with a 256-pair LDS block, the mixed forward with
more than four 32-wide latitude slices (eight waves) and the mixed inverse
with four degree slices, and the staged FP64 path in two K parts at 128 pairs
with N = 255.  Every FP64 shape must meet the FP64 bar against the oracle
(1e-12) with gradients, every mixed shape its form's bound, and the shapes
that share an MFMA sequence must agree bitwise.
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


def inverse(plan, spec_c, nf, grad=False):
    s = dev(shtamd.spec_to_real(spec_c))
    outs = [torch.empty((nf, plan.jm, plan.im), dtype=torch.float64, device="cuda") for _ in range(3)]
    if grad:
        plan.spec_to_grad(s, *outs, nf)
    else:
        plan.spec_to_grid(s, outs[0], nf)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in (outs if grad else outs[:1])]


def forward(plan, grid, nf):
    s = torch.empty((plan.K, nf, 2), dtype=torch.float64, device="cuda")
    plan.grid_to_spec(dev(grid), s, nf)
    torch.cuda.synchronize()
    return shtamd.real_to_spec(s.cpu().numpy())


@pytest.mark.parametrize("N,im,jm,nf", [(255, 512, 512, 9), (255, 512, 256, 17)])
def test_largest_plans_every_fp64_shape(N, im, jm, nf):
    spec = synth.spectra(N, nf, cfg=21)
    ref = oracle.Plan(N, im, jm)
    refs = [ref.spec_to_grid(spec, kind=k) for k in range(3)]
    s_ref = ref.grid_to_spec(refs[0])
    p = shtamd.Plan(N, im, jm, nf, flags=shtamd.FLAG_GRAD)
    for sh in range(shtamd.NSHAPE):
        say("largest plan, FP64 shape", shtamd.SHAPES[sh], N, im, jm, nf)
        p.set_shapes(0, sh, sh)
        outs = inverse(p, spec, nf, grad=True)
        for kind, o in enumerate(outs):
            assert rel(o, refs[kind]) < TOL, (sh, kind)
        assert rel(inverse(p, spec, nf)[0], refs[0]) < TOL, sh
        assert rel(forward(p, refs[0], nf), s_ref) < TOL, sh
    p.close()


@pytest.mark.parametrize("prec,tol", [(shtamd.PREC_SPLIT_BF16X6, 1e-6), (shtamd.PREC_SCALED_FP16, 1e-6)])
@pytest.mark.parametrize("N,im,jm,nf", [(255, 512, 512, 9), (255, 512, 256, 17)])
def test_largest_plans_mixed_shapes(prec, tol, N, im, jm, nf):
    spec = synth.spectra(N, nf, cfg=22)
    ref = oracle.Plan(N, im, jm)
    g_ref = ref.spec_to_grid(spec)
    p = shtamd.Plan(N, im, jm, nf, precision=prec)
    out = []
    for sh in (shtamd.SHAPE_KSPLIT, shtamd.SHAPE_WAVE, shtamd.SHAPE_TILE):
        say("largest plan, mixed shape", prec, shtamd.SHAPES[sh], N, im, jm, nf)
        p.set_shapes(0, sh, sh)
        g = inverse(p, spec, nf)[0]
        s = forward(p, g_ref, nf)
        assert rel(g, g_ref) < tol, sh
        assert rel(s, spec) < 2 * tol, sh
        out.append((g, s))
    for g, s in out[1:]:
        assert np.array_equal(out[0][0], g) and np.array_equal(out[0][1], s)
    p.close()


def test_two_part_staged_at_n255():
    """128 latitude pairs and 200 fields: the staged FP64 kernels take the K
    range in two halves; per output tile the MFMA sequence is WAVE's, so the
    results are bitwise equal, and both meet the FP64 bar."""
    N, im, jm, nf = 255, 512, 256, 200
    say("two-part staged", N, im, jm, nf)
    spec = synth.spectra(N, nf, cfg=23)
    ref = oracle.Plan(N, im, jm)
    g_ref = ref.spec_to_grid(spec)
    p = shtamd.Plan(N, im, jm, nf)
    res = {}
    for sh in (shtamd.SHAPE_STAGED, shtamd.SHAPE_WAVE):
        p.set_shapes(0, sh, sh)
        g = inverse(p, spec, nf)[0]
        s = forward(p, g_ref, nf)
        assert rel(g, g_ref) < TOL, sh
        assert rel(s, spec) < TOL, sh
        res[sh] = (g, s)
    assert np.array_equal(res[shtamd.SHAPE_STAGED][0], res[shtamd.SHAPE_WAVE][0])
    assert np.array_equal(res[shtamd.SHAPE_STAGED][1], res[shtamd.SHAPE_WAVE][1])
    p.close()
