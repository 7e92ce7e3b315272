"""The fused inverse (SHT_SHAPE_FUSED) through the C ABI.

One persistent launch runs the mixed-precision Legendre items of the KSPLIT
shape and the ring FFTs, handing the Fourier intermediate over through
device-coherent stores and per-column-tile completion counts.  It computes
the same operations in the same order as KSPLIT followed by the FFT kernel,
so its grids must equal theirs bitwise, and meet the form's bound against
the CPU oracle.  Consecutive calls with different inputs and field counts
check that the queue and completion counters come back to zero after every
call (a stale count would let an FFT tile read an unfinished column tile);
graph replays must equal the eager run each time.
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


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def inverse(plan, spec_c, nf, noalloc=False):
    s = dev(shtamd.spec_to_real(spec_c))
    g = torch.full((nf, plan.jm, plan.im), 7.0, dtype=torch.float64, device="cuda")
    (plan.spec_to_grid_noalloc if noalloc else plan.spec_to_grid)(s, g, nf)
    torch.cuda.synchronize()
    return g.cpu().numpy()


# (N, im, jm, nf): every instantiated grid (im 64 to 512), one to three K
# slices, a latitude count that is no multiple of 16, field counts giving
# ragged column tiles and field groups, a single field
CASES = [
    (85, 256, 128, 97),
    (21, 64, 32, 49),
    (42, 128, 64, 9),
    (42, 128, 64, 1),
    (31, 64, 48, 17),
    (63, 256, 96, 33),
    (127, 256, 128, 5),
    (170, 512, 256, 13),
    (106, 512, 160, 40),
]


@pytest.mark.parametrize("prec", sorted(TOL))
@pytest.mark.parametrize("N,im,jm,nf", CASES)
def test_fused_equals_ksplit(prec, N, im, jm, nf):
    say("fused vs ksplit", prec, N, im, jm, nf)
    spec = synth.spectra(N, nf, cfg=31)
    p = shtamd.Plan(N, im, jm, nf, precision=prec)
    p.set_shapes(0, shtamd.SHAPE_KSPLIT, shtamd.SHAPE_KSPLIT)
    ref = inverse(p, spec, nf)
    p.set_shapes(0, shtamd.SHAPE_FUSED, shtamd.SHAPE_KSPLIT)
    assert p.get_shapes(nf) == (shtamd.SHAPE_FUSED, shtamd.SHAPE_KSPLIT)
    got = inverse(p, spec, nf)
    assert np.array_equal(got, ref)
    assert rel(got, oracle.Plan(N, im, jm).spec_to_grid(spec)) < TOL[prec]
    p.close()


def test_fused_consecutive_calls():
    """Different inputs and field counts back to back, fused and unfused
    interleaved, and the allocation-free entry point: every grid equals the
    unfused one for the same input."""
    N, im, jm, kmax = 85, 256, 128, 120
    prec = shtamd.PREC_SPLIT_BF16X6
    p = shtamd.Plan(N, im, jm, kmax, precision=prec)
    q = shtamd.Plan(N, im, jm, kmax, precision=prec)
    q.set_shapes(0, shtamd.SHAPE_KSPLIT, shtamd.SHAPE_KSPLIT)
    p.set_shapes(0, shtamd.SHAPE_FUSED, shtamd.SHAPE_KSPLIT)
    for i, nf in enumerate([97, 1, 120, 8, 97, 33, 2, 97]):
        say("consecutive", i, nf)
        spec = synth.spectra(N, nf, cfg=40 + i)
        ref = inverse(q, spec, nf)
        assert np.array_equal(inverse(p, spec, nf), ref), (i, nf)
        if i % 3 == 1:
            assert np.array_equal(inverse(p, spec, nf, noalloc=True), ref), (i, nf)
    p.close()
    q.close()


def test_fused_graph_replays():
    """A captured fused call replays correctly several times, with new input
    and the program's own work between the replays."""
    N, im, jm, nf = 42, 128, 64, 26
    p = shtamd.Plan(N, im, jm, nf, precision=shtamd.PREC_SPLIT_BF16X6)
    q = shtamd.Plan(N, im, jm, nf, precision=shtamd.PREC_SPLIT_BF16X6)
    p.set_shapes(0, shtamd.SHAPE_FUSED, shtamd.SHAPE_KSPLIT)
    q.set_shapes(0, shtamd.SHAPE_KSPLIT, shtamd.SHAPE_KSPLIT)
    s = torch.zeros((p.K, nf, 2), dtype=torch.float64, device="cuda")
    g = torch.zeros((nf, jm, im), dtype=torch.float64, device="cuda")
    cap = torch.cuda.Stream()
    with torch.cuda.stream(cap):
        p.spec_to_grid(s, g, nf, cap)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=cap):
        p.spec_to_grid(s, g, nf, cap)
    for r in range(4):
        say("graph replay", r)
        spec = synth.spectra(N, nf, cfg=60 + r)
        s.copy_(dev(shtamd.spec_to_real(spec)))
        g.fill_(-1.0)
        graph.replay()
        torch.cuda.synchronize()
        assert np.array_equal(g.cpu().numpy(), inverse(q, spec, nf)), r
    p.close()
    q.close()


def test_fused_availability():
    """FUSED is an inverse shape of mixed plans with 64 <= im <= 512 and
    nmax < 192; elsewhere set_shapes refuses it."""
    m = shtamd.Plan(21, 64, 32, 8, precision=shtamd.PREC_SPLIT_BF16X6)
    with pytest.raises(shtamd.ShtError):
        m.set_shapes(0, shtamd.SHAPE_KSPLIT, shtamd.SHAPE_FUSED)  # forward
    m.set_shapes(0, shtamd.SHAPE_FUSED, shtamd.SHAPE_KSPLIT)
    m.close()
    for args in [dict(N=21, im=64, jm=32, prec=shtamd.PREC_FP64),
                 dict(N=7, im=32, jm=16, prec=shtamd.PREC_SPLIT_BF16X6),
                 dict(N=255, im=1024, jm=512, prec=shtamd.PREC_SPLIT_BF16X6),
                 dict(N=200, im=512, jm=256, prec=shtamd.PREC_SPLIT_BF16X6)]:
        say("not available", args)
        p = shtamd.Plan(args["N"], args["im"], args["jm"], 4, precision=args["prec"])
        with pytest.raises(shtamd.ShtError) as e:
            p.set_shapes(0, shtamd.SHAPE_FUSED, shtamd.SHAPE_AUTO)
        assert e.value.code == shtamd.SHT_ERR_UNSUPPORTED
        p.close()


def test_fused_in_autotuner_and_gradients():
    """The autotuner times FUSED with the other inverse shapes at T85; with
    FUSED set, gradients (FP64 Legendre step) are unaffected."""
    N, im, jm, nf = 85, 256, 128, 97
    p = shtamd.Plan(N, im, jm, nf, precision=shtamd.PREC_SPLIT_BF16X6, flags=shtamd.FLAG_GRAD)
    (inv, fwd), (ti, tf) = p.autotune(nf, reps=5, shapes_only=True)
    say("autotune", inv, fwd, ti, tf)
    assert inv in (shtamd.SHAPE_KSPLIT, shtamd.SHAPE_WAVE, shtamd.SHAPE_TILE, shtamd.SHAPE_STAGED,
                   shtamd.SHAPE_FUSED)
    p.set_shapes(0, shtamd.SHAPE_FUSED, shtamd.SHAPE_KSPLIT)
    spec = synth.spectra(N, nf, cfg=70)
    s = dev(shtamd.spec_to_real(spec))
    outs = [torch.empty((nf, jm, im), dtype=torch.float64, device="cuda") for _ in range(3)]
    p.spec_to_grad(s, *outs, nf)
    torch.cuda.synchronize()
    ref = oracle.Plan(N, im, jm)
    for kind, o in enumerate(outs):
        assert rel(o.cpu().numpy(), ref.spec_to_grid(spec, kind=kind)) < 1e-12, kind
    p.close()


def test_fused_wait_timeout_is_reported(monkeypatch):
    """A fused call whose column tiles never complete (forced from the host:
    SHT_INJECT_FUSED_TIMEOUT asks for one completion more than the items make
    and bounds each wait at 10 us) must still end: every FFT tile gives up,
    sets its grid rows to NaN and counts the failure.  The counters are left
    clean for the next call, and a plan created without the injection is
    unaffected."""
    N, im, jm, nf = 42, 128, 64, 26
    prec = shtamd.PREC_SPLIT_BF16X6
    spec = synth.spectra(N, nf, cfg=80)
    monkeypatch.setenv("SHT_INJECT_FUSED_TIMEOUT", "1")
    p = shtamd.Plan(N, im, jm, nf, precision=prec)
    monkeypatch.delenv("SHT_INJECT_FUSED_TIMEOUT")
    p.set_shapes(0, shtamd.SHAPE_FUSED, shtamd.SHAPE_KSPLIT)
    assert p.fused_failures() == 0
    for call in range(2):
        say("forced timeout", call)
        g = inverse(p, spec, nf)
        assert np.isnan(g).all(), call
        # one give-up per FFT tile: latitude pairs x field groups of 16 fields
        assert p.fused_failures() == (call + 1) * (jm // 2) * ((nf + 15) // 16)
    q = shtamd.Plan(N, im, jm, nf, precision=prec)
    q.set_shapes(0, shtamd.SHAPE_FUSED, shtamd.SHAPE_KSPLIT)
    g = inverse(q, spec, nf)
    assert q.fused_failures() == 0
    assert rel(g, oracle.Plan(N, im, jm).spec_to_grid(spec)) < TOL[prec]
    p.close()
    q.close()
