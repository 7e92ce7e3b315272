"""Legendre kernel shapes and the autotuner, through the C ABI.

Every shape of both FP64 Legendre kernels must meet the FP64 bar against the
CPU oracle (relative L2 <= 1e-12) on full-size, ragged, tiny and unpadded
problems, including the derivative table; the shapes differ only in the order
of the K sums, so they agree with each other to a few ulps.  The autotuner
must keep a valid shape, write its JSON cache and take the cache on a second
plan without timing.
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle
import shtamd
from shtamd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12
SHAPES = list(range(shtamd.NSHAPE))


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run_inverse(plan, spec_c, nf, grad=False):
    s = dev(shtamd.spec_to_real(spec_c))
    outs = [torch.empty((nf, plan.jm, plan.im), dtype=torch.float64, device="cuda") for _ in range(3)]
    if grad:
        plan.spec_to_grad(s, *outs, nf)
    else:
        plan.spec_to_grid(s, outs[0], nf)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in (outs if grad else outs[:1])]


def run_forward(plan, grid, nf):
    s = torch.empty((plan.K, nf, 2), dtype=torch.float64, device="cuda")
    plan.grid_to_spec(dev(grid), s, nf)
    torch.cuda.synchronize()
    return shtamd.real_to_spec(s.cpu().numpy())


CASES = [  # (N, im, jm, nf): full configs, ragged tiles, tiny and unpadded grids
    (85, 256, 128, 97),
    (170, 512, 256, 13),
    (42, 128, 64, 1),
    (42, 128, 64, 9),
    (21, 64, 32, 49),
    (0, 16, 2, 3),
    (7, 16, 10, 5),
    (31, 64, 48, 17),
]


@pytest.mark.parametrize("N,im,jm,nf", CASES)
def test_every_shape_matches_oracle(N, im, jm, nf):
    spec = synth.spectra(N, nf, cfg=5)
    spec[0] = np.linspace(1.0, 2.0, nf)
    ref = oracle.Plan(N, im, jm)
    g_ref = ref.spec_to_grid(spec)
    s_ref = ref.grid_to_spec(g_ref)
    p = shtamd.Plan(N, im, jm, nf)
    got = {}
    for sh in SHAPES:
        say("shape", sh, "case", N, im, jm, nf)
        p.set_shapes(nf, sh, sh)
        assert p.get_shapes(nf) == (sh, sh)
        g = run_inverse(p, spec, nf)[0]
        s = run_forward(p, g_ref, nf)
        assert rel(g, g_ref) < TOL, sh
        assert rel(s, s_ref) < TOL, sh
        got[sh] = (g, s)
    for sh in SHAPES[1:]:
        assert rel(got[sh][0], got[0][0]) < 1e-14
        assert rel(got[sh][1], got[0][1]) < 1e-14
    p.close()


@pytest.mark.parametrize("N,im,jm,nlev", [(42, 128, 64, 26), (7, 16, 10, 2)])
def test_every_shape_gradients(N, im, jm, nlev):
    nf = 3 * nlev + 1
    spec = synth.state_spectra(N, nlev, cfg=1)
    ref = oracle.Plan(N, im, jm)
    p = shtamd.Plan(N, im, jm, nf, flags=shtamd.FLAG_GRAD)
    for sh in SHAPES:
        say("grad shape", sh, N, nf)
        p.set_shapes(0, sh, sh)
        outs = run_inverse(p, spec, nf, grad=True)
        for kind, o in enumerate(outs):
            assert rel(o, ref.spec_to_grid(spec, kind=kind)) < TOL, (sh, kind)
    p.close()


def test_shape_results_do_not_depend_on_batch():
    """For one shape, a field's result is bitwise the same alone and in a batch."""
    N, im, jm, nf = 42, 128, 64, 40
    spec = synth.spectra(N, nf, cfg=6)
    p = shtamd.Plan(N, im, jm, nf)
    for sh in SHAPES:
        say("batch independence", sh)
        p.set_shapes(0, sh, sh)
        full = run_inverse(p, spec, nf)[0]
        one = run_inverse(p, spec[:, 17:18], 1)[0]
        assert np.array_equal(full[17:18], one)
        sf = run_forward(p, full, nf)
        s1 = run_forward(p, full[17:18], 1)
        assert np.array_equal(sf[:, 17:18], s1)
    p.close()


@pytest.mark.parametrize("N,im,jm,nf", [(85, 256, 128, 97), (42, 128, 64, 9), (170, 512, 256, 13),
                                         (170, 512, 256, 200)])  # 200: the staged shape in two K parts
def test_wave_tile_staged_bitwise_equal(N, im, jm, nf):
    """WAVE, TILE and STAGED run the same MFMA sequence per output tile of the
    inverse (steps in order, padded steps adding exact zeros), and WAVE and
    STAGED per tile of the forward, so the default shapes may switch between
    them with the field count without changing any result."""
    say("wave vs tile vs staged", N, nf)
    spec = synth.spectra(N, nf, cfg=15)
    p = shtamd.Plan(N, im, jm, nf, flags=shtamd.FLAG_GRAD)
    out = []
    for sh in (shtamd.SHAPE_WAVE, shtamd.SHAPE_TILE, shtamd.SHAPE_STAGED):
        p.set_shapes(0, sh, shtamd.SHAPE_KSPLIT)
        out.append(run_inverse(p, spec, nf, grad=True))
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert np.array_equal(a, b)
    grid = out[0][0]
    fwd = []
    for sh in (shtamd.SHAPE_WAVE, shtamd.SHAPE_STAGED):
        p.set_shapes(0, shtamd.SHAPE_KSPLIT, sh)
        fwd.append(run_forward(p, grid, nf))
    assert np.array_equal(fwd[0], fwd[1])
    p.close()


def test_autotune_and_cache(tmp_path):
    N, im, jm, nf = 85, 256, 128, 97
    cache = str(tmp_path / "shapes.json")
    p = shtamd.Plan(N, im, jm, nf)
    say("autotune", N, nf)
    (inv, fwd), (ti, tf) = p.autotune(nf, reps=10, cache_path=cache)
    assert inv in SHAPES and fwd in SHAPES
    assert ti > 0 and tf > 0
    data = json.load(open(cache))
    entry = data[f"T{N}_{im}x{jm}_nf{nf}_p0"]
    assert entry[:2] == [inv, fwd] and len(entry) == 6
    # the wavenumber classes of the split shape are kept with the shapes
    for d, (m4, m2) in ((shtamd.DIR_INVERSE, entry[2:4]), (shtamd.DIR_FORWARD, entry[4:6])):
        if m2 >= 0:
            assert p.get_classes(nf, d) == (m4, m2) and 0 <= m4 <= m2 <= N + 1
    # other field counts keep the default; an entry for nf only
    assert p.get_shapes(nf) == (inv, fwd)
    # a second plan takes the cached choice without timing
    q = shtamd.Plan(N, im, jm, nf)
    (inv2, fwd2), (t2i, t2f) = q.autotune(nf, reps=10, cache_path=cache)
    assert (inv2, fwd2) == (inv, fwd) and t2i == 0.0 and t2f == 0.0
    for d in (shtamd.DIR_INVERSE, shtamd.DIR_FORWARD):
        assert q.get_classes(nf, d) == p.get_classes(nf, d)
    # and the tuned plan still meets the FP64 bar
    spec = synth.spectra(N, nf, cfg=7)
    ref = oracle.Plan(N, im, jm)
    assert rel(run_inverse(q, spec, nf)[0], ref.spec_to_grid(spec)) < TOL
    p.close()
    q.close()


def test_shape_argument_checks():
    p = shtamd.Plan(21, 64, 32, 8)
    with pytest.raises(shtamd.ShtError):
        p.set_shapes(0, 99, 0)
    with pytest.raises(shtamd.ShtError):
        p.set_shapes(9, 0, 0)  # nf above kmax
    p.set_shapes(4, shtamd.SHAPE_TILE_SPLIT, shtamd.SHAPE_KSPLIT)
    assert p.get_shapes(4) == (shtamd.SHAPE_TILE_SPLIT, shtamd.SHAPE_KSPLIT)
    # built-in choice: staged forward; inverse staged from 32 fields, wave below
    assert p.get_shapes(0) == (shtamd.SHAPE_WAVE, shtamd.SHAPE_STAGED)
    p.set_shapes(4, shtamd.SHAPE_AUTO, shtamd.SHAPE_AUTO)
    assert p.get_shapes(4) == p.get_shapes(0)
    p.close()
    # an AUTO half of an entry resolves at call time, for the entry's own field
    # count: the plan default if set, else the built-in choice for that count
    q = shtamd.Plan(21, 64, 32, 40)
    q.set_shapes(36, shtamd.SHAPE_AUTO, shtamd.SHAPE_KSPLIT)
    assert q.get_shapes(36) == (shtamd.SHAPE_STAGED, shtamd.SHAPE_KSPLIT)
    q.set_shapes(0, shtamd.SHAPE_TILE, shtamd.SHAPE_AUTO)
    assert q.get_shapes(36) == (shtamd.SHAPE_TILE, shtamd.SHAPE_KSPLIT)
    assert q.get_shapes(3) == (shtamd.SHAPE_TILE, shtamd.SHAPE_STAGED)
    q.close()
    m = shtamd.Plan(21, 64, 32, 8, precision=shtamd.PREC_SPLIT_BF16X6)
    with pytest.raises(shtamd.ShtError):
        m.set_shapes(0, shtamd.SHAPE_TILE_SPLIT, shtamd.SHAPE_TILE_SPLIT)
    with pytest.raises(shtamd.ShtError):  # mixed plans: the staged shape is inverse only
        m.set_shapes(0, shtamd.SHAPE_KSPLIT, shtamd.SHAPE_STAGED)
    m.set_shapes(0, shtamd.SHAPE_STAGED, shtamd.SHAPE_KSPLIT)
    assert m.get_shapes(5) == (shtamd.SHAPE_STAGED, shtamd.SHAPE_KSPLIT)
    m.set_shapes(0, shtamd.SHAPE_WAVE, shtamd.SHAPE_KSPLIT)
    assert m.get_shapes(5) == (shtamd.SHAPE_WAVE, shtamd.SHAPE_KSPLIT)
    (inv, fwd), _ = m.autotune(8, reps=2)
    assert inv in MIXED_SHAPES + (shtamd.SHAPE_STAGED,) and fwd in MIXED_SHAPES
    m.close()
    # the mixed staged (streaming) inverse holds the table of at most 128
    # latitude pairs; on the T170 grid it is the built-in inverse of the
    # six-product form, and of the two-piece forms up to 8 fields
    big = shtamd.Plan(106, 512, 288, 4, precision=shtamd.PREC_SPLIT_BF16X6)
    with pytest.raises(shtamd.ShtError) as e:
        big.set_shapes(0, shtamd.SHAPE_STAGED, shtamd.SHAPE_KSPLIT)
    assert e.value.code == shtamd.SHT_ERR_UNSUPPORTED
    assert big.get_shapes(97) == (shtamd.SHAPE_TILE, shtamd.SHAPE_TILE)
    big.close()
    t170 = shtamd.Plan(170, 512, 256, 4, precision=shtamd.PREC_SPLIT_BF16X6)
    assert [t170.get_shapes(nf)[0] for nf in (1, 97, 776)] == [shtamd.SHAPE_STAGED] * 3
    t170.close()
    t170 = shtamd.Plan(170, 512, 256, 4, precision=shtamd.PREC_SPLIT_BF16)
    assert [t170.get_shapes(nf)[0] for nf in (8, 9, 24, 25)] == [
        shtamd.SHAPE_STAGED, shtamd.SHAPE_KSPLIT, shtamd.SHAPE_KSPLIT, shtamd.SHAPE_TILE]
    t170.close()


MIXED_SHAPES = (shtamd.SHAPE_KSPLIT, shtamd.SHAPE_WAVE, shtamd.SHAPE_TILE)
MIXED_TOL = {shtamd.PREC_SPLIT_BF16: 5e-5, shtamd.PREC_SCALED_FP16: 1e-6, shtamd.PREC_SPLIT_BF16X6: 1e-6}


@pytest.mark.parametrize("prec", sorted(MIXED_TOL))
@pytest.mark.parametrize("N,im,jm,nf", [(85, 256, 128, 97), (21, 64, 32, 9), (170, 512, 256, 7),
                                        (170, 512, 256, 69), (170, 512, 256, 200), (31, 64, 48, 17),
                                        (106, 256, 240, 40)])
def test_mixed_shapes_bitwise_equal(prec, N, im, jm, nf):
    """The mixed-precision shapes (one or two column tiles per workgroup, a
    wave per tile, or for the inverse the table held in registers while the
    data stream through LDS) split the same operands into the same pieces and
    sum the same products in the same order: bitwise equal results, within the
    form's bound of the oracle.  The field counts give odd and even numbers of
    16-column tiles, so the two-tile shape also runs a lone last tile, and the
    streaming shape (up to 8 column tiles per workgroup, loads 4 tiles ahead)
    runs single tiles (7 fields), chunks shorter than its load distance (69
    fields: 3 x 3 tiles) and longer ones with a short last chunk (200 fields:
    7 + 7 + 7 + 4 tiles); 240 latitudes (120 pairs) leave its last row tile
    half filled."""
    say("mixed shapes", prec, N, nf)
    spec = synth.spectra(N, nf, cfg=14)
    ref = oracle.Plan(N, im, jm)
    g_ref = ref.spec_to_grid(spec)
    p = shtamd.Plan(N, im, jm, nf, precision=prec)
    out = []
    pairs = [(sh, sh) for sh in MIXED_SHAPES] + ([(shtamd.SHAPE_STAGED, shtamd.SHAPE_KSPLIT)] if jm <= 256 else [])
    for inv, fwd in pairs:
        p.set_shapes(0, inv, fwd)
        g = run_inverse(p, spec, nf)[0]
        s = run_forward(p, g_ref, nf)
        assert rel(g, g_ref) < MIXED_TOL[prec]
        assert rel(s, spec) < 2 * MIXED_TOL[prec]
        out.append((g, s))
    for g, s in out[1:]:
        assert np.array_equal(out[0][0], g)
        assert np.array_equal(out[0][1], s)
    p.close()


@pytest.mark.parametrize("prec", [shtamd.PREC_FP64, shtamd.PREC_SPLIT_BF16X6])
def test_size_dependent_paths_bitwise_equal(prec):
    """Two code paths depend on the batch size alone: the forward FFT feeds its
    first pass from the grid while the launch fits the device at once and
    stages rows through LDS above that, and the Legendre inverse streams the
    Fourier intermediate up to 24 MB and stores it regularly above.  Both
    pairs compute the same operations, so the first 97 fields of a 388-field
    call (staged, regular) must equal a 97-field call (direct, streamed)
    bitwise, for a fixed kernel shape."""
    N, im, jm, big, small = 85, 256, 128, 388, 97
    say("size paths", prec)
    spec = synth.spectra(N, big, cfg=15)
    p = shtamd.Plan(N, im, jm, big, precision=prec)
    shapes = [shtamd.SHAPE_KSPLIT] + ([shtamd.SHAPE_STAGED] if prec == shtamd.PREC_FP64 else [])
    for sh in shapes:
        p.set_shapes(0, sh, sh)
        g_big = run_inverse(p, spec, big)[0]
        g_small = run_inverse(p, spec[:, :small], small)[0]
        assert np.array_equal(g_big[:small], g_small)
        s_big = run_forward(p, g_big, big)
        s_small = run_forward(p, g_big[:small], small)
        assert np.array_equal(s_big[:, :small], s_small)
    p.close()


@pytest.mark.parametrize("N,im,jm,nf", [(85, 256, 128, 97), (42, 128, 64, 40)])
def test_split_classes_meet_the_bar(N, im, jm, nf):
    """Every wavenumber class split the autotuner may choose (and a few
    extremes: all four-way, all two-way, none) keeps the split shape within
    the FP64 bar in both directions."""
    spec = synth.spectra(N, nf, cfg=17)
    ref = oracle.Plan(N, im, jm)
    g_ref = ref.spec_to_grid(spec)
    s_ref = ref.grid_to_spec(g_ref)
    p = shtamd.Plan(N, im, jm, nf)
    p.set_shapes(nf, shtamd.SHAPE_TILE_SPLIT, shtamd.SHAPE_TILE_SPLIT)
    M = N + 1
    for m4, m2 in [(0, M // 5), (M // 10, M // 2), (M, M), (0, M), (0, 0), (M // 3, M // 3)]:
        say("classes", N, nf, m4, m2)
        for d in (shtamd.DIR_INVERSE, shtamd.DIR_FORWARD):
            p.set_classes(nf, d, m4, m2)
            assert p.get_classes(nf, d) == (m4, m2)
        assert rel(run_inverse(p, spec, nf)[0], g_ref) < TOL
        assert rel(run_forward(p, g_ref, nf), s_ref) < TOL
    with pytest.raises(shtamd.ShtError):
        p.set_classes(nf, shtamd.DIR_INVERSE, 5, 2)
    with pytest.raises(shtamd.ShtError):
        p.set_classes(nf, shtamd.DIR_FORWARD, 0, M + 1)
    p.set_classes(nf, shtamd.DIR_INVERSE, -1, -1)  # back to the built-in classes
    p.close()


@pytest.mark.parametrize("prec", [shtamd.PREC_FP64, shtamd.PREC_SPLIT_BF16X6])
@pytest.mark.parametrize("N,im,jm,small,big", [(21, 64, 32, 49, 600), (7, 32, 16, 5, 2100), (5, 16, 12, 3, 5600)])
def test_fft_block_size_does_not_change_results(prec, N, im, jm, small, big):
    """On grids of at most 64 longitudes, calls that give the FFT few
    workgroups run it in small blocks (512 complex values, one wave), the
    others in big ones (2048, four waves), each with its own field-group
    layout of the Fourier intermediate; every transform and every Legendre
    sum is the same computation, so a field's grid and spectra are bitwise
    the same in a small-block call and a big-block call."""
    say("fft blocks", prec, N, small, big)
    spec = synth.spectra(N, big, cfg=17)
    p = shtamd.Plan(N, im, jm, big, precision=prec)
    p.set_shapes(0, shtamd.SHAPE_KSPLIT, shtamd.SHAPE_KSPLIT)
    g_big = run_inverse(p, spec, big)[0]
    g_small = run_inverse(p, spec[:, :small], small)[0]
    assert np.array_equal(g_big[:small], g_small)
    s_big = run_forward(p, g_big, big)
    s_small = run_forward(p, g_big[:small], small)
    assert np.array_equal(s_big[:, :small], s_small)
    ref = oracle.Plan(N, im, jm)
    tol = TOL if prec == shtamd.PREC_FP64 else 1e-6
    assert rel(g_small, ref.spec_to_grid(spec[:, :small])) < tol
    p.close()
