"""Ensemble batches (sht_batch_*): members of uneven truncation and level count
in one graph launch.  Each member's result must be bitwise the same as the same
member transformed alone, and match the FP64 oracle (1e-12, or the mixed
form's tolerance); graph replays must equal the eager run every time."""
import numpy as np
import pytest

import oracle
import shtamd
from shtamd import hipmem, synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GRIDS = {21: (64, 32), 42: (128, 64), 85: (256, 128)}
MEMBERS = [(21, 16), (42, 13), (85, 10), (42, 7), (21, 5), (85, 3), (42, 1)]  # (N, nfield)


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def make_batch(prec):
    descs = [(N, *GRIDS[N], nf) for (N, nf) in MEMBERS]
    b = shtamd.Batch(descs, precision=prec)
    specs = []
    for j, (N, nf) in enumerate(MEMBERS):
        s = synth.spectra(N, nf, cfg=3, member=j)
        spec_in, grid, spec_out, ld = b.member_ptrs(j)
        hipmem.put_member_spectra(spec_in, ld, s)
        specs.append(s)
    return b, specs


def read_member(b, j):
    N, nf = MEMBERS[j]
    im, jm = GRIDS[N]
    spec_in, grid, spec_out, ld = b.member_ptrs(j)
    K = (N + 1) * (N + 2) // 2
    return hipmem.from_device(grid, (nf, jm, im)), hipmem.get_member_spectra(spec_out, ld, K, nf)


@pytest.mark.parametrize("prec,tol", [(shtamd.PREC_FP64, 1e-12), (shtamd.PREC_SPLIT_BF16X6, 1e-6)])
def test_batch_matches_alone_and_oracle(prec, tol):
    b, specs = make_batch(prec)
    b.run_eager(shtamd.DIR_ROUNDTRIP)
    torch.cuda.synchronize()
    for j, (N, nf) in enumerate(MEMBERS):
        im, jm = GRIDS[N]
        g, s_rt = read_member(b, j)
        ref = oracle.Plan(N, im, jm)
        assert rel(g, ref.spec_to_grid(specs[j])) < tol, j
        assert rel(s_rt, specs[j]) < 2 * tol, j
        # the same member alone through a plan of its own: bitwise equal
        p = shtamd.Plan(N, im, jm, nf, precision=prec)
        si = torch.from_numpy(shtamd.spec_to_real(specs[j])).cuda()
        ga = torch.empty((nf, jm, im), dtype=torch.float64, device="cuda")
        so = torch.empty_like(si)
        p.spec_to_grid(si, ga, nf)
        p.grid_to_spec(ga, so, nf)
        torch.cuda.synchronize()
        assert np.array_equal(ga.cpu().numpy(), g), j
        assert np.array_equal(shtamd.real_to_spec(so.cpu().numpy()), s_rt), j
        p.close()
    b.close()


def test_batch_graph_replays_equal_eager():
    b, specs = make_batch(shtamd.PREC_FP64)
    b.run_eager(shtamd.DIR_ROUNDTRIP)
    torch.cuda.synchronize()
    ref = [read_member(b, j) for j in range(len(MEMBERS))]
    scratch = torch.zeros(1 << 20, dtype=torch.float64, device="cuda")
    for rep in range(4):
        # the program's own eager work between replays, and a fresh output state
        scratch.add_(1.0)
        for j in range(len(MEMBERS)):
            _, grid, spec_out, ld = b.member_ptrs(j)
            N, nf = MEMBERS[j]
            K = (N + 1) * (N + 2) // 2
            hipmem.put_member_spectra(spec_out, ld, np.zeros((K, nf), dtype=complex))
        b.run(shtamd.DIR_ROUNDTRIP)
        torch.cuda.synchronize()
        for j in range(len(MEMBERS)):
            g, s = read_member(b, j)
            assert np.array_equal(g, ref[j][0]) and np.array_equal(s, ref[j][1]), (rep, j)
    b.close()


def test_batch_inverse_forward_directions():
    b, specs = make_batch(shtamd.PREC_FP64)
    b.run(shtamd.DIR_INVERSE)
    torch.cuda.synchronize()
    for j, (N, nf) in enumerate(MEMBERS):
        im, jm = GRIDS[N]
        g, _ = read_member(b, j)
        assert rel(g, oracle.Plan(N, im, jm).spec_to_grid(specs[j])) < 1e-12
    b.run(shtamd.DIR_FORWARD)
    torch.cuda.synchronize()
    for j in range(len(MEMBERS)):
        _, s = read_member(b, j)
        assert rel(s, specs[j]) < 1e-12
    b.close()


def test_batch_autotune_keeps_results():
    """Autotuning a batch re-captures its graph with the tuned shapes; results
    stay within the FP64 bar and the graph still equals the eager run."""
    b, specs = make_batch(shtamd.PREC_FP64)
    b.autotune(reps=3)
    b.run(shtamd.DIR_ROUNDTRIP)
    torch.cuda.synchronize()
    graph_out = [read_member(b, j) for j in range(len(MEMBERS))]
    b.run_eager(shtamd.DIR_ROUNDTRIP)
    torch.cuda.synchronize()
    for j, (N, nf) in enumerate(MEMBERS):
        im, jm = GRIDS[N]
        g, s = read_member(b, j)
        assert np.array_equal(g, graph_out[j][0]) and np.array_equal(s, graph_out[j][1]), j
        assert rel(g, oracle.Plan(N, im, jm).spec_to_grid(specs[j])) < 1e-12, j
        assert rel(s, specs[j]) < 1e-12, j
    b.close()
