"""The ensemble and serving workloads at their own sizes.

Ensemble: members drawn from T42/T85/T170 (128x64, 256x128, 512x256) with
16/32/48 levels of three variables plus ln p_s (3L+1 fields), two of them
T170 L48, so the T170 group holds 290 fields: the two-part staged FP64
kernels and the eight-wave mixed forward run inside the captured graph.
Every member must match the FP64 oracle (1e-12; 1e-6 for split-BF16 x6), be
bitwise equal to the same member transformed alone, and the graph replay must
be bitwise equal to the eager run.

Serving: a burst drawn from the serving mix (T21/T42/T85/T170 with
probabilities 0.3/0.3/0.3/0.1, 1..64 levels, forward or inverse); it must be
batched, each result bitwise equal to the request alone and within the
form's tolerance of the oracle.
"""
import numpy as np
import pytest

import oracle
import shtamd
from shtamd import hipmem, serve, synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GRIDS = {42: (128, 64), 85: (256, 128), 170: (512, 256)}
# (N, levels): two T170 L48 members beside T42/T85 members at L16/L32/L48
MEMBERS = [(170, 48), (42, 16), (85, 32), (170, 48), (42, 48), (85, 16), (42, 32)]


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def _read(b, j, N, nf):
    im, jm = GRIDS[N]
    _, grid, spec_out, ld = b.member_ptrs(j)
    K = shtamd.ncoef(N)
    return hipmem.from_device(grid, (nf, jm, im)), hipmem.get_member_spectra(spec_out, ld, K, nf)


@pytest.mark.parametrize("prec,tol", [(shtamd.PREC_FP64, 1e-12), (shtamd.PREC_SPLIT_BF16X6, 1e-6)])
def test_ensemble_workload(prec, tol):
    members = [(N, *GRIDS[N], 3 * L + 1) for N, L in MEMBERS]
    print(f"ensemble workload prec={prec}: T170 group {sum(m[3] for m in members if m[0] == 170)} fields",
          flush=True)
    b = shtamd.Batch(members, precision=prec)
    specs = []
    for j, (N, im, jm, nf) in enumerate(members):
        s = synth.state_spectra(N, (nf - 1) // 3, cfg=3, member=j)
        hipmem.put_member_spectra(b.member_ptrs(j)[0], b.member_ptrs(j)[3], s)
        specs.append(s)
    b.run_eager(shtamd.DIR_ROUNDTRIP)
    torch.cuda.synchronize()
    eager = [_read(b, j, m[0], m[3]) for j, m in enumerate(members)]
    # graph replays, with fresh outputs and the program's own work between them
    for rep in range(2):
        torch.zeros(1 << 20, dtype=torch.float64, device="cuda").add_(rep)
        for j, (N, im, jm, nf) in enumerate(members):
            _, grid, spec_out, ld = b.member_ptrs(j)
            hipmem.put_member_spectra(spec_out, ld, np.zeros((shtamd.ncoef(N), nf), dtype=complex))
        b.run(shtamd.DIR_ROUNDTRIP)
        torch.cuda.synchronize()
        for j, m in enumerate(members):
            g, s = _read(b, j, m[0], m[3])
            assert np.array_equal(g, eager[j][0]) and np.array_equal(s, eager[j][1]), (rep, j)
    b.close()
    plans = {}
    for j, (N, im, jm, nf) in enumerate(members):
        g, s_rt = eager[j]
        ref = plans.setdefault(N, oracle.Plan(N, im, jm))
        e_grid, e_rt = rel(g, ref.spec_to_grid(specs[j])), rel(s_rt, specs[j])
        print(f"  member {j} T{N} nf={nf}: grid {e_grid:.2e} round trip {e_rt:.2e}", flush=True)
        assert e_grid < tol and e_rt < 2 * tol, j
        # the member alone through a plan of its own: bitwise equal
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


@pytest.mark.parametrize("prec,tol", [(shtamd.PREC_FP64, 1e-12), (shtamd.PREC_SPLIT_BF16X6, 1e-6)])
def test_serving_workload_mix(prec, tol):
    reqs = serve.make_requests(48, seed=23)  # serve.MIX, 1..64 levels
    assert any(T == 170 for T, _, _ in reqs) and max(L for _, L, _ in reqs) > 48
    print(f"serving workload prec={prec}: {len(reqs)} requests, "
          f"{sum(T == 170 for T, _, _ in reqs)} at T170, levels up to {max(L for _, L, _ in reqs)}",
          flush=True)
    data = serve.build(torch, reqs)
    grids = [(T, *serve.GRIDS[T]) for T, _ in serve.MIX]
    srv = shtamd.Server(grids, slots=4, max_fields=256, precision=prec)
    lat, wall, retries = serve.run_stream(torch, srv, data, rate=20000.0)  # a burst
    srv.drain()
    st = srv.stats()
    srv.close()
    assert st["requests"] == len(reqs) and not np.isnan(lat).any() and (lat >= 0).all()
    assert st["batches"] < len(reqs)  # requests were batched on arrival
    plans = {T: shtamd.Plan(T, *serve.GRIDS[T], 64, precision=prec) for T, _ in serve.MIX}
    for d in data:
        d.alone_out = torch.empty_like(d.out)
        d.run_alone(plans[d.T])
    torch.cuda.synchronize()
    refs = {}
    for k, d in enumerate(data):
        assert torch.equal(d.alone_out, d.out), k
        ref = refs.setdefault(d.T, oracle.Plan(d.T, *serve.GRIDS[d.T]))
        if d.direction == shtamd.DIR_INVERSE:
            e = rel(d.out.cpu().numpy(), ref.spec_to_grid(d.spec))
        else:
            e = rel(shtamd.real_to_spec(d.out.cpu().numpy()), ref.grid_to_spec(d.inp.cpu().numpy()))
        assert e < tol, (k, d.T, d.nf, e)
    for p in plans.values():
        p.close()
