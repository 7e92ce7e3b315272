"""Request server: continuous batching.  A burst of mixed requests is batched
(fewer batches than requests), every result is bitwise the same as the same
request transformed alone, and within tolerance of the FP64 oracle; requests
beyond the memory budget are refused with SHT_ERR_BUSY."""
import numpy as np
import pytest

import oracle
import shtamd
from shtamd import serve

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MIX = ((21, 0.4), (42, 0.4), (85, 0.2))


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("prec,tol", [(shtamd.PREC_FP64, 1e-12), (shtamd.PREC_SCALED_FP16, 1e-6)])
def test_serving_batches_and_matches_alone(prec, tol):
    reqs = serve.make_requests(40, seed=11, max_levels=16, mix=MIX)
    data = serve.build(torch, reqs, max_levels=16)
    grids = [(T, *serve.GRIDS[T]) for T, _ in MIX]
    srv = shtamd.Server(grids, slots=2, max_fields=64, precision=prec)
    lat, wall, retries = serve.run_stream(torch, srv, data, rate=20000.0)  # a burst
    srv.drain()
    st = srv.stats()
    assert st["requests"] == len(reqs) and not np.isnan(lat).any() and (lat >= 0).all()
    assert st["batches"] < len(reqs)  # requests were batched on arrival
    plans = {T: shtamd.Plan(T, *serve.GRIDS[T], 16, precision=prec) for T, _ in MIX}
    for d in data:
        d.alone_out = torch.empty_like(d.out)
        d.run_alone(plans[d.T])
    torch.cuda.synchronize()
    for k, d in enumerate(data):
        assert torch.equal(d.alone_out, d.out), k
        im, jm = serve.GRIDS[d.T]
        ref = oracle.Plan(d.T, im, jm)
        if d.direction == shtamd.DIR_INVERSE:
            assert rel(d.out.cpu().numpy(), ref.spec_to_grid(d.spec)) < tol
        else:
            got = shtamd.real_to_spec(d.out.cpu().numpy())
            assert rel(got, ref.grid_to_spec(d.inp.cpu().numpy())) < tol
    srv.close()


def test_admission_budget_refuses():
    reqs = [(42, 16, shtamd.DIR_INVERSE)] * 3
    data = serve.build(torch, reqs, max_levels=16)
    one = 16 * 8 * (128 * 64 + 43 * 44)  # bytes of one request's data
    tight = shtamd.Server([(42, 128, 64)], slots=1, max_fields=16, mem_budget=one - 1)
    with pytest.raises(shtamd.ShtError) as e:
        tight.submit(42, 16, shtamd.DIR_INVERSE, data[0].inp.data_ptr(), data[0].out.data_ptr())
    assert e.value.code == shtamd.SHT_ERR_BUSY
    assert tight.stats()["refused"] == 1
    tight.close()
    roomy = shtamd.Server([(42, 128, 64)], slots=1, max_fields=16, mem_budget=3 * one)
    for d in data:
        roomy.submit(42, 16, shtamd.DIR_INVERSE, d.inp.data_ptr(), d.out.data_ptr())
    roomy.drain()
    assert roomy.stats()["requests"] == 3 and roomy.stats()["refused"] == 0
    roomy.close()


def test_failed_batch_fails_only_its_requests(monkeypatch):
    """A batch that fails before reaching the device (forced from the host
    with SHT_INJECT_BATCH_FAILURE) fails its own requests only: they report
    latency -1, the server keeps admitting, and the requests before and after
    it complete with correct results."""
    monkeypatch.setenv("SHT_INJECT_BATCH_FAILURE", "2")
    reqs = [(42, 8, shtamd.DIR_INVERSE)] * 3
    data = serve.build(torch, reqs, max_levels=16)
    srv = shtamd.Server([(42, 128, 64)], slots=1, max_fields=16)
    monkeypatch.delenv("SHT_INJECT_BATCH_FAILURE")
    results = {}
    for d in data:  # one batch each: drained before the next submit
        t = srv.submit(42, 8, shtamd.DIR_INVERSE, d.inp.data_ptr(), d.out.data_ptr())
        srv.drain()
        for tk, ms in srv.poll():
            results[tk] = ms
        d.ticket = t
    assert [results[d.ticket] >= 0 for d in data] == [True, False, True]
    failed, why = srv.failures()
    assert failed == 1 and "injected" in why
    st = srv.stats()
    assert st["requests"] == 2 and st["batches"] == 3
    ref = oracle.Plan(42, 128, 64)
    for d in (data[0], data[2]):
        assert rel(d.out.cpu().numpy(), ref.spec_to_grid(d.spec)) < 1e-12
    srv.close()


@pytest.mark.parametrize("k", [1, 2, 5, 9, 14])
def test_plan_creation_failure_cleans_up(monkeypatch, k):
    """The k-th device allocation of a plan fails (SHT_INJECT_PLAN_FAILURE):
    creation returns SHT_ERR_NOMEM after releasing what it had allocated, a
    server whose plan fails is not created, and without the injection the
    same plans build and transform correctly."""
    monkeypatch.setenv("SHT_INJECT_PLAN_FAILURE", str(k))
    for prec in (shtamd.PREC_FP64, shtamd.PREC_SPLIT_BF16X6):
        with pytest.raises(shtamd.ShtError) as e:
            shtamd.Plan(42, 128, 64, 8, precision=prec, flags=shtamd.FLAG_GRAD)
        assert e.value.code == shtamd.SHT_ERR_NOMEM
    if k <= 13:  # a server's FP64 plans (no derivative tables) allocate 13 times
        with pytest.raises(shtamd.ShtError) as e:
            shtamd.Server([(21, 64, 32), (42, 128, 64)], slots=2, max_fields=8)
        assert e.value.code == shtamd.SHT_ERR_NOMEM
    monkeypatch.delenv("SHT_INJECT_PLAN_FAILURE")
    p = shtamd.Plan(42, 128, 64, 8, flags=shtamd.FLAG_GRAD)
    spec = serve.synth.spectra(42, 8, cfg=9)
    s = torch.from_numpy(shtamd.spec_to_real(spec)).cuda()
    g = torch.empty((8, 64, 128), dtype=torch.float64, device="cuda")
    p.spec_to_grid(s, g, 8)
    torch.cuda.synchronize()
    assert rel(g.cpu().numpy(), oracle.Plan(42, 128, 64).spec_to_grid(spec)) < 1e-12
    p.close()
