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
