"""GPU coupler remap (sht_remap_*) against the oracle's restatement of the
reference loop (interpolate_data_latlon, common/interpolation_data_latlon_mod.f90:274-306):
bitwise equal, since every receiving point sums its triplets in their original
order with unfused multiply and add."""
import sys

import numpy as np
import pytest

import oracle
import shtamd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bilinear_latlon(src, dst):
    """Triplets of a bilinear lat-lon remap between two regular grids, listed
    in a shuffled order with every receiving point's 4 weights scattered."""
    (ims, jms), (imd, jmd) = src, dst
    si, ri, cf = [], [], []
    for j in range(jmd):
        y = (j + 0.5) / jmd * jms - 0.5
        j0 = int(np.floor(y)); fy = y - j0
        for i in range(imd):
            x = i / imd * ims
            i0 = int(np.floor(x)); fx = x - i0
            r = j * imd + i
            for dj, wy in ((0, 1 - fy), (1, fy)):
                jj = min(max(j0 + dj, 0), jms - 1)
                for di, wx in ((0, 1 - fx), (1, fx)):
                    si.append(jj * ims + (i0 + di) % ims)
                    ri.append(r)
                    cf.append(wx * wy)
    perm = np.random.default_rng(7).permutation(len(cf))
    return np.array(si)[perm], np.array(ri)[perm], np.array(cf)[perm]


@pytest.mark.parametrize("case", ["random", "bilinear"])
@pytest.mark.parametrize("ndata", [1, 7])
def test_remap_bitwise_vs_oracle(case, ndata):
    say("remap", case, ndata)
    if case == "random":
        rng = np.random.default_rng(3)
        nsend, nrecv, nnz = 5000, 4000, 30000  # duplicates, empty rows, any order
        si = rng.integers(0, nsend, nnz)
        ri = rng.integers(0, nrecv, nnz)
        cf = rng.standard_normal(nnz)
    else:
        nsend, nrecv = 128 * 64, 256 * 128  # T42 grid to T85 grid
        si, ri, cf = bilinear_latlon((128, 64), (256, 128))
    send = np.random.default_rng(4).standard_normal((ndata, nsend))
    want = oracle.remap(si, ri, cf, nrecv, send)
    r = shtamd.Remap(si, ri, cf, nsend, nrecv)
    s_dev = torch.from_numpy(send).cuda()
    out = torch.full((ndata, nrecv), np.nan, dtype=torch.float64, device="cuda")
    r.apply(s_dev, out, ndata)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), want)
    r.close()


def test_remap_argument_checks():
    with pytest.raises(shtamd.ShtError):
        shtamd.Remap([0, 5], [0, 1], [1.0, 1.0], 5, 2)  # send index out of range
    with pytest.raises(shtamd.ShtError):
        shtamd.Remap([0], [3], [1.0], 5, 2)  # recv index out of range
    r = shtamd.Remap([], [], [], 4, 3)  # no triplets: zeros
    out = torch.ones((2, 3), dtype=torch.float64, device="cuda")
    r.apply(torch.ones((2, 4), dtype=torch.float64, device="cuda"), out, 2)
    torch.cuda.synchronize()
    assert torch.count_nonzero(out) == 0
    r.close()
