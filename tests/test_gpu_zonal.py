"""The zonal (m = 0) transform of SPML's l_module at the energy-balance
model's own size, l_initial(1024, 1024) (exp/APEI07Couple/ebm/EBM_main.f90:5-8,
:64), through the C ABI on the GPU.

Known answers: calc_Hn_exact (EBM_main.f90:250-299) at x_s = 1 (no ice) makes
the insolation-times-coalbedo H a degree-4 polynomial in mu, so
sqrt(2n+1) l_y(H) must equal its closed-form coefficients l_H(1), l_H(3),
l_H(5), every other coefficient vanishes, and y_l(l_y(H)) returns H
(EBM_main.f90:170-174, :242-246).  Random fields: l_y and y_l against the
oracle's m = 0 Legendre column and Gauss weights (relative L2 1e-12, the FP64
bar), and the band-limited round trip.
"""
import numpy as np
import pytest

import oracle
import shtamd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _ebm_exact_lH(x_s, C0, C2, CI, SR2):
    """calc_Hn_exact, EBM_main.f90:264-296 (coefficients in the P_n basis)."""
    G002 = 0.5 * (x_s ** 3 - x_s)
    G022 = 0.25 * (9 / 5 * x_s ** 5 - 2 * x_s ** 3 + x_s)
    G222 = 0.125 * (-x_s + 3 * x_s ** 3 - 0.2 * (27 * x_s ** 5) + (27 * x_s ** 7) / 7)
    G004 = 0.125 * (7 * x_s ** 5 - 10 * x_s ** 3 + 3 * x_s)
    G024 = 0.0625 * (15 * x_s ** 7 - 25 * x_s ** 5 + 13 * x_s ** 3 - 3 * x_s)
    G224 = x_s * (35 / 32 * x_s ** 8 - 15 / 7 * x_s ** 6 + 121 / 80 * x_s ** 4 - 0.5 * x_s ** 2 + 3 / 32)
    l1 = (C0 - CI) * (x_s + SR2 * G002) + C2 * (G002 + SR2 * G022) + CI
    l3 = 5 * (C0 - CI) * (G002 + SR2 * G022) + 5 * C2 * (G022 + SR2 * G222) + SR2 * CI
    l5 = 9 * (C0 - CI) * (G004 + SR2 * G024) + 9 * C2 * (G024 + SR2 * G224)
    return l1, l3, l5


@pytest.mark.parametrize("params", [
    (1.0, 0.0, 0.5, 0.0),            # the shipped constants (EBM_main.f90:34-38), SR_P2 = 0
    (1.0, 0.0, 0.5, -0.482),         # SR_P2 of EBM_main.f90:32
    (0.697, -0.0779, 0.38, -0.482),  # the parameter set of EBM_main.f90:24-32
])
def test_ebm_closed_form_at_1024(params):
    C0, C2, CI, SR2 = params
    jm = nm = 1024
    z = shtamd.Zonal(nm, jm)
    lat, w = z.grid()
    mu, w_ref, cl = oracle.gauss(jm)
    assert np.abs(lat - np.arctan2(mu, cl)).max() < 1e-15 and np.abs(w - w_ref).max() < 1e-15
    x = np.sin(lat)
    P2 = 0.5 * (3 * x ** 2 - 1)
    H = (C0 + C2 * P2) * (1.0 + SR2 * P2)  # y_S * y_CoAlbedo, ice-free
    l_y = z.l_y(H)  # host drop-in, EBM_main.f90:170
    assert l_y.shape == (nm + 1,)
    l_H = np.sqrt(2 * np.arange(nm + 1) + 1.0) * l_y
    l1, l3, l5 = _ebm_exact_lH(1.0, C0, C2, CI, SR2)
    print(f"EBM {params}: l_H[0,2,4] errors {abs(l_H[0] - l1):.1e} {abs(l_H[2] - l3):.1e} {abs(l_H[4] - l5):.1e}",
          flush=True)
    assert abs(l_H[0] - l1) < 1e-14 and abs(l_H[2] - l3) < 1e-14 and abs(l_H[4] - l5) < 1e-14
    assert np.abs(np.delete(l_y, [0, 2, 4])).max() < 2e-14
    # y_l(l_y(H)) = H (the inverse of EBM_main.f90:127)
    assert np.abs(z.y_l(l_y) - H).max() < 5e-12
    z.close()


@pytest.mark.parametrize("nm,jm,nf", [(1024, 1024, 1), (1023, 1024, 3), (42, 64, 40), (7, 10, 17), (0, 2, 2)])
def test_zonal_vs_oracle(nm, jm, nf):
    rng = np.random.default_rng(nm + jm + nf)
    mu, w, cl = oracle.gauss(jm)
    Pt = oracle.legendre_column(nm, 0, mu, cl)  # (n, j)
    y = rng.standard_normal((nf, jm))
    l = rng.standard_normal((nf, nm + 1)) / np.sqrt(1.0 + np.arange(nm + 1))
    z = shtamd.Zonal(nm, jm, kmax=nf)
    # device entry points
    yd, ld = torch.from_numpy(y).cuda(), torch.from_numpy(l).cuda()
    lo, yo = torch.empty_like(ld), torch.empty_like(yd)
    z.l_y_dev(yd, lo, nf)
    z.y_l_dev(ld, yo, nf)
    torch.cuda.synchronize()
    l_ref = 0.5 * (y * w) @ Pt.T
    y_ref = l @ Pt
    # relative L2 against the oracle's table and weights summed by numpy (the
    # FP64 bar of the transform); the sums have up to 1025 terms of size up to
    # sqrt(2n+1), so the two summation orders differ by ~1e-13 at T1024
    e_l = np.linalg.norm(lo.cpu().numpy() - l_ref) / np.linalg.norm(l_ref)
    e_y = np.linalg.norm(yo.cpu().numpy() - y_ref) / np.linalg.norm(y_ref)
    print(f"zonal T{nm} jm={jm} nf={nf}: l_y {e_l:.1e} y_l {e_y:.1e}", flush=True)
    assert e_l < 1e-12 and e_y < 1e-12
    # host drop-in equals the device path bitwise
    assert np.array_equal(z.l_y(y), lo.cpu().numpy())
    assert np.array_equal(z.y_l(l), yo.cpu().numpy())
    # band-limited round trip (degrees < jm are exact on jm Gauss nodes)
    if nm < jm:
        assert np.abs(z.l_y(z.y_l(l)) - l).max() < 1e-12
    z.close()
