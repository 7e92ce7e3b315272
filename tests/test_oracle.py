"""Pins the CPU oracle (oracle/sht_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible slip (a dropped term, a flipped sign,
a transposed index, a wrong normalisation) fails at least one of them:
  * the mpmath golden fixtures (tests/golden/make_golden.py, independent code),
  * library routines (numpy leggauss / rfft, scipy assoc. Legendre),
  * the reference's own quadrature identities and closed forms
    (EBM_main.f90:113-117, :250-299; gmapgen_main.f90:364-368; DCPAMUtil.rb:54,67-68),
  * invariants of the transform (exact round trip for band-limited input,
    Parseval, global mean, longitude rotation, idempotence).
"""
import math
import os

import numpy as np
import pytest
import scipy.special as sps

import oracle
from shtamd import synth

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b))


# ---------------------------------------------------------------- Gauss grid
@pytest.mark.parametrize("jm", [2, 8, 32, 64, 128, 256, 1024])
def test_gauss_vs_numpy_and_invariants(jm):
    mu, w, cl = oracle.gauss(jm)
    x, wx = np.polynomial.legendre.leggauss(jm)
    assert np.abs(mu - x).max() < 2e-15
    # numpy's weights lose ~1e-11 relative near the poles at jm >= 128 (checked
    # against mpmath below), so numpy is only a coarse check here
    assert np.abs(w - wx).max() < 1e-10 * wx.max()
    assert np.all(np.diff(mu) > 0)  # S -> N
    assert abs(w.sum() - 2.0) < 1e-14
    assert np.abs(mu + mu[::-1]).max() == 0.0 and np.abs(w - w[::-1]).max() == 0.0
    # sqrt(1-mu^2) from a rounded mu carries ~eps/cos(lat) error near the poles;
    # the oracle's cos(lat) comes from the colatitude and is checked against mpmath
    assert np.abs(cl - np.sqrt((1 - mu) * (1 + mu))).max() < 4e-16 / cl.min()
    # exact moments int mu^k dmu for k <= 2 jm - 1
    for k in range(0, min(2 * jm, 40)):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, mu ** k) - exact) < 1e-14


@pytest.mark.parametrize("case", ["golden_T5.npz", "golden_T10.npz"])
def test_gauss_vs_mpmath_golden(case):
    g = np.load(os.path.join(GOLD, case))
    mu, w, cl = oracle.gauss(int(g["jm"]))
    assert np.abs(mu - g["mu"]).max() < 2e-16
    assert np.abs(w - g["w"]).max() < 2e-16
    assert np.abs(cl - g["coslat"]).max() < 2e-16


def test_gauss_high_degree_vs_mpmath():
    import mpmath as mp
    mp.mp.dps = 40
    for jm in (128, 256):
        mu, w, cl = oracle.gauss(jm)
        for j in (0, 1, jm // 4, jm // 2 - 1):
            xr = mp.findroot(lambda t: mp.legendre(jm, t), mp.mpf(mu[j]))
            dp = mp.diff(lambda t: mp.legendre(jm, t), xr)
            wr = 2 / ((1 - xr * xr) * dp * dp)
            assert abs(mu[j] - float(xr)) < 2e-16
            assert abs(w[j] - float(wr)) < 2e-14 * float(wr)
            assert abs(cl[j] - float(mp.sqrt(1 - xr * xr))) < 5e-16 * cl[j]


def test_gauss_cell_edges_reference_construction():
    """gmapgen_main.f90:364-368 / EBM_main.f90:200-204 build cell edges
    phi_j = asin(w_j + sin phi_{j-1}) from -pi/2; they close at +pi/2 only if
    the weights sum to 2 in S->N order, and each Gauss node lies inside its cell."""
    for jm in (32, 64, 1024):
        mu, w, _ = oracle.gauss(jm)
        edge = -1.0 + np.cumsum(w)
        assert abs(edge[-1] - 1.0) < 1e-13
        lo = np.concatenate([[-1.0], edge[:-1]])
        assert np.all(mu > lo) and np.all(mu < edge + 1e-15)
        # sum over lon_w * lat_w = 4 pi (DCPAMUtil.rb:54, :67-68)
        im = 2 * jm
        assert abs(np.sum(np.full(im, 2 * np.pi / im)) * w.sum() - 4 * np.pi) < 1e-12


def test_ebm_quadrature_checks():
    """The commented checks at EBM_main.f90:113-117: (1/2) sum w P2 = 0,
    sum w P2^2 = 2/5, (1/2) sum w P2^3 = 2/35."""
    mu, w, _ = oracle.gauss(1024)
    P2 = 0.5 * (3 * mu ** 2 - 1)
    assert abs(0.5 * np.dot(w, P2)) < 1e-15
    assert abs(np.dot(w, P2 * P2) - 2 / 5) < 1e-14
    assert abs(0.5 * np.dot(w, P2 ** 3) - 2 / 35) < 1e-14


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
    (1.0, 0.0, 0.5, 0.0),          # shipped constants, SR_P2 taken as 0
    (1.0, 0.0, 0.5, -0.482),       # SR_P2 of EBM_main.f90:32
    (0.697, -0.0779, 0.38, -0.482),  # parameter set of EBM_main.f90:24-32
])
def test_ebm_zonal_legendre_closed_form(params):
    """l_y at jm = nm = 1024 (EBM_main.f90:5-8, :242-246), ice-free (x_s = 1):
    sqrt(2n+1) * l_y(H) equals calc_Hn_exact's coefficients; all others vanish."""
    C0, C2, CI, SR2 = params
    jm = 1024
    mu, w, cl = oracle.gauss(jm)
    P2 = 0.5 * (3 * mu ** 2 - 1)
    H = (C0 + C2 * P2) * (1.0 + SR2 * P2)
    Pt = oracle.legendre_column(jm - 1, 0, mu, cl)  # (n, j)
    l_y = 0.5 * Pt @ (w * H)
    l_H = np.sqrt(2 * np.arange(jm) + 1.0) * l_y
    l1, l3, l5 = _ebm_exact_lH(1.0, C0, C2, CI, SR2)
    assert abs(l_H[0] - l1) < 1e-14
    assert abs(l_H[2] - l3) < 1e-14
    assert abs(l_H[4] - l5) < 1e-14
    # every other l_y coefficient vanishes (l_H = sqrt(2n+1) l_y would scale
    # the ~1e-14 rounding of the degree-1000 quadratures by up to 45)
    assert np.abs(np.delete(l_y, [0, 2, 4])).max() < 2e-14
    # y_l(l_y) reconstructs H on the grid (the inverse used at EBM_main.f90:127)
    assert np.abs(Pt.T @ l_y - H).max() < 5e-12


def test_ebm_bonnet_recurrence_matches_zonal_table():
    """EBM_main.f90:97-103 evaluates P_n(x_s) by Bonnet's recurrence and pairs
    it with sqrt(2n+1)-scaled coefficients: Pt_n = sqrt(2n+1) P_n."""
    x = np.array([-0.9, -0.3, 0.0, 0.41, 0.77, 0.999])
    nm = 60
    P = np.zeros((nm + 1, x.size))
    P[0] = 1.0
    P[1] = x
    for n in range(2, nm + 1):
        P[n] = ((2 * n - 1) * x * P[n - 1] - (n - 1) * P[n - 2]) / n
    Pt = oracle.legendre_column(nm, 0, x)
    assert np.abs(Pt - np.sqrt(2 * np.arange(nm + 1.0) + 1)[:, None] * P).max() < 1e-12


# ---------------------------------------------------------------- Legendre
def test_legendre_closed_forms():
    mu = np.linspace(-0.99, 0.99, 23)
    c = np.sqrt(1 - mu ** 2)
    L0 = oracle.legendre_column(3, 0, mu)
    L1 = oracle.legendre_column(3, 1, mu)
    L2 = oracle.legendre_column(3, 2, mu)
    L3 = oracle.legendre_column(3, 3, mu)
    assert np.allclose(L0[0], 1.0, atol=1e-15)
    assert np.allclose(L0[1], math.sqrt(3) * mu, atol=1e-15)
    assert np.allclose(L0[2], math.sqrt(5) * 0.5 * (3 * mu ** 2 - 1), atol=1e-14)
    assert np.allclose(L0[3], math.sqrt(7) * 0.5 * (5 * mu ** 3 - 3 * mu), atol=1e-14)
    assert np.allclose(L1[0], math.sqrt(1.5) * c, atol=1e-15)
    assert np.allclose(L1[1], math.sqrt(7.5) * mu * c, atol=1e-14)
    assert np.allclose(L2[0], math.sqrt(15 / 8) * c ** 2, atol=1e-14)
    assert np.allclose(L3[0], math.sqrt(35 / 16) * c ** 3, atol=1e-14)
    # derivative table, closed forms: (1-mu^2) d/dmu
    D0 = oracle.dlegendre_column(2, 0, mu)
    D1 = oracle.dlegendre_column(2, 1, mu)
    assert np.allclose(D0[0], 0.0, atol=1e-15)
    assert np.allclose(D0[1], math.sqrt(3) * c ** 2, atol=1e-14)
    assert np.allclose(D0[2], math.sqrt(5) * 3 * mu * c ** 2, atol=1e-14)
    assert np.allclose(D1[0], -math.sqrt(1.5) * mu * c, atol=1e-14)


def test_legendre_vs_scipy_moderate_degree():
    mu = np.linspace(-0.95, 0.95, 37)
    nmax = 60
    for m in (0, 1, 7, 30, 60):
        ours = oracle.legendre_column(nmax, m, mu)
        for n in range(m, nmax + 1, 3):
            # scipy lpmv includes the Condon-Shortley phase
            ref = (-1) ** m * sps.lpmv(m, n, mu) * math.sqrt((2 * n + 1) * math.exp(
                math.lgamma(n - m + 1) - math.lgamma(n + m + 1)))
            assert np.abs(ours[n - m] - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("case", ["golden_T5.npz", "golden_T10.npz"])
def test_legendre_tables_vs_mpmath_golden(case):
    g = np.load(os.path.join(GOLD, case))
    N = int(g["nmax"])
    ns, ms = synth.coef_index(N)
    for m in range(N + 1):
        sel = ms == m
        P = oracle.legendre_column(N, m, g["mu"], g["coslat"])
        H = oracle.dlegendre_column(N, m, g["mu"], g["coslat"])
        assert np.abs(P - g["P"][sel]).max() < 5e-15
        assert np.abs(H - g["H"][sel]).max() < 5e-14


def test_legendre_high_degree_spot_checks_mpmath():
    """Spot values up to n = 171 on the T170 grid (jm = 256) against mpmath at
    50 digits, including the polar rows where Pt_m^m = prod sqrt((2k+1)/2k) s^m
    falls below the double range (closed form; mpmath's legenp does not
    converge there)."""
    import mpmath as mp
    mp.mp.dps = 50
    mu, w, cl = oracle.gauss(256)

    def ref_legenp(n, m, x):
        return mp.sqrt((2 * n + 1) * mp.factorial(n - m) / mp.factorial(n + m)) * (-1) ** m * mp.legenp(n, m, x)

    def ref_sectoral(n, m, x, s):
        pmm = mp.mpf(1)
        for k in range(1, m + 1):
            pmm *= mp.sqrt(mp.mpf(2 * k + 1) / (2 * k)) * s
        return pmm if n == m else mp.sqrt(2 * m + 3) * x * pmm  # n = m or m + 1

    def ref_recur(n, m, x, s):
        # the three-term recurrence at 50 digits (no exponent range to run out of)
        p0, p1 = mp.mpf(0), ref_sectoral(m, m, x, s)
        for k in range(m + 1, n + 1):
            e_k = mp.sqrt(mp.mpf(k * k - m * m) / (4 * k * k - 1))
            e_k1 = mp.sqrt(mp.mpf((k - 1) ** 2 - m * m) / (4 * (k - 1) ** 2 - 1)) if k - 1 > m else 0
            p0, p1 = p1, (x * p1 - e_k1 * p0) / e_k
        return p1

    cases = [(170, 0, 200, ref_legenp), (170, 85, 140, ref_legenp), (120, 100, 255, ref_recur),
             (170, 150, 255, ref_recur), (171, 140, 254, ref_recur), (171, 120, 253, ref_recur),
             (171, 3, 255, ref_legenp), (150, 140, 200, ref_legenp),
             (170, 170, 255, ref_sectoral), (171, 170, 255, ref_sectoral), (170, 169, 255, ref_sectoral),
             (160, 160, 240, ref_sectoral), (151, 150, 250, ref_sectoral)]
    for (n, m, j, fn) in cases:
        x = mp.mpf(mu[j])
        s = mp.mpf(cl[j])
        ref = fn(n, m, x) if fn is ref_legenp else fn(n, m, x, s)
        ours = oracle.legendre_column(n, m, mu[j:j + 1], cl[j:j + 1])[n - m, 0]
        refd = float(ref)
        if abs(ref) > mp.mpf("1e-290"):
            assert abs(ours - refd) <= 1e-12 * abs(refd), (n, m, j)
        else:  # below the double range: the oracle returns 0 or a subnormal
            assert abs(ours) <= 1e-280, (n, m, j)


@pytest.mark.parametrize("N,jm", [(21, 32), (85, 128), (170, 256)])
def test_discrete_orthonormality(N, jm):
    mu, w, cl = oracle.gauss(jm)
    for m in (0, 1, N // 3, N // 2, N - 1, N):
        P = oracle.legendre_column(N, m, mu, cl)
        G = 0.5 * (P * w) @ P.T
        assert np.abs(G - np.eye(G.shape[0])).max() < 2e-13


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("im,mmax", [(16, 5), (64, 21), (128, 42), (256, 85), (512, 170)])
def test_ring_transforms_vs_numpy(im, mmax):
    rng = np.random.default_rng(im)
    x = rng.standard_normal(im)
    ref = np.fft.rfft(x)[: mmax + 1] / im
    ref[0] = ref[0].real
    assert np.abs(oracle.ring_fwd(x, mmax) - ref).max() < 1e-15 * im
    assert np.abs(oracle.ring_fwd(x, mmax, direct=True) - ref).max() < 1e-15 * im
    F = (rng.standard_normal(mmax + 1) + 1j * rng.standard_normal(mmax + 1))
    F[0] = F[0].real
    full = np.zeros(im // 2 + 1, dtype=complex)
    full[: mmax + 1] = F * im
    ref_x = np.fft.irfft(full, n=im)
    assert np.abs(oracle.ring_inv(F, im) - ref_x).max() < 1e-14 * np.sqrt(im)
    assert np.abs(oracle.ring_inv(F, im, direct=True) - ref_x).max() < 1e-14 * np.sqrt(im)


# ---------------------------------------------------------------- full transform
@pytest.mark.parametrize("case", ["golden_T5.npz", "golden_T10.npz"])
def test_transforms_vs_mpmath_golden(case):
    g = np.load(os.path.join(GOLD, case))
    N, im, jm = int(g["nmax"]), int(g["im"]), int(g["jm"])
    p = oracle.Plan(N, im, jm)
    spec = g["spec"]
    assert rel(p.spec_to_grid(spec), g["grid"]) < 1e-14
    assert rel(p.spec_to_grid(spec, kind=1), g["gradlon"]) < 1e-13
    assert rel(p.spec_to_grid(spec, kind=2), g["gradlat"]) < 1e-13
    assert rel(p.grid_to_spec(g["noisy"]), g["fwd"]) < 1e-14
    assert rel(p.grid_to_spec(g["grid"]), spec) < 1e-14


def test_brute_force_scipy_sph_harm():
    """T8 synthesis against scipy.special.sph_harm_y (orthonormal, with
    Condon-Shortley): Pt_n^m e^{im lam} = sqrt(4 pi) (-1)^m Y_n^m."""
    N, im, jm = 8, 32, 16
    p = oracle.Plan(N, im, jm)
    mu, w, cl = oracle.gauss(jm)
    theta = np.arccos(mu)[:, None]
    lam = (2 * np.pi * np.arange(im) / im)[None, :]
    spec = synth.spectra(N, 2, cfg=5)
    ns, ms = synth.coef_index(N)
    ref = np.zeros((2, jm, im))
    for k in range(ns.size):
        n, m = int(ns[k]), int(ms[k])
        Y = np.sqrt(4 * np.pi) * (-1) ** m * sps.sph_harm_y(n, m, theta, lam)
        for f in range(2):
            c = spec[k, f]
            ref[f] += (c * Y).real if m == 0 else 2 * (c * Y).real
    assert rel(p.spec_to_grid(spec), ref) < 1e-13


@pytest.mark.parametrize("N,im,jm,nf", [(21, 64, 32, 49), (42, 128, 64, 79), (85, 256, 128, 97), (170, 512, 256, 4)])
def test_roundtrip_parseval_mean_rotation(N, im, jm, nf):
    p = oracle.Plan(N, im, jm)
    s = synth.state_spectra(N, (nf - 1) // 3) if nf > 4 else synth.spectra(N, nf)
    g = p.spec_to_grid(s)
    assert rel(p.grid_to_spec(g), s) < 1e-13
    ns, ms = synth.coef_index(N)
    par = (np.abs(s) ** 2 * np.where(ms == 0, 1, 2)[:, None]).sum(0)
    assert np.abs(p.avr_lonlat(g ** 2) - par).max() < 1e-13 * par.max()
    # global mean is the (0,0) coefficient
    s2 = s.copy()
    s2[0] = np.arange(nf) + 0.5
    assert np.abs(p.avr_lonlat(p.spec_to_grid(s2)) - s2[0].real).max() < 1e-13
    # rotating by one longitude step multiplies s_n^m by e^{-i m 2pi/im}
    rolled = np.roll(g, 1, axis=2)
    rot = s * np.exp(-1j * ms * 2 * np.pi / im)[:, None]
    assert rel(p.grid_to_spec(rolled), rot) < 1e-13


def test_noisy_grid_idempotence():
    N, im, jm = 42, 128, 64
    p = oracle.Plan(N, im, jm)
    mu, w, cl = oracle.gauss(jm)
    g = synth.temperature_grid(cl, im, 4)
    assert g.min() > 180 and g.max() < 320
    s1 = p.grid_to_spec(g)
    s2 = p.grid_to_spec(p.spec_to_grid(s1))
    assert rel(s2, s1) < 1e-13


def test_gradient_closed_forms():
    """f = sin(lat) -> df/dlat = cos(lat); f = cos(lat) cos(lon) ->
    (1/cos) df/dlon = -sin(lon), df/dlat = -sin(lat) cos(lon)."""
    N, im, jm = 21, 64, 32
    p = oracle.Plan(N, im, jm)
    mu, w, cl = oracle.gauss(jm)
    lam = 2 * np.pi * np.arange(im) / im
    f1 = np.broadcast_to(mu[:, None], (jm, im))
    f2 = cl[:, None] * np.cos(lam)[None, :]
    s = p.grid_to_spec(np.stack([f1, f2]))
    gl = p.spec_to_grid(s, kind=1)
    gt = p.spec_to_grid(s, kind=2)
    assert np.abs(gt[0] - cl[:, None]).max() < 1e-13
    assert np.abs(gl[0]).max() < 1e-13
    assert np.abs(gl[1] + np.sin(lam)[None, :]).max() < 1e-13
    assert np.abs(gt[1] + mu[:, None] * np.cos(lam)[None, :]).max() < 1e-13


def test_spml_packing_orthonormal_real_basis():
    """wa(l), l = n(n+1)+m (0-based), holds coefficients of the real harmonics
    sqrt2 Pt cos(m lam) (m>0), Pt (m=0), sqrt2 Pt sin(|m| lam) (m<0)."""
    N, im, jm = 5, 16, 8
    p = oracle.Plan(N, im, jm)
    mu, w, cl = oracle.gauss(jm)
    lam = 2 * np.pi * np.arange(im) / im
    fields = np.stack([
        np.ones((jm, im)),                                   # (0,0)
        math.sqrt(3) * np.broadcast_to(mu[:, None], (jm, im)),  # (1,0)
        math.sqrt(3) * cl[:, None] * np.cos(lam)[None, :],   # (1,1)
        math.sqrt(3) * cl[:, None] * np.sin(lam)[None, :],   # (1,-1)
    ])
    wa = oracle.spml_pack(p.grid_to_spec(fields), N)
    expect = np.zeros_like(wa)
    expect[0, 0] = expect[1, 2] = expect[2, 3] = expect[3, 1] = 1.0
    assert np.abs(wa - expect).max() < 1e-14
    s = synth.spectra(N, 3)
    assert np.abs(oracle.spml_unpack(oracle.spml_pack(s, N), N) - s).max() < 1e-16
    # Parseval in the real packing: mean square = sum wa^2
    g = p.spec_to_grid(s)
    assert np.abs(p.avr_lonlat(g ** 2) - (oracle.spml_pack(s, N) ** 2).sum(1)).max() < 1e-14


# ---- flux divergence and Laplacian powers (wa_DivLon_xya, wa_DivLat_xya,
#      wa_Lapla_wa, wa_LaplaInv_wa; DCPAM dynamics, atm/dcpam_main_mod.f90:2173-2180)

def _cidx(N, n, m):
    return m * (N + 1) - m * (m - 1) // 2 + n - m


def test_divergence_closed_forms():
    """DivLat(cos lat) = -2 sin lat = -(2/sqrt3) Pt_1^0; DivLon(cos^2 lat sin lon) =
    cos lat cos lon = Pt_1^1 cos lon / sqrt(3/2), i.e. s_1^1 = 1/sqrt6."""
    N, im, jm = 21, 64, 32
    p = oracle.Plan(N, im, jm)
    mu, w, cl = oracle.gauss(jm)
    lon = 2 * np.pi * np.arange(im) / im
    s = p.grid_to_div(None, np.broadcast_to(cl[:, None], (jm, im))[None].copy())
    want = np.zeros_like(s)
    want[_cidx(N, 1, 0), 0] = -2 / np.sqrt(3)
    assert np.abs(s - want).max() < 1e-14
    s = p.grid_to_div((cl[:, None] ** 2 * np.sin(lon)[None, :])[None].copy(), None)
    want = np.zeros_like(s)
    want[_cidx(N, 1, 1), 0] = 1 / np.sqrt(6)
    assert np.abs(s - want).max() < 1e-14


@pytest.mark.parametrize("N,im,jm", [(5, 16, 8), (21, 64, 32), (42, 128, 64), (85, 256, 128)])
def test_divergence_of_gradient_is_laplacian(N, im, jm):
    """div(grad f) = -n(n+1) f_n exactly on the Gauss grid: the lon and lat
    integrands are polynomials of degree < 2 jm for band-limited f."""
    p = oracle.Plan(N, im, jm)
    spec = synth.spectra(N, 3, cfg=8)
    gl, gt = p.spec_to_grid(spec, kind=1), p.spec_to_grid(spec, kind=2)
    lap = oracle.spec_lapla(spec, N, 1)
    assert rel(p.grid_to_div(gl, gt), lap) < 1e-13
    # each term alone is exact too: div_lon(grad_lon f) + div_lat(grad_lat f)
    assert rel(p.grid_to_div(gl, None) + p.grid_to_div(None, gt), lap) < 1e-13


def test_rotational_flow_vorticity_and_divergence():
    """u = -dpsi/dlat, v = (1/cos) dpsi/dlon: divergence 0, vorticity
    (1/cos) dv/dlon - (1/cos) d(u cos)/dlat = Laplacian of psi."""
    N, im, jm = 42, 128, 64
    p = oracle.Plan(N, im, jm)
    psi = synth.spectra(N, 4, cfg=9)
    u, v = -p.spec_to_grid(psi, kind=2), p.spec_to_grid(psi, kind=1)
    lap = oracle.spec_lapla(psi, N, 1)
    scale = np.abs(lap).max()
    assert np.abs(p.grid_to_div(u, v)).max() < 1e-13 * scale
    assert rel(p.grid_to_div(v, -u), lap) < 1e-13


def test_laplacian_powers():
    N = 21
    spec = synth.spectra(N, 2, cfg=10)
    n = np.concatenate([np.arange(m, N + 1) for m in range(N + 1)])
    e = -(n * (n + 1.0))
    for power in (1, 2, 4):
        assert np.allclose(oracle.spec_lapla(spec, N, power), spec * (e ** power)[:, None], rtol=1e-15, atol=0)
    inv = oracle.spec_lapla(oracle.spec_lapla(spec, N, 1), N, -1)
    want = spec.copy()
    want[0] = 0.0
    assert rel(inv, want) < 1e-15
    assert np.all(oracle.spec_lapla(spec, N, -2)[0] == 0.0)


# ---- coupler remap (interpolate_data_latlon, common/interpolation_data_latlon_mod.f90:274-306)

def _remap_case(seed=0, nsend=300, nrecv=200, nnz=1500):
    rng = np.random.default_rng(seed)
    si = rng.integers(0, nsend, nnz)
    ri = rng.integers(0, nrecv, nnz)
    cf = rng.standard_normal(nnz)
    return si, ri, cf


def test_remap_matches_brute_force_loop():
    si, ri, cf = _remap_case(nsend=17, nrecv=11, nnz=60)
    send = np.random.default_rng(1).standard_normal((3, 17))
    want = np.zeros((3, 11))
    for d in range(3):
        for i in range(len(cf)):
            want[d, ri[i]] = want[d, ri[i]] + send[d, si[i]] * cf[i]
    assert np.array_equal(oracle.remap(si, ri, cf, 11, send), want)


def test_remap_vs_scipy_sparse_and_partition_of_unity():
    import scipy.sparse as sp
    si, ri, cf = _remap_case()
    send = np.random.default_rng(2).standard_normal((4, 300))
    A = sp.coo_matrix((cf, (ri, si)), shape=(200, 300)).tocsr()
    got = oracle.remap(si, ri, cf, 200, send)
    assert np.abs(got - (A @ send.T).T).max() < 1e-13
    # weights normalised per receiving point reproduce a constant field
    rows = np.bincount(ri, weights=np.abs(cf), minlength=200)
    w = np.abs(cf) / rows[ri]
    const = oracle.remap(si, ri, w, 200, np.full((1, 300), 3.5))
    hit = np.bincount(ri, minlength=200) > 0
    assert np.allclose(const[0, hit], 3.5, rtol=0, atol=1e-14)
    assert np.all(const[0, ~hit] == 0.0)
