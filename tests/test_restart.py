"""Restart state (sht_restart_write / sht_restart_read): the ten grid fields
RestartFileOutput takes (atm/dcpam_main_mod.f90:3063-3066) and
RestartFileGet returns (:4134-4137), at the T21L16 grid of
dcpam_APEI07Couple_T21L16.conf (64 x 32 x 16) and with several tracers.
Checked by a bitwise round trip through the library, by scipy's independent
netCDF-3 reader (dimensions, variable names and shapes, attributes, values),
and by the error paths (wrong grid, truncated file, missing file)."""
import os

import numpy as np
import pytest

import oracle
import shtamd

netcdf = pytest.importorskip("scipy.io").netcdf_file


def _state(im, jm, km, nq, seed):
    rng = np.random.default_rng(seed)
    return {nm: rng.standard_normal(shtamd._restart_shape(nm, im, jm, km, nq)) * (1e5 if nm.startswith("Ps") else 10)
            for nm in shtamd.RESTART_FIELDS}


@pytest.mark.parametrize("im,jm,km,nq", [(64, 32, 16, 1), (128, 64, 26, 3), (16, 2, 1, 1)])
def test_restart_round_trip_and_independent_reader(tmp_path, im, jm, km, nq):
    mu, w, cl = oracle.gauss(jm)
    lon = 2 * np.pi * np.arange(im) / im
    lat = np.arctan2(mu, cl)
    sig = np.linspace(0.99, 0.01, km)
    st = _state(im, jm, km, nq, seed=im + km)
    path = tmp_path / "rst.nc"
    shtamd.restart_write(path, 86400.0 * 30, lon, lat, sig, st, nq=nq)
    t, back = shtamd.restart_read(path, im, jm, km, nq)
    assert t == 86400.0 * 30
    for nm in shtamd.RESTART_FIELDS:
        assert np.array_equal(back[nm], st[nm]), nm
    with netcdf(str(path), "r", mmap=False) as f:
        assert f.version_byte == 2
        assert dict(f.dimensions) == {"lon": im, "lat": jm, "sig": km, "nq": nq}
        assert float(f.time) == 86400.0 * 30
        assert np.array_equal(f.variables["lon"][:], lon) and np.array_equal(f.variables["lat"][:], lat)
        assert np.array_equal(f.variables["sig"][:], sig)
        for nm in shtamd.RESTART_FIELDS:
            v = f.variables[nm]
            assert v.dimensions == {2: ("lat", "lon"), 3: ("sig", "lat", "lon"),
                                    4: ("nq", "sig", "lat", "lon")}[st[nm].ndim]
            assert np.array_equal(v[:], st[nm]), nm
            assert v.units and v.long_name


def test_restart_errors(tmp_path):
    im, jm, km = 64, 32, 16
    st = _state(im, jm, km, 1, seed=3)
    path = tmp_path / "rst.nc"
    shtamd.restart_write(path, 0.0, np.zeros(im), np.zeros(jm), np.zeros(km), st)
    with pytest.raises(shtamd.ShtError):
        shtamd.restart_read(path, im, jm, km + 1)  # another grid
    with pytest.raises(shtamd.ShtError):
        shtamd.restart_read(tmp_path / "absent.nc", im, jm, km)
    data = open(path, "rb").read()
    for cut in (3, 40, 400, len(data) // 2, len(data) - 8):
        bad = tmp_path / f"cut{cut}.nc"
        open(bad, "wb").write(data[:cut])
        with pytest.raises(shtamd.ShtError):
            shtamd.restart_read(bad, im, jm, km)
    with pytest.raises(ValueError):
        shtamd.restart_write(path, 0.0, np.zeros(im), np.zeros(jm), np.zeros(km), {**st, "UB": st["UB"][:-1]})
    assert not os.path.exists(str(path) + ".tmp")
