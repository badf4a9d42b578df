"""Host-side forcing readers, the MJO call sequence and the history writers
(lbm_isca_amd.forcing / output) on the CPU.

Files are written in the documented layouts (Fortran sequential records:
big-endian int32 length, big-endian real*4 values, the length again) and
read back; the MJO sequence is checked against the oracle's restatement of
the reference's counter state machine and against closed forms."""

import os
import types

import numpy as np
import pytest
from scipy.io import netcdf_file

from lbm_isca_amd import forcing as F
from lbm_isca_amd import output as O
from lbm_isca_amd.constants import GRAV_GT2GR, RHO00
from oracle.history import MJOSequence


def write_records(path, recs):
    with open(path, "wb") as fh:
        for r in recs:
            data = np.ascontiguousarray(r, dtype=">f4")
            mark = np.array([data.size * 4], dtype=">i4").tobytes()
            fh.write(mark + data.tobytes() + mark)


def test_read_fortran_records(tmp_path):
    nlat, nlon = 4, 8
    recs = [np.full((nlat, nlon), k, np.float32) + np.arange(nlon) for k in range(5)]
    p = str(tmp_path / "r.bin")
    write_records(p, recs)
    got = F.read_fortran_records(p, nlon, nlat)
    assert got.shape == (5, nlat, nlon)
    np.testing.assert_array_equal(got, np.array(recs))
    np.testing.assert_array_equal(F.read_fortran_records(p, nlon, nlat, count=2, offset=3),
                                  np.array(recs[3:5]))
    with pytest.raises(ValueError):
        F.read_fortran_records(p, nlon + 1, nlat)        # not a whole number of records
    raw = bytearray(open(p, "rb").read())
    raw[0:4] = np.array([7], dtype=">i4").tobytes()       # corrupt the first marker
    open(p, "wb").write(bytes(raw))
    with pytest.raises(ValueError, match="length word"):
        F.read_fortran_records(p, nlon, nlat)


def test_steady_forcing_from_file(tmp_path):
    nlat, nlon, nlev = 4, 8, 3
    per = 3 * nlev + 1
    recs = [np.full((nlat, nlon), 100 * c + r, np.float32) + np.arange(nlat)[:, None]
            for c in range(2) for r in range(per)]
    p = str(tmp_path / "s.bin")
    write_records(p, recs)
    sf = F.SteadyForcing.from_file(p, nlev=nlev, ncase=2, nlon=nlon, nlat=nlat)
    for k in range(nlev):
        np.testing.assert_array_equal(sf.vor[..., k], recs[per + k])
        np.testing.assert_array_equal(sf.div[..., k], recs[per + nlev + k])
        np.testing.assert_array_equal(sf.t[..., k], recs[per + 2 * nlev + k])
    np.testing.assert_array_equal(sf.ps, recs[per + 3 * nlev])
    flip = F.SteadyForcing.from_file(p, nlev=nlev, ncase=1, nlon=nlon, nlat=nlat, yflip=True)
    np.testing.assert_array_equal(flip.ps, recs[3 * nlev][::-1])
    assert F.autodetect(p, nlev=nlev, nlon=nlon, nlat=nlat) == {
        "nrec": 2 * per, "steady_cases": 2}


@pytest.mark.parametrize("ndays,nt,spinup", [(5, 7, 3), (3, 4, 1), (2, 6, 10)])
def test_mjo_sequence_matches_the_reference_state_machine(ndays, nt, spinup):
    rng = np.random.default_rng(ndays)
    g32 = rng.standard_normal((ndays, 3, 4, 2)).astype(np.float32)
    for g in (g32, g32.astype(np.float64)):
        host, ref = F.MJOForcing(gtf=g, nt=nt, spinup=spinup), MJOSequence(g, nt, spinup)
        for c in range(1, spinup + nt * (ndays + 2)):
            a, b = host.step(), ref.step()
            assert a.dtype == b.dtype and np.array_equal(a, b), (c, a.dtype, b.dtype)
            assert host.day == (ref.it if not ref.spinning else 1), c


def test_mjo_closed_forms():
    g = np.arange(4, dtype=np.float64)[:, None, None, None] * np.ones((4, 2, 3, 2))
    m = F.MJOForcing(gtf=g, nt=8, spinup=5)
    assert all(np.all(m.field_at(c) == 0.0) for c in range(1, 5))       # day 1 held
    for c in range(5, 5 + 8 * 6):
        q = c - 5
        want = min(q / 8.0, 3.0)           # day index in [0, 3], clamped at the last day
        np.testing.assert_allclose(m.field_at(c), want, rtol=0, atol=1e-15)


def test_mjo_from_file(tmp_path):
    nlat, nlon, nlev, ndays = 4, 8, 2, 3
    recs = [np.full((nlat, nlon), 10 * d + k, np.float32) for d in range(ndays) for k in range(nlev)]
    p = str(tmp_path / "m.bin")
    write_records(p, recs)
    m = F.MJOForcing(p, nlev=nlev, ndays=ndays, nt=4, spinup=2, nlon=nlon, nlat=nlat, scale=0.5)
    assert m.gtf.dtype == np.float32 and m.gtf.shape == (ndays, nlat, nlon, nlev)
    np.testing.assert_array_equal(m.gtf[2, :, :, 1], 0.5 * recs[2 * nlev + 1])
    with pytest.raises(ValueError, match="do not hold"):
        F.MJOForcing(p, nlev=nlev, ndays=ndays + 1, nlon=nlon, nlat=nlat)
    assert F.autodetect(p, nlev=nlev, nlon=nlon, nlat=nlat)["mjo_days"] == ndays


def fake_grids(nlat=4, nlon=8, nlev=3):
    sp = types.SimpleNamespace(lat_deg=np.linspace(60, -60, nlat), lon_deg=np.arange(nlon) * 45.0,
                               nlon=nlon, nlat=nlat)
    vg = types.SimpleNamespace(sig=np.linspace(0.9, 0.1, nlev), kmax=nlev)
    return sp, vg


def fields_for(day, nlat=4, nlon=8, nlev=3, lead=()):
    rng = np.random.default_rng(day)
    f = {k: rng.standard_normal(lead + (nlat, nlon, nlev)) for k in O.VARS if k != "p"}
    f["p"] = rng.standard_normal(lead + (nlat, nlon))
    return f


def test_grads_writer_layout(tmp_path):
    sp, vg = fake_grids()
    p = str(tmp_path / "out.grd")
    w = O.GradsWriter(p, sp, vg)
    days = [fields_for(1), fields_for(2)]
    for d, f in enumerate(days, 1):
        w.add(d, f)
    w.close()
    rec = F.read_fortran_records(p, 8, 4)
    assert rec.shape[0] == 2 * (7 * 3 + 1)
    for t, f in enumerate(days):
        base = t * 22
        for v, name in enumerate(O.VARS[:-1]):
            want = f[name]
            if name == "z":
                want = want + f["p"][..., None] * 1.0e2 / (RHO00 * GRAV_GT2GR)
            for k in range(3):
                np.testing.assert_array_equal(rec[base + 3 * v + k], want[..., k].astype(np.float32))
        np.testing.assert_array_equal(rec[base + 21], f["p"].astype(np.float32))
    ctl = open(str(tmp_path / "out.ctl")).read().splitlines()
    assert ctl[0] == "DSET ^out.grd" and ctl[1] == "OPTIONS SEQUENTIAL YREV BIG_ENDIAN"
    assert "XDEF 8 LINEAR 0. 45" in ctl and "YDEF 4 LEVELS" in ctl and "ZDEF 3 LEVELS" in ctl
    assert "TDEF 2 LINEAR 01jan0001 1dy" in ctl and "VARS 8" in ctl and ctl[-1] == "ENDVARS"
    assert ctl[-2].startswith("p      0 99 surface pressure")
    # latitudes listed south -> north (YREV)
    ylat = [float(x) for x in ctl[ctl.index("YDEF 4 LEVELS") + 1].split()]
    assert ylat == sorted(ylat)


def test_netcdf_writer_and_members(tmp_path):
    sp, vg = fake_grids()
    p = str(tmp_path / "out.nc")
    ens = [fields_for(1, lead=(3,)), fields_for(2, lead=(3,))]
    w = O.NetCDFWriter(p, sp, vg, member=2)
    m = O.MultiWriter(w, None)
    for d, f in enumerate(ens, 1):
        m.add(d, f)
    m.close()
    with netcdf_file(p, "r", mmap=False) as nc:
        assert nc.variables["u"].shape == (2, 3, 4, 8)
        assert nc.variables["p"].shape == (2, 4, 8)
        np.testing.assert_array_equal(nc.variables["time"][:], [1.0, 2.0])
        np.testing.assert_array_equal(nc.variables["u"][1], np.moveaxis(ens[1]["u"][2], -1, 0)
                                      .astype(np.float32))
        z = ens[0]["z"][2] + ens[0]["p"][2][..., None] * 1.0e2 / (RHO00 * GRAV_GT2GR)
        np.testing.assert_array_equal(nc.variables["z"][0], np.moveaxis(z, -1, 0).astype(np.float32))
        assert nc.variables["w"].units == b"hPa/s"
