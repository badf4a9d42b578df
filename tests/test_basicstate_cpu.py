"""Host basic-state construction (lbm_isca_amd.basicstate) on the CPU:
the natural spline against scipy's, z2ps and p2s on profiles they must
reproduce exactly, the pressure-level file reader on a file written in the
documented record layout, and the balance of the synthetic jets."""

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from lbm_isca_amd import basicstate as B
from lbm_isca_amd.constants import ER, GRAV, OMEGA, PLEV17, RDGAS


def test_spline_matches_scipy_natural_spline():
    rng = np.random.default_rng(0)
    x = np.log(PLEV17)                       # decreasing
    z = rng.standard_normal((5, x.size)).cumsum(1)
    xi = rng.uniform(x[-1], x[0], (5, 30))
    got = B.spline(x, z, xi)
    for r in range(5):
        cs = CubicSpline(x[::-1], z[r, ::-1], bc_type="natural")
        np.testing.assert_allclose(got[r], cs(xi[r]), rtol=1e-12, atol=1e-12)
    # knots are reproduced, outside the range the end values are held
    np.testing.assert_allclose(B.spline(x, z, np.tile(x, (5, 1))), z, rtol=1e-13, atol=1e-13)
    out = B.spline(x, z, np.array([[x[0] + 0.3, x[-1] - 0.3]] * 5))
    np.testing.assert_array_equal(out[:, 0], z[:, 0])
    np.testing.assert_array_equal(out[:, 1], z[:, -1])
    miss = B.spline(x, z, np.array([[x[0] + 0.3, x[-1] - 0.3]] * 5), omisb=True, omist=True)
    assert np.all(miss == B.XMISS)


def test_z2ps_exact_for_log_linear_heights():
    H = 7500.0
    z = H * np.log(1000.0 / PLEV17)          # z linear in ln p
    zs = np.array([[-200.0, 0.0, 350.0], [1500.0, 3000.0, 9000.0]])
    ps = B.z2ps(np.broadcast_to(z, zs.shape + z.shape), zs)
    np.testing.assert_allclose(ps, 1000.0 * np.exp(-zs / H), rtol=1e-13)


def test_p2s_exact_for_fields_linear_in_ln_p():
    sig = np.linspace(0.975, 0.025, 20)
    ps = np.array([[1000.0, 950.0], [1010.0, 870.0]])
    g = 3.0 + 2.0 * np.log(PLEV17)
    out = B.p2s(np.broadcast_to(g, ps.shape + g.shape), ps, sig)
    lnp = np.log(ps[..., None] * sig)
    inside = (lnp <= np.log(PLEV17[0])) & (lnp >= np.log(PLEV17[-1]))
    want = np.where(inside, 3.0 + 2.0 * lnp, np.where(lnp > np.log(PLEV17[0]), g[0], g[-1]))
    np.testing.assert_allclose(out, want, rtol=1e-12, atol=1e-12)


def test_zmsw_splits_zonal_mean_and_waves():
    rng = np.random.default_rng(1)
    f = rng.standard_normal((4, 8, 3))
    zm = B.zmsw(f, ozm=True)
    sw = B.zmsw(f, osw=True)
    np.testing.assert_allclose(zm + sw, f, rtol=1e-14)
    np.testing.assert_allclose(sw.mean(axis=1), 0.0, atol=1e-14)
    assert B.zmsw(f) is f


def _write_records(path, months, nlat, nlon):
    """months: list of dicts var -> (nlev, nlat_south_first, nlon) float32."""
    npts = nlat * nlon
    mark = np.array([npts * 4], dtype=">i4").view(">f4")
    with open(path, "wb") as fh:
        for mo in months:
            for var in B.BS_VARS:
                for lev in mo[var]:
                    fh.write(np.concatenate([mark, lev.astype(">f4").ravel(), mark]).astype(">f4").tobytes())


def test_read_pressure_levels_averages_months(tmp_path):
    nlat, nlon, nlev = 4, 8, 17
    rng = np.random.default_rng(2)
    months = [{v: rng.standard_normal((nlev, nlat, nlon)).astype(np.float32) for v in B.BS_VARS}
              for _ in range(3)]
    path = tmp_path / "bs.bin"
    _write_records(path, months, nlat, nlon)
    got = B.read_pressure_levels(str(path), kmo=2, navg=2, nlon=nlon, nlat=nlat, nlev=nlev)
    for v in B.BS_KEEP:
        want = sum(np.transpose(months[m][v], (1, 2, 0)).astype(np.float64) / 2.0 for m in (1, 2))
        np.testing.assert_allclose(got[v], want[::-1], rtol=1e-14)      # north first
    # wrapping past the last month
    got = B.read_pressure_levels(str(path), kmo=3, navg=2, nlon=nlon, nlat=nlat, nlev=nlev)
    want = sum(np.transpose(months[m]["t"], (1, 2, 0)).astype(np.float64) / 2.0 for m in (2, 0))
    np.testing.assert_allclose(got["t"], want[::-1], rtol=1e-14)
    with pytest.raises(ValueError):
        B.read_pressure_levels(str(path), nlon=nlon + 1, nlat=nlat, nlev=nlev)


def test_balanced_jets_are_geostrophic_and_hydrostatic():
    lat = np.radians(np.linspace(80, -80, 161))
    lon = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    f, par = B.balanced_jets(lat, lon, 2, seed=3)
    assert np.all(par.U >= 20) and np.all(par.U <= 50)
    x = np.log(1000.0 / PLEV17)
    for e in range(2):
        zbar = f["z"][e].mean(axis=1)          # zonal mean (J, K): waves average out
        ubar = f["u"][e].mean(axis=1)
        tbar = f["t"][e].mean(axis=1)
        # geostrophic zonal-mean wind: u = -(g / (f a)) dz/dlat where |lat| > 20
        dz = np.gradient(zbar, lat, axis=0)
        fc = 2 * OMEGA * np.sin(lat)
        sel = np.abs(np.degrees(lat)) > 20
        with np.errstate(divide="ignore", invalid="ignore"):
            ug = -(GRAV / (fc[:, None] * ER)) * dz
        assert np.max(np.abs(ug[sel] - ubar[sel])) < 0.02 * np.max(np.abs(ubar))
        # hydrostatic: T = (g/R) dz/dx between levels in the troposphere
        tmid = GRAV / RDGAS * np.diff(zbar, axis=1) / np.diff(x)
        lay = PLEV17[1:] > 250
        tav = 0.5 * (tbar[:, 1:] + tbar[:, :-1])
        assert np.max(np.abs(tmid[:, lay] - tav[:, lay])) < 1.0
    assert np.all(f["q"] == 0)


def test_balanced_jets_are_statically_stable():
    """No layer of any member is steeper than 80 % of the dry adiabatic
    lapse rate g/cp on the pressure levels, and potential temperature rises
    with height on the sigma levels after the spline (an unstable layer
    would give the linear model convective modes)."""
    from lbm_isca_amd.constants import CP, KAPPA
    lat = np.radians(np.linspace(88, -88, 89))
    lon = np.linspace(0, 2 * np.pi, 32, endpoint=False)
    f, _ = B.balanced_jets(lat, lon, 16, seed=1)
    lapse = -np.diff(f["t"], axis=-1) / np.diff(f["z"], axis=-1)
    assert lapse.max() < 0.8 * GRAV / CP, lapse.max()
    sig = 0.5 * (np.linspace(1, 0, 21)[1:] + np.linspace(1, 0, 21)[:-1])
    ps = B.z2ps(f["z"], 0.0)
    t = B.p2s(f["t"], ps, sig)
    theta = t * (ps[..., None] * sig / 1000.0) ** (-KAPPA)
    assert np.diff(theta, axis=-1).min() > 0.5
    s, ds = B.jet_profile(PLEV17)
    assert s[0] == 0.0 and abs(s[-1] - 1.0) < 1e-12 and np.all(np.diff(s) >= 0)


def test_balanced_jet_waves_are_geostrophic_and_smooth_at_the_equator():
    """The stationary waves' winds satisfy f u_w = -(g/a) dz_w/dlat and
    f v_w = (g/(a cos)) dz_w/dlon with the true Coriolis parameter at every
    latitude, the equator included, and are continuous across it (no vortex
    sheet): on a 0.25 deg grid no wave wind changes by more than 0.5 m/s
    between neighbouring latitudes."""
    lat = np.radians(np.linspace(60, -60, 481))         # includes lat = 0
    lon = np.linspace(0, 2 * np.pi, 96, endpoint=False)
    f, par = B.balanced_jets(lat, lon, 8, seed=1)
    assert np.all(np.isfinite(f["u"])) and np.all(np.isfinite(f["v"]))
    fc = 2 * OMEGA * np.sin(lat)
    for e in range(8):
        k = -1                                              # strongest waves: s(p) = 1
        zw = f["z"][e][..., k] - f["z"][e][..., k].mean(axis=1, keepdims=True)
        uw = f["u"][e][..., k] - f["u"][e][..., k].mean(axis=1, keepdims=True)
        vw = f["v"][e][..., k]
        dz_lat = np.gradient(zw, lat, axis=0, edge_order=2)
        kk = np.arange(lon.size // 2 + 1)
        dz_lon = np.fft.irfft(1j * kk * np.fft.rfft(zw, axis=1), n=lon.size, axis=1)
        scale = np.abs(uw).max() * np.abs(fc).max()
        inner = slice(2, -2)
        np.testing.assert_allclose((fc[:, None] * uw)[inner], (-(GRAV / ER) * dz_lat)[inner],
                                   atol=2e-3 * scale)
        np.testing.assert_allclose(fc[:, None] * vw,
                                   (GRAV / ER) * dz_lon / np.cos(lat)[:, None],
                                   atol=2e-3 * np.abs(vw).max() * np.abs(fc).max())
        # an untapered envelope with a floor on |f| jumps by up to ~25 m/s here
        assert np.abs(np.diff(uw, axis=0)).max() < 0.5
        assert np.abs(np.diff(vw, axis=0)).max() < 0.5
    # the taper leaves the waves near full strength at the jets
    assert np.all(par.lat_jet >= np.radians(25.0))
