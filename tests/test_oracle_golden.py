"""Pin the numpy FP64 oracle against reference-generated vectors
(``tests/golden/reference_vectors.npz``, written by
``tests/golden/make_reference_vectors.py`` from the reference package) and
against the reference-held Gill golden (``analysis/gill_isca.json:4-14`` and
``examples/example_1_gill.ipynb:129``).
"""

import types

import numpy as np
import pytest

import cases
from oracle import OracleConfig, OracleModel, tendencies
from oracle.dyn import column_sum, sigma_dot
from oracle_helpers import oracle_setup, oracle_state, rel_err

LV = list(cases.SAMPLE_LEVELS)
NAMES = ("vor", "div", "t", "pi", "q")


@pytest.fixture(scope="module")
def setup():
    return oracle_setup()


def test_vertical_operators(setup, golden):
    _, vg, _, _ = setup
    for name in ("alpha", "sig", "cpg", "w3", "wP", "dsigm", "ww"):
        np.testing.assert_allclose(getattr(vg, name), golden["vg_" + name],
                                   rtol=1e-13, atol=1e-13, err_msg=name)


def test_basic_state_terms(setup, golden):
    sht, vg, bs, bst = setup
    np.testing.assert_allclose(vg.tbar, golden["bs_tbar"], rtol=1e-13)
    np.testing.assert_allclose(vg.tmtx, golden["bs_tmtx"], rtol=1e-13, atol=1e-16)
    arrs = {"x5": bst.x5[0], "sdot": bst.sdot[0], "ssum": bst.ssum[0],
            "padv": bst.padv[0], "vor_g": bst.vor[0], "div": bst.div[0]}
    for name, arr in arrs.items():
        idx = golden["bs_%s_idx" % name]
        ref = golden["bs_%s_val" % name]
        got = arr.ravel()[idx]
        assert rel_err(got, ref) < 1e-10, name
    assert rel_err(bst.pix[0, :, :, 0], golden["bs_pix"]) < 1e-11
    assert rel_err(bst.piy[0, :, :, 0], golden["bs_piy"]) < 1e-11


@pytest.mark.parametrize("lin", [False, True])
def test_tendencies_match_reference(setup, golden, lin):
    sht, vg, bs, bst = setup
    st = oracle_state(sht, bs, vg.kmax, 0)
    tend, diag = tendencies(sht, vg, bst, st, strictly_linear=lin)
    tag = "lin" if lin else "full"
    for name, x in zip(NAMES, tend):
        x = x[0]
        ref = golden["tend_%s_%s" % (tag, name)]
        got = x if x.ndim == 2 else x[:, :, LV]
        assert rel_err(got, ref) < 1e-10, (name, rel_err(got, ref))
        norm = np.sqrt((np.abs(x) ** 2).sum(axis=(0, 1)))
        np.testing.assert_allclose(norm, golden["tend_%s_%s_norm" % (tag, name)], rtol=1e-10)
    if not lin:
        om = diag["omega"][0]
        assert rel_err(om.ravel()[golden["diag_omega_idx"]], golden["diag_omega_val"]) < 1e-10


def test_semi_implicit_step_matches_reference(setup, golden):
    sht, vg, bs, _ = setup
    model = OracleModel(bs, 1, OracleConfig())
    st1 = oracle_state(sht, bs, vg.kmax, 0)
    st2 = oracle_state(sht, bs, vg.kmax, 1)
    tend, _ = tendencies(model.sht, model.vg, model.bst, st1)
    prev = st2.to_spectral(model.sht)
    for dt in (1200.0, 300.0):
        new = model.stepper.step(dt, prev, tend)
        for name in NAMES:
            x = new[name][0]
            got = x if x.ndim == 2 else x[:, :, LV]
            ref = golden["step%d_%s" % (int(dt), name)]
            assert rel_err(got, ref) < 1e-10, (dt, name, rel_err(got, ref))
    cond = model.stepper.matrices(1200.0)["cond"]
    np.testing.assert_allclose(cond, golden["semi_cond_1200"], rtol=1e-8)


@pytest.mark.parametrize("nsteps", [1, 10, 72])
def test_short_runs_match_reference(setup, golden, nsteps):
    """Start-up half steps, raw (non-band-limited) initial T, RA filter and the
    exact ps' <-> pi' round trip, against the reference's own run."""
    sht, vg, bs, _ = setup
    st1 = oracle_state(sht, bs, vg.kmax, 0)
    rng = np.random.default_rng(int(golden["run_t_raw_seed"]))
    t_raw = 0.05 * rng.standard_normal((sht.nlat, sht.nlon, vg.kmax))
    heat = cases.heating(sht.lat, sht.lon, vg.sig, lat0=10.0, lon0=150.0)
    m = OracleModel(bs, 1, OracleConfig())
    m.set_initial_perturbation(u=st1.u, v=st1.v, t=st1.t + t_raw, q=st1.q, ps=st1.ps)
    m.run(nsteps=nsteps, frc_t=heat[None])
    for f in ("u", "v", "t", "q", "ps"):
        arr = getattr(m.ga, f)[0]
        got = arr.ravel()[golden["run%d_%s_idx" % (nsteps, f)]]
        ref = golden["run%d_%s_val" % (nsteps, f)]
        assert rel_err(got, ref) < 1e-10, (nsteps, f, rel_err(got, ref))
        np.testing.assert_allclose(np.sqrt((arr ** 2).sum()),
                                   golden["run%d_%s_norm" % (nsteps, f)], rtol=1e-10)


def test_column_sum_and_sigma_dot_closed_form():
    dsig = np.array([0.1, 0.2, 0.3, 0.4])
    x = np.array([1.0, 2.0, 3.0, 4.0])
    s = column_sum(x, dsig)
    np.testing.assert_allclose(s, [3.0, 2.9, 2.5, 1.6, 0.0])
    sigm = np.array([1.0, 0.9, 0.7, 0.4, 0.0])
    sd = sigma_dot(s, sigm)
    np.testing.assert_allclose(sd, [0.0, 0.9 * 3.0 - 2.9, 0.7 * 3.0 - 2.5, 0.4 * 3.0 - 1.6, 0.0])


# -- the reference-held end-to-end golden ------------------------------------
GILL_JSON = {  # analysis/gill_isca.json:4-14 ("cpu")
    "kelvin_peak_lon_offset_deg": 42.1875,
    "kelvin_peak_u_ms": -0.5286720150965126,
    "rossby_peak_lon_offset_deg": -16.875,
    "rossby_peak_u_ms": 1.0353251982927991,
    "L_K_kelvin_efold_km": 14114.396173583584,
    "L_R_rossby_efold_km": 3632.343572635994,
}


def gill_case(sht, vg):
    """Resting 250 K, 1 K/day heating at 180E, 20/10 deg widths,
    sin(pi(1-sigma)) above sigma=0.1 (``analysis/gill_isca.py:30-47``)."""
    shape = (sht.nlat, sht.nlon, vg.kmax)
    bs = types.SimpleNamespace(u=np.zeros(shape), v=np.zeros(shape),
                               t=np.full(shape, 250.0), q=np.zeros(shape),
                               ps=np.full(shape[:2], 1000.0))
    dl = (sht.lon_deg[None, :] - 180.0 + 180.0) % 360.0 - 180.0
    hxy = np.exp(-(dl / 20.0) ** 2 - (sht.lat_deg[:, None] / 10.0) ** 2)
    vz = np.sin(np.pi * (1.0 - vg.sig)) * (vg.sig > 0.1)
    return bs, hxy[:, :, None] * vz[None, None, :] / 86400.0


def gill_scales(u_grid, lat_deg, lon_deg, sig, lon0=180.0, a=6.371e6):
    """Kelvin/Rossby peaks and e-folding lengths (``analysis/gill_isca.py:50-98``)."""
    klo = int(np.argmin(np.abs(sig - 0.85)))
    eq = int(np.argmin(np.abs(lat_deg)))
    u = u_grid[eq, :, klo]
    dlon = (lon_deg - lon0 + 180.0) % 360.0 - 180.0
    order = np.argsort(dlon)
    dlon, u = dlon[order], u[order]
    dist = np.radians(dlon) * a / 1000.0
    east, west = dlon > 0, dlon < 0
    ie = int(np.nanargmax(np.abs(np.where(east, u, np.nan))))
    iw = int(np.nanargmax(np.abs(np.where(west, u, np.nan))))

    def efold(d, amp, lo, hi):
        sel = (d >= lo) & (d <= hi) & (amp > 0)
        slope, _ = np.polyfit(d[sel], np.log(amp[sel]), 1)
        return -1.0 / slope

    lk = efold(dist, np.where(east, np.abs(u), 0.0), dist[ie], np.radians(150) * a / 1e3)
    lr = efold(-dist, np.where(west, np.abs(u), 0.0), -dist[iw], np.radians(85) * a / 1e3)
    return {"kelvin_peak_lon_offset_deg": float(dlon[ie]), "kelvin_peak_u_ms": float(u[ie]),
            "rossby_peak_lon_offset_deg": float(dlon[iw]), "rossby_peak_u_ms": float(u[iw]),
            "L_K_kelvin_efold_km": float(lk), "L_R_rossby_efold_km": float(lr)}


@pytest.mark.slow
def test_gill_golden():
    from oracle import SHT, VerticalOps
    sht, vg = SHT(), VerticalOps()
    bs, heat = gill_case(sht, vg)
    m = OracleModel(bs, 1, OracleConfig(ndays=6.0))
    m.run(frc_t=heat[None])
    got = gill_scales(m.ga.u[0], sht.lat_deg, sht.lon_deg, vg.sig)
    for k, v in GILL_JSON.items():
        assert abs(got[k] - v) <= 1e-9 * max(1.0, abs(v)), (k, got[k], v)
