"""Pins the oracle's history, diagnostics and regridding (oracle/history.py)
against what the mathematics fixes:

* the trapezoidal interval mean of a sample sequence linear in time has a
  closed form (interval centre), in both output modes;
* psi and chi are the inverse Laplacians of the analysed vorticity and
  divergence, and (psi, chi) give back the model's own winds;
* z of an isothermal column is (R T / g) ln(1/sigma) exactly on the
  Simmons-Burridge levels;
* a band-limited field survives T42 -> T127 -> T42 unchanged and a
  spherical harmonic regridded to T127 equals its direct evaluation there.

(omega is pinned with the reference-generated vectors in
test_oracle_golden.py.)"""

import types

import numpy as np

from oracle import OracleConfig, OracleModel, SHT, VerticalOps
from oracle import history as H

from lbm_isca_amd import cases


def test_trapezoid_interval_means_closed_form():
    alpha, beta, K, n = 3.0, 0.25, 6, 30
    f = lambda a: {"x": np.array([alpha + beta * a])}          # noqa: E731
    for avg in (True, False):
        h = H.History(K, time_average=avg)
        h.start_step(f(0))
        for a in range(1, n + 1):
            h.full_step(a, f(a))
        assert [d for d, _ in h.records] == list(range(1, n // K + 1))
        for d, rec in h.records:
            # mean of 0.5 (f(a) + f(a-1)) over a in ((d-1)K, dK] = alpha + beta((d-1)K + K/2);
            # the instantaneous record is the last trapezoid alpha + beta (dK - 1/2)
            want = alpha + beta * ((d - 1) * K + K / 2.0) if avg else alpha + beta * (d * K - 0.5)
            np.testing.assert_allclose(rec["x"], want, rtol=1e-14)


def small_run(nsteps=6, out_every=2):
    sht, vg = SHT(), VerticalOps()
    bs = cases.solid_body_state(sht.lat, sht.lon, vg.sig)
    heat = cases.heating(sht.lat, sht.lon, vg.sig)
    m = OracleModel(bs, 1, OracleConfig())
    hist = H.History(out_every)
    m.run(nsteps=nsteps, frc_t=heat[None], history=hist)
    return m, hist


def test_diagnostics_psi_chi_and_winds():
    m, _ = small_run()
    sht, d = m.sht, m.last_diag
    diag = H.diagnostics(m)
    n = np.arange(sht.nmax + 1)
    keep = (n >= 1)[None, None, :, None]
    # lap(psi) = vor_s and lap(chi) = div_s for n >= 1
    for name, spec in (("psi", d["vor_s"]), ("chi", d["div_s"])):
        lap = sht.laplacian(sht.grid_to_wave(diag[name]))
        err = np.abs(np.where(keep, lap - spec, 0)).max() / np.abs(spec).max()
        assert err < 2e-11, (name, err)
    # the winds of the current (band-limited) state from its own vorticity and divergence
    vor, div = sht.uv_to_vordiv(diag["u"], diag["v"])
    u, v = sht.vordiv_to_uv(vor, div)
    np.testing.assert_allclose(u, diag["u"], atol=1e-12 * np.abs(diag["u"]).max())
    np.testing.assert_allclose(v, diag["v"], atol=1e-12 * np.abs(diag["v"]).max())


def test_isothermal_z_closed_form():
    vg = VerticalOps()
    t0 = 250.0
    t = np.full((1, 2, 3, vg.kmax), t0)
    z = vg.hydrostatic(t) / H.GRAV
    np.testing.assert_allclose(z[0, 0, 0], 287.04 * t0 / H.GRAV * np.log(1.0 / vg.sig), rtol=1e-13)


def test_history_records_are_trapezoids_of_top_samples():
    # the run's records against the diagnostics recomputed from states saved
    # at the top of every step (a second path through the same oracle)
    m1, hist = small_run(nsteps=4, out_every=2)
    sht, vg = SHT(), VerticalOps()
    bs = cases.solid_body_state(sht.lat, sht.lon, vg.sig)
    heat = cases.heating(sht.lat, sht.lon, vg.sig)
    m = OracleModel(bs, 1, OracleConfig())
    tops = []
    for istep in range(1, 2 + 4 + 1):
        oinit = istep <= 2
        dt = 1200.0 / (2.0 ** (2 - istep + 1)) if oinit else 1200.0
        save = types.SimpleNamespace(**{k: getattr(m.ga, k).copy() for k in ("u", "v", "t", "ps")})
        m.step(dt, oinit, heat[None])
        cur = m.ga
        m.ga = save
        tops.append(H.diagnostics(m))
        m.ga = cur
    # full steps a = 1..4 are tops[2..5]; T_0 is the last start-up step's top (tops[1])
    T = tops[1:]
    for d, rec in hist.records:
        want = {k: sum(0.5 * (T[a][k] + T[a - 1][k]) for a in range(2 * d - 1, 2 * d + 1)) / 2
                for k in rec}
        for k in rec:
            np.testing.assert_allclose(rec[k], want[k], rtol=0, atol=1e-13 * (np.abs(want[k]).max() + 1e-300))


def test_regrid_band_limited_round_trip_and_harmonic():
    lo, hi = SHT(), SHT(384, 192, 127)
    rng = np.random.default_rng(3)
    f = lo.wave_to_grid(lo.grid_to_wave(rng.standard_normal((1, 64, 128, 2))))
    up = H.regrid_scalar(lo, hi, f)
    back = H.regrid_scalar(hi, lo, up)
    np.testing.assert_allclose(back, f, atol=1e-12 * np.abs(f).max())
    # a single spherical harmonic (m=3, n=7) evaluated directly on the fine grid
    xi = np.zeros((1, 43, 43, 1), complex)
    xi[0, 3, 7, 0] = 1.0 + 0.5j
    g_lo = lo.wave_to_grid(xi)
    xi_hi = np.zeros((1, 128, 128, 1), complex)
    xi_hi[0, 3, 7, 0] = 1.0 + 0.5j
    np.testing.assert_allclose(H.regrid_scalar(lo, hi, g_lo), hi.wave_to_grid(xi_hi), atol=1e-12)
    # winds through vorticity/divergence: a band-limited wind comes back
    psi = np.zeros((1, 43, 43, 1), complex)
    psi[0, 2, 5, 0] = 1e6
    chi = np.zeros_like(psi)
    chi[0, 1, 4, 0] = 5e5j
    u, v = lo.psichi_to_uv(psi, chi)
    u2, v2 = H.regrid_uv(hi, lo, *H.regrid_uv(lo, hi, u, v))
    np.testing.assert_allclose(u2, u, atol=1e-10 * np.abs(u).max())
    np.testing.assert_allclose(v2, v, atol=1e-10 * np.abs(v).max())


def test_steady_forcings_enter_only_their_own_tendencies():
    """frc_vor, frc_div and frc_ps change the vorticity, divergence and pi
    tendencies by their analyses (+, in their own field) and nothing else."""
    from oracle import BasicTerms
    from oracle.dyn import tendencies
    from oracle.model import GridState
    sht, vg = SHT(), VerticalOps()
    bs = cases.solid_body_state(sht.lat, sht.lon, vg.sig)
    bst = BasicTerms(sht, vg, bs)
    st = GridState(bs.ps, 1, vg.kmax)
    rng = np.random.default_rng(4)
    st.u, st.v, st.t = (sht.wave_to_grid(sht.grid_to_wave(rng.standard_normal((1, 64, 128, 20))))
                        for _ in range(3))
    fv, fd = rng.standard_normal((2, 1, 64, 128, 20)) * 1e-10
    fp = rng.standard_normal((1, 64, 128)) * 1e-7
    base, _ = tendencies(sht, vg, bst, st)
    forced, _ = tendencies(sht, vg, bst, st, frc_vor=fv, frc_div=fd, frc_ps=fp)
    want = {0: sht.grid_to_wave(fv), 1: sht.grid_to_wave(fd), 3: sht.grid_to_wave(fp[..., None])[..., 0]}
    for k, (a, b) in enumerate(zip(base, forced)):
        d = b - a
        if k in want:
            np.testing.assert_allclose(d, want[k], rtol=0, atol=1e-12 * np.abs(want[k]).max())
        else:
            assert np.abs(d).max() <= 1e-14 * max(np.abs(a).max(), 1e-300)
