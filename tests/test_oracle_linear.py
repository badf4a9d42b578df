"""The oracle's linear mode (the tangent-linear model the adjoint transposes)
and the response weights, on the CPU.

Linear mode differs from the reference's strictly linear model only in the
pi' <-> ps' pair (ps' = ps_bar pi' instead of ps_bar (exp(pi') - 1)), so for a
small perturbation the two agree to first order in its amplitude; and linear
mode obeys superposition exactly."""

import numpy as np

import cases
from oracle import OracleConfig, OracleModel, SHT, VerticalOps
from oracle_helpers import oracle_state, rel_err


def _run(bs, st, cfg, nsteps, frc=None):
    m = OracleModel(bs, 1, cfg)
    m.set_initial_perturbation(u=st.u, v=st.v, t=st.t, q=st.q, ps=st.ps)
    m.run(nsteps=nsteps, frc_t=frc)
    return m.ga


def _scaled(st, a):
    out = type(st).__new__(type(st))
    out.__dict__.update(st.__dict__)
    for f in ("u", "v", "t", "q", "ps"):
        setattr(out, f, a * getattr(st, f))
    return out


def test_linear_mode_is_first_order_strictly_linear():
    sht, vg = SHT(), VerticalOps()
    bs = cases.jet_basic_state(sht.lat, sht.lon, vg.sig)
    st = oracle_state(sht, bs, vg.kmax, 2)
    errs = []
    for a in (1e-2, 1e-3):
        s = _scaled(st, a)
        lin = _run(bs, s, OracleConfig(linear=True), 4)
        ref = _run(bs, s, OracleConfig(strictly_linear=True), 4)
        errs.append(rel_err(lin.ps, ref.ps))
        for f in ("u", "v", "t", "ps"):
            assert rel_err(getattr(lin, f), getattr(ref, f)) < 50 * a, (a, f)
    # the difference is second order: it shrinks with the amplitude
    assert errs[1] < 0.2 * errs[0], errs


def test_linear_mode_superposition():
    sht, vg = SHT(), VerticalOps()
    bs = cases.jet_basic_state(sht.lat, sht.lon, vg.sig)
    x, y = oracle_state(sht, bs, vg.kmax, 4), oracle_state(sht, bs, vg.kmax, 5)
    heat = cases.heating(sht.lat, sht.lon, vg.sig)[None]
    cfg = OracleConfig(linear=True)
    comb = _scaled(x, 2.0)
    for f in ("u", "v", "t", "q", "ps"):
        setattr(comb, f, getattr(comb, f) - 0.5 * getattr(y, f))
    gx = _run(bs, x, cfg, 3, heat)
    gy = _run(bs, y, cfg, 3, heat * 0.0)
    g0 = _run(bs, _scaled(x, 0.0), cfg, 3, heat)
    gc = _run(bs, comb, cfg, 3, heat)
    # L(2x - y/2) + f == 2 (L x + f) - (L y)/2 - f
    for f in ("u", "v", "t", "q", "ps"):
        want = 2.0 * getattr(gx, f) - 0.5 * getattr(gy, f) - getattr(g0, f)
        assert rel_err(getattr(gc, f), want) < 1e-11, f


def test_response_weights():
    from lbm_isca_amd import adjoint
    sht, vg = SHT(), VerticalOps()
    w = adjoint.box_weights(sht.lat_deg, sht.lon_deg, sht.gw, (30.0, 60.0), (150.0, 210.0))
    assert abs(w.sum() - 1.0) < 1e-14
    lat_in = (sht.lat_deg >= 30) & (sht.lat_deg <= 60)
    lon_in = (sht.lon_deg >= 150) & (sht.lon_deg <= 210)
    assert np.all(w[~lat_in] == 0) and np.all(w[:, ~lon_in] == 0)
    # area weighting: each row of the box carries gw_j / sum(gw in box)
    rows = w.sum(1)[lat_in]
    np.testing.assert_allclose(rows, sht.gw[lat_in] / sht.gw[lat_in].sum(), rtol=1e-13)
    # a field equal to 1 has mean 1; sin(lat) has the area mean of sin(lat)
    sw = adjoint.sigma_weights(vg.sig, 0.2)
    assert abs(sw.sum() - 1.0) < 1e-14 and abs(np.dot(sw, vg.sig) - 0.2) < 1e-14
    assert np.count_nonzero(sw) == 2
    ws = adjoint.step_weights(2160, 2, 72, 20.0, 30.0)
    assert ws.size == 2162 and abs(ws.sum() - 1.0) < 1e-12
    k = np.nonzero(ws)[0]
    assert k[0] == 2 + 1440 - 1 and k[-1] == 2 + 2160 - 1 and k.size == 721


def test_refined_gauss_weights_against_mpmath():
    """gauss_grid(refine=True) to ~1e-15 at J = 192, where leggauss's
    weights are off by ~6e-11 (what separates the oracle's reference weights
    from the device's)."""
    import mpmath
    from oracle import gauss_grid
    J = 192
    mpmath.mp.dps = 40
    mu_r, gw_r = gauss_grid(J, refine=True)
    mu_n, gw_n = gauss_grid(J)
    ref_w = []
    for x0 in mu_r[: J // 2]:
        z = mpmath.mpf(x0)
        for _ in range(4):
            p0, p1 = mpmath.mpf(1), z
            for k in range(2, J + 1):
                p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
            dp = J * (z * p1 - p0) / (z * z - 1)
            z = z - p1 / dp
        p0, p1 = mpmath.mpf(1), z
        for k in range(2, J + 1):
            p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
        dp = J * (z * p1 - p0) / (z * z - 1)
        ref_w.append(float(1 / ((1 - z * z) * dp * dp)))   # halved weight
    ref_w = np.array(ref_w)
    assert np.max(np.abs(gw_r[: J // 2] / ref_w - 1)) < 5e-15
    assert np.max(np.abs(gw_n[: J // 2] / ref_w - 1)) > 1e-12
    assert abs(gw_r.sum() - 1.0) < 1e-15
    np.testing.assert_allclose(mu_r, mu_n, atol=3e-16)
