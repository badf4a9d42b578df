"""The grids of the reference's resolution sweep other than T42 and T127
(``analysis/benchmark_perf_isca.py:43-45``: T21 64 x 32, T63 192 x 96, T85
256 x 128, T106 320 x 160), whose FFTs take the Stockham path (radices 4,
2, 3, 5) and whose grid kernels run in blocks of 64 longitudes where the
row is not a multiple of 128:

* every transform against the oracle's, FP64 (1e-12) and FP32 (1e-6);
* a balanced-jet member with heating, 12 steps of the default model,
  against the oracle's run (1e-10);
* the tangent-linear response's adjoint identity at T21 (the transposed
  Fourier scalings of the Stockham kernels).
"""

import numpy as np
import pytest

import cases
from oracle import OracleConfig, OracleModel, SHT, VerticalOps
from oracle_helpers import rel_err

pytestmark = pytest.mark.gpu

RES = {"T21": (64, 32, 21), "T63": (192, 96, 63), "T85": (256, 128, 85), "T106": (320, 160, 106)}


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name", list(RES))
def test_transforms_match_oracle(name, prec):
    import lbm_isca_amd as L
    nlon, nlat, nmax = RES[name]
    sht = SHT(nlon, nlat, nmax, refine_gauss=True)
    sp = L.Spectral(nlon, nlat, nmax, precision=prec)
    tol = 1e-12 if prec == "fp64" else 1e-6
    rng = np.random.default_rng(7)
    g = rng.standard_normal((nlat, nlon, 3))
    w = sht.grid_to_wave(g[None])[0]
    errs = {
        "g2w": rel_err(np.asarray(L.asnumpy(sp.grid_to_wave(g))), w),
        "w2g": rel_err(np.asarray(L.asnumpy(sp.wave_to_grid(w))), sht.wave_to_grid(w[None])[0]),
    }
    u, v = rng.standard_normal((2, nlat, nlon, 2))
    vor, div = sp.uv_to_vordiv(u, v)
    rv, rd = sht.uv_to_vordiv(u[None], v[None])
    errs["uv2vd"] = max(rel_err(L.asnumpy(vor), rv[0]), rel_err(L.asnumpy(div), rd[0]))
    pu, pv = sp.psichi_to_uv(w, 0.5 * w)
    ru, rv2 = sht.psichi_to_uv(w[None], 0.5 * w[None])
    errs["pc2uv"] = max(rel_err(L.asnumpy(pu), ru[0]), rel_err(L.asnumpy(pv), rv2[0]))
    gx, gy = sp.hgrad(w)
    rx, ry = sht.hgrad(w[None])
    errs["hgrad"] = max(rel_err(L.asnumpy(gx), rx[0]), rel_err(L.asnumpy(gy), ry[0]))
    print(name, prec, {k: f"{e:.1e}" for k, e in errs.items()})
    assert max(errs.values()) < tol, errs


@pytest.mark.parametrize("name", ["T21", "T106"])
def test_member_run_matches_oracle(name):
    import lbm_isca_amd as L
    nlon, nlat, nmax = RES[name]
    sht, vg = SHT(nlon, nlat, nmax, refine_gauss=True), VerticalOps()
    bs = cases.jet_basic_state(sht.lat, sht.lon, vg.sig)
    heat = cases.heating(sht.lat, sht.lon, vg.sig, lat0=10.0, lon0=140.0, amp_k_day=3.0)
    nsteps = 12
    m = L.LBM(bs, L.Spectral(nlon, nlat, nmax), L.VerticalGrid(), L.Config())
    m.plan.set_forcing(heat[None])
    m.plan.run(nsteps)
    got = [x.cpu().numpy()[0] for x in m.plan.get_state()]
    ref = OracleModel(bs, 1, OracleConfig(), nlon=nlon, nlat=nlat, nmax=nmax, refine_gauss=True)
    ref.run(nsteps=nsteps, frc_t=heat[None])
    for f, a in zip(("u", "v", "t", "q", "ps"), got):
        want = getattr(ref.ga, f)[0]
        if np.abs(want).max() == 0:
            continue
        e = rel_err(a, want)
        print(f"{name} {f}: {e:.2e}")
        assert e < 1e-10, (name, f, e)


def test_t21_adjoint_identity():
    import lbm_isca_amd as L
    from lbm_isca_amd import adjoint
    nlon, nlat, nmax = RES["T21"]
    sht, vg = SHT(nlon, nlat, nmax), VerticalOps()
    bs = cases.jet_basic_state(sht.lat, sht.lon, vg.sig)
    E = 2
    m, _, _ = L.make_ensemble_model(bs, E, L.Config(tangent_linear=True),
                                    spectral=L.BatchedSpectral(nlon, nlat, nmax))
    plan = m.plan
    lat = np.degrees(plan.get("lat", nlat))
    lon = np.degrees(plan.get("lon", nlon))
    plan.set_response(adjoint.box_weights(lat, lon, plan.get("gw", nlat), (30.0, 60.0), (150.0, 210.0)),
                      adjoint.sigma_weights(plan.get("sig", vg.kmax), 0.2), np.r_[0.0, 0.0, np.ones(6)])
    rng = np.random.default_rng(2)
    frc = 1e-5 * rng.standard_normal((E, nlat, nlon, vg.kmax))
    vor = np.zeros((E, nmax + 1, nmax + 1, vg.kmax), complex)
    mask = np.tril(np.ones((nmax + 1, nmax + 1)), 0).T[None, :, :, None] > 0
    vor[:] = 1e-6 * (rng.standard_normal(vor.shape) + 1j * rng.standard_normal(vor.shape)) * mask
    vor[:, 0] = vor[:, 0].real
    plan.set_forcing(frc)
    plan.set_spectral_state(vor=vor)
    plan.run(6)
    J = plan.get_response()
    gx, gf = plan.adjoint_run(6)
    g = gx[0].cpu().numpy()
    terms = [(g.real * vor.real + g.imag * vor.imag).reshape(E, -1).sum(1),
             (gf.cpu().numpy() * frc).reshape(E, -1).sum(1)]
    pred = np.sum(terms, axis=0)
    print("T21 adjoint:", J, pred)
    assert np.all(np.abs(J - pred) <= 1e-11 * np.sum(np.abs(terms), axis=0)), (J, pred)
