"""Level counts other than 20 and 40: the spectral step and its adjoint run
the register-capacity kernels (16, 32, 48, 64 levels, count at run time).

* 12 steps of a balanced-jet member with heating on 10, 26, 45 and 60
  equally spaced sigma levels (one per capacity) against the oracle's run,
  FP64 within 1e-10, and FP32 on 26 levels within 1e-5;
* the tangent-linear response's adjoint identity on 26 levels.
"""

import numpy as np
import pytest

import cases
from oracle import OracleConfig, OracleModel, SHT, VerticalOps
from oracle.model import DDRAG_DEFAULT
from oracle_helpers import rel_err

pytestmark = pytest.mark.gpu


def drag(nlev):
    """The reference's 20-level drag profile (days) resampled to nlev levels."""
    x = np.linspace(0.0, len(DDRAG_DEFAULT) - 1.0, nlev)
    return list(np.interp(x, np.arange(len(DDRAG_DEFAULT)), DDRAG_DEFAULT))


def setup(nlev):
    sigm = np.linspace(1.0, 0.0, nlev + 1)
    sht, vg = SHT(), VerticalOps(sigm)
    bs = cases.jet_basic_state(sht.lat, sht.lon, vg.sig)
    heat = cases.heating(sht.lat, sht.lon, vg.sig, lat0=10.0, lon0=140.0, amp_k_day=3.0)
    d = drag(nlev)
    return sigm, sht, vg, bs, heat, dict(ddragv=d, ddragd=d, ddragt=d)


@pytest.mark.parametrize("nlev,prec", [(10, "fp64"), (26, "fp64"), (45, "fp64"), (60, "fp64"),
                                       (26, "fp32")])
def test_member_run_matches_oracle(nlev, prec):
    import lbm_isca_amd as L
    sigm, sht, vg, bs, heat, dk = setup(nlev)
    nsteps = 12
    m = L.LBM(bs, L.Spectral(), L.VerticalGrid(sigm), L.Config(precision=prec, **dk))
    m.plan.set_forcing(heat[None])
    m.plan.run(nsteps)
    got = [x.cpu().numpy()[0] for x in m.plan.get_state()]
    ref = OracleModel(bs, 1, OracleConfig(**dk), sigm=sigm)
    ref.run(nsteps=nsteps, frc_t=heat[None])
    tol = 1e-10 if prec == "fp64" else 1e-5
    for f, a in zip(("u", "v", "t", "q", "ps"), got):
        want = getattr(ref.ga, f)[0]
        if np.abs(want).max() == 0:
            continue
        e = rel_err(a, want)
        print(f"L{nlev} {prec} {f}: {e:.2e}")
        assert e < tol, (nlev, prec, f, e)


def test_adjoint_identity_26_levels():
    import lbm_isca_amd as L
    from lbm_isca_amd import adjoint
    nlev, E = 26, 2
    sigm, sht, vg, bs, heat, dk = setup(nlev)
    m, _, _ = L.make_ensemble_model(bs, E, L.Config(tangent_linear=True, **dk), sigm=sigm)
    plan = m.plan
    lat = np.degrees(plan.get("lat", sht.nlat))
    lon = np.degrees(plan.get("lon", sht.nlon))
    plan.set_response(adjoint.box_weights(lat, lon, plan.get("gw", sht.nlat), (30.0, 60.0), (150.0, 210.0)),
                      adjoint.sigma_weights(plan.get("sig", nlev), 0.2), np.r_[0.0, 0.0, np.ones(6)])
    rng = np.random.default_rng(4)
    frc = 1e-5 * rng.standard_normal((E, sht.nlat, sht.nlon, nlev))
    M = sht.nmax + 1
    mask = (np.arange(M)[None, :] >= np.arange(M)[:, None])[None, :, :, None]
    t0 = 1e-2 * (rng.standard_normal((E, M, M, nlev)) + 1j * rng.standard_normal((E, M, M, nlev))) * mask
    t0[:, 0] = t0[:, 0].real
    plan.set_forcing(frc)
    plan.set_spectral_state(t=t0)
    plan.run(6)
    J = plan.get_response()
    gx, gf = plan.adjoint_run(6)
    g = gx[2].cpu().numpy()
    terms = [(g.real * t0.real + g.imag * t0.imag).reshape(E, -1).sum(1),
             (gf.cpu().numpy() * frc).reshape(E, -1).sum(1)]
    pred = np.sum(terms, axis=0)
    print("L26 adjoint:", J, pred)
    assert np.all(np.abs(J - pred) <= 1e-11 * np.sum(np.abs(terms), axis=0)), (J, pred)
