"""T127 L40: the 384-point FFTs and N = 127 Legendre transforms against the
oracle's, a balanced-jet member with a vorticity perturbation against the
oracle's run, FP32 against FP64, and the adjoint identity at this size.

The oracle runs with refined Gauss weights here: the reference's numpy
leggauss weights are off by up to 6e-11 at 192 latitudes (the device's are
exact to rounding), and the temperature of these runs amplifies basic-state
differences several hundred times."""

import types

import numpy as np
import pytest

from oracle import OracleConfig, OracleModel, SHT, VerticalOps
from oracle_helpers import rel_err

pytestmark = pytest.mark.gpu

NLON, NLAT, NMAX = 384, 192, 127


@pytest.fixture(scope="module")
def sht():
    return SHT(NLON, NLAT, NMAX, refine_gauss=True)


def test_t127_transforms_match_oracle(sht):
    import lbm_isca_amd as L
    sp = L.Spectral(NLON, NLAT, NMAX)
    np.testing.assert_allclose(sp.gw, sht.gw, rtol=3e-14)  # leggauss: 6e-11
    np.testing.assert_allclose(sp.mu, sht.mu, rtol=0, atol=3e-16)
    rng = np.random.default_rng(0)
    g = rng.standard_normal((NLAT, NLON, 3))
    assert rel_err(np.asarray(L.asnumpy(sp.grid_to_wave(g))), sht.grid_to_wave(g[None])[0]) < 1e-12
    w = sht.grid_to_wave(g[None])[0]
    assert rel_err(L.asnumpy(sp.wave_to_grid(w)), sht.wave_to_grid(w[None])[0]) < 1e-12
    u, v = rng.standard_normal((2, NLAT, NLON, 2))
    vor, div = sp.uv_to_vordiv(u, v)
    rv, rd = sht.uv_to_vordiv(u[None], v[None])
    assert rel_err(L.asnumpy(vor), rv[0]) < 1e-12 and rel_err(L.asnumpy(div), rd[0]) < 1e-12
    pu, pv = sp.psichi_to_uv(w, 0.5 * w)
    ru, rv2 = sht.psichi_to_uv(w[None], 0.5 * w[None])
    assert rel_err(L.asnumpy(pu), ru[0]) < 1e-12 and rel_err(L.asnumpy(pv), rv2[0]) < 1e-12


def _jet_case(E, precision="fp64", tangent_linear=False):
    import lbm_isca_amd as L
    from lbm_isca_amd import cases
    sp = L.BatchedSpectral(NLON, NLAT, NMAX)
    vg = L.VerticalGrid(cases.EQUAL_SIGM_L40)
    bs, _ = cases.jet_states(sp, vg.sig, E, seed=11)
    cfg = L.Config(**cases.t127l40_config(precision=precision, tangent_linear=tangent_linear))
    m, _, _ = L.make_ensemble_model(bs, E, cfg, sigm=cases.EQUAL_SIGM_L40, per_member_basic=True,
                                    spectral=sp)
    vor = cases.vorticity_perturbation(sp, vg.kmax, E, seed=4)
    return m, sp, vg, bs, vor


@pytest.mark.parametrize("with_t", [True, False])
def test_t127l40_member_matches_oracle(with_t):
    """A balanced-jet member with the n^-1.5 vorticity perturbation (plus, in
    one case, divergence and temperature perturbations of the same
    spectrum): start-up and one full step against the oracle, FP64 and FP32.
    The oracle with the reference's own (leggauss) weights is printed beside
    it: its temperature differs at ~1e-8, the weights' effect."""
    from lbm_isca_amd import cases
    nsteps = 1
    out = {}
    for prec in ("fp64", "fp32"):
        m, sp, vg, bs, vor = _jet_case(1, prec)
        z = np.zeros_like(vor)
        spec = {"vor": vor, "div": 0.3 * vor if with_t else z, "t": 1.0e4 * vor if with_t else z,
                "q": z, "pi": z[..., 0]}
        m.plan.set_spectral_state(**spec)
        m.plan.run(nsteps, use_graph=False)
        out[prec] = [x.cpu().numpy() for x in m.plan.get_state()]
    sigm = cases.EQUAL_SIGM_L40
    be = types.SimpleNamespace(**{k: getattr(bs, k)[0] for k in ("u", "v", "t", "q", "ps")})
    refs = {}
    for refine in (True, False):
        ref = OracleModel(be, 1, OracleConfig(**cases.t127l40_config()), sigm=sigm, nlon=NLON,
                          nlat=NLAT, nmax=NMAX, refine_gauss=refine)
        ref.ga.load_spectral(ref.sht, spec)
        ref.gb.load_spectral(ref.sht, spec)
        ref.run(nsteps=nsteps)
        refs[refine] = ref.ga
    for f, a64, a32 in zip(("u", "v", "t", "q", "ps"), out["fp64"], out["fp32"]):
        want = getattr(refs[True], f)
        e64, e32 = rel_err(a64, want), rel_err(a32, want)
        eleg = rel_err(a64, getattr(refs[False], f))
        print(f"T127L40 with_t={with_t} {f}: |ref| {np.abs(want).max():.3e} "
              f"fp64 {e64:.2e} fp32 {e32:.2e} (leggauss oracle {eleg:.2e})")
        if f == "q":
            continue
        assert e64 < 1e-9, (f, e64)
        assert e32 < 1e-4, (f, e32)


def test_t127l40_adjoint_identity():
    from lbm_isca_amd import adjoint
    E = 2
    m, sp, vg, bs, vor = _jet_case(E, tangent_linear=True)
    plan = m.plan
    wg = adjoint.box_weights(sp.lat_deg, sp.lon_deg, sp.gw, (30.0, 60.0), (150.0, 210.0))
    ws = np.zeros(2 + 4)
    ws[[1, 3, 5]] = 1.0
    plan.set_response(wg, adjoint.sigma_weights(vg.sig, 0.2), ws)
    rng = np.random.default_rng(3)
    frc = 1e-5 * rng.standard_normal((E, NLAT, NLON, vg.kmax))
    plan.set_forcing(frc)
    plan.set_spectral_state(vor=vor, t=1e3 * vor)
    plan.run(4, use_graph=True)
    J = plan.get_response()
    gx, gf = plan.adjoint_run(4)
    gvor, gt = gx[0].cpu().numpy(), gx[2].cpu().numpy()
    ip = lambda a, b: (a.real * b.real + a.imag * b.imag).reshape(E, -1).sum(1)  # noqa: E731
    terms = [ip(gvor, vor), ip(gt, 1e3 * vor), (gf.cpu().numpy() * frc).reshape(E, -1).sum(1)]
    pred = np.sum(terms, axis=0)
    assert np.all(np.abs(J - pred) <= 1e-11 * np.sum(np.abs(terms), axis=0)), (J, pred)
