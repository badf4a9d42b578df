"""The grid phase in one kernel (tuning choice 4 = 3 or 4: inverse FFT,
tendencies and forward FFT per latitude and member group, k_grid) computes
the numbers of the three separate launches bit for bit, in every path the
tendency kernel serves:

* a raw initial perturbation (its first step takes the separate kernels,
  the rest the fused one), history sums of interval means, an ensemble of
  5 members (a partial member group) with uneven lengths;
* the checkpointed adjoint of the default model, whose forward segments
  store the trajectory from the tendency kernel;
* FP64 and FP32.
"""

import numpy as np
import pytest

import cases
from oracle import SHT, VerticalOps

pytestmark = pytest.mark.gpu

HFIELDS = ("psi", "chi", "u", "v", "w", "t", "z", "p")


class Collect:
    def __init__(self):
        self.records = []

    def add(self, day, fields):
        self.records.append((day, {k: np.array(v) for k, v in fields.items()}))

    def close(self):
        pass


@pytest.fixture(scope="module")
def env():
    sht, vg = SHT(), VerticalOps()
    bs = cases.jet_basic_state(sht.lat, sht.lon, vg.sig)
    return sht, vg, bs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_raw_start_history_and_ensemble(env, precision):
    import lbm_isca_amd as L
    sht, vg, bs = env
    E, nsteps, K = 5, 12, 4
    heat = np.stack([cases.heating(sht.lat, sht.lon, vg.sig, lon0=40.0 * e) for e in range(E)])
    shape = (E, sht.nlat, sht.nlon, vg.kmax)
    rng = np.random.default_rng(5)
    u0 = rng.standard_normal(shape)
    t0 = 0.5 * rng.standard_normal(shape)
    out = {}
    for choice in (0, 3, 4):
        m, _, _ = L.make_ensemble_model(bs, E, L.Config(ndays=nsteps / 72.0, precision=precision,
                                                         out_interval_days=K / 72.0))
        m.plan.set_tuning([0, 0, 0, 0, choice])
        m.set_initial_perturbation(u=u0, t=t0)
        w = Collect()
        m.run(mjo_forcing=L.EnsembleForcing(heat), writer=w, verbose=False)
        state = [x.cpu().numpy() for x in m.plan.get_state() + m.plan.get_prev_state()]
        out[choice] = (state, w.records)
    ref_state, ref_rec = out[0]
    assert np.abs(ref_state[2]).max() > 0 and len(ref_rec) == nsteps // K
    for choice in (3, 4):
        state, rec = out[choice]
        for a, b in zip(state, ref_state):
            assert np.array_equal(a, b), (precision, choice)
        for (d1, r1), (d2, r2) in zip(rec, ref_rec):
            assert d1 == d2
            for k in HFIELDS:
                assert np.array_equal(r1[k], r2[k]), (precision, choice, d1, k)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_uneven_lengths(env, precision):
    import lbm_isca_amd as L
    sht, vg, bs = env
    E = 11
    heat = np.stack([cases.heating(sht.lat, sht.lon, vg.sig, lon0=30.0 * e) for e in range(E)])
    m, _, _ = L.make_ensemble_model(bs, E, L.Config(precision=precision))
    m.plan.set_lengths([9, 3, 9, 6, 1, 9, 2, 9, 4, 5, 7])
    out = {}
    for choice in (0, 3, 4):
        m.plan.set_tuning([0, 0, 0, 0, choice])
        m.plan.set_forcing(heat)
        m.plan.set_state()
        m.plan.run(0)
        out[choice] = [x.cpu().numpy() for x in m.plan.get_state()]
    assert np.abs(out[0][2]).max() > 0
    for choice in (3, 4):
        for a, b in zip(out[0], out[choice]):
            assert np.array_equal(a, b), (precision, choice)


def test_adjoint_of_the_default_model(env):
    import lbm_isca_amd as L
    from lbm_isca_amd import adjoint
    sht, vg, bs = env
    E, nsteps = 2, 24
    heat = np.stack([cases.heating(sht.lat, sht.lon, vg.sig, lat0=10.0, lon0=140.0),
                     cases.heating(sht.lat, sht.lon, vg.sig, lat0=-5.0, lon0=250.0)])
    spec = {}
    for f in ("vor", "div", "t", "q", "pi"):
        spec[f] = []
    for e in range(E):
        psi, chi, t, q, pi = cases.perturbation_spectra(20, 3 + 17 * e)
        vor, div = sht.laplacian(psi[None])[0], sht.laplacian(chi[None])[0]
        vor[0, 0] = div[0, 0] = 0.0
        for f, x in zip(("vor", "div", "t", "q", "pi"), (vor, div, t, q, pi)):
            spec[f].append(2.0 * x)
    spec = {f: np.stack(v) for f, v in spec.items()}
    res = {}
    for choice in (0, 3, 4):
        m, _, _ = L.make_ensemble_model(bs, E, L.Config())
        plan = m.plan
        plan.set_tuning([0, 0, 0, 0, choice])
        lat = np.degrees(plan.get("lat", plan.nlat))
        lon = np.degrees(plan.get("lon", plan.nlon))
        wgrid = adjoint.box_weights(lat, lon, plan.get("gw", plan.nlat), (30.0, 60.0), (150.0, 210.0))
        wstep = np.zeros(2 + nsteps)
        wstep[8:] = 1.0 / (wstep.size - 8)
        plan.set_response(wgrid, adjoint.sigma_weights(plan.get("sig", plan.nlev), 0.2), wstep)
        plan.set_checkpoints(8)
        plan.set_forcing(heat)
        plan.set_spectral_state(*[spec[f] for f in ("vor", "div", "t", "q", "pi")])
        plan.run(nsteps)
        J = plan.get_response()
        gx, gf = plan.adjoint_run(nsteps)
        res[choice] = (np.asarray(J), [x.cpu().numpy() for x in gx], gf.cpu().numpy())
    J0, gx0, gf0 = res[0]
    assert np.abs(gf0).max() > 0
    for choice in (3, 4):
        J, gx, gf = res[choice]
        assert np.array_equal(J, J0), choice
        for a, b in zip(gx + [gf], gx0 + [gf0]):
            assert np.array_equal(a, b), choice


# Skipped: the 384-longitude k_grid (now switched off in grid_fits) faulted
# the GPU on its first run here (FP64, memory aperture violation), cause not
# found; T127 takes the three-launch grid phase.
@pytest.mark.skip(reason="384-longitude k_grid switched off after an unexplained GPU fault")
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_t127l40_members(precision):
    """384 longitudes (T127 L40, 3 members on their own balanced jets, a
    vorticity and temperature perturbation): k_grid with one or two members
    per block against the three launches, bit for bit."""
    import lbm_isca_amd as L
    from lbm_isca_amd import cases
    E = 3
    sp = L.BatchedSpectral(384, 192, 127)
    vg = L.VerticalGrid(cases.EQUAL_SIGM_L40)
    bs, _ = cases.jet_states(sp, vg.sig, E, seed=11)
    m, _, _ = L.make_ensemble_model(bs, E, L.Config(**cases.t127l40_config(precision=precision)),
                                    sigm=cases.EQUAL_SIGM_L40, per_member_basic=True, spectral=sp)
    vor = cases.vorticity_perturbation(sp, vg.kmax, E, seed=4)
    m.plan.set_lengths([6, 4, 6])
    out = {}
    for choice in (0, 3, 4):
        m.plan.set_tuning([0, 0, 0, 0, choice])
        m.plan.set_spectral_state(vor=vor, t=1.0e3 * vor)
        m.plan.run(0)
        out[choice] = [x.cpu().numpy() for x in m.plan.get_state()]
    assert np.abs(out[0][2]).max() > 0
    for choice in (3, 4):
        for a, b in zip(out[0], out[choice]):
            assert np.array_equal(a, b), (precision, choice)
