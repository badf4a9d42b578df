"""History output, LBM.diagnostics, the device-interpolated MJO forcing,
segmented runs and spectral regridding on the GPU against the FP64 oracle
(oracle/history.py restates model.py:150-277, forcing.py:99-161 and
regrid.py:66-124)."""

import numpy as np
import pytest

import cases
from oracle import OracleConfig, OracleModel, SHT, VerticalOps
from oracle import history as H
from oracle_helpers import rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("psi", "chi", "u", "v", "w", "t", "z", "p")


class Collect:
    def __init__(self):
        self.records, self.closed = [], False

    def add(self, day, fields):
        self.records.append((day, {k: np.array(v) for k, v in fields.items()}))

    def close(self):
        self.closed = True


@pytest.fixture(scope="module")
def env():
    sht, vg = SHT(), VerticalOps()
    bs = cases.jet_basic_state(sht.lat, sht.lon, vg.sig)
    heat = cases.heating(sht.lat, sht.lon, vg.sig)
    return sht, vg, bs, heat


def hip_model(bs, nmem=1, **cfgkw):
    import lbm_isca_amd as L
    cfg = L.Config(**cfgkw)
    if nmem == 1:
        return L.LBM(bs, L.Spectral(), L.VerticalGrid(), cfg)
    m, _, _ = L.make_ensemble_model(bs, nmem, cfg)
    return m


def compare_records(got, want, tol=1e-10):
    assert [d for d, _ in got] == [d for d, _ in want]
    for (d, g), (_, w) in zip(got, want):
        for k in FIELDS:
            e = rel_err(g[k], w[k])
            assert e < tol, (d, k, e)


@pytest.mark.parametrize("time_average", [True, False])
@pytest.mark.parametrize("use_graph", [True, False])
def test_history_matches_oracle(env, time_average, use_graph):
    import lbm_isca_amd as L
    sht, vg, bs, heat = env
    nsteps, K = 12, 4
    m = hip_model(bs, ndays=nsteps / 72.0, out_interval_days=K / 72.0, time_average=time_average)
    w = Collect()
    m.run(mjo_forcing=L.ConstForcing(heat), writer=w, verbose=False, use_graph=use_graph)
    assert w.closed
    ref = OracleModel(bs, 1, OracleConfig())
    hist = H.History(K, time_average)
    ref.run(nsteps=nsteps, frc_t=heat[None], history=hist)
    compare_records(w.records, [(d, {k: v[0] for k, v in r.items()}) for d, r in hist.records])
    # the run's state is the same as without history
    m2 = hip_model(bs, ndays=nsteps / 72.0)
    m2.run(mjo_forcing=L.ConstForcing(heat), verbose=False)
    for a, b in zip(m.plan.get_state(), m2.plan.get_state()):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_history_ensemble_and_partial_interval(env):
    import lbm_isca_amd as L
    sht, vg, bs, _ = env
    E, nsteps, K = 3, 10, 4          # two whole intervals and a tail of two steps
    from lbm_isca_amd import cases as C
    frc = C.random_heatings(sht.lat, sht.lon, vg.sig, E, seed=5)
    m = hip_model(bs, nmem=E, ndays=nsteps / 72.0, out_interval_days=K / 72.0)
    w = Collect()
    m.run(mjo_forcing=L.EnsembleForcing(frc), writer=w, verbose=False)
    ref = OracleModel(bs, E, OracleConfig())
    hist = H.History(K)
    ref.run(nsteps=nsteps, frc_t=frc, history=hist)
    assert len(w.records) == 2 and w.records[0][1]["u"].shape[0] == E
    compare_records(w.records, hist.records)
    for a, f in zip(m.plan.get_state(), ("u", "v", "t", "q", "ps")):
        assert rel_err(a.cpu().numpy(), getattr(ref.ga, f)) < 1e-10


def test_diagnostics_after_run(env):
    import lbm_isca_amd as L
    sht, vg, bs, heat = env
    m = hip_model(bs, ndays=5 / 72.0)
    m.run(mjo_forcing=L.ConstForcing(heat), verbose=False)
    got = {k: L.asnumpy(v) for k, v in m.diagnostics().items()}
    ref = OracleModel(bs, 1, OracleConfig())
    ref.run(nsteps=5, frc_t=heat[None])
    want = {k: v[0] for k, v in H.diagnostics(ref).items()}
    for k in FIELDS:
        assert rel_err(got[k], want[k]) < 1e-10, k


@pytest.mark.parametrize("single", [True, False])
def test_mjo_forcing_matches_oracle(env, single):
    import lbm_isca_amd as L
    sht, vg, bs, heat = env
    rng = np.random.default_rng(7)
    amp = rng.uniform(0.5, 2.0, 4)
    gtf = np.stack([a * np.roll(heat, 8 * d, axis=1) for d, a in enumerate(amp)])
    gtf = gtf.astype(np.float32 if single else np.float64)
    nt, spinup = 3, 4
    mjo = L.MJOForcing(gtf=gtf, nt=nt, spinup=spinup)
    m = hip_model(bs, ndays=10 / 72.0)
    ref = OracleModel(bs, 1, OracleConfig())
    seq = H.MJOSequence(gtf, nt, spinup)
    for _ in range(2):                 # the second run continues the forcing's calls
        m.run(mjo_forcing=mjo, verbose=False)
        ref.run(nsteps=10, mjo=_Lead(seq))
        for a, f in zip(m.plan.get_state(), ("u", "v", "t", "q", "ps")):
            assert rel_err(a.cpu().numpy()[0], getattr(ref.ga, f)[0]) < 1e-10, f
    assert mjo.counter == 2 * (2 + 10)


class _Lead:
    """Adds the member axis to a single forcing field."""

    def __init__(self, seq):
        self.seq = seq

    def step(self):
        return np.asarray(self.seq.step())[None]


def test_grads_file_from_run(env, tmp_path):
    import lbm_isca_amd as L
    from lbm_isca_amd import forcing as F
    sht, vg, bs, heat = env
    m = hip_model(bs, ndays=8 / 72.0, out_interval_days=4 / 72.0)
    p = str(tmp_path / "h.grd")
    m.run(mjo_forcing=L.ConstForcing(heat), writer=L.GradsWriter(p, m.sp, m.vg), verbose=False)
    ref = OracleModel(bs, 1, OracleConfig())
    hist = H.History(4)
    ref.run(nsteps=8, frc_t=heat[None], history=hist)
    rec = F.read_fortran_records(p)
    assert rec.shape[0] == 2 * 141
    for t, (_, r) in enumerate(hist.records):
        z = r["z"][0] + r["p"][0][..., None] * 1.0e2 / (H.RHO00 * H.GRAV_GT2GR)
        for v, name in enumerate(("psi", "chi", "u", "v", "w", "t")):
            want = r[name][0].astype(np.float32)
            got = np.moveaxis(rec[t * 141 + 20 * v: t * 141 + 20 * v + 20], 0, -1)
            assert rel_err(got, want) < 1e-6, name
        assert rel_err(np.moveaxis(rec[t * 141 + 120: t * 141 + 140], 0, -1), z) < 1e-6
        assert rel_err(rec[t * 141 + 140], r["p"][0]) < 1e-6


def test_run_to_segments_equal_one_run(env):
    import lbm_isca_amd as L
    sht, vg, bs, heat = env
    states = []
    for stops in ((0,), (3, 9, 0)):
        m = hip_model(bs, ndays=1.0)
        m.plan.set_forcing(heat)
        for s in stops:
            m.plan.run_to(20, s)
        states.append([x.cpu().numpy() for x in m.plan.get_state()])
    for a, b in zip(*states):
        assert np.array_equal(a, b)


def test_adjoint_refuses_a_time_varying_forcing(env):
    import lbm_isca_amd as L
    sht, vg, bs, heat = env
    m = hip_model(bs, ndays=1.0, tangent_linear=True)
    m.plan.set_response(np.ones((64, 128)) / (64 * 128), None, np.ones(4))
    m.plan.set_mjo_forcing(np.stack([heat, heat]), nt=3, spinup=2)
    with pytest.raises(RuntimeError, match="constant forcings"):
        m.plan.adjoint_run(2)


def test_regrid_matches_oracle():
    import lbm_isca_amd as L
    from lbm_isca_amd import regrid as R
    lo_o, hi_o = SHT(), SHT(384, 192, 127, refine_gauss=True)
    lo, hi = L.Spectral(), L.Spectral(384, 192, 127)
    rng = np.random.default_rng(2)
    f = lo_o.wave_to_grid(lo_o.grid_to_wave(rng.standard_normal((1, 64, 128, 3))))[0]
    got = R.regrid_scalar(lo, hi, f)
    want = H.regrid_scalar(lo_o, hi_o, f[None])[0]
    assert rel_err(got, want) < 1e-12
    back = R.regrid_scalar(hi, lo, got)
    assert rel_err(back, f) < 1e-12
    u, v = rng.standard_normal((2, 64, 128, 2))
    gu, gv = R.regrid_uv(lo, hi, u, v)
    wu, wv = H.regrid_uv(lo_o, hi_o, u[None], v[None])
    assert rel_err(gu, wu[0]) < 1e-12 and rel_err(gv, wv[0]) < 1e-12
    # the coefficient copy itself is exact, both directions, with and without levels
    xi = rng.standard_normal((43, 43, 5)) + 1j * rng.standard_normal((43, 43, 5))
    up = L.asnumpy(R.resize_spec(xi, 127))
    assert np.array_equal(up[:43, :43], xi) and not np.any(up[43:]) and not np.any(up[:, 43:])
    assert np.array_equal(L.asnumpy(R.resize_spec(up, 42)), xi)
    assert np.array_equal(L.asnumpy(R.resize_spec(xi[None, ..., 0], 21, batched=True)),
                          xi[None, :22, :22, 0])


@pytest.mark.parametrize("with_mjo", [False, True])
def test_steady_forcings_match_oracle(env, with_mjo):
    import lbm_isca_amd as L
    sht, vg, bs, heat = env
    rng = np.random.default_rng(11)
    smooth = lambda x: sht.wave_to_grid(sht.grid_to_wave(x[None]))[0]    # noqa: E731
    fv = smooth(rng.standard_normal((64, 128, 20))) * 2e-11
    fd = smooth(rng.standard_normal((64, 128, 20))) * 1e-11
    fp = smooth(rng.standard_normal((64, 128, 1)))[..., 0] * 3e-9
    sf = L.SteadyForcing(fv, fd, heat, fp)
    m = hip_model(bs, ndays=10 / 72.0)
    ref = OracleModel(bs, 1, OracleConfig())
    steady = {"frc_vor": fv[None], "frc_div": fd[None], "frc_ps": fp[None]}
    if with_mjo:
        gtf = np.stack([0.5 * heat, 1.5 * np.roll(heat, 16, axis=1)])
        m.run(mjo_forcing=L.MJOForcing(gtf=gtf, nt=4, spinup=3), steady_forcing=sf, verbose=False)
        ref.run(nsteps=10, mjo=_Lead(H.MJOSequence(gtf, 4, 3)), steady=steady)
    else:
        m.run(steady_forcing=sf, verbose=False)
        ref.run(nsteps=10, frc_t=heat[None], steady=steady)
    for a, f in zip(m.plan.get_state(), ("u", "v", "t", "q", "ps")):
        assert rel_err(a.cpu().numpy()[0], getattr(ref.ga, f)[0]) < 1e-10, f
    # the vorticity forcing alone moves the state (the forcing is not lost)
    m0 = hip_model(bs, ndays=10 / 72.0)
    m0.run(steady_forcing=L.SteadyForcing(None, None, heat, None), verbose=False)
    assert rel_err(m0.plan.get_state()[0].cpu().numpy(), m.plan.get_state()[0].cpu().numpy()) > 1e-6
