"""The drop-in API around the HIP step, each entry point on its own:

* ``Dynamics.tendencies`` (``lbm_tendencies``) and ``Integrator.step``
  (``lbm_step``) against the reference-generated tendency and step vectors
  (``tests/golden/reference_vectors.npz``: ``tend_full_*``, ``tend_lin_*``,
  ``step1200_*``, ``step300_*``), and the semi-implicit conditioning and
  inverse residual of ``tests/test_physics_isca.py:64-71``;
* the resting isothermal state's exactly zero tendencies
  (``tests/test_physics_isca.py:56-61``) and the tangent-linear Taylor test
  of ``tests/test_tangent_linear.py:33-69`` on ``device="hip"``;
* ``set_initial_perturbation`` with omitted fields (both levels keep their
  values, ``model.py:142-148``) and a forcing whose ``.step()`` changes
  every call (``model.py:222-229``), against the FP64 oracle;
* ``State`` (``state.py:26-148``), ``Spectral.laplacian`` and
  ``VerticalGrid.hydrostatic`` element-wise operators against numpy.
"""

import types

import numpy as np
import pytest

import cases
from oracle import GridState, OracleConfig, OracleModel, SHT, VerticalOps
from oracle_helpers import oracle_state, rel_err

pytestmark = pytest.mark.gpu

LV = list(cases.SAMPLE_LEVELS)
REF_ORDER = ("vor", "div", "t", "pi", "q")


def np_(x):
    from lbm_isca_amd import asnumpy
    return np.asarray(asnumpy(x))


@pytest.fixture(scope="module")
def env():
    sht, vg = SHT(), VerticalOps()
    bs = cases.jet_basic_state(sht.lat, sht.lon, vg.sig)
    return sht, vg, bs


def hip_state(st):
    """A HIP State from a one-member oracle GridState."""
    import lbm_isca_amd as L
    return L.State(st.ps_bar, st.u[0], st.v[0], st.t[0], st.q[0], st.ps[0])


@pytest.mark.parametrize("lin", [False, True])
def test_tendencies_match_reference_vectors(env, golden, lin):
    import lbm_isca_amd as L
    from lbm_isca_amd.dynamics import BasicStateTerms, Dynamics
    sht, vg_o, bs = env
    sp, vg = L.Spectral(), L.VerticalGrid()
    bst = BasicStateTerms(sp, vg, bs)
    np.testing.assert_allclose(vg.tbar, golden["bs_tbar"], rtol=1e-12)
    np.testing.assert_allclose(vg.tmtx, golden["bs_tmtx"], rtol=1e-12, atol=1e-16)
    dyn = Dynamics(sp, vg, bst, strictly_linear=lin)
    tend = dyn.tendencies(hip_state(oracle_state(sht, bs, vg.kmax, 0)))
    tag = "lin" if lin else "full"
    for name, x in zip(REF_ORDER, tend):
        x = np_(x)
        got = x if x.ndim == 2 else x[:, :, LV]
        err = rel_err(got, golden[f"tend_{tag}_{name}"])
        print(f"tend {tag} {name}: {err:.2e}")
        assert err < 1e-10, (name, err)
        norm = np.sqrt((np.abs(x) ** 2).sum(axis=(0, 1)))
        np.testing.assert_allclose(norm, golden[f"tend_{tag}_{name}_norm"], rtol=1e-10)


def test_step_and_semi_implicit_match_reference_vectors(env, golden):
    import lbm_isca_amd as L
    sht, _, bs = env
    sp, vg = L.Spectral(), L.VerticalGrid()
    m = L.LBM(bs, sp, vg, L.Config())
    tend = m.dyn.tendencies(hip_state(oracle_state(sht, bs, vg.kmax, 0)))
    prev = hip_state(oracle_state(sht, bs, vg.kmax, 1)).to_spectral(sp)
    for dt in (1200.0, 300.0):
        new = m.integ.step(dt, prev, tend)
        for name in REF_ORDER:
            x = np_(new[name])
            got = x if x.ndim == 2 else x[:, :, LV]
            err = rel_err(got, golden[f"step{int(dt)}_{name}"])
            print(f"step dt={dt:g} {name}: {err:.2e}")
            assert err < 1e-10, (dt, name, err)
    # a dt that is neither a start-up step nor delt is prepared on demand
    new = m.integ.step(700.0, prev, tend)
    assert all(np.isfinite(np_(x)).all() for x in new.values())
    m.semi.prepare(1200.0)
    np.testing.assert_allclose(m.semi.cond, golden["semi_cond_1200"], rtol=1e-8)
    assert m.semi.cond.max() < 100.0
    resid = np.einsum("nij,njk->nik", m.semi.m, m.semi.minv) - np.eye(vg.kmax)
    assert np.abs(resid).max() < 1e-10


def test_resting_isothermal_state_gives_zero_tendencies():
    """tests/test_physics_isca.py:56-61 on the HIP path: exactly zero."""
    import lbm_isca_amd as L
    sp, vg = L.Spectral(), L.VerticalGrid()
    shape = (sp.nlat, sp.nlon, vg.kmax)
    rest = types.SimpleNamespace(u=np.zeros(shape), v=np.zeros(shape), t=np.full(shape, 250.0),
                                 q=np.zeros(shape), ps=np.full(shape[:2], 1000.0))
    m = L.LBM(rest, sp, vg, L.Config(ndays=0.1))
    for name, x in zip(REF_ORDER, m.dyn.tendencies(m.ga)):
        assert np.abs(np_(x)).max() == 0.0, f"{name} tendency is not zero"


def _tl_jet(sp, vg):
    """The synthetic baroclinic jet of analysis/tl_verification.py:58-77."""
    nlat, nlon, k = sp.nlat, sp.nlon, vg.kmax
    lat = np.arcsin(sp.mu)[:, None, None]
    sig = np.asarray(vg.sig)[None, None, :]
    jet = 30.0 * np.exp(-((lat - np.deg2rad(40)) / np.deg2rad(15)) ** 2)
    one = np.ones((nlat, nlon, k))
    return types.SimpleNamespace(
        u=jet * sig ** 0.5 * one, v=np.zeros((nlat, nlon, k)),
        t=(250.0 - 40.0 * np.sin(lat) ** 2 + 5.0 * sig) * one,
        q=1e-3 * np.exp(-(lat / np.deg2rad(30)) ** 2) * one, ps=1.0e5 * np.ones((nlat, nlon)))


def test_tangent_linear_taylor_slope():
    """tests/test_tangent_linear.py:33-69 on device="hip": the divergence
    remainder of F(a d) - F(0) - a L d falls as a^2 (the kept kinetic energy
    term), every other tendency is linear to round-off."""
    import lbm_isca_amd as L
    from lbm_isca_amd.dynamics import BasicStateTerms, Dynamics
    sp, vg = L.Spectral(), L.VerticalGrid()
    bs = _tl_jet(sp, vg)
    bst = BasicStateTerms(sp, vg, bs)
    full = Dynamics(sp, vg, bst, strictly_linear=False)
    lin = Dynamics(sp, vg, bst, strictly_linear=True)
    rng = np.random.default_rng(0)

    def smooth(scale, shape):
        g = np_(sp.wave_to_grid(sp.grid_to_wave(rng.standard_normal(shape))))
        return scale * g / (np.max(np.abs(g)) or 1.0)

    shp = (sp.nlat, sp.nlon, vg.kmax)
    d = (smooth(1.0, shp), smooth(1.0, shp), smooth(1.0, shp), smooth(1e-3, shp),
         smooth(1e-3, shp[:2]))

    def state(a):
        st = L.State.zeros(np.asarray(bs.ps), vg.kmax)
        st.u, st.v, st.t, st.q = (L.state._f64(a * x) for x in d[:4])
        st.set_from_lnps(a * d[4])
        return st

    def vec(dyn, st):
        return [np_(x).ravel() for x in dyn.tendencies(st)]

    f0 = vec(full, state(0.0))
    ld = vec(lin, state(1.0))
    amps = np.logspace(-1, -4, 7)
    rem = {v: [] for v in REF_ORDER}
    rel = {v: [] for v in REF_ORDER}
    for a in amps:
        fa = vec(full, state(a))
        for i, v in enumerate(REF_ORDER):
            r = np.linalg.norm(fa[i] - f0[i] - a * ld[i])
            rem[v].append(r)
            rel[v].append(r / (np.linalg.norm(a * ld[i]) or 1.0))
    slope = float(np.polyfit(np.log(amps), np.log(np.array(rem["div"]) + 1e-300), 1)[0])
    print(f"div slope {slope:.3f}; others {max(max(rel[v]) for v in ('vor', 't', 'q', 'pi')):.1e}")
    assert 1.8 <= slope <= 2.2, slope
    for v in ("vor", "t", "q", "pi"):
        assert max(rel[v]) < 1e-6, (v, max(rel[v]))


def _compare(m, ref, tol, what):
    got = m.ga.to_numpy()
    for f in ("u", "v", "t", "q", "ps"):
        b = getattr(ref.ga, f)[0] if got[f].ndim == getattr(ref.ga, f).ndim - 1 else getattr(ref.ga, f)
        err = rel_err(got[f], b)
        print(f"{what} {f}: {err:.2e}")
        assert err < tol, (what, f, err)


def _compare_prev(m, ref, tol, what):
    gb = m.gb.to_numpy()
    errs = {f: rel_err(gb[f], getattr(ref.gb, f)[0]) for f in ("u", "v", "t", "q", "ps")}
    print(what, " ".join(f"{f} {e:.2e}" for f, e in errs.items()))
    assert max(errs.values()) < tol, (what, errs)


def test_partial_perturbation_keeps_omitted_fields(env):
    """Two partial set_initial_perturbation calls before a run, and one after
    a run, against the oracle (which changes only the given fields on both
    ga and gb, model.py:142-148)."""
    import lbm_isca_amd as L
    sht, vg_o, bs = env
    st = oracle_state(sht, bs, vg_o.kmax, 0)
    heat = cases.heating(sht.lat, sht.lon, vg_o.sig, lat0=10.0, lon0=150.0)
    m = L.LBM(bs, L.Spectral(), L.VerticalGrid(), L.Config(ndays=5 / 72.0))
    ref = OracleModel(bs, 1, OracleConfig())
    m.set_initial_perturbation(u=st.u[0], ps=st.ps[0])
    m.set_initial_perturbation(t=st.t[0])
    ref.set_initial_perturbation(u=st.u, ps=st.ps)
    ref.set_initial_perturbation(t=st.t)
    np.testing.assert_allclose(m.ga.to_numpy()["u"], st.u[0], rtol=0, atol=0)
    frc = L.SteadyForcing.zeros()
    frc.t = heat
    m.run(steady_forcing=frc, verbose=False)
    ref.run(nsteps=5, frc_t=heat[None])
    _compare(m, ref, 1e-9, "two partial sets, 5 steps")
    _compare_prev(m, ref, 1e-9, "previous level after the run")
    # after a run the two levels differ: v (and q) set, the rest kept per level
    m.set_initial_perturbation(v=st.v[0] * 0.5, q=st.q[0])
    ref.set_initial_perturbation(v=st.v * 0.5, q=st.q)
    _compare(m, ref, 1e-9, "current level after the partial set")
    _compare_prev(m, ref, 1e-9, "previous level after the partial set")
    m.run(steady_forcing=frc, verbose=False)
    ref.run(nsteps=5, frc_t=heat[None])
    _compare(m, ref, 1e-9, "partial set after a run, 5 more steps")


class RampForcing:
    """A temperature forcing whose amplitude changes every call (no
    ``constant`` declaration: the run must call .step() before every step)."""

    def __init__(self, heat):
        self.heat, self.calls = heat, 0

    def step(self):
        self.calls += 1
        return self.heat * (1.0 + 0.25 * np.sin(0.3 * self.calls))


def test_time_varying_step_forcing(env):
    import lbm_isca_amd as L
    sht, vg_o, bs = env
    heat = cases.heating(sht.lat, sht.lon, vg_o.sig, lat0=5.0, lon0=90.0, amp_k_day=3.0)
    m = L.LBM(bs, L.Spectral(), L.VerticalGrid(), L.Config(ndays=6 / 72.0))
    f_hip = RampForcing(heat)
    m.run(mjo_forcing=f_hip, verbose=False)
    assert f_hip.calls == 8                      # 2 start-up + 6 full steps

    class OracleRamp(RampForcing):
        def step(self):
            return super().step()[None]

    ref = OracleModel(bs, 1, OracleConfig())
    ref.run(nsteps=6, mjo=OracleRamp(heat))
    _compare(m, ref, 1e-9, "ramp forcing")


def test_writer_failure_does_not_wedge_the_plan(env):
    import lbm_isca_amd as L
    _, _, bs = env

    class Broken:
        def add(self, day, fields):
            raise KeyError("writer broke")

        def close(self):
            pass

    cfg = L.Config(ndays=4 / 72.0, out_interval_days=2 / 72.0)
    m = L.LBM(bs, L.Spectral(), L.VerticalGrid(), cfg)
    with pytest.raises(KeyError, match="writer broke"):
        m.run(writer=Broken(), verbose=False)

    class Count:
        n = 0

        def add(self, day, fields):
            Count.n += 1

        def close(self):
            pass

    m.run(writer=Count(), verbose=False)
    assert Count.n == 2


def test_state_operators_match_numpy(env):
    import lbm_isca_amd as L
    sht, vg_o, bs = env
    sp = L.Spectral()
    st = oracle_state(sht, bs, vg_o.kmax, 0)
    s = hip_state(st)
    pi = np_(s.lnps_pert())
    # log1p(ps'/ps_bar) exactly; the reference's ln(ps_bar + ps') - ln(ps_bar)
    # agrees to its cancellation error (a few ulp of ln ps_bar ~ 7)
    np.testing.assert_allclose(pi, np.log1p(st.ps[0] / bs.ps), rtol=1e-14, atol=1e-19)
    np.testing.assert_allclose(pi, np.log(bs.ps + st.ps[0]) - np.log(bs.ps), rtol=0, atol=4e-15)
    s2 = s.copy()
    s2.set_from_lnps(pi)
    np.testing.assert_allclose(np_(s2.ps), st.ps[0], rtol=1e-12, atol=1e-14)
    # to_spectral / load_spectral against the oracle's GridState
    spec = s.to_spectral(sp)
    ospec = st.to_spectral(sht)
    for k in REF_ORDER:
        assert rel_err(np_(spec[k]), ospec[k][0]) < 1e-12, k
    back = L.State.zeros(bs.ps, vg_o.kmax).load_spectral(sp, spec)
    og = GridState(bs.ps, 1, vg_o.kmax)
    og.load_spectral(sht, ospec)
    for f in ("u", "v", "t", "q", "ps"):
        assert rel_err(np_(getattr(back, f)), getattr(og, f)[0]) < 1e-12, f
    # Robert-Asselin filter
    a, b, c = s.copy(), s.copy(), back
    rng = np.random.default_rng(3)
    b.t = L.state._f64(np_(b.t) + rng.standard_normal(b.t.shape))
    prev = {f: np_(getattr(a, f)) for f in ("u", "v", "t", "q", "ps")}
    a.robert_asselin(b, c, 0.04)
    for f in ("u", "v", "t", "q", "ps"):
        cur, new = np_(getattr(b, f)), np_(getattr(c, f))
        np.testing.assert_allclose(np_(getattr(a, f)), cur + 0.04 * (prev[f] - 2 * cur + new),
                                   rtol=1e-14, atol=1e-14)
    # max_abs / isfinite, zeros_ensemble, assign
    mx = s.max_abs()
    for f in ("u", "v", "t", "q", "ps"):
        assert mx[f] == np.abs(getattr(st, f)[0]).max()
    assert s.isfinite()
    bad = s.copy()
    t = np_(bad.t)
    t[3, 4, 5] = np.nan
    bad.t = L.state._f64(t)
    assert not bad.isfinite()
    ens = L.State.zeros_ensemble(bs.ps, vg_o.kmax, 3)
    assert tuple(ens.u.shape) == (3, sht.nlat, sht.nlon, vg_o.kmax)
    assert tuple(ens.ps.shape) == (3, sht.nlat, sht.nlon)
    ens2 = L.State.zeros_ensemble(bs.ps, vg_o.kmax, 3).assign(ens)
    assert ens2.u.data_ptr() != ens.u.data_ptr()


def test_laplacian_and_hydrostatic_match_numpy(env):
    import lbm_isca_amd as L
    sht, vg_o, _ = env
    sp, vg = L.Spectral(), L.VerticalGrid()
    rng = np.random.default_rng(5)
    xi = cases.random_spectral(rng, 42, vg.kmax, 1.0)
    n = np.arange(43.0)
    lap = -n * (n + 1) / sp.a ** 2
    inv = np.zeros(43)
    inv[1:] = 1.0 / lap[1:]
    np.testing.assert_allclose(np_(sp.laplacian(xi)), xi * lap[None, :, None], rtol=1e-14)
    np.testing.assert_allclose(np_(sp.inverse_laplacian(xi)), xi * inv[None, :, None], rtol=1e-14)
    np.testing.assert_allclose(np_(sp.laplacian(xi[:, :, 0])), xi[:, :, 0] * lap[None, :],
                               rtol=1e-14)
    tv = rng.standard_normal((sht.nlat, sht.nlon, vg.kmax))
    np.testing.assert_allclose(np_(vg.hydrostatic(tv)), tv @ vg_o.ww.T, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(np_(vg.hydrostatic(xi)), xi @ vg_o.ww.T, rtol=1e-12, atol=1e-9)
