"""The benchmark workloads themselves against the FP64 oracle (fixtures of
``tests/golden/make_oracle_runs.py``):

* the single-member solid-body case (20 equal sigma levels, 2 K/day heating
  at 0N 180E, 10 days);
* the 64-member heating ensemble the throughput line times: all 64 members
  run together, the first four compared member by member with the oracle
  after 10 days, with hipGraph replay and with eager launches (the
  per-member semantics of ``analysis/benchmark_ensemble_isca.py:87-125``);
* member 0 of that ensemble over 30 days (2160 steps), FP64 within 1e-9;
* the tangent-linear response and its adjoint gradient on two balanced-jet
  members with their own basic states, J averaged over days 20-30 of a
  30-day run, against the oracle's linear-mode runs (when the fixture
  exists).
"""

import os

import numpy as np
import pytest

from oracle_helpers import rel_err

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("u", "v", "t", "q", "ps")


def fixture(part):
    path = os.path.join(HERE, "golden", f"oracle_{part}.npz")
    if not os.path.exists(path):
        pytest.skip(f"{path} not generated")
    return dict(np.load(path))


def solid_body_model(nmem, ndays, use_graph=True):
    import lbm_isca_amd as L
    from lbm_isca_amd import cases
    sigm = cases.EQUAL_SIGM_L20
    sp = L.BatchedSpectral()
    vg = L.VerticalGrid(sigm)
    bs = cases.solid_body_state(sp.lat, sp.lon, vg.sig)
    m, _, _ = L.make_ensemble_model(bs, nmem, L.Config(ndays=ndays), sigm=sigm)
    return m, sp, vg


def check(ga, fx, tag, members, tol=1e-9):
    for f in FIELDS:
        idx = fx[f"{tag}_{f}_idx"]
        for k, e in enumerate(members):
            arr = ga[f][e]
            err = rel_err(arr.ravel()[idx], fx[f"{tag}_{f}_val"][k])
            nrm = np.sqrt((arr ** 2).sum())
            ref = fx[f"{tag}_{f}_norm"][k]
            nerr = abs(nrm - ref) / ref if ref > 0 else nrm
            print(f"{tag} member {e} {f}: samples {err:.2e} norm {nerr:.2e}")
            assert err < tol and nerr < tol, (tag, e, f, err, nerr)


def test_single_member_solid_body_10_days():
    import lbm_isca_amd as L
    from lbm_isca_amd import cases
    fx = fixture("cfg0")
    m, sp, vg = solid_body_model(1, 10.0)
    heat = cases.heating(sp.lat, sp.lon, vg.sig, 0.0, 180.0, 2.0)
    m.run(mjo_forcing=L.ConstForcing(heat[None]), verbose=False)
    check(m.ga.to_numpy(), fx, "cfg0", [0])


@pytest.mark.parametrize("use_graph", [True, False])
def test_64_member_ensemble_per_member(use_graph):
    import lbm_isca_amd as L
    from lbm_isca_amd import cases
    fx = fixture("cfg1")
    m, sp, vg = solid_body_model(64, 10.0)
    frc = cases.random_heatings(sp.lat, sp.lon, vg.sig, 64, seed=0)
    m.run(mjo_forcing=L.EnsembleForcing(frc), verbose=False, use_graph=use_graph)
    check(m.ga.to_numpy(), fx, "cfg1", [0, 1, 2, 3])


def test_member_0_over_30_days():
    import lbm_isca_amd as L
    from lbm_isca_amd import cases
    fx = fixture("month")
    m, sp, vg = solid_body_model(1, 30.0)
    frc = cases.random_heatings(sp.lat, sp.lon, vg.sig, 64, seed=0)[:1]
    m.run(mjo_forcing=L.EnsembleForcing(frc), verbose=False)
    check(m.ga.to_numpy(), fx, "month", [0])


def _inner(g, x):
    """Euclidean inner product of a complex gradient (dJ/dRe + i dJ/dIm) or a
    real one with the matching input, over the n >= m coefficients."""
    g, x = np.asarray(g), np.asarray(x)
    if np.iscomplexobj(g):
        return float((g.real * x.real + g.imag * x.imag).sum())
    return float((g * x).sum())


def test_adjoint_on_jet_members_days_20_30():
    import lbm_isca_amd as L
    from lbm_isca_amd import adjoint, cases
    import make_oracle_runs as gen
    fx = fixture("adj")
    nmem, nsteps = 2, int(fx["nsteps"])
    sigm = cases.EQUAL_SIGM_L20
    sp = L.BatchedSpectral()
    vg = L.VerticalGrid(sigm)
    bs, _ = cases.jet_states(sp, vg.sig, nmem, seed=1)
    frc = cases.random_heatings(sp.lat, sp.lon, vg.sig, nmem, seed=0)
    cfg = L.Config(tangent_linear=True)
    m, _, _ = L.make_ensemble_model(bs, nmem, cfg, sigm=sigm, per_member_basic=True)
    plan = m.plan
    resp = adjoint.StreamfunctionResponse(
        plan, np.degrees(plan.get("lat", plan.nlat)), np.degrees(plan.get("lon", plan.nlon)),
        plan.get("gw", plan.nlat), plan.get("sig", plan.nlev), 0.2, (30.0, 60.0), (150.0, 210.0),
        20.0, 30.0, nsteps, 72, cfg.inistp)
    np.testing.assert_allclose(resp.wstep, fx["wstep"], rtol=0, atol=0)
    np.testing.assert_allclose(resp.wgrid, fx["wgrid"], rtol=1e-12, atol=1e-18)
    from oracle import SHT, VerticalOps
    spec, dirn = gen.adj_inputs(SHT(), VerticalOps(sigm), nmem)
    order = ("vor", "div", "t", "q", "pi")
    plan.set_forcing(frc)
    plan.set_spectral_state(*[spec[k] for k in order])
    plan.run(nsteps)
    J = plan.get_response()
    gx, gf = plan.adjoint_run(nsteps)
    gx = [np.asarray(L.asnumpy(x)) for x in gx]
    gf = np.asarray(L.asnumpy(gf))
    dfrc = 0.5 * np.roll(frc, 17, axis=2)
    for e in range(nmem):
        err_j = abs(J[e] - fx["J"][e]) / abs(fx["J"][e])
        pred = sum(_inner(g[e], spec[k][e]) for g, k in zip(gx, order)) + _inner(gf[e], frc[e])
        err_g = abs(pred - fx["J"][e]) / abs(fx["J"][e])
        pdir = sum(_inner(g[e], dirn[k][e]) for g, k in zip(gx, order)) + _inner(gf[e], dfrc[e])
        err_d = abs(pdir - fx["Jdir"][e]) / abs(fx["Jdir"][e])
        print(f"member {e}: J {err_j:.2e}, <g,(x0,f)> {err_g:.2e}, <g,d> {err_d:.2e}")
        assert err_j < 1e-9 and err_g < 1e-9 and err_d < 1e-9, (e, err_j, err_g, err_d)
