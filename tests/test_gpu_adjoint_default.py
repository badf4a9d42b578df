"""Gradient of the default model (kinetic energy u'u'/2 + v'v'/2 kept,
dynamics.py:243-244; exact ps' <-> pi' pair, state.py:77-83), which is not
linear in the perturbation: the adjoint about the trajectory of the last
run, recomputed segment by segment from checkpoints.

* central finite differences of J of the nonlinear model along 16 seeded
  directions of the initial state and the forcing, relative error <= 1e-6;
* the gradient of the linearisation about rest (the tangent-linear model's
  adjoint) misses the same differences by far more, so the check sees the
  trajectory terms;
* the gradient does not depend on the checkpoint interval (one that is not
  a multiple of the 8-step graph chunk included), uneven member lengths.
"""

import numpy as np
import pytest

import cases
from oracle import SHT, VerticalOps

pytestmark = pytest.mark.gpu

FIELDS = ("vor", "div", "t", "q", "pi")
NSTEPS = 40


def np_(x):
    from lbm_isca_amd import asnumpy
    return np.asarray(asnumpy(x))


def spectral_state(sht, nmem, seed, scale=1.0):
    """Seeded spectral (vor, div, t, q, pi) of nmem members: winds of a few
    m/s, O(1 K) temperature, ps' of O(1 hPa)."""
    out = {f: [] for f in FIELDS}
    for e in range(nmem):
        psi, chi, t, q, pi = cases.perturbation_spectra(20, seed + 17 * e)
        vor, div = sht.laplacian(psi[None])[0], sht.laplacian(chi[None])[0]
        for f, x in zip(FIELDS, (vor, div, t, q, pi)):
            x = scale * x.copy()
            if f in ("vor", "div"):
                x[0, 0] = 0.0       # the model keeps no global mean of these
            out[f].append(x)
    return {f: np.stack(v) for f, v in out.items()}


def inner(g, x):
    g, x = np.asarray(g), np.asarray(x)
    if np.iscomplexobj(g):
        return float((g.real * x.real + g.imag * x.imag).sum())
    return float((g * x).sum())


def model(bs, nmem, tangent_linear=False, lengths=None):
    import lbm_isca_amd as L
    from lbm_isca_amd import adjoint
    cfg = L.Config(tangent_linear=tangent_linear)
    m, _, _ = L.make_ensemble_model(bs, nmem, cfg)
    plan = m.plan
    lat = np.degrees(plan.get("lat", plan.nlat))
    lon = np.degrees(plan.get("lon", plan.nlon))
    wgrid = adjoint.box_weights(lat, lon, plan.get("gw", plan.nlat), (30.0, 60.0), (150.0, 210.0))
    wlev = adjoint.sigma_weights(plan.get("sig", plan.nlev), 0.2)
    wstep = np.zeros(2 + NSTEPS)
    wstep[12:] = 1.0 / (wstep.size - 12)
    plan.set_response(wgrid, wlev, wstep)
    if lengths is not None:
        plan.set_lengths(lengths)
    return plan


def J_of(plan, spec, frc, nsteps):
    plan.set_forcing(frc)
    plan.set_spectral_state(*[spec[f] for f in FIELDS])
    plan.run(nsteps)
    return plan.get_response()


def gradient(plan, spec, frc, nsteps, every=None):
    if every is not None:
        plan.set_checkpoints(every)
    J = J_of(plan, spec, frc, nsteps)
    gx, gf = plan.adjoint_run(nsteps)
    return J, [np_(x) for x in gx], np_(gf)


@pytest.fixture(scope="module")
def setup():
    sht, vg = SHT(), VerticalOps()
    bs = cases.jet_basic_state(sht.lat, sht.lon, vg.sig)
    heat = np.stack([cases.heating(sht.lat, sht.lon, vg.sig, lat0=10.0, lon0=140.0, amp_k_day=2.0),
                     cases.heating(sht.lat, sht.lon, vg.sig, lat0=-5.0, lon0=250.0, amp_k_day=3.0)])
    x0 = spectral_state(sht, 2, seed=3, scale=2.0)
    return sht, vg, bs, heat, x0


def test_gradient_matches_central_differences(setup):
    sht, vg, bs, heat, x0 = setup
    plan = model(bs, 2)
    J0, gx, gf = gradient(plan, x0, heat, NSTEPS, every=16)
    # the trajectory-free adjoint (linearisation about rest) of the same response
    tl = model(bs, 2, tangent_linear=True)
    _, gx_tl, gf_tl = gradient(tl, x0, heat, NSTEPS)
    rng = np.random.default_rng(11)
    worst, worst_tl = 0.0, 0.0
    for k in range(16):
        d = spectral_state(sht, 2, seed=100 + k, scale=rng.uniform(0.5, 2.0))
        df = heat * rng.standard_normal((2, 1, 1, vg.kmax)) if k % 2 else np.zeros_like(heat)
        h = 1e-3
        jp = J_of(plan, {f: x0[f] + h * d[f] for f in FIELDS}, heat + h * df, NSTEPS)
        jm = J_of(plan, {f: x0[f] - h * d[f] for f in FIELDS}, heat - h * df, NSTEPS)
        fd = (jp - jm) / (2 * h)
        for e in range(2):
            pred = sum(inner(g[e], d[f][e]) for g, f in zip(gx, FIELDS)) + inner(gf[e], df[e])
            pred_tl = sum(inner(g[e], d[f][e]) for g, f in zip(gx_tl, FIELDS)) + \
                inner(gf_tl[e], df[e])
            err = abs(pred - fd[e]) / abs(fd[e])
            err_tl = abs(pred_tl - fd[e]) / abs(fd[e])
            worst, worst_tl = max(worst, err), max(worst_tl, err_tl)
            assert err < 1e-6, (k, e, err, fd[e], pred)
    print(f"central differences: worst rel err {worst:.2e}; linearisation about rest {worst_tl:.2e}")
    assert worst_tl > 1e-4, "the check does not see the trajectory terms"


def test_gradient_independent_of_checkpoint_interval(setup):
    sht, vg, bs, heat, x0 = setup
    lengths = [NSTEPS, NSTEPS - 15]
    plan = model(bs, 2, lengths=lengths)
    res = {}
    for every in (48, 16, 7):
        J, gx, gf = gradient(plan, x0, heat, 0, every=every)
        res[every] = (J, gx, gf)
    J, gx, gf = res[48]
    for every in (16, 7):
        J2, gx2, gf2 = res[every]
        np.testing.assert_allclose(J2, J, rtol=1e-13)
        for a, b in zip(gx2 + [gf2], gx + [gf]):
            scale = np.abs(b).max()
            assert np.abs(a - b).max() <= 1e-12 * scale, every
    # the shorter member: its own central difference along one direction
    d = spectral_state(sht, 2, seed=500)
    h = 1e-3
    jp = J_of(plan, {f: x0[f] + h * d[f] for f in FIELDS}, heat, 0)
    jm = J_of(plan, {f: x0[f] - h * d[f] for f in FIELDS}, heat, 0)
    for e in range(2):
        pred = sum(inner(g[e], d[f][e]) for g, f in zip(gx, FIELDS))
        fd = (jp[e] - jm[e]) / (2 * h)
        assert abs(pred - fd) / abs(fd) < 1e-6, (e, pred, fd)


def test_default_model_adjoint_needs_checkpoints(setup):
    sht, vg, bs, heat, x0 = setup
    plan = model(bs, 2)
    plan.set_checkpoints(0)
    J_of(plan, x0, heat, 8)
    with pytest.raises(RuntimeError, match="checkpoints"):
        plan.adjoint_run(8)


def test_autograd_default_model(setup):
    """torch.autograd through ``adjoint.response`` on the default model
    (checkpoints turned on by it) == the sweep's gradient weighted by the
    cotangent; a backward after the response changed is refused."""
    import torch
    from lbm_isca_amd import adjoint
    sht, vg, bs, heat, x0 = setup
    plan = model(bs, 2)
    plan.set_checkpoints(0)
    x = {f: torch.tensor(x0[f], device="cuda", requires_grad=True) for f in FIELDS}
    frc = torch.tensor(heat, device="cuda", requires_grad=True)
    J = adjoint.response(plan, NSTEPS, frc, *[x[f] for f in FIELDS], checkpoint_every=16)
    c = torch.tensor([0.7, -1.3], dtype=torch.float64, device="cuda")
    (J * c).sum().backward()
    J2, gx, gf = gradient(plan, x0, heat, NSTEPS)
    np.testing.assert_allclose(np_(J), J2, rtol=1e-14)
    cn = np_(c)
    for g, f in zip(gx, FIELDS):
        want = g * cn.reshape((2,) + (1,) * (g.ndim - 1))
        assert np.abs(np_(x[f].grad) - want).max() <= 1e-13 * np.abs(want).max(), f
    want = gf * cn.reshape(2, 1, 1, 1)
    assert np.abs(np_(frc.grad) - want).max() <= 1e-13 * np.abs(want).max()
    # the plan changes between the forward and the backward: refused
    J = adjoint.response(plan, NSTEPS, frc, *[x[f] for f in FIELDS])
    plan.set_lengths([NSTEPS, NSTEPS])
    with pytest.raises(RuntimeError, match="changed"):
        J.sum().backward()


def test_adjoint_refuses_a_forcing_changed_during_the_run(setup):
    """A forcing uploaded between the steps of the run (a time-varying
    .step() forcing) is not the constant forcing the sweep differentiates:
    refused."""
    sht, vg, bs, heat, x0 = setup
    plan = model(bs, 2)
    plan.set_checkpoints(16)
    plan.set_forcing(heat)
    plan.set_spectral_state(*[x0[f] for f in FIELDS])
    plan.run_steps(8, 2)
    plan.set_forcing(0.5 * heat)
    plan.run_to(8, 0)
    with pytest.raises(RuntimeError, match="forcing changed"):
        plan.adjoint_run(8)
