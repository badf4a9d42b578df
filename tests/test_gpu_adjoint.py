"""The adjoint of the tangent-linear model: its forward model against the
oracle's linear mode, the response functional against the oracle, the
dot-product identity J = <g_x, x0> + <g_f, f>, central finite differences
along seeded directions, the oracle's tangent-linear run along directions,
uneven lengths with per-member basic states, FP32, the autograd Function and
graph replay."""

import types

import numpy as np
import pytest

import cases
from oracle import OracleConfig, OracleModel, SHT, VerticalOps
from oracle_helpers import oracle_response, oracle_state, rel_err

pytestmark = pytest.mark.gpu

NSTEPS = 8
FIELDS = ("vor", "div", "t", "q", "pi")


def np_(x):
    from lbm_isca_amd import asnumpy
    return asnumpy(x)


@pytest.fixture(scope="module")
def env():
    sht, vg = SHT(), VerticalOps()
    bs = cases.jet_basic_state(sht.lat, sht.lon, vg.sig)
    return sht, vg, bs


def tl_model(bs, nmem=1, precision="fp64", per_member=False):
    import lbm_isca_amd as L
    cfg = L.Config(precision=precision, tangent_linear=True)
    if nmem == 1 and not per_member:
        return L.LBM(bs, L.Spectral(), L.VerticalGrid(), cfg)
    m, _, _ = L.make_ensemble_model(bs, nmem, cfg, per_member_basic=per_member)
    return m


def spectral_state(sht, nmem, seed, imag_m0=False):
    """Seeded spectral (vor, div, t, q, pi) of nmem members; physical (real
    m = 0) unless imag_m0."""
    out = {f: [] for f in FIELDS}
    for e in range(nmem):
        psi, chi, t, q, pi = cases.perturbation_spectra(20, seed + 17 * e)
        vor, div = sht.laplacian(psi[None])[0], sht.laplacian(chi[None])[0]
        for f, x in zip(FIELDS, (vor, div, t, q, pi)):
            x = x.copy()
            if imag_m0:
                rng = np.random.default_rng(seed + 1000 + e)
                x[0] = x[0] + 1j * np.abs(x[0]) * rng.standard_normal(x[0].shape)
            out[f].append(x)
    return {f: np.stack(v) for f, v in out.items()}


def response(plan, nsteps, steps=None):
    from lbm_isca_amd import adjoint
    lat = np.degrees(plan.get("lat", plan.nlat))
    lon = np.degrees(plan.get("lon", plan.nlon))
    wgrid = adjoint.box_weights(lat, lon, plan.get("gw", plan.nlat), (30.0, 60.0), (150.0, 210.0))
    wlev = adjoint.sigma_weights(plan.get("sig", plan.nlev), 0.2)
    wstep = np.zeros(2 + nsteps)
    for k in (steps if steps is not None else (1, 4, 2 + nsteps - 1)):
        wstep[k] = 1.0 + 0.25 * k
    plan.set_response(wgrid, wlev, wstep)
    return wgrid, wlev, wstep


def forward_J(plan, nsteps, spec, frc, use_graph=True):
    plan.set_forcing(frc)
    plan.set_spectral_state(*[spec[f] for f in FIELDS])
    plan.run(nsteps, use_graph=use_graph)
    return plan.get_response()


def inner(a, b):
    """sum Re(a) Re(b) + Im(a) Im(b) per member (leading axis)."""
    a, b = np.asarray(a), np.asarray(b)
    return (a.real * b.real + a.imag * b.imag).reshape(a.shape[0], -1).sum(1)


def heatings(sht, vg, nmem, seed=0):
    rng = np.random.default_rng(seed)
    return np.stack([cases.heating(sht.lat, sht.lon, vg.sig, lat0=rng.uniform(-20, 20),
                                   lon0=rng.uniform(0, 360), amp_k_day=rng.uniform(0.5, 3))
                     for _ in range(nmem)])


def test_tangent_linear_forward_matches_oracle(env):
    """Device tangent-linear run from a raw grid state == the oracle's linear
    mode (pi' = ps'/ps_bar), start-up and Robert-Asselin included."""
    import lbm_isca_amd as L
    sht, vg, bs = env
    st1 = oracle_state(sht, bs, vg.kmax, 3)
    rng = np.random.default_rng(5)
    t_raw = st1.t[0] + 0.05 * rng.standard_normal((sht.nlat, sht.nlon, vg.kmax))
    heat = cases.heating(sht.lat, sht.lon, vg.sig, lat0=10.0, lon0=150.0)
    nsteps = 12
    m = tl_model(bs)
    m.cfg.ndays = nsteps / 72.0
    m.set_initial_perturbation(u=st1.u[0], v=st1.v[0], t=t_raw, q=st1.q[0], ps=st1.ps[0])
    m.run(mjo_forcing=L.ConstForcing(heat), verbose=False)
    ga = m.ga.to_numpy()
    ref = OracleModel(bs, 1, OracleConfig(linear=True))
    ref.set_initial_perturbation(u=st1.u, v=st1.v, t=t_raw[None], q=st1.q, ps=st1.ps)
    ref.run(nsteps=nsteps, frc_t=heat[None])
    for f in ("u", "v", "t", "q", "ps"):
        err = rel_err(ga[f], getattr(ref.ga, f)[0])
        print(f"tl {f}: {err:.2e}")
        assert err < 1e-9, (f, err)


def test_response_matches_oracle(env):
    sht, vg, bs = env
    E = 2
    m = tl_model(bs, nmem=E)
    plan = m.plan
    spec = spectral_state(sht, E, 11)
    frc = heatings(sht, vg, E, 1)
    wgrid, wlev, wstep = response(plan, NSTEPS, steps=(0, 1, 3, 6, 9))
    J = forward_J(plan, NSTEPS, spec, frc)
    ref = oracle_response(bs, E, spec, frc, NSTEPS, wgrid, wlev, wstep)
    print("J", J, ref)
    assert rel_err(J, ref) < 1e-9


@pytest.mark.parametrize("E", [3, 12])
@pytest.mark.parametrize("use_graph", [True, False])
def test_dot_product_identity(env, use_graph, E):
    """J_e(x0, f) = <g_x, x0_e> + <g_f, f_e> for every member, the imaginary
    parts of m = 0 included (carried by the step, invisible to the grid).
    12 members: the forward and adjoint steps fork their second stream."""
    sht, vg, bs = env
    m = tl_model(bs, nmem=E)
    plan = m.plan
    response(plan, NSTEPS)
    spec = spectral_state(sht, E, 21, imag_m0=True)
    frc = heatings(sht, vg, E, 2)
    J = forward_J(plan, NSTEPS, spec, frc, use_graph)
    gx, gf = plan.adjoint_run(NSTEPS, use_graph=use_graph)
    terms = [inner(np_(g), spec[f]) for g, f in zip(gx, FIELDS)] + [inner(np_(gf), frc)]
    pred = np.sum(terms, axis=0)
    scale = np.sum(np.abs(terms), axis=0)
    print("J", J, "pred", pred, "terms", np.array(terms))
    assert np.all(np.abs(J - pred) <= 1e-11 * scale), (J, pred)
    # the m = 0 imaginary parts never reach the grid: zero gradient there
    for g in gx:
        assert np.abs(np_(g)[:, 0].imag).max() <= 1e-12 * (np.abs(np_(g)).max() + 1e-300)


def test_gradient_matches_central_differences(env):
    """16 seeded directions in (x0, f): (J(x+h d) - J(x-h d)) / 2h = <g, d>."""
    sht, vg, bs = env
    E = 2
    m = tl_model(bs, nmem=E)
    plan = m.plan
    response(plan, NSTEPS)
    spec = spectral_state(sht, E, 31)
    frc = heatings(sht, vg, E, 3)
    J0 = forward_J(plan, NSTEPS, spec, frc)
    gx, gf = plan.adjoint_run(NSTEPS)
    gx = [np_(g) for g in gx]
    gf = np_(gf)
    rng = np.random.default_rng(2024)
    for k in range(16):
        dspec = spectral_state(sht, E, 100 + k)
        dfrc = heatings(sht, vg, E, 200 + k) * rng.standard_normal()
        h = 0.5
        Jp = forward_J(plan, NSTEPS, {f: spec[f] + h * dspec[f] for f in FIELDS}, frc + h * dfrc)
        Jm = forward_J(plan, NSTEPS, {f: spec[f] - h * dspec[f] for f in FIELDS}, frc - h * dfrc)
        fd = (Jp - Jm) / (2 * h)
        pred = np.sum([inner(g, dspec[f]) for g, f in zip(gx, FIELDS)], axis=0) + inner(gf, dfrc)
        err = np.abs(fd - pred) / np.maximum(np.abs(fd), 1e-300)
        assert np.all(err < 1e-6), (k, fd, pred)
    assert np.all(np.isfinite(J0))


def test_gradient_matches_oracle_tangent_linear(env):
    """<g, d> == the oracle's linear-mode response along two directions."""
    sht, vg, bs = env
    E = 1
    m = tl_model(bs)
    plan = m.plan
    wgrid, wlev, wstep = response(plan, NSTEPS)
    forward_J(plan, NSTEPS, spectral_state(sht, E, 41), heatings(sht, vg, E, 4))
    gx, gf = plan.adjoint_run(NSTEPS)
    gx = [np_(g) for g in gx]
    gf = np_(gf)
    for k in range(2):
        d = spectral_state(sht, E, 300 + k)
        df = heatings(sht, vg, E, 400 + k)
        pred = np.sum([inner(g, d[f]) for g, f in zip(gx, FIELDS)], axis=0) + inner(gf, df)
        ref = oracle_response(bs, E, d, df, NSTEPS, wgrid, wlev, wstep)
        print("pred", pred, "oracle", ref)
        assert rel_err(pred, ref) < 1e-9


def test_uneven_lengths_per_member_basic_state(env):
    """Members with their own basic state and length: each member's gradient
    satisfies the identity, and a member that stops before the response
    window has J = 0 and a zero gradient."""
    sht, vg, bs = env
    bs2 = cases.jet_basic_state(sht.lat - 0.1, sht.lon, vg.sig)
    bs3 = cases.jet_basic_state(sht.lat + 0.2, sht.lon, vg.sig)
    stack = types.SimpleNamespace(**{k: np.stack([getattr(b, k) for b in (bs, bs2, bs3)])
                                     for k in ("u", "v", "t", "q", "ps")})
    E = 3
    m = tl_model(stack, nmem=E, per_member=True)
    plan = m.plan
    lengths = [NSTEPS, 3, 1]
    response(plan, NSTEPS, steps=(1, 4, 6, 9))
    plan.set_lengths(lengths)
    spec = spectral_state(sht, E, 51)
    frc = heatings(sht, vg, E, 5)
    J = forward_J(plan, 0, spec, frc)
    gx, gf = plan.adjoint_run(0)
    terms = [inner(np_(g), spec[f]) for g, f in zip(gx, FIELDS)] + [inner(np_(gf), frc)]
    pred = np.sum(terms, axis=0)
    scale = np.sum(np.abs(terms), axis=0)
    print("J", J, "pred", pred)
    assert np.all(np.abs(J - pred) <= 1e-11 * np.maximum(scale, 1e-300))
    # member 2 runs 1 full step (ctl steps 0..2): only weight 1 applies
    # member 1 (3 full steps, ctl 0..4) sees weights 1 and 4
    for e in range(E):
        ref = oracle_response([bs, bs2, bs3][e], 1, {f: spec[f][e:e + 1] for f in FIELDS},
                              frc[e:e + 1], lengths[e], *plan._resp[:2],
                              plan._resp[2][:2 + lengths[e]])
        assert rel_err(J[e:e + 1], ref) < 1e-9, (e, J[e], ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_identity_members_with_own_basic_states(env, precision):
    """Sixteen members on eight jets shifted in latitude, their own basic
    states (the spectral step and its transpose take one total wavenumber
    and a group of 8 (FP64) or 16 (FP32) members per wave): the dot-product
    identity per member."""
    sht, vg, bs = env
    E = 16
    jets = [cases.jet_basic_state(sht.lat + 0.02 * (k % 8 - 4), sht.lon, vg.sig) for k in range(E)]
    stack = types.SimpleNamespace(**{k: np.stack([getattr(b, k) for b in jets])
                                     for k in ("u", "v", "t", "q", "ps")})
    m = tl_model(stack, nmem=E, precision=precision, per_member=True)
    plan = m.plan
    response(plan, NSTEPS, steps=(1, 4, 6, 9))
    spec = spectral_state(sht, E, 61)
    frc = heatings(sht, vg, E, 6)
    J = forward_J(plan, NSTEPS, spec, frc)
    gx, gf = plan.adjoint_run(NSTEPS)
    terms = [inner(np_(g), spec[f]) for g, f in zip(gx, FIELDS)] + [inner(np_(gf), frc)]
    pred = np.sum(terms, axis=0)
    scale = np.sum(np.abs(terms), axis=0)
    tol = 1e-11 if precision == "fp64" else 1e-4
    print(precision, "J", J, "pred", pred)
    assert np.all(np.abs(J - pred) <= tol * np.maximum(scale, 1e-300))


def test_fp32_adjoint(env):
    """FP32 gradient against the FP64 gradient and the identity at FP32."""
    sht, vg, bs = env
    E = 2
    spec = spectral_state(sht, E, 61)
    frc = heatings(sht, vg, E, 6)
    out = {}
    for prec in ("fp64", "fp32"):
        m = tl_model(bs, nmem=E, precision=prec)
        plan = m.plan
        response(plan, NSTEPS)
        J = forward_J(plan, NSTEPS, spec, frc)
        gx, gf = plan.adjoint_run(NSTEPS)
        out[prec] = (J, [np_(g) for g in gx], np_(gf))
    J64, g64, f64 = out["fp64"]
    J32, g32, f32 = out["fp32"]
    assert rel_err(J32, J64) < 1e-4
    for a, b in zip(g32, g64):
        assert rel_err(a, b) < 1e-4
    assert rel_err(f32, f64) < 1e-4
    terms = [inner(g, spec[f]) for g, f in zip(g32, FIELDS)] + [inner(f32, frc)]
    assert np.all(np.abs(J32 - np.sum(terms, axis=0)) <= 1e-4 * np.sum(np.abs(terms), axis=0))


def test_autograd_function(env):
    """torch.autograd through LinearResponse == the adjoint gradients weighted
    by the output cotangent; a shared forcing gets the member sum."""
    import torch
    from lbm_isca_amd import adjoint
    sht, vg, bs = env
    E = 2
    m = tl_model(bs, nmem=E)
    plan = m.plan
    response(plan, NSTEPS)
    spec = spectral_state(sht, E, 71)
    x = {f: torch.tensor(spec[f], device="cuda", requires_grad=True) for f in FIELDS}
    frc = torch.tensor(heatings(sht, vg, 1, 7), device="cuda", requires_grad=True)
    J = adjoint.linear_response(plan, NSTEPS, frc, *[x[f] for f in FIELDS])
    c = torch.tensor([0.7, -1.3], dtype=torch.float64, device="cuda")
    (J * c).sum().backward()
    gx, gf = plan.adjoint_run(NSTEPS)
    for g, f in zip(gx, FIELDS):
        want = np_(g) * np_(c).reshape((E,) + (1,) * (g.dim() - 1))
        assert rel_err(np_(x[f].grad), want) < 1e-13, f
    want = (np_(gf) * np_(c).reshape(E, 1, 1, 1)).sum(0, keepdims=True)
    assert rel_err(np_(frc.grad), want) < 1e-13
    # J is linear: J(x) = <grad, x> summed over inputs
    pred = sum(float(np.sum(inner(np_(x[f].grad), spec[f]))) for f in FIELDS)
    pred += float(np.sum(np_(frc.grad) * np_(frc)))
    assert abs(pred - float((J * c).sum().detach())) <= 1e-10 * abs(pred)


def test_adjoint_needs_response_and_checkpoints(env):
    """The default model's adjoint needs the checkpoints of a finished run
    (test_gpu_adjoint_default.py covers its gradient); every adjoint needs a
    response."""
    import lbm_isca_amd as L
    sht, vg, bs = env
    m = L.LBM(bs, L.Spectral(), L.VerticalGrid(), L.Config())
    m.plan.set_response(np.ones((sht.nlat, sht.nlon)) / (sht.nlat * sht.nlon), None, np.ones(6))
    with pytest.raises(RuntimeError, match="checkpoints"):
        m.plan.adjoint_run(4)
    m2 = tl_model(bs)
    with pytest.raises(RuntimeError, match="response"):
        m2.plan.adjoint_run(4)


@pytest.mark.parametrize("nlon,nlat,nmax", [(128, 64, 42), (384, 192, 127)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_transform_adjoint_twins(nlon, nlat, nmax, precision):
    """<T x, y> = <x, T^T y> for the five API transforms and their twins
    (the transposed Legendre pair with the other table set, the FFT pair
    with its transposed scalings); y spectral has zeros for n < m."""
    from lbm_isca_amd.plan import Plan
    p = Plan(nlon, nlat, nmax, precision=precision)
    rng = np.random.default_rng(7)
    M, B, K = nmax + 1, 2, 3
    mask = (np.arange(M)[None, :] >= np.arange(M)[:, None])[None, :, :, None]

    def spec():
        return (rng.standard_normal((B, M, M, K)) + 1j * rng.standard_normal((B, M, M, K))) * mask

    def grid():
        return rng.standard_normal((B, nlat, nlon, K))

    def ip(a, b):
        a, b = np.asarray(a), np.asarray(b)
        return float(np.sum(a.real * b.real + a.imag * b.imag))

    tol = 1e-12 if precision == "fp64" else 2e-5
    cases_ = []
    g, w = grid(), spec()
    cases_.append(("g2w", ip(np_(p.grid_to_wave(g)), w), ip(g, np_(p.grid_to_wave_adj(w)))))
    cases_.append(("w2g", ip(np_(p.wave_to_grid(w)), g), ip(w, np_(p.wave_to_grid_adj(g)))))
    u, v, vo, dv = grid(), grid(), spec(), spec()
    a, b = p.uv_to_vordiv(u, v)
    ua, va = p.uv_to_vordiv_adj(vo, dv)
    cases_.append(("uv2vd", ip(np_(a), vo) + ip(np_(b), dv), ip(u, np_(ua)) + ip(v, np_(va))))
    ps_, ch = spec(), spec()
    a, b = p.psichi_to_uv(ps_, ch)
    pa, ca = p.psichi_to_uv_adj(u, v)
    cases_.append(("pc2uv", ip(np_(a), u) + ip(np_(b), v), ip(ps_, np_(pa)) + ip(ch, np_(ca))))
    a, b = p.hgrad(w)
    wa = p.hgrad_adj(u, v)
    cases_.append(("hgrad", ip(np_(a), u) + ip(np_(b), v), ip(w, np_(wa))))
    for name, lhs, rhs in cases_:
        print(f"{precision} {nlon}: {name} {lhs:.12e} {rhs:.12e}")
        assert abs(lhs - rhs) <= tol * max(abs(lhs), abs(rhs)), (name, lhs, rhs)


def test_adjoint_live_member_prefix_launches_are_exact(env):
    """The adjoint of sixteen members ordered by decreasing length (shrunken
    launches) equals that of the reversed order, bitwise."""
    sht, vg, bs = env
    E = 16
    lengths = np.array([8, 8, 7, 7, 6, 6, 5, 5, 4, 4, 3, 3, 2, 2, 1, 1])
    spec = spectral_state(sht, E, 91)
    frc = heatings(sht, vg, E, 9)
    res = []
    for order in (np.arange(E), np.arange(E)[::-1]):
        m = tl_model(bs, nmem=E)
        plan = m.plan
        response(plan, 8, steps=(1, 3, 5, 7, 9))
        plan.set_lengths(lengths[order])
        J = forward_J(plan, 0, {f: spec[f][order] for f in FIELDS}, frc[order])
        gx, gf = plan.adjoint_run(0)
        inv = np.argsort(order)
        res.append((J[inv], [np_(g)[inv] for g in gx], np_(gf)[inv]))
    assert np.array_equal(res[0][0], res[1][0])
    for a, b in zip(res[0][1], res[1][1]):
        assert np.array_equal(a, b)
    assert np.array_equal(res[0][2], res[1][2])
