"""The precision modes against the FP64 oracle: FP32 and split-BF16x3
Legendre contractions (FP32 data, each operand split into two BF16 parts,
three BF16 products accumulated in FP32) for the transforms and one model
day at T42 L20, and three steps at T127 L40."""

import types

import numpy as np
import pytest

import cases
from oracle import OracleConfig, OracleModel, SHT, VerticalOps
from oracle_helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_transforms_low_precision(precision):
    import lbm_isca_amd as L
    sht = SHT()
    sp = L.Spectral(precision=precision)
    rng = np.random.default_rng(1)
    g = sht.wave_to_grid(sht.grid_to_wave(rng.standard_normal((1, 64, 128, 4))))[0]
    w = sht.grid_to_wave(g[None])[0]
    e1 = rel_err(L.asnumpy(sp.grid_to_wave(g)), w)
    e2 = rel_err(L.asnumpy(sp.wave_to_grid(w)), g)
    u, v = rng.standard_normal((2, 64, 128, 3))
    vor, div = sp.uv_to_vordiv(u, v)
    rv, rd = sht.uv_to_vordiv(u[None], v[None])
    e3 = max(rel_err(L.asnumpy(vor), rv[0]), rel_err(L.asnumpy(div), rd[0]))
    print(f"{precision}: g2w {e1:.2e} w2g {e2:.2e} uv2vd {e3:.2e}")
    bound = 1e-6 if precision == "fp32" else 5e-5
    assert max(e1, e2, e3) < bound


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_one_day_low_precision_against_oracle(precision):
    import lbm_isca_amd as L
    sht, vg = SHT(), VerticalOps()
    bs = cases.jet_basic_state(sht.lat, sht.lon, vg.sig)
    heat = cases.heating(sht.lat, sht.lon, vg.sig)
    m = L.LBM(bs, L.Spectral(), L.VerticalGrid(), L.Config(ndays=1.0, precision=precision))
    m.run(mjo_forcing=L.ConstForcing(heat), verbose=False)
    ga = m.ga.to_numpy()
    ref = OracleModel(bs, 1, OracleConfig())
    ref.run(nsteps=72, frc_t=heat[None])
    errs = {f: rel_err(ga[f], getattr(ref.ga, f)[0]) for f in ("u", "v", "t", "ps")}
    print(precision, {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) < 1e-4


def test_t127l40_bf16x3_against_oracle():
    import lbm_isca_amd as L
    from lbm_isca_amd import cases as C
    sp = L.BatchedSpectral(384, 192, 127)
    vg = L.VerticalGrid(C.EQUAL_SIGM_L40)
    bs, _ = C.jet_states(sp, vg.sig, 1, seed=11)
    vor = C.vorticity_perturbation(sp, vg.kmax, 1, seed=4)
    z = np.zeros_like(vor)
    spec = {"vor": vor, "div": 0.3 * vor, "t": 1.0e4 * vor, "q": z, "pi": z[..., 0]}
    m, _, _ = L.make_ensemble_model(bs, 1, L.Config(**C.t127l40_config(precision="bf16x3")),
                                    sigm=C.EQUAL_SIGM_L40, per_member_basic=True, spectral=sp)
    m.plan.set_spectral_state(**spec)
    m.plan.run(1, use_graph=False)
    got = [x.cpu().numpy() for x in m.plan.get_state()]
    be = types.SimpleNamespace(**{k: getattr(bs, k)[0] for k in ("u", "v", "t", "q", "ps")})
    ref = OracleModel(be, 1, OracleConfig(**C.t127l40_config()), sigm=C.EQUAL_SIGM_L40, nlon=384,
                      nlat=192, nmax=127, refine_gauss=True)
    ref.ga.load_spectral(ref.sht, spec)
    ref.gb.load_spectral(ref.sht, spec)
    ref.run(nsteps=1)
    for f, a in zip(("u", "v", "t", "ps"), (got[0], got[1], got[2], got[4])):
        e = rel_err(a, getattr(ref.ga, f))
        print(f"T127L40 bf16x3 {f}: {e:.2e}")
        assert e < 1e-4, (f, e)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_t127l40_days_low_precision_against_fp64(precision):
    """Five days of two T127 L40 balanced-jet members (their own basic
    states, n^-1.5 vorticity perturbation) in FP32 and split-BF16x3 against
    the same run in FP64 (itself within 1e-9 of the oracle over its first
    steps, test_gpu_t127.py): every field within 1e-4."""
    import lbm_isca_amd as L
    from lbm_isca_amd import cases as C
    sp = L.BatchedSpectral(384, 192, 127)
    vg = L.VerticalGrid(C.EQUAL_SIGM_L40)
    bs, _ = C.jet_states(sp, vg.sig, 2, seed=1)
    vor = C.vorticity_perturbation(sp, vg.kmax, 2, seed=4)
    out = {}
    for prec in ("fp64", precision):
        cfg = L.Config(**C.t127l40_config(precision=prec))
        m, _, _ = L.make_ensemble_model(bs, 2, cfg, sigm=C.EQUAL_SIGM_L40, per_member_basic=True,
                                        spectral=sp)
        m.plan.set_spectral_state(vor=vor)
        m.plan.run(5 * 192)
        out[prec] = [x.cpu().numpy() for x in m.plan.get_state()]
    errs = {f: rel_err(a, b) for f, a, b in zip(("u", "v", "t", "q", "ps"), out[precision],
                                                out["fp64"]) if np.abs(b).max() > 0}
    print(precision, {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) < 1e-4, errs
