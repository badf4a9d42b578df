"""The basic-state pipeline's device part (spectral smoothing) against the
oracle's transforms, and a balanced-jet ensemble with per-member basic
states against separate oracle runs of its members."""

import types

import numpy as np
import pytest

import cases as golden_cases
from oracle import OracleConfig, OracleModel, SHT, VerticalOps
from oracle_helpers import rel_err

pytestmark = pytest.mark.gpu


def test_specsm_matches_oracle_truncation():
    import lbm_isca_amd as L
    from lbm_isca_amd import basicstate
    sht = SHT()
    sp = L.BatchedSpectral()
    rng = np.random.default_rng(0)
    f = rng.standard_normal((3, 64, 128, 5))
    got = basicstate.specsm(f, sp)
    want = sht.wave_to_grid(sht.grid_to_wave(f))
    assert rel_err(got, want) < 1e-12
    ps = 1000.0 + rng.standard_normal((3, 64, 128))
    got = basicstate.specsm(ps, sp)
    want = sht.wave_to_grid(sht.grid_to_wave(ps[..., None]))[..., 0]
    assert rel_err(got, want) < 1e-12


def test_balanced_jet_ensemble_matches_oracle_members():
    """Three balanced-jet members (their own T_ref, M(n)^-1, stationary
    waves: v != 0, zonally varying T) with heatings and uneven lengths."""
    import lbm_isca_amd as L
    from lbm_isca_amd import cases
    sigm = cases.EQUAL_SIGM_L20
    sp = L.BatchedSpectral()
    vg = L.VerticalGrid(sigm)
    E = 3
    bs, par = cases.jet_states(sp, vg.sig, E, seed=7)
    for k in ("u", "v", "t", "ps"):
        assert np.isfinite(getattr(bs, k)).all()
    assert np.abs(bs.v).max() > 0.1 and 15.0 < bs.u.max() < 60.0
    frc = cases.random_heatings(sp.lat, sp.lon, vg.sig, E, seed=0)
    lengths = [10, 4, 7]
    m, _, _ = L.make_ensemble_model(bs, E, L.Config(), sigm=sigm, per_member_basic=True)
    m.run(mjo_forcing=L.EnsembleForcing(frc), lengths=lengths, verbose=False)
    ga = m.ga.to_numpy()
    for e in range(E):
        be = types.SimpleNamespace(**{k: getattr(bs, k)[e] for k in ("u", "v", "t", "q", "ps")})
        ref = OracleModel(be, 1, OracleConfig(), sigm=sigm)
        ref.run(nsteps=lengths[e], frc_t=frc[e:e + 1])
        for f in ("u", "v", "t", "ps"):
            err = rel_err(ga[f][e], getattr(ref.ga, f)[0])
            print(e, f, f"{err:.2e}")
            assert err < 1e-9, (e, f, err)
