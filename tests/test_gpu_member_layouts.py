"""Members are independent: a member's numbers do not depend on how many
members share its plan.  Five members with one basic state and five
distinct heatings run alone and as the first five of a 64-member ensemble
(the rest other heatings) must end bit for bit equal.  The two plans take
different launch shapes for the same arithmetic:

* the spectral step maps one coefficient and 64 members to a wave at 64
  members and the (m, member) pairs of one total wavenumber at 5;
* the grid phase is k_grid at 64 members and the three launches at 5
  members in FP32 (the built-in choice by ensemble size);
* the Legendre and FFT kernels cover 64 or 8 (padded) member columns.

And the built-in grid phase follows the ensemble size.
"""

import types

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_five_members_alone_and_in_64(precision):
    import lbm_isca_amd as L
    from lbm_isca_amd import cases
    sigm = cases.EQUAL_SIGM_L20
    sp = L.BatchedSpectral()
    vg = L.VerticalGrid(sigm)
    bs = cases.solid_body_state(sp.lat, sp.lon, vg.sig)
    frc = cases.random_heatings(sp.lat, sp.lon, vg.sig, 64, seed=3)
    nsteps = 30
    out = {}
    for E in (64, 5):
        m, _, _ = L.make_ensemble_model(bs, E, L.Config(ndays=nsteps / 72.0, precision=precision),
                                        sigm=sigm)
        m.plan.set_forcing(frc[:E])
        m.plan.set_state()
        m.plan.run(nsteps)
        out[E] = ([x.cpu().numpy()[:5] for x in m.plan.get_state()],
                  [x.cpu().numpy()[:5] for x in m.plan.get_spectral(0)],
                  m.plan.get_tuning())
    big, small = out[64], out[5]
    print(precision, "tunings", big[2], small[2])
    assert np.abs(small[0][2]).max() > 0
    for a, b in zip(big[0] + big[1], small[0] + small[1]):
        assert a.shape == b.shape
        assert np.array_equal(a, b), (precision, float(np.abs(a - b).max()))


def test_builtin_grid_phase_by_ensemble_size():
    """The built-in grid phase (tuning knob 4): the three launches (0) for
    up to 4 FP64 or 8 FP32 members, k_grid with 4 FP32 members per block (3)
    for 9-16, with the larger member group (4) above; every other knob at 0."""
    import lbm_isca_amd as L
    from lbm_isca_amd import cases
    sp = L.BatchedSpectral()
    vg = L.VerticalGrid()
    bs = cases.solid_body_state(sp.lat, sp.lon, vg.sig)
    want = {("fp64", 1): 0, ("fp64", 4): 0, ("fp64", 5): 4, ("fp64", 64): 4,
            ("fp32", 8): 0, ("fp32", 9): 3, ("fp32", 16): 3, ("fp32", 17): 4}
    for (precision, E), choice in want.items():
        m, _, _ = L.make_ensemble_model(bs, E, L.Config(precision=precision))
        assert m.plan.get_tuning() == [0, 0, 0, 0, choice], (precision, E, m.plan.get_tuning())
        del m


@pytest.mark.parametrize("uneven", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_members_with_own_basic_states_alone_and_together(precision, uneven):
    """Members on their own balanced jets in one plan (8 FP64, 16 FP32: the
    spectral step's lanes take one total wavenumber and that many members'
    matrices per wave; the grid phase is k_grid) end bit for bit equal to
    each member run alone in a one-member plan (one basic state, the three
    launches).  Uneven: 16 members in four lengths, so the launches shrink
    to the live members (16, then 8) and change lane mapping within the
    run."""
    import lbm_isca_amd as L
    from lbm_isca_amd import cases
    sigm = cases.EQUAL_SIGM_L20
    sp = L.BatchedSpectral()
    vg = L.VerticalGrid(sigm)
    E, nsteps = (16 if uneven or precision == "fp32" else 8), 24
    lengths = np.repeat([24, 18, 12, 6], E // 4) if uneven else np.full(E, nsteps)
    bs, _ = cases.jet_states(sp, vg.sig, E, seed=7)
    frc = cases.random_heatings(sp.lat, sp.lon, vg.sig, E, seed=5)
    cfg = L.Config(ndays=nsteps / 72.0, precision=precision)
    m, _, _ = L.make_ensemble_model(bs, E, cfg, sigm=sigm, per_member_basic=True, spectral=sp)
    m.plan.set_forcing(frc)
    m.plan.set_state()
    if uneven:
        m.plan.set_lengths(lengths)
        m.plan.run(0)
    else:
        m.plan.run(nsteps)
    together = [x.cpu().numpy() for x in m.plan.get_state() + m.plan.get_spectral(0)]
    assert np.abs(together[2]).max() > 0
    for e in range(E):
        one = types.SimpleNamespace(**{k: np.ascontiguousarray(getattr(bs, k)[e])
                                       for k in ("u", "v", "t", "q", "ps")})
        m1, _, _ = L.make_ensemble_model(one, 1, cfg, sigm=sigm, spectral=sp)
        m1.plan.set_forcing(frc[e:e + 1])
        m1.plan.set_state()
        m1.plan.run(int(lengths[e]))
        alone = [x.cpu().numpy() for x in m1.plan.get_state() + m1.plan.get_spectral(0)]
        for a, b in zip(together, alone):
            assert np.array_equal(a[e], b[0]), (precision, e, float(np.abs(a[e] - b[0]).max()))
        del m1
