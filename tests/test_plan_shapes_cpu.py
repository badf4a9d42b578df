"""The Plan wrappers reject arrays whose extents are not the plan's with
ValueError before any device call (the C ABI takes J, I, L and M from the
plan, so a wrong-shaped buffer would otherwise be read out of bounds).
Runs without a GPU: the plan object is built without its device half, and
every call must raise before it would reach the device."""

import numpy as np
import pytest

from lbm_isca_amd.plan import Plan


@pytest.fixture
def plan():
    p = Plan.__new__(Plan)           # no lbm_plan_create: nothing may reach the device
    p.nlon, p.nlat, p.nmax, p.nlev, p.nmem, p.nbasic = 128, 64, 42, 20, 2, 1
    p._h = None
    return p


J, I, L, M = 64, 128, 20, 43


def grid(*shape):
    return np.zeros(shape)


def spec(*shape):
    return np.zeros(shape, dtype=np.complex128)


CASES = [
    ("grid_to_wave", lambda p: p.grid_to_wave(grid(1, 32, I, 1))),
    ("grid_to_wave lon", lambda p: p.grid_to_wave(grid(1, J, 64, 1))),
    ("wave_to_grid", lambda p: p.wave_to_grid(spec(1, 22, 22, 1))),
    ("uv_to_vordiv", lambda p: p.uv_to_vordiv(grid(1, J, I, 3), grid(1, J, 96, 3))),
    ("psichi_to_uv", lambda p: p.psichi_to_uv(spec(1, M, M, 2), spec(1, 40, 40, 2))),
    ("hgrad", lambda p: p.hgrad(spec(J, I, L))),
    ("grid_to_wave_adj", lambda p: p.grid_to_wave_adj(spec(1, 22, 22, 1))),
    ("wave_to_grid_adj", lambda p: p.wave_to_grid_adj(grid(1, 32, I, 1))),
    ("uv_to_vordiv_adj", lambda p: p.uv_to_vordiv_adj(spec(1, M, M, 1), spec(1, 22, 22, 1))),
    ("psichi_to_uv_adj", lambda p: p.psichi_to_uv_adj(grid(1, J, I, 1), grid(1, 32, I, 1))),
    ("hgrad_adj", lambda p: p.hgrad_adj(grid(1, 32, I, 1), grid(1, 32, I, 1))),
    # an L20 basic state / forcing on a plan of other levels or members
    ("basic state levels", lambda p: p.set_basic_state(grid(1, J, I, 40), grid(1, J, I, 40),
                                                       grid(1, J, I, 40), grid(1, J, I, 40),
                                                       grid(1, J, I))),
    ("basic state members", lambda p: p.set_basic_state(grid(3, J, I, L), grid(3, J, I, L),
                                                        grid(3, J, I, L), grid(3, J, I, L),
                                                        grid(3, J, I))),
    ("basic state ps", lambda p: p.set_basic_state(grid(1, J, I, L), grid(1, J, I, L),
                                                   grid(1, J, I, L), grid(1, J, I, L),
                                                   grid(1, 32, I))),
    ("forcing levels", lambda p: p.set_forcing(grid(J, I, 40))),
    ("forcing members", lambda p: p.set_forcing(grid(3, J, I, L))),
    ("steady forcing", lambda p: p.set_steady_forcing(vor=grid(J, I, 40))),
    ("steady ln ps", lambda p: p.set_steady_forcing(ps=grid(32, I))),
    ("state", lambda p: p.set_state(u=grid(2, J, I, 40))),
    ("state keep", lambda p: p.set_state_keep(t=grid(3, J, I, L))),
    ("spectral state", lambda p: p.set_spectral_state(vor=spec(2, 22, 22, L))),
    ("tendencies", lambda p: p.tendencies(u=grid(2, 32, I, L))),
    ("step", lambda p: p.step(1200.0, [spec(2, M, M, L)] * 4 + [spec(2, 22, 22)],
                              [spec(2, M, M, L)] * 5)),
    ("mjo forcing", lambda p: p.set_mjo_forcing(grid(5, J, I, 40), 72, 10)),
    ("spectral scaling", lambda p: p.spec_scale(spec(1, 22, 22, 1), 0)),
]


@pytest.mark.parametrize("name,call", CASES, ids=[c[0] for c in CASES])
def test_wrong_shape_raises_value_error(plan, name, call):
    with pytest.raises(ValueError):
        call(plan)


@pytest.mark.parametrize("nlon,nlat,msg", [(100, 64, "nlon=100"), (448, 224, "nlon=448"),
                                           (2048, 64, "nlon=2048"), (128, 60, "nlat=60")])
def test_unsupported_grids_refused_before_any_device_call(nlon, nlat, msg):
    """Grids whose FFT length has a radix other than 2, 3, 5 (or exceeds
    1024 longitudes), or whose latitudes are not a multiple of 8, are
    refused by lbm_plan_create before it touches the device."""
    from lbm_isca_amd.plan import Plan
    with pytest.raises(ValueError, match=msg):
        Plan(nlon=nlon, nlat=nlat, nmax=21)
