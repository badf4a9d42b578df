"""HIP spherical-harmonic transforms against the FP64 oracle and against the
analytic checks of the reference's ``tests/test_spectral.py``."""

import numpy as np
import pytest

import cases
from oracle import SHT
from oracle_helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sp():
    import lbm_isca_amd as L
    return L.Spectral()


@pytest.fixture(scope="module")
def sp32():
    import lbm_isca_amd as L
    return L.Spectral(precision="fp32")


@pytest.fixture(scope="module")
def spb():
    import lbm_isca_amd as L
    return L.BatchedSpectral()


@pytest.fixture(scope="module")
def osht():
    return SHT()


def np_(x):
    from lbm_isca_amd import asnumpy
    return asnumpy(x)


def rand_spec(seed, nlev, nmin=0):
    return cases.random_spectral(np.random.default_rng(seed), 42, nlev, 1.0, nmin=nmin)


def test_grid_matches_oracle(sp, osht):
    np.testing.assert_allclose(sp.mu, osht.mu, rtol=0, atol=2e-16)
    # numpy's leggauss weights carry ~1e-12 relative error near the poles;
    # the library's are computed in extended precision
    np.testing.assert_allclose(sp.gw, osht.gw, rtol=2e-12)
    assert abs(sp.gw.sum() - 1.0) < 1e-13
    assert sp.lat_deg[0] > 87.0 and sp.lat_deg[-1] < -87.0


@pytest.mark.parametrize("nlev", [1, 7])
def test_wave_to_grid_and_back(sp, osht, nlev):
    xi = rand_spec(0, nlev)
    g = np_(sp.wave_to_grid(xi))
    gref = osht.wave_to_grid(xi[None])[0]
    assert rel_err(g, gref) < 1e-13
    xi2 = np_(sp.grid_to_wave(gref))
    assert rel_err(xi2, osht.grid_to_wave(gref[None])[0]) < 1e-12
    assert rel_err(xi2[sp.mask], xi[sp.mask]) < 1e-12


def test_grid_to_wave_rough_field(sp, osht):
    f = np.random.default_rng(3).standard_normal((64, 128, 5))
    assert rel_err(np_(sp.grid_to_wave(f)), osht.grid_to_wave(f[None])[0]) < 1e-12


def test_vector_transforms(sp, osht):
    psi, chi = rand_spec(1, 4, nmin=1) * 1e6, rand_spec(2, 4, nmin=1) * 5e5
    u, v = (np_(x) for x in sp.psichi_to_uv(psi, chi))
    ur, vr = osht.psichi_to_uv(psi[None], chi[None])
    assert rel_err(u, ur[0]) < 1e-13 and rel_err(v, vr[0]) < 1e-13
    vor, div = (np_(x) for x in sp.uv_to_vordiv(ur[0], vr[0]))
    vo, dv = osht.uv_to_vordiv(ur, vr)
    assert rel_err(vor, vo[0]) < 1e-12 and rel_err(div, dv[0]) < 1e-12
    gx, gy = (np_(x) for x in sp.hgrad(psi))
    gxr, gyr = osht.hgrad(psi[None])
    assert rel_err(gx, gxr[0]) < 1e-13 and rel_err(gy, gyr[0]) < 1e-13


def test_batched_members(spb, osht):
    xi = np.stack([rand_spec(10 + e, 3) for e in range(3)])
    g = np_(spb.wave_to_grid(xi))
    assert g.shape == (3, 64, 128, 3)
    assert rel_err(g, osht.wave_to_grid(xi)) < 1e-13
    xs = xi[..., 0]
    assert np_(spb.wave_to_grid(xs)).shape == (3, 64, 128)


# -- analytic checks of the reference (tests/test_spectral.py:41-154) -------
def test_roundtrip_bandlimited(sp):
    rng = np.random.default_rng(0)
    xi = np.zeros((43, 43), dtype=complex)
    xi[sp.mask] = rng.normal(size=sp.mask.sum()) + 1j * rng.normal(size=sp.mask.sum())
    xi[0] = xi[0].real
    xi2 = np_(sp.grid_to_wave(sp.wave_to_grid(xi)))
    assert np.abs(xi2 - xi)[sp.mask].max() / np.abs(xi)[sp.mask].max() < 1e-11


def test_laplacian_eigenvalue(sp):
    for (m, n) in [(0, 1), (1, 1), (3, 7), (10, 20)]:
        xi = np.zeros((43, 43), dtype=complex)
        xi[m, n] = 1.0 + 0.5j
        g = np_(sp.wave_to_grid(xi))
        lap = np_(sp.wave_to_grid(sp.laplacian(np.asarray(xi))))
        expect = -n * (n + 1) / sp.a ** 2 * g
        assert np.abs(lap - expect).max() / np.abs(expect).max() < 1e-10


def test_gradient_against_analytic(sp):
    f = np.repeat(sp.mu[:, None], sp.nlon, axis=1)
    dfdx, dfdy = (np_(x) for x in sp.hgrad(sp.grid_to_wave(f)))
    assert np.abs(dfdx).max() < 1e-16
    expect = np.sqrt(1 - sp.mu ** 2)[:, None] / sp.a
    assert np.abs(dfdy - expect).max() / expect.max() < 1e-10


def test_gradient_zonal(sp):
    lam = sp.lon[None, :]
    cl = np.sqrt(1 - sp.mu ** 2)[:, None]
    f = np.cos(lam) * cl
    dfdx, dfdy = (np_(x) for x in sp.hgrad(sp.grid_to_wave(f)))
    ex = -np.sin(lam) / sp.a * np.ones_like(cl)
    ey = -np.cos(lam) * sp.mu[:, None] / sp.a
    assert np.abs(dfdx - ex).max() / np.abs(ex).max() < 1e-10
    assert np.abs(dfdy - ey).max() / np.abs(ey).max() < 1e-10


def test_solid_body_rotation(sp):
    u = 20.0 * np.sqrt(1 - sp.mu ** 2)[:, None] * np.ones((1, sp.nlon))
    vor, div = sp.uv_to_vordiv(u, np.zeros_like(u))
    zg, dg = np_(sp.wave_to_grid(vor)), np_(sp.wave_to_grid(div))
    expect = 2.0 * 20.0 * sp.mu[:, None] / sp.a
    assert np.abs(dg).max() < 1e-18
    assert np.abs(zg - expect).max() / np.abs(expect).max() < 1e-10


def test_divergence_of_gradient_is_laplacian(sp):
    rng = np.random.default_rng(2)
    xi = np.zeros((43, 43), dtype=complex)
    sel = sp.mask.copy()
    sel[:, 41:] = False
    xi[sel] = rng.normal(size=sel.sum()) + 1j * rng.normal(size=sel.sum())
    xi[0] = xi[0].real
    gx, gy = sp.hgrad(xi)
    divg = -np_(sp.neg_divergence(gx, gy))
    lap = np_(sp.laplacian(np.asarray(xi)))
    assert np.abs(divg - lap)[sel].max() / np.abs(lap)[sel].max() < 1e-9


def test_vordiv_roundtrip(sp):
    rng = np.random.default_rng(1)
    xi = np.zeros((43, 43), dtype=complex)
    sel = sp.mask.copy()
    sel[:, 0] = False
    xi[sel] = rng.normal(size=sel.sum()) + 1j * rng.normal(size=sel.sum())
    xi[0] = xi[0].real
    psi, chi = xi * 1e6, xi.conj() * 5e5
    chi[0] = chi[0].real
    u, v = sp.psichi_to_uv(psi, chi)
    vor, div = sp.uv_to_vordiv(u, v)
    psi2 = np_(sp.inverse_laplacian(vor))
    chi2 = np_(sp.inverse_laplacian(div))
    assert np.abs(psi2 - psi)[sel].max() / np.abs(psi)[sel].max() < 1e-10
    assert np.abs(chi2 - chi)[sel].max() / np.abs(chi)[sel].max() < 1e-10


def test_fp32_transforms(sp32, osht):
    xi = rand_spec(4, 6)
    g = np_(sp32.wave_to_grid(xi))
    assert rel_err(g, osht.wave_to_grid(xi[None])[0]) < 2e-6
    f = np.random.default_rng(5).standard_normal((64, 128, 6))
    assert rel_err(np_(sp32.grid_to_wave(f)), osht.grid_to_wave(f[None])[0]) < 2e-6
