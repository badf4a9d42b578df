"""Pins of the oracle's spherical-harmonic transform that do not go through
its own recurrences, at T42 (128 x 64) and T127 (384 x 192):

* Gauss-Legendre quadrature: weights sum to one and integrate mu^k exactly
  for every k < 2J (the closed form (1/2) int mu^k = 1/(k+1), k even);
* the fully normalised P_n^m and H_n^m = (1-mu^2) dP_n^m/dmu against
  closed forms built from numpy's Legendre polynomials (n <= 7: Rodrigues
  polynomial differentiated m times);
* P_n^m for every n, m of the truncation against scipy's spherical
  harmonics (P = sqrt(4 pi) (-1)^m Y_n^m(theta, 0): no Condon-Shortley
  phase, (1/2) int P^2 dmu = 1), and H_n^m against the textbook identity
  (1-mu^2) dP_n^m/dmu = -n mu P_n^m + (n+m) P_{n-1}^m (unnormalised) on
  scipy's values -- not the oracle's eps recurrence;
* discrete orthonormality sum_j gw_j P_n^m P_n'^m = delta_nn';
* the Fourier stage against a brute-force DFT of a few rows (forward
  rfft/nlon truncated to m <= N, inverse with the m = 0 imaginary part
  dropped).

A dropped factor, a sign convention or a transposed index in
``oracle/sht.py`` fails one of these.
"""

import math

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg
from scipy import special

from oracle import SHT, gauss_grid

GRIDS = [(128, 64, 42, False), (384, 192, 127, True)]
IDS = ["T42", "T127"]


@pytest.fixture(scope="module", params=GRIDS, ids=IDS)
def sht(request):
    nlon, nlat, nmax, refine = request.param
    return SHT(nlon, nlat, nmax, refine_gauss=refine)


@pytest.mark.parametrize("nlat,refine", [(64, False), (192, True)], ids=IDS)
def test_gauss_quadrature_exact_on_monomials(nlat, refine):
    mu, gw = gauss_grid(nlat, refine)
    assert np.all(np.diff(mu) < 0), "north first"
    assert abs(gw.sum() - 1.0) < 1e-14
    assert np.all(gw > 0)
    for k in range(2 * nlat):
        want = 1.0 / (k + 1) if k % 2 == 0 else 0.0
        got = float(np.sum(gw * mu ** k))
        assert abs(got - want) < 2e-15 * max(1.0, 1.0 / (k + 1)) * 8, (k, got, want)
    # the nodes are the roots of P_J
    pj = npleg.legval(mu, [0] * nlat + [1])
    assert np.abs(pj).max() < 1e-12


def closed_form(n, m, mu):
    """Normalised P_n^m and H_n^m from the Rodrigues polynomial: P_n^m =
    N (1-mu^2)^(m/2) Q, Q = d^m P_n / dmu^m, N = sqrt((2n+1)(n-m)!/(n+m)!);
    H = N (1-mu^2)^(m/2) ((1-mu^2) Q' - m mu Q)."""
    q = npleg.Legendre.basis(n).deriv(m) if m else npleg.Legendre.basis(n)
    qv = q(mu)
    dq = q.deriv(1)(mu) if n > m else np.zeros_like(mu)
    s = (1.0 - mu * mu) ** (0.5 * m)
    norm = math.sqrt((2 * n + 1) * math.factorial(n - m) / math.factorial(n + m))
    return norm * s * qv, norm * s * ((1.0 - mu * mu) * dq - m * mu * qv)


def test_legendre_small_n_closed_forms(sht):
    for n in range(8):
        for m in range(n + 1):
            p, h = closed_form(n, m, sht.mu)
            assert np.abs(sht.pnm[m, n] - p).max() < 1e-13, (n, m)
            assert np.abs(sht.hnm[m, n] - h).max() < 1e-13, (n, m)


def scipy_tables(nmax, mu):
    """P[m, n, j] (n <= nmax) from scipy's spherical harmonics and H[m, n, j]
    from the identity above."""
    theta = np.arccos(mu)
    n = np.arange(nmax + 1)[None, :, None]
    m = np.arange(nmax + 1)[:, None, None]
    ok = n >= m
    y = special.sph_harm_y(np.where(ok, n, m), m, theta[None, None, :], 0.0)
    p = np.where(ok, np.sqrt(4 * np.pi) * (-1.0) ** m * y.real, 0.0)
    h = np.zeros_like(p)
    mu3 = mu[None, None, :]
    for nn in range(nmax + 1):
        mm = np.arange(nn + 1)[:, None]
        h[:nn + 1, nn] = -nn * mu3[0] * p[:nn + 1, nn]
        if nn >= 1:
            c = np.sqrt((2 * nn + 1) / (2 * nn - 1) * (nn * nn - mm * mm))
            h[:nn + 1, nn] += c * p[:nn + 1, nn - 1]
    return p, h


def test_legendre_all_n_against_scipy(sht):
    p, h = scipy_tables(sht.nmax, sht.mu)
    ep = np.abs(sht.pnm - p).max()
    eh = np.abs(sht.hnm - h).max() / max(1.0, np.abs(h).max())
    print(f"nmax {sht.nmax}: max |P - scipy| {ep:.2e}, max rel |H - identity| {eh:.2e}")
    assert ep < 1e-11
    assert eh < 1e-11


def test_discrete_orthonormality(sht):
    worst = 0.0
    for m in range(sht.nmax + 1):
        pm = sht.pnm[m, m:]
        g = (pm * sht.gw) @ pm.T
        worst = max(worst, np.abs(g - np.eye(g.shape[0])).max())
    print(f"nmax {sht.nmax}: max |G - I| {worst:.2e}")
    assert worst < 1e-12


def test_fourier_stage_against_brute_force_dft(sht):
    rng = np.random.default_rng(5)
    N, M = sht.nlon, sht.nmax + 1
    f = rng.standard_normal((1, 3, N, 2))
    i = np.arange(N)
    w = np.exp(-2j * np.pi * np.outer(np.arange(M), i) / N)   # (M, N)
    want = np.einsum("mi,ejik->emjk", w, f) / N
    got = sht.fourier(f)
    assert got.shape == (1, M, 3, 2)
    assert np.abs(got - want).max() < 1e-14 * np.abs(want).max() * 10
    c = rng.standard_normal((1, M, 3, 2)) + 1j * rng.standard_normal((1, M, 3, 2))
    # f_i = Re c_0 + 2 sum_{m >= 1} Re(c_m e^{2 pi i m i / N})
    waves = np.einsum("mi,emjk->ejik", np.conj(w[1:]), c[:, 1:])       # (E, J, I, K)
    back = np.real(c[:, 0])[:, :, None, :] + 2.0 * np.real(waves)
    got = sht.fourier_inv(c)
    assert np.abs(got - back).max() < 1e-12 * np.abs(back).max()
