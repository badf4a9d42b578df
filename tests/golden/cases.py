"""Deterministic inputs shared by the golden-vector generator and the tests.

Everything here is a closed form or a seeded draw on the T42 grid, so the
same inputs are rebuilt bit-for-bit (up to the transform round-off of a
band-limiting synthesis) wherever they are needed.
"""

import types

import numpy as np

NLAT, NLON, NMAX = 64, 128, 42
SAMPLE_LEVELS = (0, 7, 13, 19)


def jet_basic_state(lat, lon, sig):
    """A zonally asymmetric moist baroclinic basic state.

    Exercises every basic-state term of the tendencies: a sheared jet with a
    wavenumber-2 meander (u, v), a meridional and vertical temperature
    gradient, moisture, and a surface pressure with a zonal wave (so
    grad ln ps != 0).  ``lat``/``lon`` in radians (north first), ``sig`` the
    full-level sigma.  Returns a namespace with u, v, t, q (J, I, L) and ps
    (J, I) in hPa.
    """
    la = lat[:, None, None]
    lo = lon[None, :, None]
    s = sig[None, None, :]
    jet = 30.0 * np.exp(-((la - np.deg2rad(40.0)) / np.deg2rad(15.0)) ** 2)
    u = jet * np.sqrt(s) * (1.0 + 0.2 * np.cos(2.0 * lo))
    v = 3.0 * np.cos(la) ** 2 * np.sin(2.0 * lo) * np.sqrt(s)
    t = 250.0 - 40.0 * np.sin(la) ** 2 + 30.0 * s + 2.0 * np.cos(la) * np.cos(lo)
    q = 1e-3 * np.exp(-(la / np.deg2rad(30.0)) ** 2) * s ** 2 * (1.0 + 0.1 * np.sin(lo))
    ps = 1000.0 + 8.0 * np.cos(lat)[:, None] ** 2 * np.cos(2.0 * lon)[None, :]
    shape = (lat.size, lon.size, sig.size)
    b = lambda x: np.ascontiguousarray(np.broadcast_to(x, shape))  # noqa: E731
    return types.SimpleNamespace(u=b(u), v=b(v), t=b(t), q=b(q), ps=ps)


def random_spectral(rng, nmax, nlev, scale, nmin=0):
    """Seeded complex coefficients (M, M, nlev) with n>=m, real m=0, n>=nmin."""
    m1 = nmax + 1
    xi = np.zeros((m1, m1, nlev), dtype=np.complex128)
    mm, nn = np.meshgrid(np.arange(m1), np.arange(m1), indexing="ij")
    sel = (nn >= mm) & (nn >= nmin)
    cnt = int(sel.sum())
    xi[sel] = (rng.standard_normal((cnt, nlev)) + 1j * rng.standard_normal((cnt, nlev)))
    xi[0] = xi[0].real
    # red spectrum so the grid field is smooth
    nfac = 1.0 / (1.0 + nn) ** 1.5
    return scale * xi * nfac[:, :, None]


def perturbation_spectra(nlev, seed=0):
    """Spectral (psi, chi, t, q, pi) of a seeded perturbation direction."""
    rng = np.random.default_rng(seed)
    psi = random_spectral(rng, NMAX, nlev, 2.0e6, nmin=1)
    chi = random_spectral(rng, NMAX, nlev, 5.0e5, nmin=1)
    t = random_spectral(rng, NMAX, nlev, 1.0)
    q = random_spectral(rng, NMAX, nlev, 1.0e-4)
    pi = random_spectral(rng, NMAX, 1, 1.0e-3)[..., 0]
    return psi, chi, t, q, pi


def heating(lat, lon, sig, lat0=0.0, lon0=180.0, amp_k_day=2.0,
            wlat=10.0, wlon=30.0):
    """Gaussian heating with sin(pi sigma) profile, K/s, (J, I, L)."""
    latd = np.degrees(lat)[:, None]
    lond = np.degrees(lon)[None, :]
    dl = ((lond - lon0) + 180.0) % 360.0 - 180.0
    hxy = np.exp(-((latd - lat0) / wlat) ** 2 - (dl / wlon) ** 2)
    return (amp_k_day / 86400.0) * hxy[:, :, None] * np.sin(np.pi * sig)[None, None, :]


def sample_index(shape, count, seed):
    """Seeded flat indices into an array of ``shape``."""
    rng = np.random.default_rng(seed)
    return rng.choice(int(np.prod(shape)), size=count, replace=False)
