"""Synthetic workloads of the T42 L20 benchmark configurations.

* ``solid_body_state``: u = 20 m/s cos(lat), v = q = 0, ps = 1000 hPa,
  T(sigma) = max(288 sigma^(R Gamma / g), 220) K with Gamma = 6.5 K/km
  (the standard-atmosphere lapse rate written in sigma), on 20 equally
  spaced pure-sigma levels.
* ``heating``: Gaussian heating, e-folding widths 10 deg latitude x 30 deg
  longitude, vertical profile sin(pi sigma), amplitude in K/day.
* ``random_heatings``: per-member centres lat U(-20, 20), lon U(0, 360) and
  amplitude U(0.5, 3) K/day from ``numpy.random.default_rng(seed)`` (three
  length-E draws in that order).
* ``jet_states``: per-member zonal jets u = U exp(-((lat - lat_j)/12 deg)^2),
  U ~ U(20, 50) m/s, lat_j ~ U(25N, 45N), with stationary waves 1-3 in the
  geopotential, built on pressure levels in thermal-wind balance and taken
  to sigma by the basic-state pipeline (``basicstate``).
"""

import types

import numpy as np

from .constants import GRAV, RDGAS

EQUAL_SIGM_L20 = np.linspace(1.0, 0.0, 21)
EQUAL_SIGM_L40 = np.linspace(1.0, 0.0, 41)

# Per-level drag (days) of 40 equally spaced levels: each entry of the
# reference's 20-level profile (model.py:30-31) covers two levels, so the
# strong surface drag, the free-atmosphere 20 days and the weak top drag sit
# at the same sigma ranges.
DDRAG_L40 = list(np.repeat([0.5, 0.5, 0.5, 5, 15, 20, 20, 20, 20, 20,
                            20, 20, 20, 20, 20, 20, 20, 20, 0.5, 0.5], 2).astype(float))


def t127l40_config(**kw):
    """Config keywords of the T127 L40 workload: the 40-level drag for
    vorticity, divergence and temperature, and dt = 450 s.  The leapfrog
    advection needs u N dt / (a cos lat) < 1: at N = 127 the 50 m/s jets
    give about 1.7 at 1200 s (the T42 step; runs blow up by day 3) and 0.85
    at 600 s (blow-up by day 10); 450 s gives 0.64."""
    out = dict(ddragv=list(DDRAG_L40), ddragd=list(DDRAG_L40), ddragt=list(DDRAG_L40),
               delt=450.0)
    out.update(kw)
    return out


def vorticity_perturbation(sp, nlev, nmem, seed=4, rms=1.0e-6, slope=-1.5):
    """Seeded spectral vorticity (E, M, M, L) with amplitude spectrum
    proportional to n^slope (energy ~ n^(2 slope)), random phases, real
    m = 0, zero at n = 0, scaled to the given grid rms (s^-1) per member."""
    rng = np.random.default_rng(seed)
    M = sp.nmax + 1
    mm, nn = np.meshgrid(np.arange(M), np.arange(M), indexing="ij")
    sel = (nn >= mm) & (nn >= 1)
    amp = np.where(sel, np.maximum(nn, 1).astype(float) ** slope, 0.0)
    out = np.zeros((nmem, M, M, nlev), dtype=np.complex128)
    for e in range(nmem):
        ph = rng.uniform(0.0, 2.0 * np.pi, (M, M, nlev))
        x = amp[:, :, None] * np.exp(1j * ph)
        x[0] = x[0].real
        out[e] = np.where(sel[:, :, None], x, 0.0)
    g = np.asarray(_host(sp.wave_to_grid(out)))
    r = np.sqrt((g ** 2).mean(axis=(1, 2, 3)))
    return out * (rms / r)[:, None, None, None]


def _host(x):
    from .backend import asnumpy
    return asnumpy(x)


def solid_body_state(lat, lon, sig, u0=20.0, ps=1000.0):
    shape = (lat.size, lon.size, sig.size)
    gamma = 6.5e-3
    tprof = np.maximum(288.0 * sig ** (RDGAS * gamma / GRAV), 220.0)
    u = u0 * np.cos(lat)[:, None, None] * np.ones(shape)
    return types.SimpleNamespace(
        u=u, v=np.zeros(shape), t=np.broadcast_to(tprof, shape).copy(), q=np.zeros(shape),
        ps=np.full(shape[:2], ps))


def heating(lat, lon, sig, lat0=0.0, lon0=180.0, amp_k_day=2.0, wlat=10.0, wlon=30.0):
    latd = np.degrees(lat)[:, None]
    lond = np.degrees(lon)[None, :]
    dl = ((lond - lon0) + 180.0) % 360.0 - 180.0
    hxy = np.exp(-((latd - lat0) / wlat) ** 2 - (dl / wlon) ** 2)
    return (amp_k_day / 86400.0) * hxy[:, :, None] * np.sin(np.pi * sig)[None, None, :]


def random_heatings(lat, lon, sig, nmem, seed=0):
    rng = np.random.default_rng(seed)
    lat0 = rng.uniform(-20.0, 20.0, nmem)
    lon0 = rng.uniform(0.0, 360.0, nmem)
    amp = rng.uniform(0.5, 3.0, nmem)
    return np.stack([heating(lat, lon, sig, a, b, c) for a, b, c in zip(lat0, lon0, amp)])


def jet_states(sp, sig, nmem, seed=1):
    """Per-member basic states of the uneven ensemble: ``balanced_jets`` on
    the 17 pressure levels taken to sigma by the basic-state pipeline
    (surface pressure from z, spectral smoothing through ``sp`` -- a
    BatchedSpectral -- and the natural-spline p -> sigma interpolation).
    Returns the state (E, J, I, L) / ps (E, J, I) and the drawn parameters."""
    from . import basicstate
    fields, params = basicstate.balanced_jets(sp.lat, sp.lon, nmem, seed=seed)
    bs = basicstate.sigma_basic_state(fields, sp, np.asarray(sig), source="balanced jets")
    return bs, params


def member_lengths(nmem, seed=1, steps_per_day=72, lo=5.0, hi=60.0):
    rng = np.random.default_rng(seed + 1000)
    days = rng.uniform(lo, hi, nmem)
    return np.round(days * steps_per_day).astype(np.int64)
