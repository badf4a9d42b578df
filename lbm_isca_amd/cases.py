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
  geopotential and temperature in thermal-wind balance with the jet.
"""

import types

import numpy as np

from .constants import GRAV, OMEGA, RDGAS, ER

EQUAL_SIGM_L20 = np.linspace(1.0, 0.0, 21)


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


def jet_states(lat, lon, sig, nmem, seed=1, tsurf=288.0):
    """Per-member jets with a balanced temperature (E, J, I, L) + ps (E, J, I).

    The zonal-mean temperature follows from thermal wind in sigma
    coordinates, f du/dln(sigma) = (R/a) dT/dlat, with u growing linearly in
    -ln(sigma) to the jet level; stationary waves m = 1..3 of random phase
    and amplitude U(0, 100) m are added to the geopotential and their
    geostrophic winds and hydrostatic temperature to u, v, T.
    """
    rng = np.random.default_rng(seed)
    J, I, L = lat.size, lon.size, sig.size
    U = rng.uniform(20.0, 50.0, nmem)
    latj = np.radians(rng.uniform(25.0, 45.0, nmem))
    amps = rng.uniform(0.0, 100.0, (nmem, 3))
    phases = rng.uniform(0.0, 2 * np.pi, (nmem, 3))
    out = {k: np.zeros((nmem, J, I, L)) for k in ("u", "v", "t", "q")}
    out["ps"] = np.full((nmem, J, I), 1000.0)
    f = 2.0 * OMEGA * np.sin(lat)
    zeta = np.log(sig)                 # ln sigma, negative aloft
    shape_v = np.clip(-zeta / 1.6, 0.0, 1.0)           # 0 at the surface, 1 at ~200 hPa
    dshape = np.where(-zeta < 1.6, 1.0 / 1.6, 0.0)     # d shape / d(-ln sigma)
    tbase = np.maximum(tsurf * sig ** (RDGAS * 6.5e-3 / GRAV), 215.0)
    for e in range(nmem):
        ujet = U[e] * np.exp(-((lat - latj[e]) / np.radians(12.0)) ** 2)
        u = ujet[:, None] * shape_v[None, :]
        # thermal wind: dT/dlat = -(a f / R) du/dln(sigma)... du/dln(sigma) = -ujet*dshape
        dtdlat = (ER * f[:, None] / RDGAS) * ujet[:, None] * dshape[None, :]
        order = np.argsort(lat)
        tz = np.zeros((J, L))
        latg = lat[order]
        cum = np.concatenate([[0.0], np.cumsum(0.5 * (dtdlat[order][1:] + dtdlat[order][:-1])
                                               * np.diff(latg)[:, None], axis=0)[:, :]]) \
            if J > 1 else np.zeros((1, L))
        tz[order] = cum
        tz -= np.sum(tz * np.cos(lat)[:, None], axis=0) / np.sum(np.cos(lat))
        t = tbase[None, :] + tz
        # stationary waves in the geopotential, geostrophic wind, hydrostatic T
        z = np.zeros((J, I, L))
        for m in range(3):
            env = np.exp(-((lat - latj[e]) / np.radians(20.0)) ** 2)
            z += amps[e, m] * env[:, None, None] * np.cos((m + 1) * lon[None, :, None]
                                                         + phases[e, m]) * shape_v[None, None, :]
        fcl = np.where(np.abs(f) < 2e-5, np.sign(f + 1e-30) * 2e-5, f)
        dzdx = np.gradient(z, lon, axis=1) / (ER * np.cos(lat)[:, None, None])
        dzdy = np.gradient(z, lat, axis=0) / ER
        out["u"][e] = u[:, None, :] - GRAV / fcl[:, None, None] * dzdy
        out["v"][e] = GRAV / fcl[:, None, None] * dzdx
        dz = np.gradient(z, zeta, axis=2)
        out["t"][e] = t[:, None, :] - GRAV / RDGAS * dz
    return types.SimpleNamespace(**out), U, latj


def member_lengths(nmem, seed=1, steps_per_day=72, lo=5.0, hi=60.0):
    rng = np.random.default_rng(seed + 1000)
    days = rng.uniform(lo, hi, nmem)
    return np.round(days * steps_per_day).astype(np.int64)
