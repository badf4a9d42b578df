"""Basic-state construction on the model's sigma levels.

The preprocessing of ``pylbm_isca/basicstate.py:1-324``: read and average
the pressure-level file (``read_pressure_levels``), surface pressure from
geopotential height (``z2ps``), spectral smoothing (``specsm``, through the
HIP transforms), zonal-mean / stationary-wave filtering (``zmsw``) and
pressure -> sigma interpolation by a natural cubic spline in ln p (``p2s``,
``spline``), chained as in ``load_basic_state``.  All of it is one-off host
setup in FP64 numpy except the smoothing round trip; every function accepts
leading batch axes (members).

``balanced_jets`` is the synthetic input of the uneven ensemble: zonal jets
with stationary waves on the 17 pressure levels, geostrophic winds and a
hydrostatic temperature, so that the temperature the pipeline produces is in
thermal-wind balance with the jet.
"""

import types

import numpy as np

from .constants import ER, GRAV, OMEGA, PLEV17, RDGAS

UNDEF_IN = -9.999e20      # undefined value of the pressure-level file
XMISS = -999.0            # missing-value marker after reading

# variables of the pressure-level file in record order; rhum and omega are
# skipped (basicstate.py:46-49)
BS_VARS = ("z", "rhum", "q", "t", "u", "v", "omega")
BS_KEEP = ("z", "q", "t", "u", "v")


# ---------------------------------------------------------------------------
# Reading
# ---------------------------------------------------------------------------
def read_pressure_levels(path, kmo=1, navg=3, nlon=128, nlat=64, nlev=17, nvar=7, nmonth=None,
                         yflip=True):
    """Mean over ``navg`` months from month ``kmo`` (wrapping past the last
    month) of the big-endian float32 Fortran sequential file of (nlon, nlat)
    records ordered month -> variable -> level (basicstate.py:56-108).
    ``yflip`` turns the file's south-first rows north-first.  A point
    undefined in any month becomes ``XMISS``.  Returns (nlat, nlon, nlev)
    arrays of z, q, t, u, v."""
    npts = nlon * nlat
    words = np.fromfile(path, dtype=">f4")
    reclen = npts + 2
    if words.size % reclen:
        raise ValueError(f"{path}: not a whole number of {npts}-point Fortran records")
    recs = words.reshape(-1, reclen)
    marks = recs[:, [0, -1]].copy().view(">i4")
    if not np.all(marks == npts * 4):
        raise ValueError(f"{path}: record markers are not {npts * 4} bytes "
                         f"(wrong byte order or grid size)")
    data = recs[:, 1:-1].reshape(-1, nlat, nlon)
    per_month = nvar * nlev
    if data.shape[0] % per_month:
        raise ValueError(f"{path}: {data.shape[0]} records is not a whole number of months "
                         f"of {nvar} variables x {nlev} levels")
    nmo = data.shape[0] // per_month if nmonth is None else int(nmonth)
    acc = {v: np.zeros((nlat, nlon, nlev)) for v in BS_KEEP}
    bad = {v: np.zeros((nlat, nlon, nlev), dtype=bool) for v in BS_KEEP}
    month = int(kmo)
    for _ in range(int(navg)):
        first = (month - 1) * per_month
        if first + per_month > data.shape[0]:
            raise ValueError(f"{path}: month {month} is past the end of the file")
        for iv, name in enumerate(BS_VARS[:nvar]):
            if name not in acc:
                continue
            blk = np.transpose(data[first + iv * nlev:first + (iv + 1) * nlev], (1, 2, 0))
            if yflip:
                blk = blk[::-1]
            blk = blk.astype(np.float64)
            # the sentinel is compared in float64, as the reference does
            bad[name] |= blk == UNDEF_IN
            acc[name] += blk / float(navg)
        month = month + 1 if month < nmo else 1
    for name in BS_KEEP:
        acc[name][bad[name]] = XMISS
    return acc


# ---------------------------------------------------------------------------
# Surface pressure
# ---------------------------------------------------------------------------
def z2ps(z, zs, plev=PLEV17):
    """Surface pressure [hPa] where the geopotential height profile z (...,
    nlev; level 0 = the highest pressure) reaches the surface height zs:
    linear in ln p between the first level above the surface and the one
    below it, the index clamped to the second..last level (so a surface
    outside the profile extrapolates from the end pair) (basicstate.py:115-148)."""
    z = np.asarray(z, dtype=np.float64)
    nlev = z.shape[-1]
    zs = np.broadcast_to(np.asarray(zs, dtype=np.float64), z.shape[:-1])
    above = z > zs[..., None]
    k = np.where(above.any(axis=-1), np.argmax(above, axis=-1), nlev - 1)
    k = np.clip(k, 1, nlev - 1)
    lnp = np.log(np.asarray(plev, dtype=np.float64))
    z1 = np.take_along_axis(z, k[..., None], axis=-1)[..., 0]
    z0 = np.take_along_axis(z, (k - 1)[..., None], axis=-1)[..., 0]
    frac = (zs - z0) / (z1 - z0)
    return np.exp(lnp[k - 1] + (lnp[k] - lnp[k - 1]) * frac)


# ---------------------------------------------------------------------------
# Smoothing and filtering
# ---------------------------------------------------------------------------
def specsm(field, sp):
    """Grid -> wave -> grid round trip at the model's truncation
    (basicstate.py:154-163), on the GPU through ``sp`` (a Spectral, or a
    BatchedSpectral for a leading member axis)."""
    from .backend import asnumpy
    return asnumpy(sp.truncate(field))


def zmsw(field, ozm=False, osw=False, lon_axis=1):
    """Zonal mean (``ozm``) or its complement, the stationary waves
    (``osw``), along ``lon_axis`` (1 for (nlat, nlon[, nlev]) fields;
    basicstate.py:166-170)."""
    if not (ozm or osw):
        return field
    f = np.asarray(field)
    zm = f.mean(axis=lon_axis, keepdims=True)
    return np.broadcast_to(zm, f.shape).copy() if ozm else f - zm


# ---------------------------------------------------------------------------
# Vertical interpolation
# ---------------------------------------------------------------------------
def spline(x, z, xi, omisb=False, omist=False, xmiss=XMISS):
    """Natural cubic spline through (x, z[r]) evaluated at xi[r] for every
    row r (Press et al., Numerical Recipes, 3rd ed., sect. 3.3; the
    interpolation of basicstate.py:176-236).

    x (k,) is shared and strictly decreasing (ln p from the bottom up); z
    (n, k); xi (n, l).  Targets beyond x[0] take z[:, 0] (``xmiss`` if
    ``omisb``) and targets beyond x[-1] take z[:, -1] (``xmiss`` if
    ``omist``).
    """
    x = np.asarray(x, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    xi = np.asarray(xi, dtype=np.float64)
    k = x.size
    # second derivatives: tridiagonal system with zero end curvature; the
    # elimination factors depend on x only
    fac = np.zeros(k)
    rhs = np.zeros(z.shape)
    for i in range(1, k - 1):
        w = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1])
        piv = w * fac[i - 1] + 2.0
        fac[i] = (w - 1.0) / piv
        slope = (z[:, i + 1] - z[:, i]) / (x[i + 1] - x[i]) - (z[:, i] - z[:, i - 1]) / (x[i] - x[i - 1])
        rhs[:, i] = (6.0 * slope / (x[i + 1] - x[i - 1]) - w * rhs[:, i - 1]) / piv
    d2 = np.zeros(z.shape)
    for i in range(k - 2, -1, -1):
        d2[:, i] = fac[i] * d2[:, i + 1] + rhs[:, i]
    # interval: the first knot strictly below the target (x decreasing)
    nge = np.searchsorted(-x, -xi.ravel(), side="right").reshape(xi.shape)
    hi = np.clip(nge, 1, k - 1)
    lo = hi - 1
    h = x[hi] - x[lo]
    a = (x[hi] - xi) / h
    b = (xi - x[lo]) / h
    take = lambda arr, idx: np.take_along_axis(arr, idx, axis=1)  # noqa: E731
    val = (a * take(z, lo) + b * take(z, hi)
           + (a * (a * a - 1.0) * take(d2, lo) + b * (b * b - 1.0) * take(d2, hi)) * h * h / 6.0)
    val = np.where(nge == 0, xmiss if omisb else z[:, :1], val)
    val = np.where(nge == k, xmiss if omist else z[:, -1:], val)
    return val


def p2s(gdi, ps, sig, plev=PLEV17, xmiss=XMISS):
    """Pressure -> sigma: the natural spline in ln p of every column of gdi
    (..., nlat, nlon, nlevp) at ln(ps sig) (basicstate.py:239-256)."""
    gdi = np.asarray(gdi, dtype=np.float64)
    lead = gdi.shape[:-1]
    cols = gdi.reshape(-1, gdi.shape[-1])
    tgt = np.log(np.asarray(ps, dtype=np.float64).reshape(-1, 1) * np.asarray(sig)[None, :])
    out = spline(np.log(np.asarray(plev, dtype=np.float64)), cols, tgt, xmiss=xmiss)
    return out.reshape(lead + (len(sig),))


# ---------------------------------------------------------------------------
# Pipeline
# ---------------------------------------------------------------------------
def sigma_basic_state(fields, sp, sig, zs=None, smooth=True, ozm=False, osw=False, plev=PLEV17,
                      source=""):
    """z, u, v, t, q on pressure levels (..., nlat, nlon, nlevp) -> u, v, t, q
    on sigma levels and ps [hPa]: ps from z, smoothed and filtered before the
    interpolation, the other fields after it (basicstate.py:290-324)."""
    z = np.asarray(fields["z"], dtype=np.float64)
    if zs is None:
        zs = np.zeros(z.shape[:-1])
    ps = z2ps(z, zs, plev)
    if smooth:
        ps = specsm(ps, sp)
    ps = zmsw(ps, ozm, osw, lon_axis=ps.ndim - 1)
    out = {}
    for name in ("u", "v", "t", "q"):
        f = p2s(fields[name], ps, sig, plev)
        if smooth:
            f = specsm(f, sp)
        out[name] = zmsw(f, ozm, osw, lon_axis=f.ndim - 2)
    return types.SimpleNamespace(u=out["u"], v=out["v"], t=out["t"], q=out["q"], ps=ps,
                                 lat_deg=sp.lat_deg, lon_deg=sp.lon_deg,
                                 sig=np.asarray(sig), source=source)


def load_basic_state(path, sp, vgrid, kmo=1, navg=3, zs=None, ozm=False, osw=False, smooth=True,
                     yflip=True, plev=PLEV17):
    """Read -> average -> ps -> smooth -> p2s of ``load_basic_state``
    (basicstate.py:290-324)."""
    raw = read_pressure_levels(path, kmo=kmo, navg=navg, nlon=sp.nlon, nlat=sp.nlat,
                               nlev=len(plev), yflip=yflip)
    from .backend import asnumpy
    return sigma_basic_state(raw, sp, asnumpy(vgrid.sig), zs=zs, smooth=smooth, ozm=ozm, osw=osw,
                             plev=plev, source=path)


# ---------------------------------------------------------------------------
# Synthetic balanced jets on pressure levels
# ---------------------------------------------------------------------------
def standard_profile(plev, t0=288.0, lapse=6.5e-3, tfloor=220.0):
    """Temperature and geopotential height of a 6.5 K/km troposphere from t0
    at 1000 hPa over an isothermal tfloor stratosphere, hydrostatic."""
    p = np.asarray(plev, dtype=np.float64)
    kexp = RDGAS * lapse / GRAV
    p_trop = 1000.0 * (tfloor / t0) ** (1.0 / kexp)
    z_trop = t0 / lapse * (1.0 - (p_trop / 1000.0) ** kexp)
    trop = p >= p_trop
    t = np.where(trop, t0 * (p / 1000.0) ** kexp, tfloor)
    z = np.where(trop, t0 / lapse * (1.0 - (p / 1000.0) ** kexp),
                 z_trop + RDGAS * tfloor / GRAV * np.log(p_trop / p))
    return t, z


def jet_profile(plev, p_jet=200.0, taper=1.5):
    """Vertical structure s(p) of the jets and its derivative in
    x = ln(1000/p): s' is constant from the surface (s = 0) to ``p_jet`` and
    falls to zero over ``taper`` units of x by a half cosine, so s levels off
    at 1 above.  The temperature gradient of thermal wind, proportional to
    s', is then height independent in the troposphere and fades out in the
    isothermal stratosphere, where the static stability absorbs the change
    of lapse rate (proportional to s'')."""
    x = np.log(1000.0 / np.asarray(plev, dtype=np.float64))
    xj = np.log(1000.0 / p_jet)
    t = np.clip(x - xj, 0.0, taper)
    norm = xj + 0.5 * taper
    s = (np.minimum(x, xj) + 0.5 * (t + taper / np.pi * np.sin(np.pi * t / taper))) / norm
    ds = np.where(x <= xj, 1.0, 0.5 * (1.0 + np.cos(np.pi * t / taper))) / norm
    return s, ds


def balanced_jets(lat, lon, nmem, seed=1, plev=PLEV17, width_deg=12.0, wave_width_deg=20.0,
                  p_jet=200.0, taper=1.5, eq_taper_deg=20.0):
    """Per-member jets u = U exp(-((lat - lat_j)/12 deg)^2) with U ~ U(20, 50)
    m/s and lat_j ~ U(25N, 45N), plus stationary waves m = 1, 2, 3 of
    amplitude U(0, 100) m and random phase in the geopotential, on the
    pressure levels; ``numpy.random.default_rng(seed)`` draws U, lat_j, the
    amplitudes and the phases (in that order).

    The jets and waves have the vertical structure of ``jet_profile`` (zero
    at 1000 hPa, linear in ln p to ``p_jet``, levelling off at U by about 45
    hPa).  The jet's geopotential solves geostrophic
    balance exactly, z = z_std - (a/g) s(p) int f u dlat, and the
    temperature is hydrostatic, T = (g/R) dz/dln(1000/p), so T is in thermal
    wind balance with u.  The waves' envelope exp(-((lat - lat_j)/20 deg)^2)
    is tapered by mu = 1 - exp(-(lat/20 deg)^2) toward the equator; they carry
    the geostrophic winds of the true Coriolis parameter and their
    hydrostatic temperature.  Since mu ~ lat^2, dz/dlat / f and
    dz/dlon / f stay finite at the equator and the winds are smooth there
    (geostrophic winds of an untapered envelope, with |f| held at a floor,
    jump at the equator: a vortex sheet whose shear instability T127
    resolves).  q = 0.
    Returns dict of (nmem, nlat, nlon, nlev) arrays z, u, v, t, q and the
    drawn parameters.
    """
    rng = np.random.default_rng(seed)
    U = rng.uniform(20.0, 50.0, nmem)
    latj = np.radians(rng.uniform(25.0, 45.0, nmem))
    amp = rng.uniform(0.0, 100.0, (nmem, 3))
    phase = rng.uniform(0.0, 2.0 * np.pi, (nmem, 3))
    lat = np.asarray(lat, dtype=np.float64)
    lon = np.asarray(lon, dtype=np.float64)
    J, I, K = lat.size, lon.size, len(plev)
    s, ds = jet_profile(plev, p_jet, taper)
    t_std, z_std = standard_profile(plev)
    wj = np.radians(width_deg)
    ww = np.radians(wave_width_deg)
    coslat = np.cos(lat)
    # equatorial taper mu and the ratios mu / sin(lat), mu' / sin(lat),
    # written so that lat = 0 needs no division by zero
    l0 = np.radians(eq_taper_deg)
    y = (lat / l0) ** 2
    lat_over_sin = 1.0 / np.sinc(lat / np.pi)
    frac = np.where(y > 1e-12, -np.expm1(-y) / np.maximum(y, 1e-300), 1.0 - 0.5 * y)
    mu = -np.expm1(-y)
    mu_sin = (lat / l0 ** 2) * frac * lat_over_sin
    dmu_sin = (2.0 / l0 ** 2) * np.exp(-y) * lat_over_sin
    gf = GRAV / (2.0 * OMEGA * ER)            # g / (2 Omega a): u = -gf dz/dlat / sin(lat)
    # fine quadrature for F(lat) = int_{-pi/2}^{lat} f u_jet dlat
    fine = np.linspace(-np.pi / 2, np.pi / 2, 4001)
    out = {k: np.zeros((nmem, J, I, K)) for k in ("z", "u", "v", "t", "q")}
    for e in range(nmem):
        uj_f = U[e] * np.exp(-((fine - latj[e]) / wj) ** 2)
        integrand = 2.0 * OMEGA * np.sin(fine) * uj_f
        F_f = np.concatenate([[0.0], np.cumsum(0.5 * (integrand[1:] + integrand[:-1])
                                                * np.diff(fine))])
        F = np.interp(lat, fine, F_f)
        F -= np.sum(F * coslat) / np.sum(coslat)          # no change of the mean height
        ujet = U[e] * np.exp(-((lat - latj[e]) / wj) ** 2)
        zbar = z_std[None, :] - (ER / GRAV) * s[None, :] * F[:, None]          # (J, K)
        tbar = t_std[None, :] - (ER / RDGAS) * ds[None, :] * F[:, None]
        ubar = ujet[:, None] * s[None, :]
        env = np.exp(-((lat - latj[e]) / ww) ** 2)
        denv = -2.0 * (lat - latj[e]) / ww ** 2 * env
        # z_w = sum_m A_m cos(arg_m) env mu; the winds use dz_w/dlat / sin(lat)
        # = sum A cos (env' mu/sin + env mu'/sin) and dz_w/dlon / sin(lat)
        zw = np.zeros((J, I))
        dlat_sin = np.zeros((J, I))
        dlon_sin = np.zeros((J, I))
        for m in range(3):
            arg = (m + 1) * lon[None, :] + phase[e, m]
            zw += amp[e, m] * (env * mu)[:, None] * np.cos(arg)
            dlat_sin += amp[e, m] * (denv * mu_sin + env * dmu_sin)[:, None] * np.cos(arg)
            dlon_sin += -amp[e, m] * (m + 1) * (env * mu_sin)[:, None] * np.sin(arg)
        out["z"][e] = zbar[:, None, :] + zw[:, :, None] * s[None, None, :]
        out["u"][e] = ubar[:, None, :] - gf * dlat_sin[:, :, None] * s[None, None, :]
        out["v"][e] = (gf / coslat)[:, None, None] * dlon_sin[:, :, None] * s[None, None, :]
        out["t"][e] = tbar[:, None, :] + (GRAV / RDGAS) * zw[:, :, None] * ds[None, None, :]
    params = types.SimpleNamespace(U=U, lat_jet=latj, amp=amp, phase=phase)
    return out, params
