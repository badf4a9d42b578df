"""Oracle: damping, semi-implicit gravity-wave solve, implicit vertical
diffusion and the leapfrog update (numpy FP64).
TEST INFRASTRUCTURE -- see ``oracle/__init__.py``.

Restates ``pylbm_isca/integrator.py:33-315``.  Spectral fields are
``(E, M, M, L)`` complex (surface ``(E, M, M)``).
"""

import numpy as np

_UNIT = {"SEC": 1.0, "MIN": 60.0, "HOU": 3600.0, "DAY": 86400.0, "MON": 30 * 86400.0}


def to_seconds(value, unit):
    """``integrator.py:115-118``."""
    return float(value) * _UNIT[unit.strip().upper()[:3]]


def drag_rate(profile, kmax, unit):
    """Per-level timescales -> s^-1, non-positive disables (``integrator.py:101-112``)."""
    out = np.zeros(kmax)
    if profile is None:
        return out
    d = np.asarray(profile, dtype=np.float64)
    if d.size != kmax:
        raise ValueError(f"damping profile must have {kmax} entries")
    sec = np.array([to_seconds(x, unit) for x in d])
    out[sec > 0] = 1.0 / sec[sec > 0]
    return out


class Damping:
    """Implicit hyperdiffusion + sponge + drag coefficients, (N+1, L) in s^-1.

    Reference: ``integrator.py:41-98``.
    """

    def __init__(self, nmax, vg, order=4, tefold=2.0, tunit_hdif="HOUR",
                 ddragv=None, ddragd=None, ddragt=None, ddragq=None,
                 tunit_damp="DAY", ray0=30.0, zsclh=8.0e3, rzsclh=7.0e3, sponge=True):
        kmax = vg.kmax
        n = np.arange(nmax + 1, dtype=np.float64)
        nn1 = n * (n + 1.0)
        p = order // 2
        tau = to_seconds(tefold, tunit_hdif)
        difmax = (nmax * (nmax + 1.0) / nmax) ** p
        dift = (nn1 / nmax) ** p
        difv = dift - (2.0 / nmax) ** p
        kt = dift / difmax / tau
        kv = difv / difmax / tau
        kt[0] = 0.0
        kv[0] = 0.0
        kv[1] = max(kv[1], 0.0)
        self.sponge_k = np.zeros(kmax)
        if sponge:
            ray0_s = to_seconds(ray0, tunit_hdif)
            z = -zsclh * np.log(vg.sig)
            zofs = -zsclh * np.log(float(vg.sig[-1]))
            self.sponge_k = np.maximum(1.0 + np.tanh((z - zofs) / rzsclh), 1e-20) / ray0_s
        dv = drag_rate(ddragv, kmax, tunit_damp)
        dd = drag_rate(ddragd, kmax, tunit_damp)
        dt_ = drag_rate(ddragt, kmax, tunit_damp)
        dq = drag_rate(ddragq, kmax, tunit_damp)
        self.kv = kv[:, None] + self.sponge_k[None, :] + dv[None, :]
        self.kd = kv[:, None] + self.sponge_k[None, :] + dd[None, :]
        self.ktt = kt[:, None] + dt_[None, :]
        self.kq = kt[:, None] + dq[None, :]


def vdif_rate(v, kmax):
    """1/(86400 v) per level; None or any v<=0 disables (``integrator.py:243-253``)."""
    if v is None:
        return None
    v = np.atleast_1d(np.asarray(v, dtype=np.float64))
    if v.size == 1:
        v = np.full(kmax, v[0])
    if np.any(v <= 0):
        return None
    return 1.0 / (86400.0 * v)


def vdiff_operator(vg):
    """Tridiagonal sigma Laplacian with zero-flux ends (``integrator.py:182-206``)."""
    k = vg.kmax
    dsig, dsigm = vg.dsig, vg.dsigm
    lower = np.zeros(k)
    upper = np.zeros(k)
    lower[1:] = 1.0 / (dsig[1:] * dsigm[1:k])
    upper[:-1] = 1.0 / (dsig[:-1] * dsigm[1:k])
    diag = -(lower + upper)
    upper[0] = (dsigm[0] + dsigm[1]) / (dsig[0] * dsigm[0] * dsigm[1])
    lower[0] = 0.0
    diag[0] = -upper[0]
    lower[-1] = (dsigm[k - 1] + dsigm[k]) / (dsig[-1] * dsigm[k - 1] * dsigm[k])
    upper[-1] = 0.0
    diag[-1] = -lower[-1]
    lop = np.diag(diag) + np.diag(lower[1:], -1) + np.diag(upper[:-1], 1)
    return lop


class Stepper:
    """Semi-implicit leapfrog step (``integrator.py:121-315``)."""

    def __init__(self, sht, vg, damping, vdifv=None, vdifd=None, vdift=None, vdifq=None):
        self.sht, self.vg, self.dmp = sht, vg, damping
        self.lop = vdiff_operator(vg)
        self.vrate = {name: vdif_rate(v, vg.kmax) for name, v in
                      (("vor", vdifv), ("div", vdifd), ("t", vdift), ("q", vdifq))}
        self._cache = {}

    def matrices(self, dt):
        """M(n) = diag(fd ft) - dt^2 lap_n (whmtx + gcmtx ft), its inverse and
        the vertical-diffusion inverses (I - dt nu lop)^-1 for this dt.
        Reference: ``integrator.py:130-163, 216-228``.
        """
        if dt in self._cache:
            return self._cache[dt]
        vg, dmp = self.vg, self.dmp
        nmax, k = self.sht.nmax, vg.kmax
        dt2 = 2.0 * dt
        m = np.empty((nmax + 1, k, k))
        for n in range(nmax + 1):
            ft = 1.0 + dt2 * dmp.ktt[n]
            fd = 1.0 + dt2 * dmp.kd[n]
            mat = -dt * dt * self.sht.laplace[n] * (vg.whmtx + vg.gcmtx * ft[:, None])
            mat[np.arange(k), np.arange(k)] += fd * ft
            m[n] = mat
        minv = np.linalg.inv(m)
        vinv = {}
        for name, nu in self.vrate.items():
            vinv[name] = None if nu is None else np.linalg.inv(np.eye(k) - dt * nu[:, None] * self.lop)
        out = {"m": m, "minv": minv, "vinv": vinv,
               "cond": np.array([np.linalg.cond(x) for x in m])}
        self._cache[dt] = out
        return out

    def _vdiff(self, x, name, mats, skip_low_n=False):
        vi = mats["vinv"][name]
        if vi is None:
            return x
        out = np.einsum("ij,...j->...i", vi, x)
        if skip_low_n:
            out[..., 0, :] = x[..., 0, :]
            out[..., 1, :] = x[..., 1, :]
        return out

    def step(self, dt, prev, tend):
        """One leapfrog step; ``prev``/returned dicts have vor, div, t, pi, q.
        Reference: ``integrator.py:264-315``.
        """
        vg, dmp = self.vg, self.dmp
        mats = self.matrices(dt)
        dt2 = 2.0 * dt
        wt_vor, wt_div, wt_t, wt_pi, wt_q = tend
        ft = 1.0 + dt2 * dmp.ktt
        fd = 1.0 + dt2 * dmp.kd
        wdx = prev["t"] * (1.0 + dt * dmp.ktt) + wt_t * dt
        wdphix = vg.hydrostatic(wdx)
        wdpix = prev["pi"] + wt_pi * dt
        wdphiy = wdphix + ft * (wdpix[..., None] * vg.gg)
        lap = self.sht.laplace[None, :, None]
        rhs = prev["div"] * (1.0 + dt * dmp.kd) * ft + dt * (wt_div * ft - lap * wdphiy)
        dbar = np.einsum("nij,...nj->...ni", mats["minv"], rhs)
        wt_t = wt_t - np.einsum("kl,...l->...k", vg.tmtx, dbar)
        wt_pi = wt_pi - (dbar * vg.dsig).sum(axis=-1)
        new = {
            "vor": (prev["vor"] + dt2 * wt_vor) / (1.0 + dt2 * dmp.kv),
            "div": 2.0 * dbar - prev["div"],
            "t": (prev["t"] + dt2 * wt_t) / ft,
            "q": (prev["q"] + dt2 * wt_q) / (1.0 + dt2 * dmp.kq),
            "pi": prev["pi"] + dt2 * wt_pi,
        }
        new["vor"] = self._vdiff(new["vor"], "vor", mats, skip_low_n=True)
        new["div"] = self._vdiff(new["div"], "div", mats, skip_low_n=True)
        new["t"] = self._vdiff(new["t"], "t", mats)
        new["q"] = self._vdiff(new["q"], "q", mats)
        return new
