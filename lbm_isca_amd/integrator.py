"""Damping, semi-implicit matrices, vertical diffusion and the leapfrog step
with the reference's interface (``pylbm_isca/integrator.py:33-315``).

The coefficients and matrices are FP64 host setup, as in the reference
(which assembles them on the host and inverts them once per dt).  The step
itself, ``Integrator.step(dt, prev, tend)``, runs the library's spectral
step kernel on the caller's spectral fields (``lbm_step``): T estimate,
geopotential, M(n)^-1 solve, restored gravity-wave coupling, damped leapfrog
update and implicit vertical diffusion in one launch.  ``SemiImplicit.solve``
and ``VerticalDiffusion.apply`` are host (numpy) applications of the same
FP64 inverses, for inspection; the time loop never calls them.

An Integrator needs the library plan that holds the basic state's
reference temperature: the model's own plan (``LBM.integ``), or, for one
built by hand, a plan made from the ``BasicStateTerms`` that set the
vertical grid's reference temperature.
"""

import numpy as np

from .backend import asnumpy
from .constants import S2D

_UNIT = {"SEC": 1.0, "MIN": 60.0, "HOU": 3600.0, "DAY": 86400.0, "MON": 30 * 86400.0}


def to_seconds(value, unit):
    """A time in the namelist unit (SEC, MIN, HOUR, DAY, MON) in seconds."""
    return float(value) * _UNIT[unit.strip().upper()[:3]]


def drag_rates(profile, kmax, unit):
    """Per-level damping timescales -> rates in s^-1; entries <= 0 (or no
    profile) mean no damping (``integrator.py:101-112``)."""
    out = np.zeros(kmax)
    if profile is None:
        return out
    d = np.asarray(profile, dtype=np.float64)
    if d.size != kmax:
        raise ValueError(f"damping profile must have {kmax} entries")
    sec = np.array([to_seconds(x, unit) for x in d])
    out[sec > 0] = 1.0 / sec[sec > 0]
    return out


def vdif_rate(v, kmax):
    """1/(86400 v) per level; None or any v <= 0 switches the variable off
    (``integrator.py:243-253``)."""
    if v is None:
        return None
    v = np.atleast_1d(np.asarray(v, dtype=np.float64))
    if v.size == 1:
        v = np.full(kmax, v[0])
    if np.any(v <= 0):
        return None
    return 1.0 / (S2D * v)


class Damping:
    """Implicit hyperdiffusion, top sponge and Rayleigh/Newtonian drag as
    (nmax+1, nlev) rates in s^-1 (``integrator.py:33-98``): the hyperdiffusion
    [n(n+1)/(N(N+1))]^p / tau (zero at n = 0, and for vorticity/divergence at
    n = 1), a tanh sponge of timescale ray0 above the top scale height and the
    per-level drag profiles, all applied as 1 / (1 + 2 dt K) factors."""

    def __init__(self, sp, vgrid, order=4, tefold=2.0, tunit_hdif="HOUR", ddragv=None,
                 ddragd=None, ddragt=None, ddragq=None, tunit_damp="DAY", ray0=30.0,
                 zsclh=8.0e3, rzsclh=7.0e3, sponge=True):
        self.sp, self.vg = sp, vgrid
        nmax, kmax = int(sp.nmax), int(vgrid.kmax)
        sig = np.asarray(asnumpy(vgrid.sig), dtype=np.float64)
        n = np.arange(nmax + 1, dtype=np.float64)
        p = order // 2
        tau = to_seconds(tefold, tunit_hdif)
        scale = (n * (n + 1.0) / nmax) ** p
        top = (nmax * (nmax + 1.0) / nmax) ** p
        kt = scale / top / tau
        kv = (scale - (2.0 / nmax) ** p) / top / tau
        kt[0] = kv[0] = 0.0
        kv[1] = max(kv[1], 0.0)
        self.kt_h, self.kv_h = kt, kv
        self.sponge_k = np.zeros(kmax)
        if sponge:
            height = -zsclh * np.log(sig)
            base = -zsclh * np.log(float(sig[-1]))
            self.sponge_k = (np.maximum(1.0 + np.tanh((height - base) / rzsclh), 1e-20)
                             / to_seconds(ray0, tunit_hdif))
        self.drag_v = drag_rates(ddragv, kmax, tunit_damp)
        self.drag_d = drag_rates(ddragd, kmax, tunit_damp)
        self.drag_t = drag_rates(ddragt, kmax, tunit_damp)
        self.drag_q = drag_rates(ddragq, kmax, tunit_damp)
        self.kv = kv[:, None] + self.sponge_k + self.drag_v
        self.kd = kv[:, None] + self.sponge_k + self.drag_d
        self.ktt = kt[:, None] + self.drag_t
        self.kq = kt[:, None] + self.drag_q

    @classmethod
    def from_config(cls, sp, vgrid, cfg):
        """The damping of a :class:`~lbm_isca_amd.model.Config` (``model.py:109-115``)."""
        return cls(sp, vgrid, order=cfg.hdif_order, tefold=cfg.hdif_tefold,
                   tunit_hdif=cfg.hdif_tunit, ddragv=cfg.ddragv, ddragd=cfg.ddragd,
                   ddragt=cfg.ddragt, ddragq=cfg.ddragq, tunit_damp=cfg.damp_tunit,
                   ray0=cfg.ray0, sponge=cfg.sponge)

    def arrays(self):
        return self.kv, self.kd, self.ktt, self.kq


class VerticalDiffusion:
    """Implicit sigma-space vertical diffusion (``integrator.py:178-240``):
    the zero-flux tridiagonal sigma Laplacian ``lop`` and, per variable and
    dt, (I - dt nu lop)^-1."""

    NAMES = ("vor", "div", "t", "q")

    def __init__(self, vgrid, vdifv=None, vdifd=None, vdift=None, vdifq=None):
        self.vg = vgrid
        kmax = int(vgrid.kmax)
        ds = np.asarray(asnumpy(vgrid.dsig), dtype=np.float64)
        dm = np.asarray(asnumpy(vgrid.dsigm), dtype=np.float64)
        below = np.zeros(kmax)
        above = np.zeros(kmax)
        below[1:] = 1.0 / (ds[1:] * dm[1:kmax])
        above[:-1] = 1.0 / (ds[:-1] * dm[1:kmax])
        above[0] = (dm[0] + dm[1]) / (ds[0] * dm[0] * dm[1])
        below[0] = 0.0
        below[-1] = (dm[kmax - 1] + dm[kmax]) / (ds[-1] * dm[kmax - 1] * dm[kmax])
        above[-1] = 0.0
        self.lop = (np.diag(-(below + above)) + np.diag(below[1:], -1) + np.diag(above[:-1], 1))
        self.coef = {k: vdif_rate(v, kmax) for k, v in zip(self.NAMES, (vdifv, vdifd, vdift, vdifq))}
        self._dt = None
        self.minv = {}

    def rates(self):
        """Per-variable nu (or None) in the order vor, div, t, q."""
        return tuple(self.coef[k] for k in self.NAMES)

    def prepare(self, dt):
        if self._dt == dt:
            return
        eye = np.eye(int(self.vg.kmax))
        self.minv = {k: (None if nu is None else np.linalg.inv(eye - dt * nu[:, None] * self.lop))
                     for k, nu in self.coef.items()}
        self._dt = dt

    def apply(self, field, name, skip_low_n=False):
        """Host application of (I - dt nu lop)^-1 to a spectral field (..., m,
        n, k); ``skip_low_n`` keeps n = 0, 1 (vorticity and divergence)."""
        minv = self.minv.get(name)
        f = np.asarray(asnumpy(field))
        if minv is None:
            return f
        out = np.einsum("ij,...j->...i", minv, f)
        if skip_low_n:
            out[..., 0, :] = f[..., 0, :]
            out[..., 1, :] = f[..., 1, :]
        return out


class _PlanUser:
    """Finds (or builds) the library plan whose basic state sets the
    reference temperature of the vertical grid."""

    def _plan_for(self, nmem):
        plans = self.__dict__.setdefault("_plans", {})
        if self._plan is not None and self._plan.nmem == nmem:
            return self._plan
        if nmem not in plans:
            bst = getattr(self.vg, "basic_terms", None)
            if bst is None:
                raise RuntimeError("the vertical grid has no reference temperature: build "
                                   "BasicStateTerms(sp, vgrid, basic_state) first")
            plans[nmem] = bst.make_plan(nmem, damping=self.dmp.arrays(),
                                        vdif=self._vdif_rates(), delt=self._delt)
        return plans[nmem]

    def _vdif_rates(self):
        return (None, None, None, None)


class SemiImplicit(_PlanUser):
    """The per-wavenumber gravity-wave matrices M(n) = diag(fd ft) - dt^2
    lap_n (whmtx + gcmtx ft) and their FP64 inverses
    (``integrator.py:120-172``), built by the library for the basic state's
    reference temperature."""

    def __init__(self, sp, vgrid, damping, plan=None):
        self.sp, self.vg, self.dmp = sp, vgrid, damping
        self._plan = plan
        self._delt = 1200.0
        self._dt = None
        self.m = self.minv = self.cond = None

    def prepare(self, dt):
        """M(n), M(n)^-1 and the 2-norm condition numbers for dt, (nbasic,)
        nmax+1 matrices of nlev x nlev (the member axis only with per-member
        basic states)."""
        if self._dt == dt:
            return
        plan = self._plan_for(self._plan.nmem if self._plan is not None else 1)
        plan.prepare_dt(dt)
        nb, M, L = plan.nbasic, plan.nmax + 1, plan.nlev
        nset = plan.inistp + 2          # start-up dts, delt, the free dt
        minv = plan.get("minv", nset * nb * M * L * L).reshape(nset, nb, M, L, L)[-1]
        m = plan.get("mmat", nb * M * L * L).reshape(nb, M, L, L)
        if nb == 1:
            m, minv = m[0], minv[0]
        self.m, self.minv = m, minv
        self.cond = np.linalg.cond(m.reshape(-1, L, L)).reshape(m.shape[:-2])
        self._dt = dt

    def solve(self, rhs):
        """M(n)^-1 applied to a spectral field (..., m, n, k) on the host."""
        r = np.asarray(asnumpy(rhs))
        return np.einsum("nij,...nj->...ni", self.minv, r)


class Integrator(_PlanUser):
    """Semi-implicit leapfrog step (``integrator.py:255-315``)."""

    def __init__(self, sp, vgrid, damping, semi, vdiff, plan=None):
        self.sp, self.vg, self.dmp, self.semi, self.vdiff = sp, vgrid, damping, semi, vdiff
        self._plan = plan
        self._delt = 1200.0

    def _vdif_rates(self):
        return self.vdiff.rates()

    def step(self, dt, prev, tend):
        """One step: ``prev`` a dict of the spectral fields at t-dt (vor, div,
        t, pi, q), ``tend`` the explicit tendencies (wt_vor, wt_div, wt_t,
        wt_pi, wt_q); returns the dict of the t+dt level.  Single-member
        fields (m, n[, k]) or ensemble fields with a leading member axis."""
        import torch
        single = torch.as_tensor(prev["pi"]).dim() == 2
        lead = (lambda x: torch.as_tensor(x)[None]) if single else (lambda x: torch.as_tensor(x))
        wt_vor, wt_div, wt_t, wt_pi, wt_q = tend
        p5 = [lead(prev[k]) for k in ("vor", "div", "t", "q", "pi")]
        t5 = [lead(x) for x in (wt_vor, wt_div, wt_t, wt_q, wt_pi)]
        plan = self._plan_for(int(p5[4].shape[0]))
        out = plan.step(dt, p5, t5)
        if single:
            out = [x[0] for x in out]
        return dict(zip(("vor", "div", "t", "q", "pi"), out))
