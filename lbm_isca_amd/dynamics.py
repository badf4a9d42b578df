"""Basic-state terms and the explicit linear tendencies with the reference's
interface (``pylbm_isca/dynamics.py:51-271``).

The basic-state terms (f + vorticity, divergence, grad ln ps, Tv, the
column sums, sigma-dot and x5 of the basic flow) are formed on the device
by the library when a plan receives the basic state
(``lbm_set_basic_state``), which also sets the reference temperature
(the plain grid mean of the basic Tv, ``vertical.py:206-216``).
``Dynamics.tendencies(st)`` evaluates the tendencies of a grid
:class:`~lbm_isca_amd.state.State` with the same kernels as the model's
time step (``lbm_tendencies``).
"""

import numpy as np

from .backend import asnumpy
from .plan import Plan, zero_damping


def _host(x):
    return np.ascontiguousarray(asnumpy(x), dtype=np.float64)


class BasicStateTerms:
    """The time-constant basic state a :class:`Dynamics` linearises about
    (``dynamics.py:51-93``): ``bs`` has u, v, t, q (nlat, nlon, nlev) and ps
    (nlat, nlon) in hPa, or the same with a leading member axis (one basic
    state per member).  Sets the vertical grid's reference temperature."""

    def __init__(self, sp, vgrid, bs, plan=None):
        self.sp, self.vg = sp, vgrid
        self.u, self.v, self.t, self.q = (_host(getattr(bs, k)) for k in ("u", "v", "t", "q"))
        self.ps = _host(bs.ps)
        self.per_member = self.u.ndim == 4
        self.nbasic = self.u.shape[0] if self.per_member else 1
        self._plans = {}
        # a model's own plan already holds this basic state
        plan = plan if plan is not None else self.make_plan(self.nbasic if self.per_member else 1)
        self.tbar = plan.get("tbar", self.nbasic * vgrid.kmax).reshape(self.nbasic, vgrid.kmax)
        vgrid.set_reference_temperature(self.tbar[0])
        vgrid.basic_terms = self
        self.coriol = 2.0 * 7.2921150e-5 * np.asarray(sp.mu)[:, None]

    def fields(self):
        """u, v, t, q (nbasic, J, I, L) and ps (nbasic, J, I)."""
        f = [self.u, self.v, self.t, self.q, self.ps]
        return f if self.per_member else [x[None] for x in f]

    def make_plan(self, nmem, strictly_linear=False, damping=None, vdif=(None,) * 4, delt=1200.0,
                  precision="fp64"):
        """A library plan of ``nmem`` members holding this basic state; plans
        without damping and vertical diffusion are kept and shared."""
        key = (int(nmem), bool(strictly_linear), precision)
        plain = damping is None and all(v is None for v in vdif)
        if plain and key in self._plans:
            return self._plans[key]
        sp, vg = self.sp, self.vg
        if self.per_member and nmem != self.nbasic:
            raise ValueError(f"{self.nbasic} basic states for {nmem} members")
        plan = Plan(sp.nlon, sp.nlat, sp.nmax, sp.a, vg.sigm, nmem=nmem,
                    nbasic=self.nbasic if self.per_member else 1, precision=precision,
                    damping=damping if damping is not None else zero_damping(sp.nmax, vg.kmax),
                    vdif=vdif, delt=delt, strictly_linear=strictly_linear)
        plan.set_basic_state(*self.fields())
        if plain:
            self._plans[key] = plan
        return plan


class Dynamics:
    """Explicit (non-gravity-wave) tendencies of the linear model
    (``dynamics.py:146-271``)."""

    def __init__(self, sp, vgrid, bst, strictly_linear=False, plan=None):
        self.sp, self.vg, self.bs = sp, vgrid, bst
        self.strictly_linear = bool(strictly_linear)
        self._plan = plan
        self.diag = None

    def _plan_for(self, nmem):
        if self._plan is not None:
            if self._plan.nmem != nmem:
                raise ValueError(f"the model's plan has {self._plan.nmem} members, the state {nmem}")
            return self._plan
        return self.bs.make_plan(nmem, strictly_linear=self.strictly_linear)

    def tendencies(self, st, frc_t=None, frc_vor=None, frc_div=None, frc_ps=None, frc_q=None):
        """Spectral tendencies (wt_vor, wt_div, wt_t, wt_pi, wt_q) of the grid
        state ``st`` (u, v, t, q, ps'), with an optional temperature forcing
        (K/s), vorticity / divergence forcings (s^-2), ln ps forcing (s^-1)
        and humidity forcing on the grid (each None: none, as in the
        reference).  Single-member states give (m, n, k) / (m, n) fields,
        ensemble states a leading member axis.  On a model's own plan
        (``LBM.dyn``) this replaces the forcings the last run uploaded; the
        next ``LBM.run`` uploads its own."""
        import torch
        single = torch.as_tensor(st.ps).dim() == 2
        lead = (lambda x: None if x is None else torch.as_tensor(x)[None]) if single else \
            (lambda x: None if x is None else torch.as_tensor(x))
        u, v, t, q, ps = (lead(getattr(st, k)) for k in ("u", "v", "t", "q", "ps"))
        plan = self._plan_for(int(ps.shape[0]))
        plan.set_forcing(lead(frc_t))
        plan.set_steady_forcing(vor=lead(frc_vor), div=lead(frc_div), ps=lead(frc_ps))
        tv, td, tt, tq, tp = plan.tendencies(u, v, t, q, ps)
        if frc_q is not None:
            # humidity forcing (dynamics.py:260-261): its analysis added to
            # the humidity tendency (not used by the LBM time loop)
            tq = tq + self.sp.grid_to_wave(lead(frc_q)[0] if single else lead(frc_q)) \
                .reshape(tq.shape)
        out = (tv, td, tt, tp, tq)
        if single:
            out = tuple(x[0] for x in out)
        vor_s, div_s = self.sp.uv_to_vordiv(st.u, st.v)
        self.diag = {"vor_s": vor_s, "div_s": div_s, "vor": self.sp.wave_to_grid(vor_s),
                     "div": self.sp.wave_to_grid(div_s), "pi": st.lnps_pert()
                     if hasattr(st, "lnps_pert") else None}
        return out
