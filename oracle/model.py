"""Oracle: grid state, Robert-Asselin filter and the time loop (numpy FP64).
TEST INFRASTRUCTURE -- see ``oracle/__init__.py``.

Restates ``pylbm_isca/state.py:26-148`` and ``pylbm_isca/model.py:30-290``
(with the member axis of ``ensemble.py`` always present: E >= 1).
"""

from dataclasses import dataclass, field

import numpy as np

from .dyn import BasicTerms, tendencies
from .integ import Damping, Stepper
from .sht import SHT
from .vgrid import VerticalOps

DDRAG_DEFAULT = [0.5, 0.5, 0.5, 5, 15, 20, 20, 20, 20, 20,
                 20, 20, 20, 20, 20, 20, 20, 20, 0.5, 0.5]


@dataclass
class OracleConfig:
    """Namelist defaults of ``pylbm_isca/model.py:34-74``."""
    delt: float = 1200.0
    inistp: int = 2
    ndays: float = 66.0
    hdif_order: int = 4
    hdif_tefold: float = 2.0
    hdif_tunit: str = "HOUR"
    sponge: bool = True
    ray0: float = 30.0
    ddragv: list = field(default_factory=lambda: list(DDRAG_DEFAULT))
    ddragd: list = field(default_factory=lambda: list(DDRAG_DEFAULT))
    ddragt: list = field(default_factory=lambda: list(DDRAG_DEFAULT))
    ddragq: list = None
    damp_tunit: str = "DAY"
    vdifv: float = 1.0e3
    vdifd: float = 1.0e3
    vdift: float = 1.0e3
    vdifq: float = None
    robert: float = 0.04
    strictly_linear: bool = False
    # linear mode: strictly linear tendencies and the linearised change of
    # variable ps' = ps_bar * pi' (the tangent-linear model about a zero
    # perturbation); not a reference option, used to pin the adjoint's
    # forward model
    linear: bool = False

    @property
    def steps_per_day(self):
        return int(round(86400.0 / self.delt))


class GridState:
    """(u, v, t, q) (E, J, I, L) and ps' (E, J, I) perturbations (``state.py:26-71``).

    ``linear=True`` replaces the exact pi' <-> ps' pair by its linearisation.
    """

    linear = False

    def __init__(self, ps_bar, nmem, nlev):
        self.ps_bar = np.asarray(ps_bar, dtype=np.float64)
        self.lnps_bar = np.log(self.ps_bar)
        j, i = self.ps_bar.shape
        self.u = np.zeros((nmem, j, i, nlev))
        self.v = np.zeros_like(self.u)
        self.t = np.zeros_like(self.u)
        self.q = np.zeros_like(self.u)
        self.ps = np.zeros((nmem, j, i))

    FIELDS = ("u", "v", "t", "q", "ps")

    def lnps_pert(self):
        """pi' = ln(ps_bar + ps') - ln(ps_bar) (``state.py:77-79``)."""
        if self.linear:
            return self.ps / self.ps_bar
        return np.log(self.ps_bar + self.ps) - self.lnps_bar

    def set_from_lnps(self, pi):
        """ps' = ps_bar (exp(pi') - 1) (``state.py:81-83``)."""
        if self.linear:
            self.ps = self.ps_bar * pi
            return
        self.ps = self.ps_bar * (np.exp(pi) - 1.0)

    def to_spectral(self, sht):
        """``state.py:86-92``."""
        vor, div = sht.uv_to_vordiv(self.u, self.v)
        return {"vor": vor, "div": div, "t": sht.grid_to_wave(self.t),
                "pi": sht.grid_to_wave(self.lnps_pert()[..., None])[..., 0],
                "q": sht.grid_to_wave(self.q)}

    def load_spectral(self, sht, spec):
        """``state.py:94-99``."""
        self.u, self.v = sht.vordiv_to_uv(spec["vor"], spec["div"])
        self.t = sht.wave_to_grid(spec["t"])
        self.q = sht.wave_to_grid(spec["q"])
        self.set_from_lnps(sht.wave_to_grid(spec["pi"][..., None])[..., 0])

    def assign(self, other):
        for f in self.FIELDS:
            setattr(self, f, getattr(other, f).copy())

    def robert_asselin(self, cur, new, gamma):
        """X_f = X + gamma (X_prev - 2 X + X_new) (``state.py:110-128``)."""
        for f in self.FIELDS:
            p, c, n = getattr(self, f), getattr(cur, f), getattr(new, f)
            setattr(self, f, c + gamma * (p - 2.0 * c + n))

    def isfinite(self):
        return all(np.isfinite(getattr(self, f)).all() for f in self.FIELDS)


class OracleModel:
    """The LBM time loop (``model.py:92-290``), E members sharing one basic state."""

    def __init__(self, bs, nmem=1, cfg=None, sigm=None, nlon=128, nlat=64, nmax=42,
                 refine_gauss=False):
        self.cfg = cfg or OracleConfig()
        c = self.cfg
        self.sht = SHT(nlon, nlat, nmax, refine_gauss=refine_gauss)
        self.vg = VerticalOps() if sigm is None else VerticalOps(sigm)
        self.bst = BasicTerms(self.sht, self.vg, bs)
        self.dmp = Damping(nmax, self.vg, order=c.hdif_order, tefold=c.hdif_tefold,
                           tunit_hdif=c.hdif_tunit, ddragv=c.ddragv, ddragd=c.ddragd,
                           ddragt=c.ddragt, ddragq=c.ddragq, tunit_damp=c.damp_tunit,
                           ray0=c.ray0, sponge=c.sponge)
        self.stepper = Stepper(self.sht, self.vg, self.dmp, c.vdifv, c.vdifd, c.vdift, c.vdifq)
        self.nmem = nmem
        mk = lambda: GridState(bs.ps, nmem, self.vg.kmax)  # noqa: E731
        self.ga, self.gb, self.gc = mk(), mk(), mk()
        for g in (self.ga, self.gb, self.gc):
            g.linear = c.linear
        self.time = 0.0
        self.last_diag = None

    def set_initial_perturbation(self, **fields):
        """``model.py:142-148``: sets both ga and gb."""
        for name, val in fields.items():
            if val is None:
                continue
            arr = np.asarray(val, dtype=np.float64)
            tgt = getattr(self.ga, name)
            arr = np.broadcast_to(arr, tgt.shape).copy()
            setattr(self.ga, name, arr.copy())
            setattr(self.gb, name, arr.copy())

    def step(self, dt, oinit, frc_t=None, steady=None):
        """One step; ``steady`` = dict of frc_vor, frc_div (E|1, J, I, L) and
        frc_ps (E|1, J, I) grid forcings (``model.py:202-226``)."""
        tend, diag = tendencies(self.sht, self.vg, self.bst, self.ga, frc_t=frc_t,
                                strictly_linear=self.cfg.strictly_linear or self.cfg.linear,
                                **(steady or {}))
        self.last_diag = diag
        prev = self.gb.to_spectral(self.sht)
        new = self.stepper.step(dt, prev, tend)
        self.gc.load_spectral(self.sht, new)
        if oinit:
            self.gb.assign(self.ga)
        else:
            self.gb.robert_asselin(self.ga, self.gc, self.cfg.robert)
        self.ga.assign(self.gc)
        return new

    def run(self, nsteps=None, frc_t=None, check_finite=True, mjo=None, history=None,
            steady=None):
        """Start-up half steps (dt/4, dt/2, clock not advanced) then ``nsteps``
        full steps (``model.py:211-253``).  ``frc_t`` is a constant (E|1, J, I, L)
        temperature forcing in K/s; ``mjo`` an object whose ``step()`` gives
        each call's forcing (replacing ``frc_t``); ``history`` an
        ``oracle.history.History`` fed the top-of-step samples
        (``model.py:233-277``).
        """
        from .history import diagnostics
        c = self.cfg
        if nsteps is None:
            nsteps = int(round(c.ndays * c.steps_per_day))
        istep = 0
        advanced = 0
        while advanced < nsteps:
            istep += 1
            if istep > c.inistp:
                oinit, dt = False, c.delt
            else:
                oinit, dt = True, c.delt / (2.0 ** (c.inistp - istep + 1))
            frc = frc_t if mjo is None else np.asarray(mjo.step())
            ga_top = None
            if history is not None:
                ga_top = GridState(self.ga.ps_bar, self.nmem, self.vg.kmax)
                ga_top.assign(self.ga)
            self.step(dt, oinit, frc, steady)
            top = None
            if history is not None:
                cur = self.ga
                self.ga = ga_top
                top = diagnostics(self)
                self.ga = cur
            if check_finite and not self.ga.isfinite():
                raise FloatingPointError(
                    f"model blew up at step {istep} (t={self.time / 86400:.2f} d)")
            if oinit:
                if history is not None:
                    history.start_step(top)
                continue
            self.time += dt
            advanced += 1
            if history is not None:
                history.full_step(advanced, top)
        return self
