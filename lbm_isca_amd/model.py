"""``Config`` and ``LBM`` with the reference's interface (``pylbm_isca/model.py:34-290``).

``LBM.run`` executes the whole time loop on the GPU: the start-up half
steps (dt/4, dt/2, clock not advanced), the fused step kernels and the
Robert-Asselin filter, optionally replayed as a hipGraph, with an on-device
finiteness flag instead of a host sync per step.  A time-varying MJO
forcing is interpolated on the device inside the step graph, and the
history (trapezoidal interval means of the top-of-step samples) is
accumulated on the device and handed to the writer once per interval.
"""

import time as _time
from dataclasses import dataclass, field

import numpy as np

from . import tuning as _tuning
from .backend import asnumpy, hip
from .constants import S2D
from .dynamics import BasicStateTerms, Dynamics
from .forcing import MJOForcing
from .integrator import Damping, Integrator, SemiImplicit, VerticalDiffusion
from .plan import Plan
from .state import State

DDRAG_DEFAULT = [0.5, 0.5, 0.5, 5, 15, 20, 20, 20, 20, 20,
                 20, 20, 20, 20, 20, 20, 20, 20, 0.5, 0.5]


@dataclass
class Config:
    """Namelist knobs and defaults of ``pylbm_isca/model.py:34-74``."""
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
    out_interval_days: float = 1.0
    time_average: bool = True
    strictly_linear: bool = False
    # tangent-linear model about a zero perturbation (strictly linear, with
    # ps' = ps_bar pi'): exactly linear, the model the adjoint transposes
    tangent_linear: bool = False
    nmem: int = 1
    device: str = "hip"
    precision: str = "fp64"      # "fp64", "fp32" or "bf16x3" arithmetic on the GPU
    # launch variants of the step: False built-in, True timed at set-up,
    # "cache" reuse a stored pick (lbm_isca_amd.tuning)
    autotune: object = False

    @property
    def steps_per_day(self):
        return int(round(86400.0 / self.delt))


def damping_coefficients(cfg, nmax, sig):
    """(kv, kd, ktt, kq), each (nmax+1, nlev) in s^-1, of a Config
    (``integrator.py:41-98``)."""
    import types
    sp = types.SimpleNamespace(nmax=nmax)
    vg = types.SimpleNamespace(sig=np.asarray(sig), kmax=np.asarray(sig).size)
    return Damping.from_config(sp, vg, cfg).arrays()


def _field(obj, name):
    return getattr(obj, name)


def _lead(x):
    """Add the member axis to a single-member field (numpy or torch)."""
    if x is None:
        return None
    if hasattr(x, "unsqueeze"):
        return x.unsqueeze(0)
    return np.asarray(x)[None]


class LBM:
    """Linear baroclinic model on the MI355X.

    ``basic_state`` is any object with ``u, v, t, q`` (nlat, nlon, nlev) and
    ``ps`` (nlat, nlon) in hPa (``dynamics.py:59-63``), or -- as an extension
    of the reference, for ensembles whose members each have their own basic
    state -- the same fields with a leading member axis when
    ``per_member_basic=True``.
    """

    def __init__(self, basic_state, spectral, vertical, config=None, basic_spectral=None,
                 per_member_basic=False):
        self.cfg = config or Config()
        c = self.cfg
        self.sp, self.vg, self.bs = spectral, vertical, basic_state
        self.xp = hip
        self.batched = bool(getattr(spectral, "batched", False)) or c.nmem > 1
        nmem = max(int(c.nmem), 1)
        self.nmem = nmem
        self.damping = Damping.from_config(spectral, vertical, c)
        self.vdiff = VerticalDiffusion(vertical, c.vdifv, c.vdifd, c.vdift, c.vdifq)
        nbasic = nmem if per_member_basic else 1
        self.plan = Plan(spectral.nlon, spectral.nlat, spectral.nmax, spectral.a,
                         vertical.sigm, nmem=nmem, nbasic=nbasic, precision=c.precision,
                         damping=self.damping.arrays(), vdif=self.vdiff.rates(), delt=c.delt,
                         robert=c.robert, inistp=c.inistp, strictly_linear=c.strictly_linear,
                         tangent_linear=c.tangent_linear)
        fields = [np.asarray(_field(basic_state, k), dtype=np.float64) for k in ("u", "v", "t", "q", "ps")]
        if not per_member_basic:
            fields = [f[None] for f in fields]
        self.plan.set_basic_state(*fields)
        self.ps_bar = fields[4] if per_member_basic else fields[4][0]
        kmax = vertical.kmax
        tb = self.plan.get("tbar", nbasic * kmax).reshape(nbasic, kmax)
        vertical.set_reference_temperature(tb[0])
        self.tbar = tb
        # the reference's component objects (model.py:104-124) on this plan
        self.bst = BasicStateTerms(spectral, vertical, basic_state, plan=self.plan)
        self.dyn = Dynamics(spectral, vertical, self.bst, strictly_linear=c.strictly_linear,
                            plan=self.plan)
        self.semi = SemiImplicit(spectral, vertical, self.damping, plan=self.plan)
        self.integ = Integrator(spectral, vertical, self.damping, self.semi, self.vdiff,
                                plan=self.plan)
        self.plan.set_state()
        self.tuning = _tuning.autotune(self.plan, c.autotune)
        # keep the last top-of-step sample for diagnostics() after a run
        self.plan.set_history(0, keep_last=True)
        self._ga = None
        self.time = 0.0
        self.istep = 0

    @property
    def device(self):
        return "hip"

    # -- state -----------------------------------------------------------
    def set_initial_perturbation(self, u=None, v=None, t=None, q=None, ps=None):
        """Sets the given fields on both time levels (model.py:142-148); the
        fields left out keep each level's current values."""
        given = dict(u=u, v=v, t=t, q=q, ps=ps)
        if not self.batched:
            given = {k: _lead(x) for k, x in given.items()}
        self.plan.set_state_keep(**given)
        self._ga = None

    def _state(self, fields):
        u, v, t, q, ps = fields
        if not self.batched:
            u, v, t, q, ps = (x[0] for x in (u, v, t, q, ps))
        return State(self.ps_bar, u, v, t, q, ps)

    @property
    def ga(self):
        """Current time level t as a grid :class:`State` (device tensors)."""
        if self._ga is None:
            self._ga = self._state(self.plan.get_state())
        return self._ga

    @property
    def gb(self):
        """The filtered previous level t - dt as a grid :class:`State`."""
        return self._state(self.plan.get_prev_state())

    @property
    def gc(self):
        """The t + dt scratch level of the last step, which the reference's
        loop leaves equal to ga (model.py:237-243)."""
        return self.ga.copy()

    def spectral_state(self, level=0):
        """(vor, div, t, q, pi) spectral fields of the current (0) or
        previous filtered (1) level; member axis always present."""
        return self.plan.get_spectral(level)

    # -- time loop ---------------------------------------------------------
    def diagnostics(self):
        """The eight history variables on the grid (model.py:150-166): psi,
        chi and w from the last step's tendencies (its top-of-step state), u,
        v, t, z and p from the current state.  Device tensors, with the
        member axis for an ensemble."""
        d = self.plan.get_diagnostics()
        ga = self.ga
        if not self.batched:
            d = {k: v[0] for k, v in d.items()}
        return {"psi": d["psi"], "chi": d["chi"], "u": ga.u, "v": ga.v, "w": d["w"],
                "t": ga.t, "z": d["z"], "p": ga.ps}

    def _set_forcing(self, mjo_forcing, steady_forcing):
        """Upload the run's forcings (model.py:196-229): the temperature
        forcing (MJO, else the steady one) and the steady vorticity,
        divergence and ln ps forcings.  Returns ``(interpolated, stepped)``:
        the MJOForcing whose call counter the run advances on the device, or
        the general forcing object whose ``.step()`` the run calls before
        every step (start-up steps included), or None."""
        sv = {}
        if steady_forcing is not None:
            for name in ("vor", "div", "ps"):
                val = getattr(steady_forcing, name, None)
                if val is not None and np.any(np.asarray(asnumpy(val)) != 0):
                    sv[name] = val
        self.plan.set_steady_forcing(**sv)
        if mjo_forcing is not None:
            if isinstance(mjo_forcing, MJOForcing):
                self.plan.set_mjo_forcing(mjo_forcing.gtf, mjo_forcing.nt, mjo_forcing.spinup,
                                          call0=mjo_forcing.counter)
                return mjo_forcing, None
            if getattr(mjo_forcing, "constant", False):
                # declared constant: one upload serves the whole run
                self.plan.set_forcing(mjo_forcing.step())
                return None, None
            if not callable(getattr(mjo_forcing, "step", None)):
                raise TypeError("a temperature forcing needs a .step() method (model.py:224)")
            return None, mjo_forcing
        frc = None if steady_forcing is None else steady_forcing.t
        self.plan.set_forcing(frc)
        return None, None

    def _step_forcing(self, obj):
        frc = obj.step()
        self.plan.set_forcing(frc)

    def _write(self, writer, day):
        fields = {k: asnumpy(v) for k, v in self.plan.get_history(day).items()}
        if not self.batched:
            fields = {k: v[0] for k, v in fields.items()}
        writer.add(day, fields)

    def _run_stepped(self, obj, nsteps, lengths, writer, K, use_graph, check_every, verbose,
                     progress_every, t0):
        """The run with a forcing whose ``.step()`` changes every call: the
        field of each call is uploaded before its step (model.py:222-229),
        one executed step at a time."""
        c = self.cfg
        total = c.inistp + (nsteps if lengths is None else int(max(lengths)))
        for k in range(total):
            self._step_forcing(obj)
            self.plan.run_steps(nsteps, 1, use_graph=use_graph, check_every=check_every)
            done = k + 1 - c.inistp           # full steps done
            if writer is not None and done > 0 and done % K == 0:
                self._write(writer, done // K)
            if verbose and progress_every and done and done % progress_every == 0:
                print(f"  day {done / c.steps_per_day:6.2f}   [{_time.time() - t0:5.1f}s]",
                      flush=True)

    def run(self, mjo_forcing=None, steady_forcing=None, writer=None, verbose=True,
            progress_every=240, use_graph=True, lengths=None, check_every=0):
        """Integrate ``cfg.ndays`` (model.py:168-290).

        ``writer`` (``.add(day, fields)``, ``.close()``) receives, after each
        ``out_interval_days``, the trapezoidal mean (``time_average``) or the
        last trapezoid of the top-of-step history fields as host arrays, with
        the member axis for an ensemble.  ``lengths`` (extension): per-member
        numbers of full steps for an uneven ensemble; members stop when they
        reach theirs (no writer then).
        """
        c = self.cfg
        tv, stepped = self._set_forcing(mjo_forcing, steady_forcing)
        spd = c.steps_per_day
        nsteps = int(round(c.ndays * spd))
        if lengths is not None:
            if writer is not None:
                raise ValueError("history output needs one length for every member")
            self.plan.set_lengths(lengths)
            nsteps = 0
        t0 = _time.time()
        K = 0
        if writer is not None:
            K = int(round(c.out_interval_days * spd))
            if K < 1:
                raise ValueError("out_interval_days is shorter than one step")
            self.plan.set_history(K, time_average=c.time_average, keep_last=True)
        try:
            if stepped is not None:
                self._run_stepped(stepped, nsteps, lengths, writer, K, use_graph, check_every,
                                  verbose, progress_every, t0)
            elif writer is None:
                self.plan.run(nsteps, use_graph=use_graph, check_every=check_every)
            else:
                day = 0
                for stop in range(K, nsteps + 1, K):
                    self.plan.run_to(nsteps, stop, use_graph=use_graph, check_every=check_every)
                    day += 1
                    self._write(writer, day)
                    if verbose and progress_every and (stop % progress_every) < K:
                        print(f"  day {stop / spd:6.2f}   [{_time.time() - t0:5.1f}s]", flush=True)
                if nsteps % K:
                    self.plan.run_to(nsteps, 0, use_graph=use_graph, check_every=check_every)
        finally:
            # a failed writer or an interrupt leaves no half-finished run behind
            self.plan.run_abort()
            if writer is not None:
                self.plan.set_history(0, keep_last=True)
        if writer is not None:
            writer.close()
        adv = nsteps if lengths is None else int(max(lengths))
        if tv is not None:
            tv.counter += c.inistp + adv
        self.time += adv * c.delt
        self.istep += adv
        self._ga = None
        if verbose:
            import torch
            torch.cuda.synchronize()
            print(f"  done: {adv} steps, {_time.time() - t0:.1f}s", flush=True)
        return self
