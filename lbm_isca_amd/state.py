"""Perturbation state on the Gaussian grid (``pylbm_isca/state.py:26-148``).

Five device fields: u, v, t, q (nlat, nlon, nlev) and ps' (nlat, nlon) in
hPa, with a leading member axis for an ensemble.  The prognostic spectral
variable of surface pressure is the exact pi' = ln(ps_bar + ps') - ln ps_bar
and its inverse ps' = ps_bar (exp(pi') - 1) (``state.py:77-83``).

Every operation runs on the device through the C ABI: the transforms of
``to_spectral`` / ``load_spectral``, the pi' <-> ps' pair (``lbm_lnps``),
the Robert-Asselin filter (``lbm_robert_asselin``) and the finiteness /
max-abs statistics (``lbm_grid_stats``); torch tensors only hold the memory.
The LBM time loop itself does not go through this class: it keeps its state
inside the library (spectral levels and the ps' grid), and ``LBM.ga`` /
``LBM.gb`` / ``LBM.gc`` hand out ``State`` snapshots of it.
"""

import numpy as np

from . import _lib
from .backend import asnumpy, resolve

FIELDS = ("u", "v", "t", "q", "ps")


def _torch():
    import torch
    return torch


def _f64(x):
    """Contiguous float64 device tensor (a copy only when needed)."""
    torch = _torch()
    t = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x))
    return t.to(device="cuda", dtype=torch.float64).contiguous()


def lnps_from_ps(ps, ps_bar):
    """pi' = ln(ps_bar + ps') - ln ps_bar on the device (``state.py:77-79``);
    ps_bar repeats over any leading axes of ps'."""
    ps, pb = _f64(ps), _f64(ps_bar)
    out = _torch().empty_like(ps)
    _lib.check(_lib.load().lbm_lnps(ps.numel(), pb.numel(), _lib.ptr(ps), _lib.ptr(pb),
                                    _lib.ptr(out), 0, _lib.stream()))
    return out


def ps_from_lnps(pi, ps_bar):
    """ps' = ps_bar (exp(pi') - 1) on the device (``state.py:81-83``)."""
    pi, pb = _f64(pi), _f64(ps_bar)
    out = _torch().empty_like(pi)
    _lib.check(_lib.load().lbm_lnps(pi.numel(), pb.numel(), _lib.ptr(pi), _lib.ptr(pb),
                                    _lib.ptr(out), 1, _lib.stream()))
    return out


def grid_stats(x):
    """(max |x| over the finite values, number of non-finite values)."""
    x = _f64(x)
    out = np.zeros(2)
    _lib.check(_lib.load().lbm_grid_stats(x.numel(), _lib.ptr(x), _lib.dptr(out), _lib.stream()))
    return float(out[0]), int(out[1])


def robert_asselin_field(prev, cur, new, gamma):
    """cur + gamma (prev - 2 cur + new) as a new device tensor."""
    out = _f64(prev).clone()
    cur, new = _f64(cur), _f64(new)
    if cur.shape != out.shape or new.shape != out.shape:
        raise ValueError("Robert-Asselin levels differ in shape")
    _lib.check(_lib.load().lbm_robert_asselin(out.numel(), _lib.ptr(out), _lib.ptr(cur),
                                              _lib.ptr(new), float(gamma), _lib.stream()))
    return out


class State:
    """Perturbation fields on the Gaussian grid (``state.py:26-71``)."""

    __slots__ = ("xp", "u", "v", "t", "q", "ps", "ps_bar", "lnps_bar")

    def __init__(self, ps_bar, u=None, v=None, t=None, q=None, ps=None, device="hip", xp=None):
        self.xp = xp if xp is not None else resolve(device)
        torch = _torch()
        pb = np.asarray(asnumpy(ps_bar), dtype=np.float64)
        self.ps_bar = _f64(pb)
        # setup-time constant of the basic state (host log, uploaded once)
        self.lnps_bar = _f64(np.log(pb))
        shape3 = next((tuple(f.shape) for f in (u, v, t, q) if f is not None), None)
        if shape3 is None:
            raise ValueError("need at least one 3-D field to infer the shape")
        z3 = lambda: torch.zeros(shape3, dtype=torch.float64, device="cuda")  # noqa: E731
        self.u = z3() if u is None else _f64(u)
        self.v = z3() if v is None else _f64(v)
        self.t = z3() if t is None else _f64(t)
        self.q = z3() if q is None else _f64(q)
        self.ps = (torch.zeros(tuple(pb.shape), dtype=torch.float64, device="cuda")
                   if ps is None else _f64(ps))

    # -- constructors ------------------------------------------------------
    @classmethod
    def zeros(cls, ps_bar, nlev, device="hip", xp=None):
        """Zero perturbation of one member (``state.py:54-58``)."""
        nlat, nlon = np.shape(ps_bar)
        torch = _torch()
        return cls(ps_bar, u=torch.zeros((nlat, nlon, nlev), dtype=torch.float64, device="cuda"),
                   device=device, xp=xp)

    @classmethod
    def zeros_ensemble(cls, ps_bar, nlev, nmem, device="hip", xp=None):
        """Zero perturbation of ``nmem`` members: (nmem, nlat, nlon, nlev) and
        ps' (nmem, nlat, nlon); ps_bar stays 2-D (``state.py:60-71``)."""
        nlat, nlon = np.shape(ps_bar)
        torch = _torch()
        obj = cls(ps_bar, u=torch.zeros((nmem, nlat, nlon, nlev), dtype=torch.float64,
                                        device="cuda"), device=device, xp=xp)
        obj.ps = torch.zeros((nmem, nlat, nlon), dtype=torch.float64, device="cuda")
        return obj

    def copy(self):
        return State(asnumpy(self.ps_bar), *(getattr(self, k).clone() for k in FIELDS),
                     xp=self.xp)

    # -- surface pressure ------------------------------------------------
    def lnps_pert(self):
        """pi' = ln(ps_bar + ps') - ln ps_bar (``state.py:77-79``)."""
        return lnps_from_ps(self.ps, self.ps_bar)

    def set_from_lnps(self, pi):
        """ps' = ps_bar (exp(pi') - 1) (``state.py:81-83``)."""
        self.ps = ps_from_lnps(pi, self.ps_bar)

    # -- grid <-> spectral (state.py:86-99) ------------------------------
    def to_spectral(self, sp):
        vor, div = sp.uv_to_vordiv(self.u, self.v)
        return {"vor": vor, "div": div, "t": sp.grid_to_wave(self.t),
                "pi": sp.grid_to_wave(self.lnps_pert()), "q": sp.grid_to_wave(self.q)}

    def load_spectral(self, sp, spec):
        self.u, self.v = sp.vordiv_to_uv(spec["vor"], spec["div"])
        self.t = sp.wave_to_grid(spec["t"])
        self.q = sp.wave_to_grid(spec["q"])
        self.set_from_lnps(sp.wave_to_grid(spec["pi"]))
        return self

    # -- time levels -------------------------------------------------------
    def assign(self, other):
        """Copies of every field of ``other`` (``state.py:101-108``)."""
        for k in FIELDS:
            setattr(self, k, _f64(getattr(other, k)).clone())
        return self

    def robert_asselin(self, current, new, gamma=0.04):
        """Robert-Asselin filter (``state.py:110-128``): ``self`` holds the
        t-dt level on entry and the filtered t level on return,
        X_f = X + gamma (X_prev - 2 X + X_new)."""
        for k in FIELDS:
            setattr(self, k, robert_asselin_field(getattr(self, k), getattr(current, k),
                                                  getattr(new, k), gamma))
        return self

    # -- checks ------------------------------------------------------------
    def max_abs(self):
        """max |x| per field (``state.py:131-137``)."""
        return {k: grid_stats(getattr(self, k))[0] for k in ("u", "v", "t", "ps", "q")}

    def isfinite(self):
        """No NaN or infinity in any field (``state.py:139-142``)."""
        return all(grid_stats(getattr(self, k))[1] == 0 for k in FIELDS)

    def to_numpy(self):
        return {k: asnumpy(getattr(self, k)) for k in FIELDS}
