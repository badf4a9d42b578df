"""Owning wrapper of an ``lbm_plan`` (``include/lbm_hip.h``).

All device buffers passed to the library are contiguous float64 torch
tensors on the current HIP device, in the reference's layouts; the plan's
internal precision (FP64 or FP32) only changes the arithmetic inside.
"""

import ctypes

import numpy as np

from . import _lib
from .constants import ER, NMAX, IMAX, JMAX, SIGM


def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(x, dtype=None):
    """float64 (or complex128) contiguous device tensor."""
    torch = _torch()
    if isinstance(x, torch.Tensor):
        t = x.to(device="cuda")
    else:
        a = np.ascontiguousarray(x)
        t = torch.as_tensor(a if a.flags.writeable else a.copy(), device="cuda")
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


def _shape(x):
    """Shape of an array-like without moving it."""
    return tuple(int(n) for n in (x.shape if hasattr(x, "shape") else np.shape(x)))


def _lead_axis(x, nd):
    """``x`` with a leading member axis when it has ``nd`` dimensions."""
    if len(_shape(x)) != nd:
        return x
    return x.unsqueeze(0) if hasattr(x, "unsqueeze") else np.asarray(x)[None]


def _expand(x, shp, what, dtype):
    """``x`` as a contiguous device tensor of shape ``shp`` (broadcast);
    ValueError, before any device transfer, if it does not broadcast."""
    have = _shape(x)
    try:
        ok = np.broadcast_shapes(have, tuple(shp)) == tuple(shp)
    except ValueError:
        ok = False
    if not ok:
        raise ValueError(f"{what}: shape {have} does not broadcast to {tuple(shp)}")
    t = _dev(x, dtype)
    return t if tuple(t.shape) == tuple(shp) else t.expand(shp).contiguous()


def zero_damping(nmax, nlev):
    z = np.zeros((nmax + 1) * nlev)
    return z, z, z, z


class Plan:
    """One model/transform plan: grid, tables, operators, state, work buffers."""

    # bumped by every change of what a run computes (response, lengths,
    # forcings, basic state): the autograd Function checks it
    generation = 0

    def __init__(self, nlon=IMAX, nlat=JMAX, nmax=NMAX, radius=ER, sigm=SIGM, nmem=1,
                 nbasic=1, precision="fp64", damping=None, vdif=(None, None, None, None),
                 delt=1200.0, robert=0.04, inistp=2, strictly_linear=False,
                 tangent_linear=False):
        lib = _lib.load()
        self.nlon, self.nlat, self.nmax = int(nlon), int(nlat), int(nmax)
        self.sigm = np.ascontiguousarray(sigm, dtype=np.float64)
        self.nlev = self.sigm.size - 1
        self.nmem, self.nbasic = int(nmem), int(nbasic)
        self.precision = precision
        prec = {"fp64": _lib.LBM_FP64, "fp32": _lib.LBM_FP32,
                "bf16x3": _lib.LBM_BF16X3}.get(str(precision).lower())
        if prec is None:
            raise ValueError(f"unknown precision {precision!r} (fp64, fp32 or bf16x3)")
        if damping is None:
            damping = zero_damping(self.nmax, self.nlev)
        self._keep = [self.sigm]
        d = _lib.LbmDesc()
        d.nlon, d.nlat, d.nmax, d.nlev = self.nlon, self.nlat, self.nmax, self.nlev
        d.nmem, d.nbasic, d.precision = self.nmem, self.nbasic, prec
        d.strictly_linear = int(bool(strictly_linear))
        d.tangent_linear = int(bool(tangent_linear))
        self.tangent_linear = bool(tangent_linear)
        d.inistp = int(inistp)
        self.inistp, self.delt = int(inistp), float(delt)
        self.strictly_linear = bool(strictly_linear)
        d.skip_low_n_vdiff = 1
        d.radius, d.delt, d.robert = float(radius), float(delt), float(robert)
        d.h_sigm = _lib.dptr(self.sigm)
        arrs = []
        for a in damping:
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            if a.size != (self.nmax + 1) * self.nlev:
                raise ValueError("damping arrays must be (nmax+1, nlev)")
            arrs.append(a)
        self._keep += arrs
        d.h_kv, d.h_kd, d.h_ktt, d.h_kq = (_lib.dptr(a) for a in arrs)
        vd = []
        for v in vdif:
            if v is None:
                vd.append(None)
            else:
                v = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (self.nlev,)))
                self._keep.append(v)
                vd.append(_lib.dptr(v))
        d.h_vdif_vor, d.h_vdif_div, d.h_vdif_t, d.h_vdif_q = vd
        h = ctypes.c_void_p()
        _lib.check(lib.lbm_plan_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.lbm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- host setup arrays -------------------------------------------------
    def get(self, name, count):
        out = np.empty(count, dtype=np.float64)
        n = _lib.check(self._lib.lbm_plan_get(self._h, name.encode(), _lib.dptr(out), count))
        return out[:n]

    # -- shape checks (before any device call) -----------------------------
    def _grid(self, x, what, nk=None, lead=None):
        """(B, nlat, nlon, K) float64 device tensor of a grid argument; raises
        ValueError, before touching the device, when its latitude/longitude
        (or level, member) extents are not the plan's."""
        self._grid_check(x, what, nk, lead)
        return _dev(x, _torch().float64)

    def _grid_check(self, x, what, nk=None, lead=None):
        shp = _shape(x)
        if len(shp) != 4 or shp[1:3] != (self.nlat, self.nlon):
            raise ValueError(f"{what}: grid field must be (B, {self.nlat}, {self.nlon}, K), "
                             f"got {shp}")
        if nk is not None and shp[3] != nk:
            raise ValueError(f"{what}: expected {nk} levels, got {shp[3]}")
        if lead is not None and shp[0] not in lead:
            raise ValueError(f"{what}: leading axis must be one of {sorted(set(lead))}, "
                             f"got {shp[0]}")

    def _surf(self, x, what, lead):
        """(B, nlat, nlon) float64 device tensor of a surface argument."""
        self._surf_check(x, what, lead)
        return _dev(x, _torch().float64)

    def _surf_check(self, x, what, lead):
        shp = _shape(x)
        if len(shp) != 3 or shp[1:] != (self.nlat, self.nlon) or shp[0] not in lead:
            raise ValueError(f"{what}: surface field must be ({'|'.join(map(str, sorted(set(lead))))}"
                             f", {self.nlat}, {self.nlon}), got {shp}")

    def _spec(self, x, what):
        """(B, nmax+1, nmax+1, K) complex128 device tensor of a spectral argument."""
        self._spec_check(x, what)
        return _dev(x, _torch().complex128)

    def _spec_check(self, x, what):
        shp = _shape(x)
        M = self.nmax + 1
        if len(shp) != 4 or shp[1:3] != (M, M):
            raise ValueError(f"{what}: spectral field must be (B, {M}, {M}, K), got {shp}")

    def _pair(self, a, b, what, spectral):
        """Both arguments of a two-field transform, checked (same shape, the
        plan's extents) before either moves to the device."""
        chk = self._spec_check if spectral else self._grid_check
        chk(a, what)
        chk(b, what)
        if _shape(a) != _shape(b):
            raise ValueError(f"{what}: the two fields differ in shape")
        dt = _torch().complex128 if spectral else _torch().float64
        return _dev(a, dt), _dev(b, dt)

    # -- transforms --------------------------------------------------------
    def _call(self, fn, *ptrs, B, nk):
        _lib.check(fn(self._h, *ptrs, int(B), int(nk), _stream()))

    def grid_to_wave(self, g):
        torch = _torch()
        g = self._grid(g, "grid_to_wave")
        B, J, I, K = g.shape
        w = torch.empty((B, self.nmax + 1, self.nmax + 1, K), dtype=torch.complex128, device="cuda")
        self._call(self._lib.lbm_grid_to_wave, _lib.ptr(g), _lib.ptr(w), B=B, nk=K)
        return w

    def wave_to_grid(self, w):
        torch = _torch()
        w = self._spec(w, "wave_to_grid")
        B, M, _, K = w.shape
        g = torch.empty((B, self.nlat, self.nlon, K), dtype=torch.float64, device="cuda")
        self._call(self._lib.lbm_wave_to_grid, _lib.ptr(w), _lib.ptr(g), B=B, nk=K)
        return g

    def uv_to_vordiv(self, u, v):
        torch = _torch()
        u, v = self._pair(u, v, "uv_to_vordiv", False)
        B, J, I, K = u.shape
        shp = (B, self.nmax + 1, self.nmax + 1, K)
        vor = torch.empty(shp, dtype=torch.complex128, device="cuda")
        div = torch.empty(shp, dtype=torch.complex128, device="cuda")
        self._call(self._lib.lbm_uv_to_vordiv, _lib.ptr(u), _lib.ptr(v), _lib.ptr(vor),
                   _lib.ptr(div), B=B, nk=K)
        return vor, div

    def psichi_to_uv(self, psi, chi):
        torch = _torch()
        psi, chi = self._pair(psi, chi, "psichi_to_uv", True)
        B, M, _, K = psi.shape
        shp = (B, self.nlat, self.nlon, K)
        u = torch.empty(shp, dtype=torch.float64, device="cuda")
        v = torch.empty(shp, dtype=torch.float64, device="cuda")
        self._call(self._lib.lbm_psichi_to_uv, _lib.ptr(psi), _lib.ptr(chi), _lib.ptr(u),
                   _lib.ptr(v), B=B, nk=K)
        return u, v

    def hgrad(self, w):
        torch = _torch()
        w = self._spec(w, "hgrad")
        B, M, _, K = w.shape
        shp = (B, self.nlat, self.nlon, K)
        gx = torch.empty(shp, dtype=torch.float64, device="cuda")
        gy = torch.empty(shp, dtype=torch.float64, device="cuda")
        self._call(self._lib.lbm_hgrad, _lib.ptr(w), _lib.ptr(gx), _lib.ptr(gy), B=B, nk=K)
        return gx, gy

    # -- transposed transforms (adjoint twins) ------------------------------
    def _grid_out(self, B, K, n):
        torch = _torch()
        return [torch.empty((B, self.nlat, self.nlon, K), dtype=torch.float64, device="cuda")
                for _ in range(n)]

    def _spec_out(self, B, K, n):
        torch = _torch()
        M = self.nmax + 1
        return [torch.empty((B, M, M, K), dtype=torch.complex128, device="cuda") for _ in range(n)]

    def grid_to_wave_adj(self, w):
        """Transpose of grid_to_wave: spectral (B, M, M, K) -> grid."""
        w = self._spec(w, "grid_to_wave_adj")
        B, _, _, K = w.shape
        (g,) = self._grid_out(B, K, 1)
        self._call(self._lib.lbm_grid_to_wave_adj, _lib.ptr(w), _lib.ptr(g), B=B, nk=K)
        return g

    def wave_to_grid_adj(self, g):
        """Transpose of wave_to_grid: grid (B, J, I, K) -> spectral."""
        g = self._grid(g, "wave_to_grid_adj")
        B, _, _, K = g.shape
        (w,) = self._spec_out(B, K, 1)
        self._call(self._lib.lbm_wave_to_grid_adj, _lib.ptr(g), _lib.ptr(w), B=B, nk=K)
        return w

    def uv_to_vordiv_adj(self, vor, div):
        """Transpose of uv_to_vordiv: spectral (vor, div) -> grid (u, v)."""
        torch = _torch()
        vor, div = self._pair(vor, div, "uv_to_vordiv_adj", True)
        B, _, _, K = vor.shape
        u, v = self._grid_out(B, K, 2)
        self._call(self._lib.lbm_uv_to_vordiv_adj, _lib.ptr(vor), _lib.ptr(div), _lib.ptr(u),
                   _lib.ptr(v), B=B, nk=K)
        return u, v

    def psichi_to_uv_adj(self, u, v):
        """Transpose of psichi_to_uv: grid (u, v) -> spectral (psi, chi)."""
        torch = _torch()
        u, v = self._pair(u, v, "psichi_to_uv_adj", False)
        B, _, _, K = u.shape
        psi, chi = self._spec_out(B, K, 2)
        self._call(self._lib.lbm_psichi_to_uv_adj, _lib.ptr(u), _lib.ptr(v), _lib.ptr(psi),
                   _lib.ptr(chi), B=B, nk=K)
        return psi, chi

    def hgrad_adj(self, gx, gy):
        """Transpose of hgrad: grid (gx, gy) -> spectral."""
        torch = _torch()
        gx, gy = self._pair(gx, gy, "hgrad_adj", False)
        B, _, _, K = gx.shape
        (w,) = self._spec_out(B, K, 1)
        self._call(self._lib.lbm_hgrad_adj, _lib.ptr(gx), _lib.ptr(gy), _lib.ptr(w), B=B, nk=K)
        return w

    # -- model -------------------------------------------------------------
    def set_basic_state(self, u, v, t, q, ps):
        """nbasic basic states: u, v, t, q (nbasic, J, I, L), ps (nbasic, J, I) hPa."""
        self.generation += 1
        nb = (self.nbasic,)
        for k, x in (("u", u), ("v", v), ("t", t), ("q", q)):
            self._grid_check(x, f"basic state {k}", self.nlev, nb)
        shp = _shape(ps)
        if len(shp) != 3 or shp != (self.nbasic, self.nlat, self.nlon):
            raise ValueError(f"basic state ps: must be ({self.nbasic}, {self.nlat}, {self.nlon}), "
                             f"got {shp}")
        f = [_dev(x, _torch().float64) for x in (u, v, t, q, ps)]
        _lib.check(self._lib.lbm_set_basic_state(self._h, *[_lib.ptr(x) for x in f], _stream()))

    def set_forcing(self, frc):
        self.generation += 1
        torch = _torch()
        if frc is None:
            _lib.check(self._lib.lbm_set_forcing(self._h, None, 0, _stream()))
            return
        f = self._grid(_lead_axis(frc, 3), "temperature forcing", self.nlev, (1, self.nmem))
        _lib.check(self._lib.lbm_set_forcing(self._h, _lib.ptr(f), int(f.shape[0]), _stream()))

    def set_steady_forcing(self, vor=None, div=None, ps=None):
        """Steady vorticity / divergence (J, I, L) and ln ps (J, I) forcings,
        or with a leading member axis; all None removes them."""
        self.generation += 1
        torch = _torch()
        E, J, I, L = self.nmem, self.nlat, self.nlon, self.nlev
        given, nf = [], None
        for x, nd in ((vor, 3), (div, 3), (ps, 2)):
            if x is None:
                given.append(None)
                continue
            x = _lead_axis(x, nd)
            if nd == 3:
                self._grid_check(x, "steady forcing", self.nlev, (1, self.nmem))
            else:
                self._surf_check(x, "steady ln ps forcing", (1, self.nmem))
            if nf is not None and _shape(x)[0] != nf:
                raise ValueError("steady forcings need the same member count")
            nf = _shape(x)[0]
            given.append(x)
        args = [None if x is None else _dev(x, torch.float64) for x in given]
        self._steady = args
        _lib.check(self._lib.lbm_set_steady_forcing(self._h, *[_lib.ptr(x) for x in args],
                                                    int(nf or 0), _stream()))

    def set_state(self, u=None, v=None, t=None, q=None, ps=None):
        torch = _torch()
        E, J, I, L = self.nmem, self.nlat, self.nlon, self.nlev
        args = []
        for x, shp in ((u, (E, J, I, L)), (v, (E, J, I, L)), (t, (E, J, I, L)),
                       (q, (E, J, I, L)), (ps, (E, J, I))):
            if x is None:
                args.append(None)
                continue
            x = _expand(x, shp, "initial perturbation", torch.float64)
            args.append(x)
        self._state_inputs = args
        _lib.check(self._lib.lbm_set_state(self._h, *[_lib.ptr(x) for x in args], _stream()))

    def set_spectral_state(self, vor=None, div=None, t=None, q=None, pi=None):
        """Spectral initial perturbation (E, M, M, L) / (E, M, M) complex128 for
        both time levels; None means zero."""
        torch = _torch()
        E, M, L = self.nmem, self.nmax + 1, self.nlev
        args = []
        for x, shp in ((vor, (E, M, M, L)), (div, (E, M, M, L)), (t, (E, M, M, L)),
                       (q, (E, M, M, L)), (pi, (E, M, M))):
            if x is None:
                args.append(None)
                continue
            x = _expand(x, shp, "spectral initial state", torch.complex128)
            args.append(x)
        self._state_inputs = args
        _lib.check(self._lib.lbm_set_spectral_state(self._h, *[_lib.ptr(x) for x in args],
                                                    _stream()))

    def set_response(self, wgrid, wlev=None, wstep=None):
        """J_e = sum_k wstep[k] sum_l wlev[l] sum_ji wgrid[j,i] psi_l(j,i) over the
        states after steps k of the next run; wgrid None removes it."""
        self.generation += 1
        torch = _torch()
        torch.cuda.current_stream().synchronize()
        if wgrid is None:
            _lib.check(self._lib.lbm_set_response(self._h, None, None, None, 0))
            return
        wg = np.ascontiguousarray(wgrid, dtype=np.float64)
        if wg.shape != (self.nlat, self.nlon):
            raise ValueError(f"wgrid must be ({self.nlat}, {self.nlon})")
        wl = np.ascontiguousarray(np.broadcast_to(
            np.asarray(1.0 if wlev is None else wlev, np.float64), (self.nlev,)))
        ws = np.ascontiguousarray(np.asarray(wstep, dtype=np.float64).reshape(-1))
        self._resp = (wg, wl, ws)
        _lib.check(self._lib.lbm_set_response(self._h, _lib.dptr(wg), _lib.dptr(wl),
                                              _lib.dptr(ws), int(ws.size)))

    def get_response(self):
        torch = _torch()
        torch.cuda.current_stream().synchronize()
        out = np.zeros(self.nmem)
        _lib.check(self._lib.lbm_get_response(self._h, _lib.dptr(out)))
        return out

    def adjoint_run(self, nsteps, use_graph=True):
        """Gradients of J_e with respect to member e's spectral initial state
        (vor, div, t, q, pi; complex128 dJ/dRe + i dJ/dIm) and its temperature
        forcing (E, J, I, L)."""
        torch = _torch()
        E, M, L, J, I = self.nmem, self.nmax + 1, self.nlev, self.nlat, self.nlon
        g3 = [torch.empty((E, M, M, L), dtype=torch.complex128, device="cuda") for _ in range(4)]
        gpi = torch.empty((E, M, M), dtype=torch.complex128, device="cuda")
        gf = torch.empty((E, J, I, L), dtype=torch.float64, device="cuda")
        _lib.check(self._lib.lbm_adjoint_run(self._h, int(nsteps), *[_lib.ptr(x) for x in g3],
                                             _lib.ptr(gpi), _lib.ptr(gf), int(bool(use_graph)),
                                             _stream()))
        return g3 + [gpi], gf

    def set_checkpoints(self, every):
        """Keep the model state every ``every`` executed steps of the next runs
        (0: off), for the adjoint of the default model."""
        _torch().cuda.synchronize()
        _lib.check(self._lib.lbm_set_checkpoints(self._h, int(every)))

    def set_lengths(self, nsteps):
        self.generation += 1
        if nsteps is None:
            _lib.check(self._lib.lbm_set_lengths(self._h, None, None))
            return
        arr = (ctypes.c_int32 * self.nmem)(*[int(x) for x in nsteps])
        _lib.check(self._lib.lbm_set_lengths(self._h, arr, None))

    def run(self, nsteps, use_graph=True, check_every=0):
        _lib.check(self._lib.lbm_run(self._h, int(nsteps), int(bool(use_graph)),
                                     int(check_every), _stream()))

    def run_to(self, nsteps, stop, use_graph=True, check_every=0):
        """Run (or continue) the run of ``run(nsteps)`` up to full step
        ``stop`` (0: its end) without restarting the start-up steps."""
        _lib.check(self._lib.lbm_run_to(self._h, int(nsteps), int(stop), int(bool(use_graph)),
                                        int(check_every), _stream()))

    def set_history(self, interval_steps, time_average=True, keep_last=True):
        """Accumulate the top-of-step history over intervals of
        ``interval_steps`` full steps (0: off); ``keep_last`` keeps the last
        top-of-step sample for :meth:`get_diagnostics`."""
        _torch().cuda.current_stream().synchronize()
        _lib.check(self._lib.lbm_set_history(self._h, int(interval_steps), int(bool(time_average)),
                                             int(bool(keep_last)), _stream()))

    def get_history(self, interval):
        """Fields of output interval ``interval`` (1-based): dict of
        psi, chi, u, v, w, t, z (E, J, I, L) and p (E, J, I) float64 tensors."""
        torch = _torch()
        E, J, I, L = self.nmem, self.nlat, self.nlon, self.nlev
        out = {k: torch.empty((E, J, I, L), dtype=torch.float64, device="cuda")
               for k in ("psi", "chi", "u", "v", "w", "t", "z")}
        out["p"] = torch.empty((E, J, I), dtype=torch.float64, device="cuda")
        _lib.check(self._lib.lbm_get_history(
            self._h, int(interval), *[_lib.ptr(out[k]) for k in ("psi", "chi", "u", "v", "w",
                                                                  "t", "z", "p")], _stream()))
        return out

    def get_diagnostics(self):
        """psi, chi, w of the last step's tendencies and z of the current state
        (E, J, I, L) float64 tensors."""
        torch = _torch()
        E, J, I, L = self.nmem, self.nlat, self.nlon, self.nlev
        out = {k: torch.empty((E, J, I, L), dtype=torch.float64, device="cuda")
               for k in ("psi", "chi", "w", "z")}
        _lib.check(self._lib.lbm_get_diagnostics(
            self._h, *[_lib.ptr(out[k]) for k in ("psi", "chi", "w", "z")], _stream()))
        return out

    def set_mjo_forcing(self, gtf, nt, spinup, call0=0):
        """Daily temperature forcing fields (days, J, I, L) shared by every
        member or (E, days, J, I, L) per member, interpolated on the device
        (with the float32 slope of the reference when ``gtf`` is float32)."""
        self.generation += 1
        torch = _torch()
        single = (gtf.dtype == torch.float32) if isinstance(gtf, torch.Tensor) else \
            (np.asarray(gtf).dtype == np.float32)
        shp = _shape(gtf)
        if len(shp) == 4:
            shp = (1,) + shp
        if len(shp) != 5 or shp[2:] != (self.nlat, self.nlon, self.nlev) or \
                shp[0] not in (1, self.nmem):
            raise ValueError(f"MJO forcing must be (days, {self.nlat}, {self.nlon}, {self.nlev}) "
                             "or (nmem, days, ...)")
        g = _dev(gtf, torch.float64).reshape(shp)
        _lib.check(self._lib.lbm_set_mjo_forcing(self._h, _lib.ptr(g), int(g.shape[0]),
                                                 int(g.shape[1]), int(nt), int(spinup), int(call0),
                                                 int(single), _stream()))

    def get_state(self):
        torch = _torch()
        E, J, I, L = self.nmem, self.nlat, self.nlon, self.nlev
        out = [torch.empty((E, J, I, L), dtype=torch.float64, device="cuda") for _ in range(4)]
        ps = torch.empty((E, J, I), dtype=torch.float64, device="cuda")
        _lib.check(self._lib.lbm_get_state(self._h, *[_lib.ptr(x) for x in out], _lib.ptr(ps),
                                           _stream()))
        return out + [ps]

    def get_spectral(self, level=0):
        torch = _torch()
        E, M, L = self.nmem, self.nmax + 1, self.nlev
        out = [torch.empty((E, M, M, L), dtype=torch.complex128, device="cuda") for _ in range(4)]
        pi = torch.empty((E, M, M), dtype=torch.complex128, device="cuda")
        _lib.check(self._lib.lbm_get_spectral(self._h, int(level), *[_lib.ptr(x) for x in out],
                                              _lib.ptr(pi), _stream()))
        return out + [pi]

    def set_state_keep(self, u=None, v=None, t=None, q=None, ps=None):
        """set_initial_perturbation (``model.py:142-148``): given fields
        replace both time levels, omitted ones keep each level's values."""
        torch = _torch()
        E, J, I, L = self.nmem, self.nlat, self.nlon, self.nlev
        args = []
        for x, shp in ((u, (E, J, I, L)), (v, (E, J, I, L)), (t, (E, J, I, L)),
                       (q, (E, J, I, L)), (ps, (E, J, I))):
            args.append(None if x is None else
                        _expand(x, shp, "initial perturbation", torch.float64))
        self._state_inputs = args
        _lib.check(self._lib.lbm_set_state_keep(self._h, *[_lib.ptr(x) for x in args],
                                                _stream()))

    def get_prev_state(self):
        """Grid fields (u, v, t, q, ps') of the filtered previous level."""
        torch = _torch()
        E, J, I, L = self.nmem, self.nlat, self.nlon, self.nlev
        out = [torch.empty((E, J, I, L), dtype=torch.float64, device="cuda") for _ in range(4)]
        ps = torch.empty((E, J, I), dtype=torch.float64, device="cuda")
        _lib.check(self._lib.lbm_get_prev_state(self._h, *[_lib.ptr(x) for x in out],
                                                _lib.ptr(ps), _stream()))
        return out + [ps]

    def _spec5(self, fields, what):
        """(vor, div, t, q, pi) complex128 device tensors of the plan's shapes."""
        torch = _torch()
        E, M, L = self.nmem, self.nmax + 1, self.nlev
        for x, shp in zip(fields, [(E, M, M, L)] * 4 + [(E, M, M)]):
            if x is None:
                raise ValueError(f"{what}: every field is needed")
            if _shape(x) != shp:
                raise ValueError(f"{what}: expected {shp}, got {_shape(x)}")
        return [_dev(x, torch.complex128) for x in fields]

    def _spec5_check(self, fields, what):
        E, M, L = self.nmem, self.nmax + 1, self.nlev
        if len(fields) != 5:
            raise ValueError(f"{what}: five fields (vor, div, t, q, pi)")
        for x, shp in zip(fields, [(E, M, M, L)] * 4 + [(E, M, M)]):
            if x is None or _shape(x) != shp:
                raise ValueError(f"{what}: expected {shp}, got {None if x is None else _shape(x)}")

    def _spec5_out(self):
        torch = _torch()
        E, M, L = self.nmem, self.nmax + 1, self.nlev
        out = [torch.empty((E, M, M, L), dtype=torch.complex128, device="cuda") for _ in range(4)]
        return out + [torch.empty((E, M, M), dtype=torch.complex128, device="cuda")]

    def tendencies(self, u=None, v=None, t=None, q=None, ps=None):
        """Explicit spectral tendencies (vor, div, t, q, pi) of a grid state
        with the plan's basic state and forcings (``dynamics.py:156-271``)."""
        torch = _torch()
        E, J, I, L = self.nmem, self.nlat, self.nlon, self.nlev
        args = []
        for x, shp in ((u, (E, J, I, L)), (v, (E, J, I, L)), (t, (E, J, I, L)),
                       (q, (E, J, I, L)), (ps, (E, J, I))):
            args.append(None if x is None else _expand(x, shp, "state", torch.float64))
        out = self._spec5_out()
        _lib.check(self._lib.lbm_tendencies(self._h, *[_lib.ptr(x) for x in args],
                                            *[_lib.ptr(x) for x in out], _stream()))
        return out

    def prepare_dt(self, dt):
        _torch().cuda.current_stream().synchronize()
        _lib.check(self._lib.lbm_prepare_dt(self._h, float(dt)))

    def step(self, dt, prev, tend):
        """One semi-implicit leapfrog step (``integrator.py:264-315``): prev and
        tend are (vor, div, t, q, pi) spectral tensors; returns the new level."""
        self._spec5_check(prev, "previous level")
        self._spec5_check(tend, "tendencies")
        p = self._spec5(prev, "previous level")
        t = self._spec5(tend, "tendencies")
        out = self._spec5_out()
        _lib.check(self._lib.lbm_step(self._h, float(dt), *[_lib.ptr(x) for x in p],
                                      *[_lib.ptr(x) for x in t], *[_lib.ptr(x) for x in out],
                                      _stream()))
        return out

    def run_steps(self, nsteps, count, use_graph=True, check_every=0):
        """Advance the run of ``run(nsteps)`` by ``count`` executed steps."""
        _lib.check(self._lib.lbm_run_steps(self._h, int(nsteps), int(count),
                                           int(bool(use_graph)), int(check_every), _stream()))

    def run_abort(self):
        _lib.check(self._lib.lbm_run_abort(self._h))

    def run_segment_done(self):
        n = ctypes.c_int32()
        _lib.check(self._lib.lbm_run_segment_done(self._h, ctypes.byref(n)))
        return n.value

    def spec_scale(self, xi, kind):
        """(B, M, M, K) complex times -n(n+1)/a^2 (kind 0) or its inverse (1)."""
        t = self._spec(xi, "spectral scaling")
        out = _torch().empty_like(t)
        B, _, _, K = t.shape
        _lib.check(self._lib.lbm_spec_scale(self._h, _lib.ptr(t), _lib.ptr(out), int(B), int(K),
                                            int(kind), _stream()))
        return out

    # -- autotuner -----------------------------------------------------------
    TUNED = ("ifft_rows", "fft_rows", "syn_waves", "ana_waves", "tend_fft")

    def autotune(self, reps=10):
        """Time the variants of the step's tuned launches on this plan and
        keep the fastest (``lbm_autotune``); returns ``{"choice": [...],
        "report": text}``."""
        _torch().cuda.synchronize()
        buf = ctypes.create_string_buffer(2048)
        _lib.check(self._lib.lbm_autotune(self._h, int(reps), buf, len(buf)))
        return {"choice": self.get_tuning(), "report": buf.value.decode()}

    def set_tuning(self, choice):
        arr = (ctypes.c_int32 * len(choice))(*[int(c) for c in choice])
        _lib.check(self._lib.lbm_set_tuning(self._h, arr, len(choice)))

    def get_tuning(self):
        arr = (ctypes.c_int32 * len(self.TUNED))()
        _lib.check(self._lib.lbm_get_tuning(self._h, arr, len(self.TUNED)))
        return [int(x) for x in arr]

    def tuning_key(self):
        """Cache key: grid, levels, members, basic states and precision."""
        return (f"{self.nlon}x{self.nlat} N{self.nmax} L{self.nlev} E{self.nmem} "
                f"B{self.nbasic} {self.precision}")

    def steps_taken(self):
        n = ctypes.c_int64()
        _lib.check(self._lib.lbm_steps_taken(self._h, ctypes.byref(n)))
        return n.value

    # -- measurement -------------------------------------------------------
    def profile(self, enable):
        _lib.check(self._lib.lbm_profile(self._h, int(bool(enable))))

    def profile_read(self):
        """{slot: (total ms, launches)} accumulated since the last read."""
        n = len(_lib.PROFILE_SLOTS)
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        k = _lib.check(self._lib.lbm_profile_read(
            self._h, _lib.dptr(ms), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), n))
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(_lib.PROFILE_SLOTS[:k])}
