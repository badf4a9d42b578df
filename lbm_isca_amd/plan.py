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
        t = torch.as_tensor(np.ascontiguousarray(x), device="cuda")
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


def zero_damping(nmax, nlev):
    z = np.zeros((nmax + 1) * nlev)
    return z, z, z, z


class Plan:
    """One model/transform plan: grid, tables, operators, state, work buffers."""

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

    # -- transforms --------------------------------------------------------
    def _call(self, fn, *ptrs, B, nk):
        _lib.check(fn(self._h, *ptrs, int(B), int(nk), _stream()))

    def grid_to_wave(self, g):
        torch = _torch()
        g = _dev(g, torch.float64)
        B, J, I, K = g.shape
        w = torch.empty((B, self.nmax + 1, self.nmax + 1, K), dtype=torch.complex128, device="cuda")
        self._call(self._lib.lbm_grid_to_wave, _lib.ptr(g), _lib.ptr(w), B=B, nk=K)
        return w

    def wave_to_grid(self, w):
        torch = _torch()
        w = _dev(w, torch.complex128)
        B, M, _, K = w.shape
        g = torch.empty((B, self.nlat, self.nlon, K), dtype=torch.float64, device="cuda")
        self._call(self._lib.lbm_wave_to_grid, _lib.ptr(w), _lib.ptr(g), B=B, nk=K)
        return g

    def uv_to_vordiv(self, u, v):
        torch = _torch()
        u = _dev(u, torch.float64)
        v = _dev(v, torch.float64)
        B, J, I, K = u.shape
        shp = (B, self.nmax + 1, self.nmax + 1, K)
        vor = torch.empty(shp, dtype=torch.complex128, device="cuda")
        div = torch.empty(shp, dtype=torch.complex128, device="cuda")
        self._call(self._lib.lbm_uv_to_vordiv, _lib.ptr(u), _lib.ptr(v), _lib.ptr(vor),
                   _lib.ptr(div), B=B, nk=K)
        return vor, div

    def psichi_to_uv(self, psi, chi):
        torch = _torch()
        psi = _dev(psi, torch.complex128)
        chi = _dev(chi, torch.complex128)
        B, M, _, K = psi.shape
        shp = (B, self.nlat, self.nlon, K)
        u = torch.empty(shp, dtype=torch.float64, device="cuda")
        v = torch.empty(shp, dtype=torch.float64, device="cuda")
        self._call(self._lib.lbm_psichi_to_uv, _lib.ptr(psi), _lib.ptr(chi), _lib.ptr(u),
                   _lib.ptr(v), B=B, nk=K)
        return u, v

    def hgrad(self, w):
        torch = _torch()
        w = _dev(w, torch.complex128)
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
        w = _dev(w, _torch().complex128)
        B, _, _, K = w.shape
        (g,) = self._grid_out(B, K, 1)
        self._call(self._lib.lbm_grid_to_wave_adj, _lib.ptr(w), _lib.ptr(g), B=B, nk=K)
        return g

    def wave_to_grid_adj(self, g):
        """Transpose of wave_to_grid: grid (B, J, I, K) -> spectral."""
        g = _dev(g, _torch().float64)
        B, _, _, K = g.shape
        (w,) = self._spec_out(B, K, 1)
        self._call(self._lib.lbm_wave_to_grid_adj, _lib.ptr(g), _lib.ptr(w), B=B, nk=K)
        return w

    def uv_to_vordiv_adj(self, vor, div):
        """Transpose of uv_to_vordiv: spectral (vor, div) -> grid (u, v)."""
        torch = _torch()
        vor, div = _dev(vor, torch.complex128), _dev(div, torch.complex128)
        B, _, _, K = vor.shape
        u, v = self._grid_out(B, K, 2)
        self._call(self._lib.lbm_uv_to_vordiv_adj, _lib.ptr(vor), _lib.ptr(div), _lib.ptr(u),
                   _lib.ptr(v), B=B, nk=K)
        return u, v

    def psichi_to_uv_adj(self, u, v):
        """Transpose of psichi_to_uv: grid (u, v) -> spectral (psi, chi)."""
        torch = _torch()
        u, v = _dev(u, torch.float64), _dev(v, torch.float64)
        B, _, _, K = u.shape
        psi, chi = self._spec_out(B, K, 2)
        self._call(self._lib.lbm_psichi_to_uv_adj, _lib.ptr(u), _lib.ptr(v), _lib.ptr(psi),
                   _lib.ptr(chi), B=B, nk=K)
        return psi, chi

    def hgrad_adj(self, gx, gy):
        """Transpose of hgrad: grid (gx, gy) -> spectral."""
        torch = _torch()
        gx, gy = _dev(gx, torch.float64), _dev(gy, torch.float64)
        B, _, _, K = gx.shape
        (w,) = self._spec_out(B, K, 1)
        self._call(self._lib.lbm_hgrad_adj, _lib.ptr(gx), _lib.ptr(gy), _lib.ptr(w), B=B, nk=K)
        return w

    # -- model -------------------------------------------------------------
    def set_basic_state(self, u, v, t, q, ps):
        torch = _torch()
        f = [_dev(x, torch.float64) for x in (u, v, t, q, ps)]
        _lib.check(self._lib.lbm_set_basic_state(self._h, *[_lib.ptr(x) for x in f], _stream()))

    def set_forcing(self, frc):
        torch = _torch()
        if frc is None:
            _lib.check(self._lib.lbm_set_forcing(self._h, None, 0, _stream()))
            return
        f = _dev(frc, torch.float64)
        if f.dim() == 3:
            f = f.unsqueeze(0)
        _lib.check(self._lib.lbm_set_forcing(self._h, _lib.ptr(f), int(f.shape[0]), _stream()))

    def set_steady_forcing(self, vor=None, div=None, ps=None):
        """Steady vorticity / divergence (J, I, L) and ln ps (J, I) forcings,
        or with a leading member axis; all None removes them."""
        torch = _torch()
        E, J, I, L = self.nmem, self.nlat, self.nlon, self.nlev
        args, nf = [], None
        for x, nd in ((vor, 3), (div, 3), (ps, 2)):
            if x is None:
                args.append(None)
                continue
            x = _dev(x, torch.float64)
            if x.dim() == nd:
                x = x.unsqueeze(0)
            if nf is not None and x.shape[0] != nf:
                raise ValueError("steady forcings need the same member count")
            nf = int(x.shape[0])
            args.append(x.contiguous())
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
            x = _dev(x, torch.float64)
            x = x.expand(shp).contiguous() if tuple(x.shape) != shp else x
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
            x = _dev(x, torch.complex128)
            x = x.expand(shp).contiguous() if tuple(x.shape) != shp else x
            args.append(x)
        self._state_inputs = args
        _lib.check(self._lib.lbm_set_spectral_state(self._h, *[_lib.ptr(x) for x in args],
                                                    _stream()))

    def set_response(self, wgrid, wlev=None, wstep=None):
        """J_e = sum_k wstep[k] sum_l wlev[l] sum_ji wgrid[j,i] psi_l(j,i) over the
        states after steps k of the next run; wgrid None removes it."""
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

    def set_lengths(self, nsteps):
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
        torch = _torch()
        single = (gtf.dtype == torch.float32) if isinstance(gtf, torch.Tensor) else \
            (np.asarray(gtf).dtype == np.float32)
        g = _dev(gtf, torch.float64)
        if g.dim() == 4:
            g = g.unsqueeze(0)
        if g.dim() != 5 or tuple(g.shape[2:]) != (self.nlat, self.nlon, self.nlev):
            raise ValueError(f"MJO forcing must be (days, {self.nlat}, {self.nlon}, {self.nlev}) "
                             "or (nmem, days, ...)")
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
