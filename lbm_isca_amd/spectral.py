"""``Spectral`` with the reference's interface (``pylbm_isca/spectral.py:43-289``),
backed by the gfx950 transform kernels.

Grid fields are ``(nlat, nlon)`` or ``(nlat, nlon, nk)``, latitudes north to
south; spectral fields ``(nmax+1, nmax+1[, nk])`` complex indexed ``[m, n]``
with zeros for n < m (``spectral.py:24-34``).  Inputs may be numpy arrays or
torch tensors; outputs are torch tensors on the HIP device (use
:func:`lbm_isca_amd.asnumpy` for host copies).  ``BatchedSpectral`` in
:mod:`lbm_isca_amd.ensemble` carries a leading member axis.
"""

import numpy as np

from .backend import resolve
from .constants import ER, IMAX, JMAX, NMAX
from .plan import Plan


def _torch():
    import torch
    return torch


class Spectral:
    """Gaussian grid + spherical-harmonic transform operators on the GPU."""

    batched = False

    def __init__(self, nlon=IMAX, nlat=JMAX, nmax=NMAX, radius=ER, device="hip",
                 precision="fp64"):
        self.xp = resolve(device)
        self.nlon, self.nlat, self.nmax = int(nlon), int(nlat), int(nmax)
        self.a = float(radius)
        self.plan = Plan(self.nlon, self.nlat, self.nmax, self.a, precision=precision)
        J = self.nlat
        self.mu = self.plan.get("mu", J)
        self.gw = self.plan.get("gw", J)
        self.lat = self.plan.get("lat", J)
        self.lon = self.plan.get("lon", self.nlon)
        coslat = self.plan.get("coslat", J)
        self.gwdel = self.gw / self.a / (1.0 - self.mu ** 2)
        n = np.arange(self.nmax + 1, dtype=np.float64)
        self.nn1 = n * (n + 1.0)
        laplace = -self.nn1 / self.a ** 2
        inv_laplace = np.zeros_like(n)
        inv_laplace[1:] = -self.a ** 2 / self.nn1[1:]
        torch = _torch()
        self.coslat = torch.as_tensor(coslat, device="cuda")
        self.laplace = torch.as_tensor(laplace, device="cuda")
        self.inv_laplace = torch.as_tensor(inv_laplace, device="cuda")
        self.im = torch.as_tensor(1j * n, device="cuda")
        mm, nn = np.meshgrid(np.arange(self.nmax + 1), np.arange(self.nmax + 1), indexing="ij")
        self.mask = nn >= mm

    @property
    def lat_deg(self):
        return np.degrees(self.lat)

    @property
    def lon_deg(self):
        return np.degrees(self.lon)

    # -- shape handling: (.., K) with an explicit leading batch axis --------
    def _grid4(self, f):
        torch = _torch()
        t = torch.as_tensor(f, device="cuda") if not isinstance(f, torch.Tensor) else f.to("cuda")
        t = t.to(torch.float64)
        if self.batched:
            if t.dim() == 3:
                return t.unsqueeze(-1).contiguous(), (lambda x: x[..., 0])
            if t.dim() == 4:
                return t.contiguous(), (lambda x: x)
        else:
            if t.dim() == 2:
                return t[None, :, :, None].contiguous(), (lambda x: x[0, ..., 0])
            if t.dim() == 3:
                return t[None].contiguous(), (lambda x: x[0])
        raise ValueError(f"unexpected grid field shape {tuple(t.shape)}")

    def _spec4(self, xi):
        torch = _torch()
        t = torch.as_tensor(xi, device="cuda") if not isinstance(xi, torch.Tensor) else xi.to("cuda")
        t = t.to(torch.complex128)
        if self.batched:
            if t.dim() == 3:
                return t.unsqueeze(-1).contiguous(), (lambda x: x[..., 0])
            if t.dim() == 4:
                return t.contiguous(), (lambda x: x)
        else:
            if t.dim() == 2:
                return t[None, :, :, None].contiguous(), (lambda x: x[0, ..., 0])
            if t.dim() == 3:
                return t[None].contiguous(), (lambda x: x[0])
        raise ValueError(f"unexpected spectral field shape {tuple(t.shape)}")

    # -- scalar transforms (spectral.py:186-197) ----------------------------
    def grid_to_wave(self, f):
        g, back = self._grid4(f)
        return back(self.plan.grid_to_wave(g))

    def wave_to_grid(self, xi):
        w, back = self._spec4(xi)
        return back(self.plan.wave_to_grid(w))

    # -- differential operators (spectral.py:202-276) -----------------------
    def laplacian(self, xi):
        """xi times -n(n+1)/a^2 (``spectral.py:202-204``), on the device."""
        w, back = self._spec4(xi)
        return back(self.plan.spec_scale(w, 0))

    def inverse_laplacian(self, xi):
        """xi times -a^2/(n(n+1)), zero at n = 0 (``spectral.py:206-208``)."""
        w, back = self._spec4(xi)
        return back(self.plan.spec_scale(w, 1))

    def hgrad(self, xi):
        w, back = self._spec4(xi)
        gx, gy = self.plan.hgrad(w)
        return back(gx), back(gy)

    def uv_to_vordiv(self, u, v):
        gu, back = self._grid4(u)
        gv, _ = self._grid4(v)
        vor, div = self.plan.uv_to_vordiv(gu, gv)
        return back(vor), back(div)

    def neg_divergence(self, fx, fy):
        return -self.uv_to_vordiv(fx, fy)[1]

    def curl(self, fx, fy):
        return self.uv_to_vordiv(fx, fy)[0]

    def vordiv_to_uv(self, vor, div):
        return self.psichi_to_uv(self.inverse_laplacian(vor), self.inverse_laplacian(div))

    def psichi_to_uv(self, psi, chi):
        wp, back = self._spec4(psi)
        wc, _ = self._spec4(chi)
        u, v = self.plan.psichi_to_uv(wp, wc)
        return back(u), back(v)

    def truncate(self, f):
        return self.wave_to_grid(self.grid_to_wave(f))
