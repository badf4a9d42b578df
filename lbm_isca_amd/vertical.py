"""``VerticalGrid`` with the reference's interface (``pylbm_isca/vertical.py:50-216``).

The Simmons-Burridge coefficients are FP64 host setup (closed forms of the
half-level sigmas); the time-stepping kernels get the same values from the
library's own FP64 setup, built from the same ``sigm``.
"""

import numpy as np

from .backend import asnumpy, resolve
from .constants import KAPPA, RAIR, SIGM


class VerticalGrid:
    """Pure-sigma Simmons-Burridge coordinate and operator matrices."""

    def __init__(self, sigm=SIGM, kappa=KAPPA, device="hip"):
        self.xp = resolve(device)
        self.sigm = np.asarray(sigm, dtype=np.float64).copy()
        self.kmax = self.sigm.size - 1
        self.kappa = kappa
        bot, top = self.sigm[:-1], self.sigm[1:]
        if np.any(bot <= top):
            raise ValueError("sigm must decrease monotonically from the surface to the top")
        self.dsig = bot - top
        safe = np.where(top > 0, top, 1.0)
        ratio = np.where(top > 0, top * (np.log(bot) - np.log(safe)) / self.dsig, 0.0)
        self.alpha = 1.0 - ratio
        self.sig = bot * np.exp(-self.alpha)
        self.dlog1 = self.alpha.copy()
        self.dlog2 = np.where(top > 0, np.log(self.sig) - np.log(safe), 0.0)
        self.dlog3 = np.where(top > 0, np.log(bot) - np.log(safe), 0.0)
        self.cpg = (bot * self.dlog1 + top * self.dlog2) / self.dsig
        self.w3 = self.dlog3 / self.dsig
        self.wP = self.alpha - self.cpg
        k = self.kmax
        self.dsigm = np.empty(k + 1)
        self.dsigm[0] = self.sigm[0] - self.sig[0]
        self.dsigm[1:k] = self.sig[:-1] - self.sig[1:]
        self.dsigm[k] = self.sig[-1] - self.sigm[-1]
        self.ww = np.tril(np.ones((k, k)), -1) * (RAIR * self.dlog3)[None, :]
        self.ww[np.arange(k), np.arange(k)] = RAIR * self.alpha
        self.tbar = self.gg = self.tmtx = self.whmtx = self.gcmtx = None
        self._ww_dev = None
        # the BasicStateTerms that last set the reference temperature (the
        # semi-implicit operators of an Integrator built on this grid use it)
        self.basic_terms = None

    def set_reference_temperature(self, tbar):
        """gg, tmtx, whmtx, gcmtx about a reference Tv profile (vertical.py:139-193)."""
        tbar = np.asarray(asnumpy(tbar), dtype=np.float64).copy()
        k = self.kmax
        self.tbar = tbar
        self.gg = RAIR * tbar * self.cpg
        tm = np.zeros((k, k))
        for r in range(k):
            tm[r, r] = KAPPA * tbar[r] * self.alpha[r]
            tm[r, r + 1:] = KAPPA * tbar[r] * self.w3[r] * self.dsig[r + 1:]
        self.tmtx = tm
        self.whmtx = self.ww @ tm
        self.gcmtx = np.outer(self.gg, self.dsig)

    def hydrostatic(self, tv):
        """phi = Tv @ WW^T over the last (level) axis of a grid or spectral
        array (``vertical.py:196-203``): WW is R alpha on the diagonal and
        R dlog3_l below it, so phi_k is a running sum up the column, formed
        on the device (``lbm_hydrostatic``)."""
        import ctypes
        import torch
        from . import _lib
        t = tv if isinstance(tv, torch.Tensor) else torch.as_tensor(np.asarray(tv))
        t = t.to("cuda")
        if not t.is_complex():
            t = t.to(torch.float64)
        else:
            t = t.to(torch.complex128)
        t = t.contiguous()
        if t.shape[-1] != self.kmax:
            raise ValueError(f"hydrostatic: the last axis must hold the {self.kmax} levels")
        if getattr(self, "_ww_dev", None) is None:
            self._ww_dev = (torch.as_tensor(self.alpha, device="cuda"),
                            torch.as_tensor(self.dlog3, device="cuda"))
        al, dl = self._ww_dev
        out = torch.empty_like(t)
        _lib.check(_lib.load().lbm_hydrostatic(
            _lib.ptr(t), _lib.ptr(out), t.numel() // self.kmax, self.kmax, 2 if t.is_complex() else 1,
            _lib.ptr(al), _lib.ptr(dl), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return out


def compute_tbar(gft):
    """Plain grid mean over (lat, lon) per level (vertical.py:206-216)."""
    g = np.asarray(asnumpy(gft))
    nlat, nlon, _ = g.shape
    return g.sum(axis=(0, 1)) / (nlat * nlon)
