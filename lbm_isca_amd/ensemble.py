"""Ensemble batching with the reference's interface (``pylbm_isca/ensemble.py:41-165``).

The reference folds the member axis into the transform columns of every
call.  Here the member index is the outer batch dimension of every kernel of
the step (Legendre contractions, FFT rows, grid columns, spectral solves),
so E members advance in one launch sequence per step.  Extensions beyond
the reference: per-member basic states (``per_member_basic=True``) and
per-member integration lengths (``LBM.run(lengths=...)``).
"""

import numpy as np

from .model import LBM, Config
from .spectral import Spectral
from .vertical import VerticalGrid


class BatchedSpectral(Spectral):
    """A :class:`Spectral` whose fields carry a leading member axis
    (``ensemble.py:41-106``): grid ``(E, nlat, nlon[, k])``, spectral
    ``(E, m, n[, k])``."""

    batched = True


class EnsembleForcing:
    """Constant per-member temperature forcing ``(E, nlat, nlon, kmax)`` in K/s
    presented through ``.step()`` (``ensemble.py:109-121``)."""

    constant = True

    def __init__(self, gtf, xp=None):
        self.gtf = gtf

    def step(self):
        return self.gtf


def build_heating_sweep(base_heat, nmem, mode="longitude", xp=None):
    """``nmem`` heating fields from one base field (``ensemble.py:124-149``)."""
    base = np.asarray(base_heat)
    nlat, nlon, kmax = base.shape
    out = np.empty((nmem, nlat, nlon, kmax), dtype=np.float64)
    if mode == "longitude":
        for e in range(nmem):
            out[e] = np.roll(base, int(round(e * nlon / nmem)), axis=1)
    elif mode == "amplitude":
        for e, s in enumerate(np.linspace(0.5, 1.5, nmem)):
            out[e] = s * base
    else:
        raise ValueError(mode)
    return out


def make_ensemble_model(basic_state, nmem, config=None, device="hip", sigm=None,
                        per_member_basic=False, spectral=None):
    """E-member :class:`LBM` on one device; returns ``(model, spectral, vgrid)``
    (``ensemble.py:152-165``).  ``spectral`` (a BatchedSpectral) selects
    another resolution than T42 (e.g. ``BatchedSpectral(384, 192, 127)``)."""
    sp = spectral if spectral is not None else BatchedSpectral(device=device)
    vg = VerticalGrid(device=device) if sigm is None else VerticalGrid(sigm, device=device)
    cfg = config or Config()
    cfg.nmem = nmem
    model = LBM(basic_state, sp, vg, cfg, per_member_basic=per_member_basic)
    return model, sp, vg
