"""Forcing containers with the reference's protocol (``pylbm_isca/forcing.py:71-96``).

A temperature forcing object needs ``.step()`` returning a grid array in
K/s, ``(J, I, L)`` or ``(E, J, I, L)`` (``model.py:222-229``).  Objects
whose ``.step()`` always returns the same array (``SteadyForcing.t``,
``EnsembleForcing``, a constant forcing) are uploaded once and stay in HBM
for the whole run.
"""

import numpy as np

from .constants import IMAX, JMAX, KMAX


class SteadyForcing:
    """vorticity / divergence / temperature / ln(Ps) forcing, time constant."""

    def __init__(self, vor, div, t, ps):
        self.vor, self.div, self.t, self.ps = vor, div, t, ps

    @classmethod
    def zeros(cls, nlat=JMAX, nlon=IMAX, nlev=KMAX):
        z3 = np.zeros((nlat, nlon, nlev))
        return cls(z3.copy(), z3.copy(), z3.copy(), np.zeros((nlat, nlon)))


class ConstForcing:
    """A constant temperature forcing presented through ``.step()``."""

    constant = True

    def __init__(self, gtf, xp=None):
        self.gtf = gtf

    def step(self):
        return self.gtf
