"""Forcing containers and readers with the reference's protocol
(``pylbm_isca/forcing.py:1-174``).

A temperature forcing object needs ``.step()`` returning a grid array in
K/s, ``(J, I, L)`` or ``(E, J, I, L)`` (``model.py:222-229``).  Objects
whose ``.step()`` always returns the same array (``SteadyForcing.t``,
``EnsembleForcing``, a constant forcing) are uploaded once and stay in HBM
for the whole run.  An :class:`MJOForcing` is uploaded once as its daily
fields; the per-step interpolation runs on the device inside the step
graph (``lbm_set_mjo_forcing``), so no forcing crosses PCIe during a run.
"""

import os

import numpy as np

from .constants import IMAX, JMAX, KMAX

_REC = ">f4"


def read_fortran_records(path, nlon=IMAX, nlat=JMAX, count=None, offset=0):
    """Fortran sequential records of ``nlon * nlat`` big-endian real*4 values
    (``forcing.py:45-65``): returns ``(nrec, nlat, nlon)`` float32 after
    checking every record marker."""
    npts = nlon * nlat
    raw = np.fromfile(path, dtype=_REC)
    per = npts + 2
    if raw.size % per:
        raise ValueError(f"{path}: size {raw.size * 4} is not a whole number of "
                         f"{npts}-value Fortran records")
    rec = raw.reshape(-1, per)
    marks = rec[:, [0, -1]].copy().view(">i4")
    bad = np.nonzero(np.any(marks != npts * 4, axis=1))[0]
    if bad.size:
        b = int(bad[0])
        raise ValueError(f"{path}: bad record marker at record {b} ({int(marks[b, 0])} != "
                         f"{npts * 4}); wrong endianness or record length?")
    stop = None if count is None else offset + count
    return rec[offset:stop, 1:-1].reshape(-1, nlat, nlon)


def _orient(a, yflip):
    return a[:, ::-1, :] if yflip else a


class SteadyForcing:
    """vorticity / divergence / temperature / ln(Ps) forcing, time constant."""

    def __init__(self, vor, div, t, ps):
        self.vor, self.div, self.t, self.ps = vor, div, t, ps

    @classmethod
    def zeros(cls, nlat=JMAX, nlon=IMAX, nlev=KMAX):
        z3 = np.zeros((nlat, nlon, nlev))
        return cls(z3.copy(), z3.copy(), z3.copy(), np.zeros((nlat, nlon)))

    @classmethod
    def from_file(cls, path, nlev=KMAX, ncase=1, yflip=False, nlon=IMAX, nlat=JMAX):
        """Case ``ncase`` of a steady-forcing file: ``3 nlev + 1`` records
        (vorticity, divergence, temperature levels, then ln Ps), ``forcing.py:78-96``."""
        per = 3 * nlev + 1
        rec = read_fortran_records(path, nlon, nlat, count=per, offset=(ncase - 1) * per)
        if rec.shape[0] != per:
            raise ValueError(f"{path}: expected {per} records, got {rec.shape[0]}")
        rec = _orient(rec, yflip)
        lev = lambda k: np.moveaxis(rec[k * nlev:(k + 1) * nlev], 0, -1).copy()  # noqa: E731
        return cls(lev(0), lev(1), lev(2), rec[3 * nlev].copy())


class ConstForcing:
    """A constant temperature forcing presented through ``.step()``."""

    constant = True

    def __init__(self, gtf, xp=None):
        self.gtf = gtf

    def step(self):
        return self.gtf


class MJOForcing:
    """Time-varying temperature forcing read from a file of daily fields
    (``forcing.py:99-161``).

    The dynamics calls are counted from 1.  Calls before ``spinup`` return
    day 1; from call ``spinup`` on, the field moves from day ``it`` to day
    ``it + 1`` over ``nt`` calls, ``gtf[it] + (gtf[it+1] - gtf[it]) / nt * k``
    with ``k = 0 .. nt-1``, both days clamped to the last one.  ``step()``
    gives the same sequence on the host; ``LBM.run`` uploads the daily fields
    once and the device computes each call's field inside the step graph,
    continuing from this object's call counter.
    """

    def __init__(self, path=None, nlev=KMAX, ndays=73, nt=72, spinup=435, yflip=False,
                 nlon=IMAX, nlat=JMAX, scale=1.0, gtf=None):
        if gtf is None:
            rec = read_fortran_records(path, nlon, nlat)
            if rec.shape[0] != ndays * nlev:
                raise ValueError(f"{path}: expected {ndays * nlev} records ({ndays} days x {nlev} "
                                 f"levels), got {rec.shape[0]}.  Check ndays/nlev.")
            # real*4 stays float32, scale included (numpy keeps float32 for a
            # Python float factor), as forcing.py:131 leaves it
            rec = _orient(rec, yflip).astype(np.float32) * scale
            # (day, lev, lat, lon) -> (day, lat, lon, lev)
            gtf = np.moveaxis(rec.reshape(ndays, nlev, nlat, nlon), 1, -1)
        self.gtf = np.asarray(gtf)
        if self.gtf.dtype not in (np.float32, np.float64):
            self.gtf = self.gtf.astype(np.float64)
        self.ndays = int(self.gtf.shape[0])
        self.nt, self.spinup = int(nt), int(spinup)
        self.path = path
        self.reset()

    def reset(self):
        self.counter = 0

    def field_at(self, call):
        """The field of dynamics call ``call`` (1-based)."""
        if call < self.spinup:
            return self.gtf[0]
        # the slope in the fields' precision, a + ka k in float64: numpy's
        # promotion of ka times the int64 arange of forcing.py:151-153
        q = call - self.spinup
        it = 1 + q // self.nt
        k = q % self.nt
        a = self.gtf[min(it, self.ndays) - 1]
        b = self.gtf[min(it + 1, self.ndays) - 1]
        ka = (b - a) / self.nt
        return a + ka * np.float64(k)

    def step(self):
        """Advance one dynamics call and return its field."""
        self.counter += 1
        return self.field_at(self.counter)

    @property
    def day(self):
        """1-based day index after the last call (forcing.py:158-160): it
        moves on once the nt calls of a day are done."""
        if self.counter < self.spinup:
            return 1
        return 1 + (self.counter - self.spinup + 1) // self.nt


def autodetect(path, nlev=KMAX, nlon=IMAX, nlat=JMAX):
    """Which forcing layouts a file's record count fits (``forcing.py:163-174``)."""
    npts = nlon * nlat
    nrec = os.path.getsize(path) // (npts * 4 + 8)
    per = 3 * nlev + 1
    info = {"nrec": nrec}
    if nrec % per == 0:
        info["steady_cases"] = nrec // per
    if nrec % nlev == 0:
        info["mjo_days"] = nrec // nlev
    return info
