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
    """Planes of a file of Fortran sequential records, each record one
    ``nlat x nlon`` big-endian real*4 field between two 4-byte length words
    (the layout ``forcing.py:45-65`` reads).  Returns ``(nrec, nlat, nlon)``
    float32 from record ``offset`` on (``count`` of them, or all).  Every
    length word is checked, so a file of another grid or byte order raises
    ValueError instead of being misread."""
    words = nlon * nlat + 2                      # length, field, length
    flat = np.fromfile(path, dtype=_REC)
    nrec, extra = divmod(flat.size, words)
    if extra:
        raise ValueError(f"{path}: {flat.size * 4} bytes do not split into records of "
                         f"{nlon}x{nlat} real*4 values")
    table = flat.reshape(nrec, words)
    lengths = table[:, [0, -1]].copy().view(">i4")
    wrong = np.argwhere(lengths != 4 * (words - 2))
    if wrong.size:
        r = int(wrong[0, 0])
        raise ValueError(f"{path}: record {r} has length word {int(lengths[r, wrong[0, 1]])}, "
                         f"expected {4 * (words - 2)} (another grid, or little-endian data?)")
    sel = table[offset:] if count is None else table[offset:offset + count]
    return sel[:, 1:-1].reshape(-1, nlat, nlon)


def _orient(a, yflip):
    return a[:, ::-1, :] if yflip else a


class SteadyForcing:
    """Time-constant forcing of four prognostic variables (the
    ``SteadyForcing(vor, div, t, ps)`` container of ``forcing.py:71-96``):
    vorticity and divergence (J, I, L) in s^-2, temperature (J, I, L) in
    K/s and ln ps (J, I) in s^-1."""

    FIELDS = ("vor", "div", "t", "ps")

    def __init__(self, vor, div, t, ps):
        self.vor, self.div, self.t, self.ps = vor, div, t, ps

    @classmethod
    def zeros(cls, nlat=JMAX, nlon=IMAX, nlev=KMAX):
        """No forcing: three zero 3-D fields and a zero surface field."""
        shapes = [(nlat, nlon, nlev)] * 3 + [(nlat, nlon)]
        return cls(*(np.zeros(shp) for shp in shapes))

    @classmethod
    def from_file(cls, path, nlev=KMAX, ncase=1, yflip=False, nlon=IMAX, nlat=JMAX):
        """Case ``ncase`` (1-based) of a steady-forcing file, whose cases are
        consecutive blocks of ``3 nlev + 1`` records: the vorticity, then
        divergence, then temperature levels, then ln ps
        (``forcing.py:78-96``)."""
        block = 3 * nlev + 1
        planes = read_fortran_records(path, nlon, nlat, count=block, offset=(ncase - 1) * block)
        if planes.shape[0] != block:
            raise ValueError(f"{path}: case {ncase} needs {block} records, the file has "
                             f"{planes.shape[0]} from there")
        planes = _orient(planes, yflip)
        cube = np.moveaxis(planes[:3 * nlev].reshape(3, nlev, nlat, nlon), 1, -1)
        return cls(cube[0].copy(), cube[1].copy(), cube[2].copy(), planes[3 * nlev].copy())


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
                raise ValueError(f"{path}: {rec.shape[0]} records do not hold {ndays} days of "
                                 f"{nlev} levels; pass the file's ndays and nlev")
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
