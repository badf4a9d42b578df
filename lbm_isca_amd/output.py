"""History output: GrADS flat binary and netCDF (``pylbm_isca/output.py``).

The GrADS writer emits one Fortran sequential record per level,
big-endian real*4, in the order

    psi(L) chi(L) u(L) v(L) w(L) t(L) z(L) p(1)   (141 records per time at L20)

on the model's north-first grid; the ``.ctl`` file says ``OPTIONS
SEQUENTIAL YREV BIG_ENDIAN``.  Output stays on the sigma levels, and ``z``
gets the surface-pressure term ``Z' + ps'[hPa] 100 / (rho0 g)`` with
rho0 = 1.225, g = 9.80 (``output.py:1-15, 60-71``).

The fields reach a writer as one dict per output interval from
``LBM.run``; the trapezoidal averaging behind them runs on the device.
For an ensemble the dict carries a leading member axis; ``member=`` picks
the member a writer records (``None``: a single-member model).
"""

import os

import numpy as np

from .backend import asnumpy
from .constants import GRAV_GT2GR, RHO00

VARS = ["psi", "chi", "u", "v", "w", "t", "z", "p"]
VAR_INFO = {
    "psi": ("stream function", "m**2/s"),
    "chi": ("velocity potential", "m**2/s"),
    "u": ("zonal wind", "m/s"),
    "v": ("meridional wind", "m/s"),
    "w": ("p-vertical velocity", "hPa/s"),
    "t": ("temperature", "K"),
    "z": ("geopotential height", "m"),
    "p": ("surface pressure", "hPa"),
}


def _host_fields(fields, member, apply_gt2gr_z):
    f = {k: np.asarray(asnumpy(v), dtype=np.float64) for k, v in fields.items()}
    if member is not None:
        f = {k: v[member] for k, v in f.items()}
    if apply_gt2gr_z:
        f["z"] = f["z"] + f["p"][..., None] * 1.0e2 / (RHO00 * GRAV_GT2GR)
    return f


def fortran_records(planes):
    """Bytes of one big-endian Fortran sequential record per 2-D plane of
    ``planes`` (n, nlat, nlon): a 4-byte length, the plane as real*4, the
    length again -- packed in one go through a structured dtype."""
    planes = np.asarray(planes)
    n, npts = planes.shape[0], int(np.prod(planes.shape[1:]))
    rec = np.empty(n, dtype=[("head", ">i4"), ("body", ">f4", (npts,)), ("tail", ">i4")])
    rec["head"] = rec["tail"] = 4 * npts
    rec["body"] = planes.reshape(n, npts)
    return rec.tobytes()


def history_planes(f, nlev):
    """The (n, nlat, nlon) planes of one output time in file order: each
    variable of ``VARS`` level by level, 2-D variables as one plane."""
    out = []
    for name in VARS:
        a = f[name]
        out.append(a[None] if a.ndim == 2 else np.moveaxis(a, -1, 0))
    planes = np.concatenate(out, axis=0)
    if planes.shape[0] != 7 * nlev + 1:
        raise ValueError(f"history fields give {planes.shape[0]} planes, expected {7 * nlev + 1}")
    return planes


def _rows(values, fmt, per=6):
    """Descriptor continuation lines: ``per`` formatted values a line."""
    return [" " + " ".join(fmt.format(x) for x in values[i:i + per])
            for i in range(0, len(values), per)]


def grads_descriptor(data_name, title, nlon, lat_deg, sig, ntime):
    """Text of the GrADS ``.ctl`` descriptor of a history file: sequential
    big-endian records, north-first latitudes declared with YREV (listed
    south to north), sigma levels, daily steps, the eight variables."""
    lat = np.asarray(lat_deg, dtype=np.float64)[::-1]
    sig = np.asarray(sig, dtype=np.float64)
    head = [("DSET", f"^{data_name}"), ("OPTIONS", "SEQUENTIAL YREV BIG_ENDIAN"),
            ("TITLE", title), ("UNDEF", "-999."),
            ("XDEF", f"{nlon} LINEAR 0. {360.0 / nlon:.6g}")]
    lines = [f"{key} {val}" for key, val in head]
    lines += [f"YDEF {lat.size} LEVELS"] + _rows(lat, "{:9.4f}")
    lines += [f"ZDEF {sig.size} LEVELS"] + _rows(sig, "{:.6g}")
    lines += [f"TDEF {max(ntime, 1)} LINEAR 01jan0001 1dy", f"VARS {len(VARS)}"]
    lines += [f"{v:4s} {0 if v == 'p' else sig.size:3d} 99 {VAR_INFO[v][0]} [{VAR_INFO[v][1]}]"
              for v in VARS]
    return "\n".join(lines + ["ENDVARS"]) + "\n"


class GradsWriter:
    """History as a GrADS flat binary file and its ``.ctl`` descriptor
    (the byte layout of ``output.py:38-102``): every ``add`` appends the
    7 L + 1 records of one output time, ``close`` writes the descriptor."""

    def __init__(self, path, spectral, vertical, ctl_path=None, apply_gt2gr_z=True,
                 title="PyLBM-Isca T42L20", member=None):
        self.path = path
        self.ctl_path = ctl_path or (os.path.splitext(path)[0] + ".ctl")
        self.sp, self.vg = spectral, vertical
        self.apply_gt2gr_z = apply_gt2gr_z
        self.title = title
        self.member = member
        self.ntime = 0
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        self.fh = open(path, "wb")

    def add(self, day, fields):
        f = _host_fields(fields, self.member, self.apply_gt2gr_z)
        self.fh.write(fortran_records(history_planes(f, int(self.vg.kmax))))
        self.ntime += 1

    def close(self):
        if not self.fh.closed:
            self.fh.close()
        self.write_ctl()

    def write_ctl(self):
        text = grads_descriptor(os.path.basename(self.path), self.title, int(self.sp.nlon),
                                np.asarray(asnumpy(self.sp.lat_deg)),
                                np.asarray(asnumpy(self.vg.sig)), self.ntime)
        with open(self.ctl_path, "w") as fh:
            fh.write(text)


class NetCDFWriter:
    """netCDF output: (time, sigma, lat, lon) float32 variables with the
    reference's names, attributes and coordinates (``output.py:104-152``).
    Written as netCDF-3 by ``scipy.io.netcdf_file`` (no netCDF library is
    needed); ``close()`` returns a dict of the arrays written."""

    def __init__(self, path, spectral, vertical, apply_gt2gr_z=True, member=None,
                 title="LBM T42L20 dry, python port"):
        self.path = path
        self.sp, self.vg = spectral, vertical
        self.apply_gt2gr_z = apply_gt2gr_z
        self.member = member
        self.title = title
        self.days = []
        self.data = {v: [] for v in VARS}

    def add(self, day, fields):
        f = _host_fields(fields, self.member, self.apply_gt2gr_z)
        self.days.append(float(day))
        for v in VARS:
            a = f[v]
            self.data[v].append(np.moveaxis(a, -1, 0).astype(np.float32) if a.ndim == 3
                                else a.astype(np.float32))

    def close(self):
        from scipy.io import netcdf_file
        lat = np.asarray(asnumpy(self.sp.lat_deg), dtype=np.float64)
        lon = np.asarray(asnumpy(self.sp.lon_deg), dtype=np.float64)
        sig = np.asarray(asnumpy(self.vg.sig), dtype=np.float64)
        out = {}
        with netcdf_file(self.path, "w", version=2) as nc:
            nc.title = self.title
            nc.note = ("latitudes are the Gaussian grid, north first; vertical coordinate "
                       "is sigma")
            for name, vals in (("time", np.array(self.days, dtype=np.float64)), ("sigma", sig),
                               ("lat", lat), ("lon", lon)):
                nc.createDimension(name, vals.size)
                var = nc.createVariable(name, "d", (name,))
                var[:] = vals
            for v in VARS:
                arr = np.array(self.data[v], dtype=np.float32)
                dims = ("time", "lat", "lon") if arr.ndim == 3 else ("time", "sigma", "lat", "lon")
                var = nc.createVariable(v, "f", dims)
                var[:] = arr
                desc, unit = VAR_INFO[v]
                var.long_name = desc
                var.units = unit
                out[v] = arr
        return out


class MultiWriter:
    """Hands every output time to each of several writers in turn
    (``output.py:155-165``); ``None`` entries are dropped."""

    def __init__(self, *writers):
        self.writers = [w for w in writers if w is not None]

    def add(self, day, fields):
        for w in self.writers:
            w.add(day, fields)

    def close(self):
        for w in self.writers:
            w.close()
