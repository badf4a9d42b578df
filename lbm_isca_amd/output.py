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


class GradsWriter:
    """Writes ``out.grd`` + ``out.ctl`` GrADS flat binary."""

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

    def _record(self, a2d):
        """One Fortran sequential record of a (nlat, nlon) real*4 field."""
        data = np.ascontiguousarray(a2d, dtype=">f4")
        mark = np.array([data.size * 4], dtype=">i4").tobytes()
        self.fh.write(mark + data.tobytes() + mark)

    def add(self, day, fields):
        f = _host_fields(fields, self.member, self.apply_gt2gr_z)
        for name in VARS:
            a = f[name]
            if a.ndim == 2:
                self._record(a)
            else:
                for k in range(a.shape[-1]):
                    self._record(a[..., k])
        self.ntime += 1

    def close(self):
        if not self.fh.closed:
            self.fh.close()
        self.write_ctl()

    def write_ctl(self):
        lat = np.asarray(asnumpy(self.sp.lat_deg))[::-1]    # listed south -> north with YREV
        sig = np.asarray(asnumpy(self.vg.sig))
        nlon, kmax = int(self.sp.nlon), int(self.vg.kmax)
        lines = [f"DSET ^{os.path.basename(self.path)}",
                 "OPTIONS SEQUENTIAL YREV BIG_ENDIAN",
                 f"TITLE {self.title}",
                 "UNDEF -999.",
                 f"XDEF {nlon} LINEAR 0. {360.0 / nlon:.6g}",
                 f"YDEF {lat.size} LEVELS"]
        lines += [" " + " ".join(f"{x:9.4f}" for x in lat[i:i + 6]) for i in range(0, lat.size, 6)]
        lines.append(f"ZDEF {kmax} LEVELS")
        lines += [" " + " ".join(f"{x:.6g}" for x in sig[i:i + 6]) for i in range(0, kmax, 6)]
        lines.append(f"TDEF {max(self.ntime, 1)} LINEAR 01jan0001 1dy")
        lines.append(f"VARS {len(VARS)}")
        for name in VARS:
            desc, unit = VAR_INFO[name]
            lines.append(f"{name:4s} {0 if name == 'p' else kmax:3d} 99 {desc} [{unit}]")
        lines.append("ENDVARS")
        with open(self.ctl_path, "w") as fh:
            fh.write("\n".join(lines) + "\n")


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
    """Fan a history stream out to several writers."""

    def __init__(self, *writers):
        self.writers = [w for w in writers if w is not None]

    def add(self, day, fields):
        for w in self.writers:
            w.add(day, fields)

    def close(self):
        for w in self.writers:
            w.close()
