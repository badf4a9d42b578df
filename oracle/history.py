"""Oracle: history diagnostics, MJO forcing and spectral regridding (numpy FP64).
TEST INFRASTRUCTURE -- see ``oracle/__init__.py``.

Restates ``pylbm_isca/model.py:150-166,189-277`` (top-of-step diagnostics
and the trapezoidal history), ``pylbm_isca/forcing.py:99-161`` (the MJO
forcing's call sequence, as the reference's state machine of counters) and
``pylbm_isca/regrid.py:66-124`` (spectral resize and regridding).
"""

import numpy as np

GRAV = 9.80                     # constants.py:39
RHO00, GRAV_GT2GR = 1.225, 9.80  # constants.py:63-64


def diagnostics(model):
    """The eight history fields of ``model.py:150-166`` for an OracleModel:
    psi, chi, w from the last tendencies, u, v, t, z, p from ga."""
    sht, vg, d, ga = model.sht, model.vg, model.last_diag, model.ga
    psi = sht.wave_to_grid(sht.inverse_laplacian(d["vor_s"]))
    chi = sht.wave_to_grid(sht.inverse_laplacian(d["div_s"]))
    z = vg.hydrostatic(ga.t) / GRAV
    return {"psi": psi, "chi": chi, "u": ga.u.copy(), "v": ga.v.copy(), "w": d["omega"],
            "t": ga.t.copy(), "z": z, "p": ga.ps.copy()}


class History:
    """The writer branch of ``model.py:255-277``: each full step's fields are
    0.5 (this step's top sample + the previous one); with ``time_average``
    they are summed over ``out_every`` full steps and divided by the count,
    otherwise the interval's last trapezoid is written.  ``records`` holds
    ``(day, fields)``."""

    def __init__(self, out_every, time_average=True):
        self.out_every, self.time_average = int(out_every), bool(time_average)
        self.prev = None
        self.acc, self.nacc, self.day = None, 0, 0
        self.records = []

    def start_step(self, top):
        """Top sample of a start-up step: only remembered."""
        self.prev = top

    def full_step(self, advanced, top):
        if self.prev is None:
            self.prev = top
        fields = {k: 0.5 * (top[k] + self.prev[k]) for k in top}
        if self.time_average:
            if self.acc is None:
                self.acc = {k: np.zeros_like(v) for k, v in fields.items()}
            for k, v in fields.items():
                self.acc[k] += v
            self.nacc += 1
        if advanced % self.out_every == 0:
            self.day += 1
            if self.time_average:
                self.records.append((self.day, {k: v / self.nacc for k, v in self.acc.items()}))
                self.acc, self.nacc = None, 0
            else:
                self.records.append((self.day, {k: v.copy() for k, v in fields.items()}))
        self.prev = top


class MJOSequence:
    """The reference MJO forcing's call sequence as its own state machine
    (``forcing.py:115-156``): counters, the spin-up hold of day 1, then a
    table of nt interpolated fields per day rebuilt when a day starts."""

    def __init__(self, gtf, nt, spinup):
        self.gtf, self.nt, self.spinup = gtf, int(nt), int(spinup)
        self.ndays = gtf.shape[0]
        self.counter, self.counter2, self.it = 0, 0, 1
        self.spinning, self.work = True, None

    def step(self):
        self.counter += 1
        if self.spinning:
            if self.counter < self.spinup:
                return self.gtf[0]
            self.counter2, self.it, self.spinning = 0, 1, False
        if self.counter2 == 0:
            a = self.gtf[min(self.it, self.ndays) - 1]
            b = self.gtf[min(self.it + 1, self.ndays) - 1]
            slope = (b - a) / self.nt
            # numpy promotes slope * int64 arange to float64 (float32 fields)
            self.work = [a + slope * np.int64(k) for k in range(self.nt)]
        out = self.work[self.counter2]
        self.counter2 += 1
        if self.counter2 == self.nt:
            self.counter2, self.it = 0, self.it + 1
        return out


def resize_spec(xi, dst_nmax):
    """Zero-pad or truncate spectral ``[.., m, n, k]`` (``regrid.py:66-77``);
    the (m, n) axes are the two before the last."""
    ms = xi.shape[-3]
    k = min(ms, dst_nmax + 1)
    out = np.zeros(xi.shape[:-3] + (dst_nmax + 1, dst_nmax + 1) + xi.shape[-1:], dtype=xi.dtype)
    out[..., :k, :k, :] = xi[..., :k, :k, :]
    return out


def regrid_scalar(src, dst, field):
    """``regrid.py:84-96`` with oracle SHTs; field (E, J, I, K)."""
    return dst.wave_to_grid(resize_spec(src.grid_to_wave(field), dst.nmax))


def regrid_uv(src, dst, u, v):
    """``regrid.py:99-110``: through vorticity and divergence."""
    vor, div = src.uv_to_vordiv(u, v)
    return dst.vordiv_to_uv(resize_spec(vor, dst.nmax), resize_spec(div, dst.nmax))
