"""Write the FP64 oracle's results on the benchmark workloads as small
fixtures (``tests/golden/oracle_<part>.npz``) for the GPU parity tests,
whose oracle runs would take minutes on the GPU box's host.

Parts (``python tests/golden/make_oracle_runs.py <part>``):

* ``cfg0``  -- one member on the solid-body basic state, 20 equal sigma
  levels, 2 K/day Gaussian heating at (0N, 180E), 10 days (720 steps);
* ``cfg1``  -- the first 4 members of the 64-member heating ensemble (seed
  0) on the same state, 10 days;
* ``month`` -- member 0 of that ensemble for 30 days (2160 steps);
* ``adj``   -- the tangent-linear response of two balanced-jet members
  (seed 1, their own basic states) to a seeded spectral initial state and
  their heatings: J = area mean of the sigma = 0.2 streamfunction over
  30-60N, 150-210E, averaged over days 20-30 of a 30-day run, for the
  state/forcing pair and for a second seeded direction.

Grid fields are stored as 1024 seeded samples per field and member plus the
field's full L2 norm.  Test infrastructure: runs the oracle, never the
product.
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, HERE, os.path.dirname(HERE)]

import cases as gcases  # noqa: E402  (tests/golden/cases.py: seeded spectra)
from lbm_isca_amd import adjoint, basicstate, cases  # noqa: E402  (host functions only)
from oracle import OracleConfig, OracleModel, SHT, VerticalOps  # noqa: E402
from oracle_helpers import oracle_response  # noqa: E402

FIELDS = ("u", "v", "t", "q", "ps")
NSAMP = 1024


def samples(ga, tag, out, rng):
    """Seeded samples and norms of every member's grid fields."""
    for f in FIELDS:
        arr = getattr(ga, f)
        per = arr[0].size
        idx = rng.choice(per, NSAMP, replace=False).astype(np.int64)
        out[f"{tag}_{f}_idx"] = idx
        out[f"{tag}_{f}_val"] = np.stack([arr[e].ravel()[idx] for e in range(arr.shape[0])])
        out[f"{tag}_{f}_norm"] = np.sqrt((arr.reshape(arr.shape[0], -1) ** 2).sum(axis=1))


def solid_body():
    sht, vg = SHT(), VerticalOps(cases.EQUAL_SIGM_L20)
    bs = cases.solid_body_state(sht.lat, sht.lon, vg.sig)
    return sht, vg, bs


def part_cfg0():
    sht, vg, bs = solid_body()
    heat = cases.heating(sht.lat, sht.lon, vg.sig, 0.0, 180.0, 2.0)[None]
    m = OracleModel(bs, 1, OracleConfig(), sigm=cases.EQUAL_SIGM_L20)
    m.run(nsteps=720, frc_t=heat)
    out = {"nsteps": np.int64(720)}
    samples(m.ga, "cfg0", out, np.random.default_rng(100))
    return out


def part_cfg1(nmem=4, nsteps=720, tag="cfg1", members=None):
    sht, vg, bs = solid_body()
    frc = cases.random_heatings(sht.lat, sht.lon, vg.sig, 64, seed=0)
    frc = frc[:nmem] if members is None else frc[members]
    m = OracleModel(bs, frc.shape[0], OracleConfig(), sigm=cases.EQUAL_SIGM_L20)
    m.run(nsteps=nsteps, frc_t=frc)
    out = {"nsteps": np.int64(nsteps)}
    samples(m.ga, tag, out, np.random.default_rng(101))
    return out


def part_month():
    return part_cfg1(nsteps=2160, tag="month", members=[0])


class OracleTruncate:
    """The basic-state pipeline's spectral smoothing through the oracle's
    transforms (grid -> wave -> grid), for fields with a member axis."""

    def __init__(self, sht):
        self.sht = sht
        self.lat, self.lon = sht.lat, sht.lon
        self.lat_deg, self.lon_deg = np.degrees(sht.lat), np.degrees(sht.lon)
        self.nlat, self.nlon = sht.nlat, sht.nlon

    def truncate(self, f):
        f = np.asarray(f, dtype=np.float64)
        surf = f.ndim == 3
        g = f[..., None] if surf else f
        out = self.sht.wave_to_grid(self.sht.grid_to_wave(g))
        return out[..., 0] if surf else out


def adj_inputs(sht, vg, nmem=2):
    """Seeded spectral initial state and direction (vor/div without n = 0)."""
    rng = np.random.default_rng(33)
    spec, dirn = {}, {}
    for d, scale in ((spec, 1.0), (dirn, 0.7)):
        for f, amp, nmin in (("vor", 2e-6, 1), ("div", 5e-7, 1), ("t", 0.5, 0), ("q", 1e-5, 0)):
            d[f] = np.stack([gcases.random_spectral(rng, 42, vg.kmax, scale * amp, nmin)
                             for _ in range(nmem)])
        d["pi"] = np.stack([gcases.random_spectral(rng, 42, 1, scale * 1e-4)[..., 0]
                            for _ in range(nmem)])
    return spec, dirn


def part_adj(nmem=2, days=30.0, window=(20.0, 30.0)):
    sig = cases.EQUAL_SIGM_L20
    sht, vg = SHT(), VerticalOps(sig)
    fields, _ = basicstate.balanced_jets(sht.lat, sht.lon, nmem, seed=1)
    bs = basicstate.sigma_basic_state(fields, OracleTruncate(sht), vg.sig, source="balanced jets")
    frc = cases.random_heatings(sht.lat, sht.lon, vg.sig, nmem, seed=0)
    nsteps = int(round(72 * days))
    cfg = OracleConfig()
    wgrid = adjoint.box_weights(np.degrees(sht.lat), np.degrees(sht.lon), sht.gw,
                                (30.0, 60.0), (150.0, 210.0))
    wlev = adjoint.sigma_weights(vg.sig, 0.2)
    wstep = adjoint.step_weights(nsteps, cfg.inistp, 72, window[0], window[1])
    spec, dirn = adj_inputs(sht, vg, nmem)
    dfrc = 0.5 * np.roll(frc, 17, axis=2)
    out = {"nsteps": np.int64(nsteps), "wgrid": wgrid, "wlev": wlev, "wstep": wstep}
    for tag, s, f in (("J", spec, frc), ("Jdir", dirn, dfrc)):
        vals = []
        for e in range(nmem):
            be = type(bs)(**{k: (getattr(bs, k)[e] if k in ("u", "v", "t", "q", "ps")
                                 else getattr(bs, k)) for k in vars(bs)})
            se = {k: v[e:e + 1] for k, v in s.items()}
            vals.append(oracle_response(be, 1, se, f[e:e + 1], nsteps, wgrid, wlev, wstep,
                                        sigm=sig)[0])
        out[tag] = np.array(vals)
    return out


PARTS = {"cfg0": part_cfg0, "cfg1": part_cfg1, "month": part_month, "adj": part_adj}


def main():
    parts = sys.argv[1:] or list(PARTS)
    for p in parts:
        t0 = time.time()
        out = PARTS[p]()
        path = os.path.join(HERE, f"oracle_{p}.npz")
        np.savez_compressed(path, **out)
        print(f"{p}: {time.time() - t0:.0f} s -> {path} ({os.path.getsize(path)} bytes)", flush=True)


if __name__ == "__main__":
    main()
