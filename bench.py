#!/usr/bin/env python3
"""T42L20 ensemble throughput on one MI355X.

Workload (default): 64 members on the solid-body-rotation basic state
(u = 20 m/s cos(lat), T from 288 K at 6.5 K/km to a 220 K floor,
ps = 1000 hPa, q = 0) on 20 equally spaced sigma levels, each member with
its own steady Gaussian heating (centre lat U(-20,20), lon U(0,360),
amplitude U(0.5,3) K/day, 10 x 30 deg e-folding, sin(pi sigma) profile,
seed 0), FP64, default T42 dt = 1200 s and damping.  One step advances all
members by one leapfrog step; ``--steps 2160`` is the 30-day run.

Prints one JSON line: member-days/s (full steps only; the two start-up
half steps of each run are executed inside the timed region but not
counted as days), the dominant kernel's achieved HBM bandwidth against the
8 TB/s peak (HIP events on the plan's stream, eager replay of the same
steps) and the numpy FP64 oracle timed on the host cores on a bounded
single-member sample of the same workload.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "T42L20 member-days/s, FP32 vs FP64 at <=1e-4 rel err; adjoint/forward time"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_model(members, precision):
    import lbm_isca_amd as L
    from lbm_isca_amd import cases
    sigm = cases.EQUAL_SIGM_L20
    sp = L.BatchedSpectral()
    vg = L.VerticalGrid(sigm)
    bs = cases.solid_body_state(sp.lat, sp.lon, vg.sig)
    frc = cases.random_heatings(sp.lat, sp.lon, vg.sig, members, seed=0)
    cfg = L.Config(precision=precision)
    m, _, _ = L.make_ensemble_model(bs, members, cfg, sigm=sigm)
    m.plan.set_forcing(frc)
    return m, bs, frc, vg


def cpu_baseline(bs, frc, sigm, seconds=15.0):
    """numpy FP64 oracle on one member of the same workload, bounded."""
    from oracle import OracleConfig, OracleModel
    m = OracleModel(bs, 1, OracleConfig(), sigm=sigm)
    t0 = time.time()
    m.step(1200.0 / 4, True, frc[:1])        # start-up step, warm caches
    n = 0
    t0 = time.time()
    while time.time() - t0 < seconds:
        m.step(1200.0, False, frc[:1])
        n += 1
    dt = time.time() - t0
    return {"value": n / 72.0 / dt, "unit": "member-days/s", "cores": os.cpu_count(),
            "kind": "port",
            "sample": f"numpy FP64 oracle, member 0 of the workload, {n} full steps "
                      f"({dt:.1f} s) after one start-up step, BLAS threads = host default"}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=720)
    ap.add_argument("--warmup", type=int, default=72)
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--precision", default="fp64", choices=("fp64", "fp32"))
    ap.add_argument("--profile-steps", type=int, default=24)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); members are "
              "independent, so several GPUs would each run their own share", file=sys.stderr)
        sys.exit(2)

    import torch
    from lbm_isca_amd import roofline
    if not torch.cuda.is_available():
        raise RuntimeError("bench.py needs a HIP device")
    E = args.members
    m, bs, frc, vg = build_model(E, args.precision)
    plan = m.plan
    log(f"warmup {args.warmup} steps ...")
    plan.run(args.warmup, use_graph=True)
    torch.cuda.synchronize()
    log(f"timed {args.steps} steps ...")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.run(args.steps, use_graph=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = E * args.steps / 72.0 / elapsed
    ms_step = elapsed * 1e3 / args.steps
    log(f"{value:.1f} member-days/s, {ms_step:.3f} ms/step")

    # kernel durations with HIP events on the plan's stream (eager launches)
    plan.profile(True)
    plan.run(args.profile_steps, use_graph=False)
    prof = plan.profile_read()
    plan.profile(False)
    s = 8 if args.precision == "fp64" else 4
    kb = roofline.kernel_bytes(E, L=vg.kmax, s=s, nfrc=E)
    avg = {k: ms / n for k, (ms, n) in prof.items() if n and ms > 0}
    dom = max(avg, key=lambda k: prof[k][0])
    achieved = kb[dom] / (avg[dom] * 1e-3) / 1e9
    log("kernel ms/launch: " + ", ".join(f"{k} {v:.3f}" for k, v in avg.items()))
    rl = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
          "peak": roofline.MI355X_HBM_GBS, "unit": "GB/s",
          "frac": round(achieved / roofline.MI355X_HBM_GBS, 4), "traffic": None,
          "algorithmic_bytes": kb[dom], "avg_ms": round(avg[dom], 5),
          "kernels_ms": {k: round(v, 5) for k, v in avg.items()}}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(bs, frc, vg.sigm, args.cpu_seconds)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "member-days/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (solid-body basic state, seeded random heatings)",
        "config": {"workload": f"T42L20 {E}-member heating ensemble, solid-body basic state, "
                               f"{args.precision.upper()}, 20 equal sigma levels",
                   "members": E, "resolution": "T42L20", "dt_s": 1200.0,
                   "graph": True, "startup_steps_in_timed_region": 2},
        "roofline": rl, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
