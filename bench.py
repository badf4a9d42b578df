#!/usr/bin/env python3
"""T42L20 ensemble throughput on one MI355X.

Workload (default): 64 members on the solid-body-rotation basic state
(u = 20 m/s cos(lat), T from 288 K at 6.5 K/km to a 220 K floor,
ps = 1000 hPa, q = 0) on 20 equally spaced sigma levels, each member with
its own steady Gaussian heating (centre lat U(-20,20), lon U(0,360),
amplitude U(0.5,3) K/day, 10 x 30 deg e-folding, sin(pi sigma) profile,
seed 0), FP64, default T42 dt = 1200 s and damping.  One step advances all
members by one leapfrog step; ``--steps 2160`` is the 30-day run.

Prints one JSON line: member-days/s over exactly --steps full steps (the
run's two start-up half steps execute before the timed region), the
dominant kernel's achieved HBM bandwidth against the 8 TB/s peak (HIP events
on the plan's stream, eager replay of the same steps) and the numpy FP64
oracle timed on the host cores the reference's way (all threads, one
thread, and a node of concurrent single-thread processes) on the
single-member solid-body workload.

Two more parts of the metric ride along in the same line:

* ``fp32_vs_fp64``: the uneven ensemble -- 256 members, each with its own
  jet basic state and heating and a length of U(5, 60) days -- run in FP32
  and in FP64 (one hipGraph replay each); member-days/s of both, the
  speed-up, and the relative L2 error of the FP32 final state against the
  FP64 one per field (the FP64 path matches the FP64 oracle to 1e-9).
* ``adjoint``: 32 of those members in the tangent-linear model with the
  response J = area mean of the sigma = 0.2 streamfunction over 30-60N,
  150-210E averaged over days 20-30; wall time of the adjoint sweep (both
  gradients, all members) over the forward run, FP64 and FP32.
* ``t127l40``: 32 balanced-jet members at T127 L40 (384 x 192 grid, 40
  equal sigma levels) from a seeded n^-1.5 vorticity perturbation of rms
  1e-6 1/s, 20 days, with the Legendre contractions in FP64, FP32 and
  split-BF16x3: member-days/s, the relative L2 error of each low-precision
  final state against FP64, and the Legendre kernels' matrix-core
  utilisation (useful flops / time / the precision's dense peak).
"""

import argparse
import json
import os
import sys
import time
import types

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "T42L20 member-days/s, FP32 vs FP64 at <=1e-4 rel err; adjoint/forward time"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# kernel of each profiling slot (FP64 names; FP32 replaces double by float)
# kernels of each profiling slot of one step, with their launches per step
SLOT_KERNEL = {"syn_level": {"k_syn_mfma<double, lbm::SynLevel": 1},
               "syn_surf": {"k_syn_mfma<double, lbm::SynSurf": 1},
               "ifft": {"k_ifft<double, 128": 1}, "tend": {"k_tend<double": 1, "k_grid<double": 1},
               "fft": {"k_fft<double, 128": 1},
               "ana_level": {"k_ana_mfma<double, lbm::AnaTendVD": 1,
                             "k_ana_mfma<double, lbm::AnaFluxDiv": 2},
               "ana_surf": {"k_ana_mfma<double, lbm::AnaScalar": 1, "k_hist_acc<double>": 1},
               "step": {"k_step<double, 20": 1}}


def measured_traffic(slot, E, precision, grid_choice=0):
    """HBM bytes of a slot's kernels per step from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (of each kernel name,
    the variant with the most launches; the tendency kernel as the run chose
    it: separate (grid_choice 0), with the forward FFT inside (1, 2) or the
    fused grid kernel with 2 or 4 (FP64), 4 or 8 (FP32) members per block
    (3, 4)), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_T42L20_E{E}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        table = json.load(fh)["kernels"]
    total = 0
    for name, launches in SLOT_KERNEL[slot].items():
        tail = None
        if slot == "tend":
            if (grid_choice >= 3) != ("k_grid" in name):
                continue
            if grid_choice >= 3:
                # k_grid<T, 128, members per block, one basic state>
                mb = (2 if precision == "fp64" else 4) << (grid_choice - 3)
                tail = f", {mb}, true>"
            else:
                name += ", true" if grid_choice else ", false"
        if precision == "fp32":
            name = name.replace("double", "float")
        hits = [v for k, v in table.items() if name in k and (tail is None or k.endswith(tail))]
        if not hits:
            return None
        best = max(hits, key=lambda v: v["launches"])
        total += launches * int(best["bytes_per_launch"])
    return total


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


def _oracle_member_days(steps):
    """member-days/s of the numpy FP64 oracle on the single-member
    solid-body workload (2 K/day at 0N 180E, 20 equal sigma levels): the two
    start-up half steps warm the caches, then ``steps`` full steps are timed
    (analysis/benchmark_perf_isca.py:78-106 times a warmed run the same way)."""
    from lbm_isca_amd import cases
    from oracle import OracleConfig, OracleModel, SHT, VerticalOps
    sigm = cases.EQUAL_SIGM_L20
    sht, vg = SHT(), VerticalOps(sigm)
    bs = cases.solid_body_state(sht.lat, sht.lon, vg.sig)
    heat = cases.heating(sht.lat, sht.lon, vg.sig, 0.0, 180.0, 2.0)[None]
    m = OracleModel(bs, 1, OracleConfig(), sigm=sigm)
    m.step(300.0, True, heat)
    m.step(600.0, True, heat)
    t0 = time.perf_counter()
    for _ in range(steps):
        m.step(1200.0, False, heat)
    dt = time.perf_counter() - t0
    return steps / 72.0 / dt


def cpu_worker(steps):
    """One single-thread oracle run (a child process of the node figure)."""
    print(f"member_days_per_s={_oracle_member_days(steps):.9g}", flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """CPU cores this process may use, at most 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(steps=16, node_steps=12):
    """The numpy FP64 oracle timed on the host the way the reference times its
    CPU path: a warmed single-member run with BLAS/FFT threads = the cores
    and = 1 (analysis/benchmark_perf_isca.py:78-106), and the fully loaded
    node -- N concurrent single-thread processes, each one member
    (analysis/benchmark_ensemble_isca.py:225-244).  ``value`` is the node's
    aggregate member-days/s; ``cores`` the processes used for it."""
    import subprocess
    from threadpoolctl import threadpool_limits
    ncores = host_cores()
    with threadpool_limits(ncores):
        all_thr = _oracle_member_days(steps)
    with threadpool_limits(1):
        one_thr = _oracle_member_days(steps)
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1",
               NUMEXPR_NUM_THREADS="1")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", str(node_steps)]
    t0 = time.perf_counter()
    procs = [subprocess.Popen(cmd, env=env, stdout=subprocess.PIPE, text=True)
             for _ in range(ncores)]
    outs = [p.communicate()[0] for p in procs]
    wall = time.perf_counter() - t0
    per = [float(line.split("=")[1]) for o in outs for line in o.splitlines()
           if line.startswith("member_days_per_s=")]
    if len(per) != ncores:
        raise RuntimeError(f"{ncores - len(per)} CPU worker processes failed")
    node = float(sum(per))
    return {"value": round(node, 4), "unit": "member-days/s", "cores": ncores, "kind": "port",
            "cpu_model": cpu_model(),
            "threads_all": {"threads": ncores, "member_days_per_s": round(all_thr, 4)},
            "threads_1": {"threads": 1, "member_days_per_s": round(one_thr, 4)},
            "node": {"processes": ncores, "member_days_per_s": round(node, 4),
                     "per_process_median": round(float(np.median(per)), 4),
                     "wall_s": round(wall, 1)},
            "sample": f"numpy FP64 oracle, single-member solid-body workload (2 K/day at 0N "
                      f"180E, 20 equal sigma levels), two start-up steps then {steps} timed full "
                      f"steps with BLAS threads = {ncores} and = 1; node = {ncores} concurrent "
                      f"single-thread processes of {node_steps} timed steps each (value)"}


def uneven_ensemble(E, seed=1, max_days=None):
    """Per-member jet basic states, heatings and lengths of the uneven
    ensemble (E, J, I, L) on 20 equal sigma levels, members ordered by
    decreasing length (the plan skips finished members in blocks, so the
    finished ones should be neighbours); ``max_days`` caps the lengths
    (profiling runs only)."""
    import numpy as np
    import lbm_isca_amd as L
    from lbm_isca_amd import cases
    sigm = cases.EQUAL_SIGM_L20
    sp = L.BatchedSpectral()
    vg = L.VerticalGrid(sigm)
    bs, _ = cases.jet_states(sp, vg.sig, E, seed=seed)
    frc = cases.random_heatings(sp.lat, sp.lon, vg.sig, E, seed=0)
    lengths = cases.member_lengths(E, seed=seed)
    if max_days is not None:
        lengths = np.minimum(lengths, int(round(72 * max_days)))
    order = np.argsort(-lengths, kind="stable")
    for k in ("u", "v", "t", "q", "ps"):
        setattr(bs, k, np.ascontiguousarray(getattr(bs, k)[order]))
    return bs, np.ascontiguousarray(frc[order]), lengths[order], sigm


def measure_fp32_vs_fp64(E=256, max_days=None):
    """FP32 and FP64 runs of the uneven ensemble: throughput and the FP32
    error against FP64."""
    import numpy as np
    import torch
    import lbm_isca_amd as L
    from lbm_isca_amd import tuning
    bs, frc, lengths, sigm = uneven_ensemble(E, max_days=max_days)
    out, states = {}, {}
    days = float(lengths.sum()) / 72.0

    def timed_run(m):
        m.plan.run(8, use_graph=True)               # capture + warm caches
        m.plan.set_state()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.plan.run(0, use_graph=True)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    for prec in ("fp64", "fp32"):
        cfg = L.Config(precision=prec)
        m, _, _ = L.make_ensemble_model(bs, E, cfg, sigm=sigm, per_member_basic=True)
        m.plan.set_forcing(frc)
        m.plan.set_lengths(lengths)
        # built-in launch variants, then the autotuned ones (Legendre tiles'
        # register budget and FFT rows per block timed on this ensemble)
        dt_builtin = timed_run(m)
        tuned = tuning.autotune(m.plan, True)
        m.plan.set_forcing(frc)
        dt = timed_run(m)
        out[prec] = {"member_days_per_s": round(days / dt, 2), "seconds": round(dt, 3),
                     "builtin_member_days_per_s": round(days / dt_builtin, 2),
                     "autotuned": dict(zip(m.plan.TUNED, tuned["choice"])),
                     "autotune_report": tuned["report"]}
        states[prec] = [x.cpu().numpy() for x in m.plan.get_state()]
        log(f"uneven ensemble {prec}: {out[prec]}")
        del m
        torch.cuda.empty_cache()
    err = {}
    for f, a, b in zip(("u", "v", "t", "q", "ps"), states["fp32"], states["fp64"]):
        den = np.sqrt((b ** 2).sum())
        err[f] = float(np.sqrt(((a - b) ** 2).sum()) / den) if den > 0 else 0.0
    return {"workload": f"T42L20 {E}-member uneven ensemble, per-member jet basic states, "
                        "heatings of the main workload, lengths U(5,60) days, one graph replay, "
                        "autotuned launches (built-in variants timed too)",
            "members": E, "member_days": round(float(lengths.sum()) / 72.0, 2),
            "fp64": out["fp64"], "fp32": out["fp32"],
            "speedup_fp32_over_fp64": round(out["fp32"]["member_days_per_s"]
                                            / out["fp64"]["member_days_per_s"], 3),
            "fp32_rel_l2_vs_fp64": {k: float(f"{v:.3e}") for k, v in err.items()},
            "max_rel_l2": float(f"{max(err.values()):.3e}")}


def measure_adjoint(E=32, max_days=None, window=(20.0, 30.0), every=48):
    """Adjoint sweep time over forward run time for the streamfunction
    response of the uneven ensemble's first E members: the tangent-linear
    model (no trajectory) and the default model about its trajectory
    (checkpoints every ``every`` steps, segments recomputed in the sweep)."""
    import numpy as np
    import torch
    import lbm_isca_amd as L
    from lbm_isca_amd import adjoint
    bs, frc, lengths, sigm = uneven_ensemble(E, max_days=max_days)
    nmax_steps = int(lengths.max())
    res = {}
    for model in ("tangent_linear", "default"):
        for prec in ("fp64", "fp32"):
            cfg = L.Config(precision=prec, tangent_linear=model == "tangent_linear")
            m, _, _ = L.make_ensemble_model(bs, E, cfg, sigm=sigm, per_member_basic=True)
            plan = m.plan
            lat = np.degrees(plan.get("lat", plan.nlat))
            lon = np.degrees(plan.get("lon", plan.nlon))
            adjoint.StreamfunctionResponse(plan, lat, lon, plan.get("gw", plan.nlat),
                                           plan.get("sig", plan.nlev), 0.2, (30.0, 60.0),
                                           (150.0, 210.0), window[0], window[1], nmax_steps, 72,
                                           cfg.inistp)
            plan.set_forcing(frc)
            plan.set_lengths(lengths)
            rec = {}
            if model == "default":
                # the forward without checkpoints, for their cost
                plan.set_checkpoints(0)
                plan.set_spectral_state()
                plan.run(0, use_graph=True)
                torch.cuda.synchronize()
                plan.set_spectral_state()
                t0 = time.perf_counter()
                plan.run(0, use_graph=True)
                torch.cuda.synchronize()
                rec["forward_without_checkpoints_s"] = round(time.perf_counter() - t0, 4)
                plan.set_checkpoints(every)
            plan.set_spectral_state()
            plan.run(0, use_graph=True)
            plan.adjoint_run(0)
            torch.cuda.synchronize()
            plan.set_spectral_state()
            t0 = time.perf_counter()
            plan.run(0, use_graph=True)
            torch.cuda.synchronize()
            tf = time.perf_counter() - t0
            J = plan.get_response()
            t0 = time.perf_counter()
            gx, gf = plan.adjoint_run(0)
            torch.cuda.synchronize()
            ta = time.perf_counter() - t0
            gf = gf.cpu().numpy()
            sel = np.abs(J) > 0
            rec.update({"forward_s": round(tf, 4), "adjoint_s": round(ta, 4),
                        "adjoint_over_forward": round(ta / tf, 3),
                        "members_with_window": int(sel.sum())})
            if model == "tangent_linear":
                # from a zero initial state J_e = <dJ_e/df_e, f_e> exactly (linearity)
                pred = (gf * frc).reshape(E, -1).sum(1)
                lin = float(np.max(np.abs(pred[sel] - J[sel]) / np.abs(J[sel]))) if sel.any() else 0.0
                rec["max_rel_err_J_vs_grad_dot_forcing"] = float(f"{lin:.2e}")
            else:
                rec["checkpoint_every"] = every
                # central difference of J along a forcing direction
                h = 1e-3
                dfrc = 0.5 * np.roll(frc, 17, axis=2)
                out = []
                for sgn in (1.0, -1.0):
                    plan.set_forcing(frc + sgn * h * dfrc)
                    plan.set_spectral_state()
                    plan.run(0, use_graph=True)
                    out.append(plan.get_response())
                fd = (out[0] - out[1]) / (2 * h)
                pred = (gf * dfrc).reshape(E, -1).sum(1)
                ok = np.abs(fd) > 0
                err = float(np.max(np.abs(pred[ok] - fd[ok]) / np.abs(fd[ok]))) if ok.any() else 0.0
                rec["max_rel_err_grad_vs_central_difference"] = float(f"{err:.2e}")
            res[f"{model}_{prec}"] = rec
            log(f"adjoint {model} {prec}: {rec}")
            del m, plan
            torch.cuda.empty_cache()
    return {"workload": f"T42L20 {E} members of the uneven ensemble, J = area mean of "
                        "psi(sigma=0.2) over 30-60N 150-210E, days "
                        f"{window[0]:g}-{window[1]:g}; gradients w.r.t. forcing and spectral "
                        "initial state, of the tangent-linear model and of the default model "
                        "about its trajectory",
            "members": E, **res}


RESOLUTIONS = [("T21", 21, 64, 32), ("T42", 42, 128, 64), ("T63", 63, 192, 96),
               ("T85", 85, 256, 128), ("T106", 106, 320, 160)]


def measure_resolutions(E=64, single_steps=108, ens_steps=72):
    """The resolution sweep of the reference's perf benchmark
    (analysis/benchmark_perf_isca.py:43-45, 78-106): a single member about a
    resting isothermal (250 K) basic state, default Config (L20, dt 1200 s),
    ms per step over 1.5 days after a warm-up; plus the E-member ensemble's
    member-days/s on the same grid."""
    import numpy as np
    import torch
    import lbm_isca_amd as L
    out = {}
    for name, nmax, nlon, nlat in RESOLUTIONS:
        sp = L.BatchedSpectral(nlon, nlat, nmax)
        vg = L.VerticalGrid()
        shp = (nlat, nlon, vg.kmax)
        rest = types.SimpleNamespace(u=np.zeros(shp), v=np.zeros(shp), t=np.full(shp, 250.0),
                                     q=np.zeros(shp), ps=np.full((nlat, nlon), 1000.0))
        rec = {"nlon": nlon, "nlat": nlat, "ndof": (nmax + 1) * (nmax + 2) // 2}
        for nmem, steps in ((1, single_steps), (E, ens_steps)):
            m, _, _ = L.make_ensemble_model(rest, nmem, L.Config(), spectral=sp)
            plan = m.plan
            plan.set_forcing(cases_heating(sp, vg, nmem))
            plan.run(22, use_graph=True)
            plan.run_steps(steps, m.cfg.inistp, use_graph=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plan.run_to(steps, 0, use_graph=True)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if nmem == 1:
                rec["single_member_ms_per_step"] = round(dt * 1e3 / steps, 4)
            else:
                rec[f"members_{nmem}_member_days_per_s"] = round(nmem * steps / 72.0 / dt, 1)
                rec[f"members_{nmem}_ms_per_step"] = round(dt * 1e3 / steps, 4)
            del m, plan
            torch.cuda.empty_cache()
        log(f"{name}: {rec}")
        out[name] = rec
    return {"workload": "resting isothermal 250 K basic state, L20, dt 1200 s, 2 K/day heating at "
                        "0N 180E; single member (the reference sweep's case) and a "
                        f"{E}-member ensemble; full steps only",
            "members": E, **out}


def cases_heating(sp, vg, nmem):
    from lbm_isca_amd import cases
    return np.broadcast_to(cases.heating(sp.lat, sp.lon, vg.sig, 0.0, 180.0, 2.0)[None],
                           (nmem, sp.nlat, sp.nlon, vg.kmax))


def legendre_flops(E, L, nspec, Jh):
    """Useful flops of the level Legendre kernels of one step: each real
    contraction is 2 nspec Jh (2 E L); synthesis has 8 of them (P for the six
    synthesised fields, H for U and V), analysis 11 (P for seven grid
    tendencies, H for four)."""
    one = 2.0 * nspec * Jh * 2 * E * L
    return {"syn_level": 8 * one, "ana_level": 11 * one}


def measure_t127(E=32, days=20.0):
    """T127 L40 in the three precisions of the Legendre contractions."""
    import numpy as np
    import torch
    import lbm_isca_amd as L
    from lbm_isca_amd import cases, roofline, tuning
    sigm = cases.EQUAL_SIGM_L40
    sp = L.BatchedSpectral(384, 192, 127)
    vg = L.VerticalGrid(sigm)
    t0 = time.perf_counter()
    bs, _ = cases.jet_states(sp, vg.sig, E, seed=1)
    vor = cases.vorticity_perturbation(sp, vg.kmax, E, seed=4)
    log(f"t127l40 setup {time.perf_counter() - t0:.1f} s")
    delt = cases.t127l40_config()["delt"]
    spd = int(round(86400.0 / delt))
    nsteps = int(round(spd * days))
    M, Jh = 128, 96
    nspec = M * (M + 1) // 2
    flops = legendre_flops(E, vg.kmax, nspec, Jh)
    peaks = {"fp64": roofline.MI355X_FP64_TFLOPS, "fp32": roofline.MI355X_FP32_TFLOPS,
             "bf16x3": roofline.MI355X_BF16_TFLOPS}
    res, states = {}, {}
    for prec in ("fp64", "fp32", "bf16x3"):
        cfg = L.Config(**cases.t127l40_config(precision=prec))
        m, _, _ = L.make_ensemble_model(bs, E, cfg, sigm=sigm, per_member_basic=True, spectral=sp)
        plan = m.plan
        plan.set_spectral_state(vor=vor)
        tuned = tuning.autotune(plan, True, reps=3)   # bit-identical variants, timed on the plan
        plan.set_spectral_state(vor=vor)
        plan.run(8, use_graph=True)              # capture + warm caches
        plan.set_spectral_state(vor=vor)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.run(nsteps, use_graph=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        states[prec] = [x.cpu().numpy() for x in plan.get_state()]
        plan.profile(True)
        plan.run(3, use_graph=False)
        prof = plan.profile_read()
        plan.profile(False)
        avg = {k: ms / n for k, (ms, n) in prof.items() if n}
        mfma = {}
        for k in ("syn_level", "ana_level"):
            tf = flops[k] / (avg[k] * 1e-3) / 1e12
            executed = 3.0 if prec == "bf16x3" else 1.0
            mfma[k] = {"ms": round(avg[k], 4), "useful_tflops": round(tf, 2),
                       "frac_of_peak": round(executed * tf / peaks[prec], 4)}
        res[prec] = {"member_days_per_s": round(E * nsteps / spd / dt, 3),
                     "ms_per_step": round(dt * 1e3 / (nsteps + cfg.inistp), 3),
                     "legendre": mfma, "peak_tflops": peaks[prec],
                     "autotuned": dict(zip(plan.TUNED, tuned["choice"]))}
        log(f"t127l40 {prec}: {res[prec]}")
        del m, plan
        torch.cuda.empty_cache()
    err = {}
    for prec in ("fp32", "bf16x3"):
        e = {}
        for f, a, b in zip(("u", "v", "t", "q", "ps"), states[prec], states["fp64"]):
            den = np.sqrt((b ** 2).sum())
            e[f] = float(f"{np.sqrt(((a - b) ** 2).sum()) / den:.3e}") if den > 0 else 0.0
        err[prec] = e
    return {"workload": f"T127L40 {E} balanced-jet members (own basic states), n^-1.5 vorticity "
                        f"perturbation rms 1e-6 1/s, {days:g} days at dt = {delt:g} s, one graph "
                        "replay per precision",
            "members": E, "days": days, "dt_s": delt, **res, "rel_l2_vs_fp64": err,
            "note": "useful Legendre flops (executed on the matrix cores: x3 for bf16x3) over the "
                    "dense peak of the precision"}


def calibrate_copy(mib, reps=3):
    """A plain device copy of known size: each repetition reads and writes
    ``mib`` MiB once, so FETCH_SIZE / WRITE_SIZE of these dispatches give the
    factors that turn the counters into HBM bytes."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied {mib} MiB x {reps}")


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=720)
    ap.add_argument("--warmup", type=int, default=72)
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--precision", default="fp64", choices=("fp64", "fp32"))
    ap.add_argument("--profile-steps", type=int, default=24)
    ap.add_argument("--cpu-steps", type=int, default=16,
                    help="timed oracle steps of the CPU baseline's thread figures")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the FP32-vs-FP64 uneven ensemble and the adjoint timing")
    ap.add_argument("--only", choices=("uneven", "adjoint", "t127", "resolutions"), default=None,
                    help="run only one of the extra measurements (profiling)")
    ap.add_argument("--uneven-members", type=int, default=256)
    ap.add_argument("--adjoint-members", type=int, default=32)
    ap.add_argument("--max-days", type=float, default=None,
                    help="cap the uneven lengths (profiling runs only)")
    ap.add_argument("--window", type=float, nargs=2, default=(20.0, 30.0),
                    help="response averaging window in days")
    ap.add_argument("--t127-members", type=int, default=32)
    ap.add_argument("--t127-days", type=float, default=20.0)
    ap.add_argument("--cpu-worker", type=int, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--tuning", default=None, metavar="C0,C1,C2,C3,C4",
                    help="apply these launch variants instead of autotuning (profiling runs)")
    ap.add_argument("--calibrate-copy", type=int, default=None, metavar="MIB",
                    help="only copy a float64 buffer of MIB mebibytes three times (the "
                         "reference kernel that scales the HBM byte counters) and exit")
    args = ap.parse_args()
    if args.cpu_worker is not None:
        cpu_worker(args.cpu_worker)         # host only: no GPU is touched
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); members are "
              "independent, so several GPUs would each run their own share", file=sys.stderr)
        sys.exit(2)

    # the host baseline first: its worker processes start before this
    # process initialises the GPU
    cpu = None
    if args.only is None and args.calibrate_copy is None and not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(args.cpu_steps)
        log(f"cpu baseline: {cpu['value']} member-days/s on {cpu['cores']} cores "
            f"({cpu['cpu_model']})")

    import torch
    from lbm_isca_amd import roofline
    if not torch.cuda.is_available():
        raise RuntimeError("bench.py needs a HIP device")
    if args.calibrate_copy is not None:
        calibrate_copy(args.calibrate_copy)
        return
    if args.only is not None:
        if args.only == "uneven":
            res = measure_fp32_vs_fp64(args.uneven_members, args.max_days)
        elif args.only == "t127":
            res = measure_t127(args.t127_members, args.t127_days)
        elif args.only == "resolutions":
            res = measure_resolutions()
        else:
            res = measure_adjoint(args.adjoint_members, args.max_days, tuple(args.window))
        print(json.dumps({args.only: res}), flush=True)
        return
    E = args.members
    m, bs, frc, vg = build_model(E, args.precision)
    plan = m.plan
    log(f"warmup {args.warmup} steps ...")
    plan.run(args.warmup, use_graph=True)
    torch.cuda.synchronize()
    # the step's launch variants timed on this ensemble (the state is kept)
    from lbm_isca_amd import tuning
    if args.tuning is not None:
        plan.set_tuning([int(c) for c in args.tuning.split(",")])
        tuned = {"choice": plan.get_tuning(), "report": "given"}
    else:
        tuned = tuning.autotune(plan, True)
    log(f"autotuned: {tuned['report']} -> {tuned['choice']}")
    # a run's two start-up half steps go before the timed region, which then
    # holds exactly --steps full steps (model.py:211-251)
    plan.run_steps(args.steps, m.cfg.inistp, use_graph=True)
    torch.cuda.synchronize()
    log(f"timed {args.steps} steps ...")
    t0 = time.perf_counter()
    plan.run_to(args.steps, 0, use_graph=True)
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
    gchoice = tuned["choice"][4]
    kb = roofline.kernel_bytes(E, L=vg.kmax, s=s, nfrc=E, fused_tend_fft=gchoice in (1, 2),
                               grid_fused=gchoice >= 3)
    avg = {k: ms / n for k, (ms, n) in prof.items() if n and ms > 0}
    dom = max(avg, key=lambda k: prof[k][0])
    achieved = kb[dom] / (avg[dom] * 1e-3) / 1e9
    log("kernel ms/launch: " + ", ".join(f"{k} {v:.3f}" for k, v in avg.items()))
    if dom == "tend" and gchoice >= 3:
        dom_name = "grid (k_grid: inverse FFT, tendencies, forward FFT)"
    else:
        dom_name = dom
    rl = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1),
          "peak": roofline.MI355X_HBM_GBS, "unit": "GB/s",
          "frac": round(achieved / roofline.MI355X_HBM_GBS, 4),
          "traffic": measured_traffic(dom, E, args.precision, gchoice),
          "algorithmic_bytes": kb[dom], "avg_ms": round(avg[dom], 5),
          "kernels_ms": {k: round(v, 5) for k, v in avg.items()}}

    autotuned = dict(zip(plan.TUNED, tuned["choice"]))
    extras = {}
    if not args.no_extras:
        del m, plan
        torch.cuda.empty_cache()
        for key, fn in (("fp32_vs_fp64",
                         lambda: measure_fp32_vs_fp64(args.uneven_members, args.max_days)),
                        ("adjoint", lambda: measure_adjoint(args.adjoint_members, args.max_days,
                                                            tuple(args.window))),
                        ("t127l40", lambda: measure_t127(args.t127_members, args.t127_days)),
                        ("resolutions", measure_resolutions)):
            log(f"{key} ...")
            try:
                extras[key] = fn()
            except (FloatingPointError, RuntimeError, ValueError) as exc:
                # the main line stands; the failed part says why
                extras[key] = {"error": f"{type(exc).__name__}: {exc}"}
                log(f"{key} failed: {exc}")
            torch.cuda.empty_cache()


    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "member-days/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (solid-body basic state, seeded random heatings)",
        "config": {"workload": f"T42L20 {E}-member heating ensemble, solid-body basic state, "
                               f"{args.precision.upper()}, 20 equal sigma levels",
                   "members": E, "resolution": "T42L20", "dt_s": 1200.0,
                   "graph": True, "startup_steps_in_timed_region": 0,
                   "autotuned": autotuned},
        "roofline": rl, "cpu_baseline": cpu, **extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
