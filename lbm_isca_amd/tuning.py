"""Autotuned launch variants of the time step, with an on-disk cache.

``autotune(plan, mode)``: ``mode`` True times the variants of the step's
tuned launches on the plan (``lbm_autotune``: FFT rows per block in each
direction, occupancy targets of the level Legendre synthesis and tendency
analyses) and keeps the fastest; "cache" first looks the plan's key (grid,
levels, members, basic states, precision) up in the cache file and applies
a stored pick without timing, and times and stores it otherwise.  The cache
file is ``$LBM_TUNING_CACHE`` or ``~/.cache/lbm_isca_amd/tuning.json``.
Every variant computes the same numbers, so a stale entry can only cost
speed.
"""

import json
import os


def cache_path():
    return os.environ.get("LBM_TUNING_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "lbm_isca_amd", "tuning.json")


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def autotune(plan, mode=True, reps=10):
    """Apply the tuned variants to ``plan``; returns a dict with the choice
    (variant index per tuned launch), the timing report and its source."""
    if not mode:
        return None
    key = plan.tuning_key()
    path = cache_path()
    if mode == "cache":
        hit = _load(path).get(key)
        if hit is not None:
            plan.set_tuning(hit["choice"])
            return {"choice": plan.get_tuning(), "report": hit.get("report", ""), "source": path}
    res = plan.autotune(reps)
    res["source"] = "timed"
    if mode == "cache":
        table = _load(path)
        table[key] = {"choice": res["choice"], "report": res["report"]}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path + ".tmp", "w") as fh:
                json.dump(table, fh, indent=1, sort_keys=True)
            os.replace(path + ".tmp", path)
        except OSError:
            pass  # an unwritable cache only costs the timing next time
    return res
