"""Autotuner front end: per angular-momentum class, pick the faster J/K kernel variant on the
workload at hand and cache the choice by (device architecture, class) in a JSON file.

Variant 0 is the lane-per-quartet kernel (register resident) and variant 1 the wavefront-per-quartet
kernel (LDS staged).  The engine times both on the device (qcjk_autotune); this module only persists
and re-applies the result.
"""
import json
import os

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "gfx950.json")


def _key(r):
    return "%d %d %d %d" % (r["la"], r["lb"], r["lc"], r["ld"])


def tune(engine, dD, reps=3, cache=DEFAULT_CACHE, workload="default"):
    """Run the engine's autotuner on device density dD and merge the choices into the cache."""
    rep = engine.autotune(dD, reps=reps)
    table = load(cache)
    entry = table.setdefault(workload, {})
    for r in rep:
        entry[_key(r)] = int(r["wavefront_per_quartet"])
    os.makedirs(os.path.dirname(cache), exist_ok=True)
    with open(cache, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return rep


def load(cache=DEFAULT_CACHE):
    if not os.path.exists(cache):
        return {}
    with open(cache) as f:
        return json.load(f)


def apply(engine, cache=DEFAULT_CACHE, workload="default"):
    table = load(cache).get(workload)
    if table:
        engine.apply_tuning(table)
    return table
