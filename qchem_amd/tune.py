"""Autotuner front end: per angular-momentum class, pick the faster J/K kernel variant on the
workload at hand and cache the choice by (device architecture, class) in a JSON file.

Variants: 0 lane per quartet (register resident), 1 wavefront per quartet with LDS-staged transforms,
2 wavefront per quartet with register transforms, 3 lane per quartet chunked over the ket's first
component, 4 variant 2 with the factorised bra transform, 5 and 6 variants 2 and 4 with small LDS
batches, 7 and 8 variant 2 with both transforms factorised (big and small batches), 9 and 10 variant 2
with each lane building whole R tables (R batches of ~16 KB, and of ~8 KB at
<= 256 registers).  The engine times every variant compiled for a class on the device (qcjk_autotune); this
module only persists and re-applies the result.
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
