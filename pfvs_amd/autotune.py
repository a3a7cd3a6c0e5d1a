"""Autotuner for the batched search's launch parameters.

The batched form (a breadth-first split of the roots, then one persistent
launch with per-wavefront deques and donation through a frontier ring) has
four tunables: persistent wavefronts per CU, the most nodes one wavefront
donates at once, the deque capacity and the split's nodes per wavefront.
Their best values depend on the tree shapes (h11, dilation, p-vector mix)
and on the device, so `tune` times the search (its own HIP-event
duration, `stats["ms_search"]`) on a seeded sample of the caller's batch
for every setting of a small grid and keeps the fastest.  Every setting
must visit the same nodes (the search is deterministic in what it visits,
only the schedule changes); a difference raises.

Results are cached in a JSON file keyed by the device's architecture and
CU count and by a signature of the workload (h11 values, dilation range,
sample size), so a serving loop or a repeated scan tunes once.
"""

from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from . import gpu

DEFAULT_GRID = {
    "waves_per_cu": (2, 3, 4),
    "budget": (4, 16),
    "stack_cap": (0,),
    "split": (512, 2048, 8192),
}


def _cache_path():
    return os.environ.get("PFVS_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "pfvs_amd_tune.json"))


def workload_key(geoms, pgeo, n_sample, device=0):
    cus, clk = gpu.device_facts(device)
    hs = sorted({int(np.asarray(g["kappa"]).shape[0]) for g in geoms})
    dils = sorted({float(g.get("dilation", 1)) for g in geoms})
    sig = json.dumps({"arch": gpu.require().pfa_arch().decode(), "cus": cus, "h": hs,
                      "dil": [dils[0], dils[-1]], "n": int(n_sample)}, sort_keys=True)
    return hashlib.sha1(sig.encode()).hexdigest()[:16], sig


def _load_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(path, cache):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def sample_batch(ps, pgeo, n_sample, seed=0):
    """A seeded subset of the batch (all of it when n_sample >= len)."""
    n = len(ps)
    if n_sample >= n:
        return np.arange(n)
    return np.sort(np.random.default_rng(seed).choice(n, size=n_sample, replace=False))


def tune(geoms, ps, pgeo, device=0, n_sample=4096, grid=None, repeats=2, use_cache=True,
         runner=None, verbose=False):
    """Best launch parameters for this workload:
    {"waves_per_cu", "budget", "stack_cap", "split", "ms_search", "key", "table"}.

    runner(geoms, ps, pgeo, **params) -> stats dict (default: gpu.run in
    batched mode without points); it must report "ms_search" and "nodes"."""
    grid = dict(DEFAULT_GRID, **(grid or {}))
    key, sig = workload_key(geoms, pgeo, min(n_sample, len(ps)), device)
    path = _cache_path()
    if use_cache:
        hit = _load_cache(path).get(key)
        if hit:
            return dict(hit, key=key, cached=True)
    sel = sample_batch(ps, pgeo, n_sample)
    sp_ps, sg = np.ascontiguousarray(ps[sel]), np.ascontiguousarray(pgeo[sel])
    if runner is None:
        def runner(g, p, pg, **kw):
            return gpu.run(g, p, pg, mode="batched", device=device, keep_points=False,
                           filter=False, **kw)["stats"]
    runner(geoms, sp_ps, sg)                          # warm: buffers sized, code loaded
    table, nodes0, best = [], None, None
    for w in grid["waves_per_cu"]:
        for b in grid["budget"]:
            for sc in grid["stack_cap"]:
                for sp in grid["split"]:
                    params = {"waves_per_cu": int(w), "budget": int(b), "stack_cap": int(sc),
                              "split": int(sp)}
                    ms = []
                    for _ in range(repeats):
                        st = runner(geoms, sp_ps, sg, **params)
                        if nodes0 is None:
                            nodes0 = st["nodes"]
                        elif st["nodes"] != nodes0:
                            raise RuntimeError(f"settings {params} visited {st['nodes']} nodes, "
                                               f"not {nodes0}")
                        ms.append(st["ms_search"])
                    row = dict(params, ms_search=min(ms))
                    table.append(row)
                    if verbose:
                        print(f"  {params} -> {min(ms):.2f} ms", flush=True)
                    if best is None or row["ms_search"] < best["ms_search"]:
                        best = row
    out = dict(best, table=table, signature=sig, tuned_at=time.time())
    if use_cache:
        cache = _load_cache(path)
        cache[key] = out
        _store_cache(path, cache)
    return dict(out, key=key, cached=False)
