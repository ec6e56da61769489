"""Autotuner of the renderer's launch parameters.

Searches traversal block size, blocks per CU, LDS stack depth, BVH node
width (child slots per node: the world tree is collapsed again for each
candidate), the traversal kernels' refill (compaction) threshold, shading
block size and the number of paths kept in flight per pass by timing real
renders of the caller's workload: coordinate descent from the defaults; for
every candidate one untimed render first (a hipGraph is captured, a tree
collapsed again, buffers grown), then `repeats` timed renders, the median
wins.  The images do not depend on these parameters -- only the schedule
and the tree's shape do, and closest hits are the same in any tree -- so
any choice is safe.

A result can be kept per workload in a JSON file (`cache`): the key is the
device (name, compute units), the views' sizes and sample counts, the
render options and the caller's scene tag; a later tune() of the same key
applies the stored parameters without searching.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

SPACE = {
    "trace_block": [64, 128, 256],
    "trace_blocks_per_cu": [4, 5, 6],
    "stack_lds": [8, 16, 32],
    "node_width": [2, 3, 4],
    "refill_lanes": [8, 16, 32, 64],
    "shade_block": [64, 128, 256],
    "target_slots": [32, 64, 128],
}
DEFAULT = {"trace_block": 256, "trace_blocks_per_cu": 6, "stack_lds": 16, "node_width": 4, "refill_lanes": 32,
           "shade_block": 256, "target_slots": 128}


def _time(ctx, views, opts, repeats: int) -> float:
    ctx.render(views, opts)  # untimed: capture, re-collapse, allocations
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        ctx.render(views, opts)
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2]


def workload_key(views, opts, scene_tag: str = "", device: str = "") -> str:
    """Key of a workload in the tuning cache: the device, every view's
    (w, h, spp), the options that change the schedule, and a scene tag."""
    v = [(int(x["w"]), int(x["h"]), int(x["spp"])) for x in views.reshape(-1)]
    o = opts.reshape(-1)[0]
    desc = json.dumps({"device": device, "views": v, "depth": int(o["max_depth"]), "rr": int(o["rr_start"]),
                       "flags": int(o["flags"]), "spp_pass": int(o["samples_per_pass"]), "scene": scene_tag})
    return hashlib.sha1(desc.encode()).hexdigest()[:16]


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(ctx, views, opts, space: dict | None = None, repeats: int = 3, sweeps: int = 1, log=None,
         cache: str | None = None, key: str | None = None) -> dict:
    """Coordinate descent over `space`; leaves the best setting applied on
    ctx and returns {"params", "seconds", "trials", "cached"}.  With a
    cache file and a key (workload_key), a stored result of the same key
    and space is applied instead of searching, and a new one is stored."""
    space = space or SPACE
    if cache and key:
        hit = _load(cache).get(key)
        if hit and hit.get("space") == space:
            ctx.set_tuning(**hit["params"])
            return {"params": hit["params"], "seconds": hit["seconds"], "trials": [], "cached": True}
    best = dict(DEFAULT)
    ctx.set_tuning(**best)
    best_t = _time(ctx, views, opts, repeats)
    trials = [(dict(best), best_t)]
    for _ in range(sweeps):
        for k, values in space.items():
            for v in values:
                if v == best[k]:
                    continue
                cand = dict(best)
                cand[k] = v
                ctx.set_tuning(**cand)
                t = _time(ctx, views, opts, repeats)
                trials.append((cand, t))
                if log:
                    log(f"[autotune] {k}={v}: {t * 1e3:.2f} ms (best {best_t * 1e3:.2f})")
                if t < best_t:
                    best, best_t = cand, t
    ctx.set_tuning(**best)
    if cache and key:
        d = _load(cache)
        d[key] = {"params": best, "seconds": best_t, "space": space, "repeats": repeats}
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1)
        os.replace(tmp, cache)
    return {"params": best, "seconds": best_t, "trials": [{"params": p, "seconds": t} for p, t in trials],
            "cached": False}


def save(result: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(result, f, indent=1)
