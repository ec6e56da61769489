"""Autotuner of the renderer's launch parameters.

Searches traversal block size, blocks per CU, LDS stack depth, BVH node
width (child slots per node: the world tree is collapsed again for each
candidate), the traversal kernels' refill (compaction) threshold, shading
block size and the number of paths kept in flight per pass by timing real
renders of the caller's workload (coordinate descent from the defaults,
two repeats per candidate, best median wins).  The images do not depend on
these parameters -- only the schedule and the tree's shape do, and closest
hits are the same in any tree -- so any choice is safe.
"""
from __future__ import annotations

import json
import time

SPACE = {
    "trace_block": [64, 128, 256],
    "trace_blocks_per_cu": [4, 5, 6, 8],
    "stack_lds": [8, 16, 32],
    "node_width": [2, 3, 4],
    "refill_lanes": [8, 16, 32, 64],
    "shade_block": [64, 128, 256],
    "target_slots": [16, 32, 64],
}
DEFAULT = {"trace_block": 256, "trace_blocks_per_cu": 5, "stack_lds": 16, "node_width": 4, "refill_lanes": 32,
           "shade_block": 256, "target_slots": 64}


def _time(ctx, views, opts, repeats: int) -> float:
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        ctx.render(views, opts)
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2]


def tune(ctx, views, opts, space: dict | None = None, repeats: int = 2, sweeps: int = 1, log=None) -> dict:
    """Coordinate descent over `space`; leaves the best setting applied on
    ctx and returns {"params": ..., "seconds": ..., "trials": [...]}."""
    space = space or SPACE
    best = dict(DEFAULT)
    ctx.set_tuning(**best)
    ctx.render(views, opts)  # warm caches / allocations
    best_t = _time(ctx, views, opts, repeats)
    trials = [(dict(best), best_t)]
    for _ in range(sweeps):
        for key, values in space.items():
            for v in values:
                if v == best[key]:
                    continue
                cand = dict(best)
                cand[key] = v
                ctx.set_tuning(**cand)
                t = _time(ctx, views, opts, repeats)
                trials.append((cand, t))
                if log:
                    log(f"[autotune] {key}={v}: {t * 1e3:.2f} ms (best {best_t * 1e3:.2f})")
                if t < best_t:
                    best, best_t = cand, t
    ctx.set_tuning(**best)
    return {"params": best, "seconds": best_t, "trials": [{"params": p, "seconds": t} for p, t in trials]}


def save(result: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(result, f, indent=1)
