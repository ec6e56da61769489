"""The autotuner's search and its per-workload cache, on the CPU with a
stand-in context whose render time is a known function of the tuning: the
search finds the fastest coordinate-wise setting, times each candidate
after one untimed render, and a cached result is applied without renders."""
import numpy as np

from globillum_amd import autotune, scenes


class _Ctx:
    def __init__(self):
        self.t = dict(autotune.DEFAULT)
        self.renders = 0
        self.clock = 0.0

    def set_tuning(self, **kw):
        self.t.update(kw)

    def render(self, views, opts):
        self.renders += 1


def test_search_and_cache(tmp_path, monkeypatch):
    ctx = _Ctx()
    cost = {("refill_lanes", 16): -3.0, ("trace_blocks_per_cu", 6): -2.0, ("trace_blocks_per_cu", 4): 1.0}
    clock = {"now": 0.0}

    def fake_render(views, opts):
        ctx.renders += 1
        clock["now"] += 10.0 + sum(v for (k, val), v in cost.items() if ctx.t[k] == val)

    ctx.render = fake_render
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: clock["now"])
    sc = scenes.cornell_boxes(w=16, h=16, spp=2)
    space = {"refill_lanes": [16, 32, 64], "trace_blocks_per_cu": [4, 5, 6]}
    cache = str(tmp_path / "t.json")
    key = autotune.workload_key(sc.views, sc.opts, "c")
    r = autotune.tune(ctx, sc.views, sc.opts, space=space, repeats=3, cache=cache, key=key)
    assert r["params"]["refill_lanes"] == 16 and r["params"]["trace_blocks_per_cu"] == 6
    assert r["seconds"] == 5.0 and not r["cached"]
    # the defaults, refill 16 (wins before the default's turn, so 32 is a
    # candidate again), 32, 64, blocks 4 and 6: each one untimed render and
    # 3 timed ones
    assert len(r["trials"]) == 6 and ctx.renders == 6 * 4
    ctx2 = _Ctx()
    r2 = autotune.tune(ctx2, sc.views, sc.opts, space=space, cache=cache, key=key)
    assert r2["cached"] and ctx2.renders == 0 and ctx2.t["refill_lanes"] == 16
    # a different space or key is not a hit
    ctx3 = _Ctx()
    ctx3.render = fake_render
    r3 = autotune.tune(ctx3, sc.views, sc.opts, space={"refill_lanes": [16, 32]}, repeats=1, cache=cache, key=key)
    assert not r3["cached"]
    v2 = sc.views.copy()
    v2["spp"] = 4
    assert autotune.workload_key(v2, sc.opts, "c") != key
    assert autotune.workload_key(sc.views, sc.opts, "c", device="x") != key
    assert np.all(sc.views == sc.views)
