"""Tuning parameters and the autotuner on the GPU: images do not depend on
node width (2, 3, 4 child slots per BVH node, host- and device-built trees)
or on the traversal kernels' refill threshold (8..64 lanes); narrower nodes
mean more node visits; invalid values and a node-width change under placed
instances are refused; the autotuner returns a setting from its space,
leaves it applied, and the image rendered with it is the default one.  A
scene small enough for the traversal kernels' LDS copy (at most 16 KB of
nodes and triangles; LDS stack depth 16) renders and traces exactly as
through the global-memory kernels (stack depth 8 or 32)."""
import numpy as np
import pytest

from globillum_amd import autotune, scenes
from globillum_amd._lib import FLAG_BATCHED, GiError

pytestmark = pytest.mark.gpu
FMAX = np.finfo(np.float32).max


@pytest.fixture(scope="module")
def ts():
    return scenes.terrain_spheres(cells=(32, 32), n_spheres=40, w=64, h=48, spp=4, depth=4)


def test_node_width_and_refill_keep_the_image(ts, gpu_ctx_factory):
    with gpu_ctx_factory(ts) as ctx:
        ref = ctx.render(flags=FLAG_BATCHED).image(0).copy()
        c0 = ctx.counters()
        nodes = {}
        for w in (2, 3, 4):
            ctx.set_tuning(node_width=w)
            assert ctx.tuning()["node_width"] == w
            ctx.traversal_stats(True)
            assert np.array_equal(ctx.render(flags=FLAG_BATCHED).image(0), ref), w
            s = ctx.traversal_stats()
            ctx.traversal_stats(False)
            nodes[w] = s["closest_nodes"] / s["closest_rays"]
            assert ctx.counters()["rays_closest"] == c0["rays_closest"]
        assert nodes[2] > nodes[3] > nodes[4]
        for r in (8, 16, 64, 32):
            ctx.set_tuning(refill_lanes=r)
            assert np.array_equal(ctx.render(flags=FLAG_BATCHED).image(0), ref), r
        # a device-built tree at width 2 and 3
        for w in (2, 3):
            ctx.set_tuning(node_width=w)
            ctx.build_bvh_device(4, "ploc")
            assert np.array_equal(ctx.render(flags=FLAG_BATCHED).image(0), ref), ("ploc", w)


def test_tuning_errors(ts, gpu_ctx_factory):
    with gpu_ctx_factory(ts) as ctx:
        for kw in ({"node_width": 1}, {"node_width": 5}, {"refill_lanes": 12}, {"refill_lanes": 128}):
            with pytest.raises(GiError) as e:
                ctx.set_tuning(**kw)
            assert e.value.status == 1, kw
        assert ctx.tuning()["node_width"] == 4 and ctx.tuning()["refill_lanes"] == 32
    inst = scenes.instanced_terrain(cells=(16, 16), n_instances=20, w=32, h=24, spp=1, depth=2)
    with gpu_ctx_factory(inst) as ctx:
        with pytest.raises(GiError) as e:
            ctx.set_tuning(node_width=3)
        assert e.value.status == 4
        ref = ctx.render(flags=FLAG_BATCHED).image(0).copy()
        for r in (8, 64):  # launch parameters stay free, and the two-level kernels agree
            ctx.set_tuning(refill_lanes=r)
            assert np.array_equal(ctx.render(flags=FLAG_BATCHED).image(0), ref), r


def test_autotuner(ts, gpu_ctx_factory):
    space = {"node_width": [3, 4], "refill_lanes": [16, 32], "trace_blocks_per_cu": [5, 6]}
    with gpu_ctx_factory(ts) as ctx:
        ref = ctx.render(flags=FLAG_BATCHED).image(0).copy()
        opts = ts.opts.copy()
        opts["flags"] = FLAG_BATCHED
        res = autotune.tune(ctx, ts.views, opts, space=space, repeats=1)
        p = res["params"]
        for k, vals in space.items():
            assert p[k] in vals
        # one render of the defaults, then each other value of each key; a value that wins
        # before the default's turn in the list makes the default a candidate again
        n = len(res["trials"])
        assert 1 + sum(len(v) - 1 for v in space.values()) <= n <= 1 + sum(len(v) for v in space.values())
        assert all(t["seconds"] > 0 for t in res["trials"]) and res["seconds"] > 0
        applied = ctx.tuning()
        assert all(applied[k] == v for k, v in p.items())
        assert np.array_equal(ctx.render(flags=FLAG_BATCHED).image(0), ref)


def test_autotuner_cache_per_workload(ts, gpu_ctx_factory, tmp_path):
    space = {"refill_lanes": [16, 32], "trace_blocks_per_cu": [5, 6]}
    cache = str(tmp_path / "tuning.json")
    opts = ts.opts.copy()
    opts["flags"] = FLAG_BATCHED
    key = autotune.workload_key(ts.views, opts, "ts")
    with gpu_ctx_factory(ts) as ctx:
        first = autotune.tune(ctx, ts.views, opts, space=space, repeats=3, cache=cache, key=key)
        assert not first["cached"] and len(first["trials"]) >= 3
    with gpu_ctx_factory(ts) as ctx:
        again = autotune.tune(ctx, ts.views, opts, space=space, cache=cache, key=key)
        assert again["cached"] and again["params"] == first["params"] and again["trials"] == []
        assert all(ctx.tuning()[k] == v for k, v in first["params"].items())
        # another workload (view size) is searched again
        v2 = ts.views.copy()
        v2["w"] = 48
        other = autotune.tune(ctx, v2, opts, space=space, repeats=1, cache=cache,
                              key=autotune.workload_key(v2, opts, "ts"))
        assert not other["cached"]


def test_lds_scene_matches_global_path(oracle, gpu_ctx_factory):
    sc = scenes.cornell_boxes(w=96, h=64, spp=4, depth=4)
    rng = np.random.default_rng(9)
    o = rng.uniform(-0.2, 1.2, (50000, 3)).astype(np.float32)
    d = rng.normal(size=(50000, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = np.where(rng.random(50000) < 0.3, rng.uniform(0, 1, 50000), FMAX).astype(np.float32)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights)
    ref_hits, ref_occ = osc.trace_closest(o, d, t), osc.trace_any(o, d, t)
    with gpu_ctx_factory(sc) as ctx:
        out = {}
        for lds in (16, 8, 32):  # 16: the scene is copied into LDS; 8 and 32: read from global memory
            ctx.set_tuning(stack_lds=lds)
            out[lds] = (ctx.render(flags=FLAG_BATCHED).image(0).copy(), ctx.trace_closest(o, d, t),
                        ctx.trace_any(o, d, t))
        for lds in (8, 32):
            assert np.array_equal(out[lds][0], out[16][0])
            assert np.array_equal(out[lds][1], out[16][1]) and np.array_equal(out[lds][2], out[16][2])
        bad = np.nonzero(out[16][1] != ref_hits)[0]
        assert len(bad) <= 2  # FP64 ties only, as in the other parity tests
        assert np.array_equal(out[16][2].astype(bool), ref_occ.astype(bool))
