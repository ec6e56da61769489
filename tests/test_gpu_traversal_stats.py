"""Traversal statistics (gi_set_traversal_stats / gi_get_traversal_stats):
exact counts on a one-triangle tree (one node visit per ray, one triangle
test per ray that reaches the leaf's box, none otherwise), identical
results and images with counting on, ray totals equal to the renderer's
counters, repeatable totals, instance entries counted, and the state rules.

A lane's node visits depend on the other rays of its wave: with
speculative leaves, a lane that has parked a leaf goes on with its stack
until every lane of the wave has one.  The renderer's queues are filled in
block order, which varies from run to run, so a render's node and
triangle totals repeat only to about 1e-3; a primitive call on a fixed
ray array repeats exactly."""
import numpy as np
import pytest

from globillum_amd import scenes
from globillum_amd._lib import FLAG_BATCHED, GiError

pytestmark = pytest.mark.gpu
FMAX = np.finfo(np.float32).max


def _one_tri_scene():
    t = np.zeros(1, scenes.TRI)
    t["a"], t["b"], t["c"] = (100, 0, 0), (100, 1, 0), (100, 0, 1)
    t["na"] = t["nb"] = t["nc"] = (-1, 0, 0)
    return scenes.Scene(t, scenes.lambert((0.5, 0.5, 0.5)), scenes.make_light((0, 5, 0), (0, -1, 0), (1, 1, 1), 1, 1),
                        scenes.make_view((0, 0, -5), (0, 0, 0), (0, 1, 0), 40.0, 8, 8, 1), scenes.make_opts(2))


def test_exact_counts_on_one_triangle(gpu_ctx_factory):
    n = 1000
    rng = np.random.default_rng(1)
    tgt = np.stack([np.full(n, 100.0), rng.uniform(0.01, 0.4, n), rng.uniform(0.01, 0.4, n)], 1)
    o = np.zeros((n, 3), np.float32)
    d = (tgt / np.linalg.norm(tgt, axis=1, keepdims=True)).astype(np.float32)
    t = np.full(n, FMAX, np.float32)
    with gpu_ctx_factory(_one_tri_scene()) as ctx:
        with pytest.raises(GiError) as e:
            ctx.traversal_stats()
        assert e.value.status == 4
        ctx.traversal_stats(True)
        h = ctx.trace_closest(o, d, t)
        assert (h["ref"] == 0).all()
        s = ctx.traversal_stats()
        assert (s["closest_rays"], s["closest_nodes"], s["closest_triangles"]) == (n, n, n)
        assert s["closest_instances"] == 0 and s["any_rays"] == 0
        # pointing away: the root is visited, its only box is missed
        ctx.trace_any(o, -d, t)
        s = ctx.traversal_stats()
        assert (s["any_rays"], s["any_nodes"], s["any_triangles"]) == (n, n, 0)
        # counting off: the totals stay
        ctx.traversal_stats(False)
        ctx.trace_closest(o, d, t)
        assert ctx.traversal_stats() == s


def test_render_unchanged_and_totals(gpu_ctx_factory):
    sc = scenes.terrain_spheres(cells=(32, 32), n_spheres=40, w=64, h=48, spp=4, depth=4)
    with gpu_ctx_factory(sc) as ctx:
        ref = ctx.render(flags=FLAG_BATCHED).image(0).copy()
        c = ctx.counters()
        totals = []
        for _ in range(2):
            ctx.traversal_stats(True)
            img = ctx.render(flags=FLAG_BATCHED).image(0)
            assert np.array_equal(img, ref)
            totals.append(ctx.traversal_stats())
        s = totals[0]
        for k, v in s.items():
            assert abs(totals[1][k] - v) <= 5e-3 * v, k
        assert s["closest_rays"] == totals[1]["closest_rays"] and s["any_rays"] == totals[1]["any_rays"]
        assert s["closest_rays"] == c["rays_closest"] and s["any_rays"] == c["rays_shadow"]
        assert s["closest_nodes"] >= s["closest_rays"] and s["closest_triangles"] > 0
        assert s["any_nodes"] >= s["any_rays"] and s["closest_instances"] == 0


def test_instance_entries_counted(gpu_ctx_factory):
    sc = scenes.instanced_terrain(cells=(24, 24), n_instances=60, w=48, h=32, spp=2, depth=3)
    rng = np.random.default_rng(3)
    n = 20000
    o = rng.uniform(-2, 42, (n, 3)).astype(np.float32)
    o[:, 1] = rng.uniform(0.0, 6.0, n)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = np.full(n, FMAX, np.float32)
    with gpu_ctx_factory(sc) as ctx:
        h0 = ctx.trace_closest(o, d, t)
        ctx.traversal_stats(True)
        h1 = ctx.trace_closest(o, d, t)
        s = ctx.traversal_stats()
        ctx.traversal_stats(True)
        ctx.trace_closest(o, d, t)
        assert ctx.traversal_stats() == s  # a fixed ray array repeats exactly
    assert np.array_equal(h0, h1)
    assert s["closest_rays"] == n and s["closest_instances"] > 0
    # every ray that hits an instance's triangle entered at least one instance
    assert s["closest_instances"] >= int((h1["ref"][h1["ref"] != 0xFFFFFFFF] >= len(sc.tris)).sum())
