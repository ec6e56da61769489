"""Two-level instancing on the GPU through the C-ABI (gi_upload_mesh,
gi_upload_instances): closest and any hits against the oracle's
instanced traversal (ids, t and barycentrics bit-exact, any mismatch
adjudicated by the FP64 brute force), closed-form box hits, renders
bit-exact against the oracle with the host and device BVH builders and the
batched and graph schedules, the adjoint against the oracle's, and the error paths."""
import dataclasses

import numpy as np
import pytest

from globillum_amd import scenes
from globillum_amd._lib import FLAG_BATCHED, FLAG_GRAPH, GiError

pytestmark = pytest.mark.gpu
FMAX = np.finfo(np.float32).max
MISS = 0xFFFFFFFF


def _osc(oracle, sc):
    return oracle.OracleScene(sc.tris, sc.mats, sc.lights, meshes=sc.meshes, instances=sc.instances)


def _assert_hits_equal(gpu, ref, osc, o, d, tmax):
    bad = np.nonzero(gpu != ref)[0]
    if len(bad) == 0:
        return
    t64, _, _, _ = osc.brute_closest_f64(o[bad], d[bad], tmax[bad])
    for k, i in enumerate(bad):
        assert gpu["ref"][i] != MISS and ref["ref"][i] != MISS, (i, gpu[i], ref[i])
        assert abs(float(gpu["t"][i]) - t64[k]) <= 1e-6 * t64[k] and abs(float(ref["t"][i]) - t64[k]) <= 1e-6 * t64[k]
    assert len(bad) <= 2, f"{len(bad)} FP64-tie mismatches"


@pytest.fixture(scope="module")
def inst():
    return scenes.instanced_terrain(cells=(48, 48), n_instances=300, w=96, h=64, spp=2, depth=4)


def _rays(rng, n, size):
    o = rng.uniform(-2, size + 2, (n, 3)).astype(np.float32)
    o[:, 1] = rng.uniform(0.0, 6.0, n)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d


@pytest.mark.parametrize("builder", ["host", "device", "ploc", "sah"])
def test_closest_and_any_parity(builder, inst, oracle, gpu_ctx_factory):
    rng = np.random.default_rng(21)
    o, d = _rays(rng, 200_000, 40.0)
    tmax = np.where(rng.random(len(o)) < 0.25, rng.uniform(0, 10, len(o)), FMAX).astype(np.float32)
    osc = _osc(oracle, inst)
    ref = osc.trace_closest(o, d, tmax)
    assert (ref["ref"][ref["ref"] != MISS] >= len(inst.tris)).sum() > 10_000
    with gpu_ctx_factory(inst, builder=builder) as ctx:
        # the traversal's half-precision boxes (world, mesh and instance
        # trees) enclose the f32 ones
        assert ctx.check_half_nodes() == 0
        _assert_hits_equal(ctx.trace_closest(o, d, tmax), ref, osc, o, d, tmax)
        assert np.array_equal(ctx.trace_any(o, d, tmax).astype(bool), ref["ref"] != MISS)
        # camera rays
        co, cd, ct = oracle.camera_rays(inst.views, 0, 0)
        cref = osc.trace_closest(co, cd, ct)
        _assert_hits_equal(ctx.trace_closest(co, cd, ct), cref, osc, co, cd, ct)


def _world_tri(x=100.0):
    t = np.zeros(1, scenes.TRI)
    t["a"], t["b"], t["c"] = (x, 0, 0), (x, 1, 0), (x, 0, 1)
    t["na"] = t["nb"] = t["nc"] = (1, 0, 0)
    return t


def _plain(tris, meshes=None, instances=None):
    return scenes.Scene(tris, scenes.lambert((0.5, 0.5, 0.5)), scenes.make_light((0, 5, 0), (0, -1, 0), (1, 1, 1), 1, 1),
                        scenes.make_view((0, 0, -5), (0, 0, 0), (0, 1, 0), 40.0, 8, 8, 1), scenes.make_opts(2),
                        meshes=meshes, instances=instances)


def test_closed_form_box_hits(oracle, gpu_ctx_factory):
    box = scenes.unit_box()
    A = scenes.rotation((0, 0, 1), np.pi / 2) * 2.0
    insts = np.concatenate([scenes.make_instance(A, (10, 0, 0), 0),
                            scenes.make_instance(np.eye(3) * 0.5, (0, 0, 10), 0, material=0)])
    sc = _plain(_world_tri(), [box], insts)
    o = np.float32([[0, 0, 0], [0, 0, 0], [0, 0.25, 20], [0, 0, 0], [0, 0, 0]])
    d = np.float32([[1, 0, 0], [0, 0, 1], [0, 0, -1], [-1, 0, 0], [1, 0, 0]])
    tmax = np.float32([FMAX, FMAX, FMAX, FMAX, 7.9])
    osc = _osc(oracle, sc)
    with gpu_ctx_factory(sc) as ctx:
        h = ctx.trace_closest(o, d, tmax)
        assert np.array_equal(h, osc.trace_closest(o, d, tmax))
        assert h["t"][0] == pytest.approx(8.0, abs=2e-6) and h["t"][1] == pytest.approx(9.5, abs=2e-6)
        assert 1 <= h["ref"][0] <= 12 and 13 <= h["ref"][1] <= 24 and h["ref"][3] == MISS and h["ref"][4] == MISS
        assert ctx.trace_any(o, d, tmax).tolist() == [1, 1, 1, 0, 0]


@pytest.mark.parametrize("builder,flags", [("host", 0), ("device", FLAG_BATCHED), ("host", FLAG_BATCHED | FLAG_GRAPH),
                                           ("sah", FLAG_BATCHED)])
def test_render_bit_exact(builder, flags, inst, oracle, gpu_ctx_factory):
    ref, st = _osc(oracle, inst).render(inst.views, inst.opts)
    with gpu_ctx_factory(inst, builder=builder) as ctx:
        img = ctx.render(flags=flags).image(0).reshape(-1, 3)
        c = ctx.counters()
    bad = np.nonzero((img != ref).any(1))[0]
    assert len(bad) == 0, f"{len(bad)} pixels differ, first {bad[:5].tolist()}"
    assert c["rays_closest"] == st["rays_closest"] and c["rays_shadow"] == st["rays_shadow"]


def test_adjoint_matches_oracle(oracle, gpu_ctx_factory):
    sc = scenes.instanced_terrain(cells=(24, 24), n_instances=80, n_mats=6, w=48, h=32, spp=8, depth=3)
    tgt = scenes.perturbed(sc, seed=1)
    tgt = dataclasses.replace(tgt, meshes=sc.meshes, instances=sc.instances)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render().image(0).reshape(-1, 3)
        with gpu_ctx_factory(tgt) as c2:
            timg = c2.render().image(0).reshape(-1, 3)
        dL = (2.0 * (img - timg) / img.size).astype(np.float32)
        gm, ge = ctx.adjoint(dL)
    osc = _osc(oracle, sc)
    om32, oe32 = oracle.adjoint(osc, sc.views, sc.opts, dL, f64=False)
    assert np.abs(om32).max() > 0
    assert np.allclose(gm, om32, rtol=2e-3, atol=2e-4 * np.abs(om32).max())
    assert np.allclose(ge, oe32, rtol=2e-3, atol=2e-4 * np.abs(oe32).max())


def test_errors_and_state(gpu_ctx_factory):
    box = scenes.unit_box()
    ok = scenes.make_instance(np.eye(3), (0, 0, 0), 0)
    with gpu_ctx_factory() as ctx:
        lib, h = ctx.lib, ctx.h
        t = _world_tri()
        ctx._ck(lib.gi_upload_triangles(h, t.ctypes.data, 1))
        with pytest.raises(GiError, match="before gi_upload_instances"):
            ctx.upload_instances([box], ok)
    with gpu_ctx_factory(_plain(_world_tri())) as ctx:
        with pytest.raises(GiError, match="singular"):
            ctx.upload_instances([box], scenes.make_instance(np.diag([1.0, 0.0, 1.0]), (0, 0, 0), 0))
        with pytest.raises(GiError, match="mesh id"):
            ctx.upload_instances([box], scenes.make_instance(np.eye(3), (0, 0, 0), 5))
    # an override beyond the materials is refused at render time
    with gpu_ctx_factory(_plain(_world_tri(), [box], scenes.make_instance(np.eye(3), (0, 0, 0), 0, 3))) as ctx:
        with pytest.raises(GiError, match="material index"):
            ctx.render()
    # a new world drops the instances: the box at the origin is no longer hit
    sc = _plain(_world_tri(), [box], ok)
    o, d, tm = np.float32([[0, 0.25, 0.25]]), np.float32([[1, 0, 0]]), np.float32([FMAX])
    with gpu_ctx_factory(sc) as ctx:
        assert ctx.trace_closest(o, d, tm)["t"][0] == pytest.approx(1.0)
        ctx.upload(_plain(_world_tri()))
        assert ctx.trace_closest(o, d, tm)["t"][0] == pytest.approx(100.0)
        ctx.upload(sc)
        assert ctx.trace_closest(o, d, tm)["t"][0] == pytest.approx(1.0)
        ctx.build_bvh_device()
        assert ctx.trace_closest(o, d, tm)["t"][0] == pytest.approx(1.0)
