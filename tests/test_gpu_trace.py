"""GPU traversal parity through the C-ABI: gi_trace_closest / gi_trace_any
against the oracle's binary-BVH traversal on the same rays -- triangle id,
t and barycentrics bit-exact, any mismatch adjudicated by the FP64 brute
force -- plus the edge cases (empty input, axis-parallel directions,
max_t <= 0, odd sizes, external BVH upload, call-order errors)."""
import numpy as np
import pytest

from globillum_amd import scenes
from globillum_amd._lib import GiError
from globillum_amd.context import build_bvh2

pytestmark = pytest.mark.gpu
FMAX = np.finfo(np.float32).max


def _rays(rng, n, lo, hi, ylo=None, yhi=None):
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    if ylo is not None:
        o[:, 1] = rng.uniform(ylo, yhi, n)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d


def _assert_hits_equal(gpu, ref, osc, o, d, tmax):
    bad = np.nonzero(gpu != ref)[0]
    if len(bad) == 0:
        return
    # adjudicate: both candidates must be FP64 ties (same exact t within 1e-6)
    t64, _, _, r64 = osc.brute_closest_f64(o[bad], d[bad], tmax[bad])
    for k, i in enumerate(bad):
        assert gpu["ref"][i] != 0xFFFFFFFF and ref["ref"][i] != 0xFFFFFFFF, (i, gpu[i], ref[i])
        assert abs(float(gpu["t"][i]) - t64[k]) <= 1e-6 * t64[k] and abs(float(ref["t"][i]) - t64[k]) <= 1e-6 * t64[k]
    assert len(bad) <= 2, f"{len(bad)} FP64-tie mismatches"


@pytest.fixture(scope="module")
def c1():
    return scenes.cornell_boxes()


@pytest.fixture(scope="module")
def terrain():
    return scenes.terrain_spheres(cells=(96, 96), n_spheres=300, w=96, h=54, spp=4)


@pytest.mark.parametrize("which", ["c1", "terrain"])
def test_closest_hit_parity_random_rays(which, c1, terrain, oracle, gpu_ctx_factory):
    sc = c1 if which == "c1" else terrain
    rng = np.random.default_rng(11)
    if which == "c1":
        o, d = _rays(rng, 250_000, -0.2, 1.2)
    else:
        o, d = _rays(rng, 250_000, -5, 105, 0, 20)
    tmax = np.where(rng.random(len(o)) < 0.25, rng.uniform(0, 30, len(o)), FMAX).astype(np.float32)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights)
    ref = osc.trace_closest(o, d, tmax)
    with gpu_ctx_factory(sc) as ctx:
        _assert_hits_equal(ctx.trace_closest(o, d, tmax), ref, osc, o, d, tmax)
    # the GPU traversing the oracle's own (median-split) BVH gives the same
    with gpu_ctx_factory(sc, bvh2=osc.bvh2()) as ctx:
        _assert_hits_equal(ctx.trace_closest(o, d, tmax), ref, osc, o, d, tmax)


def test_closest_hit_parity_camera_and_secondary(terrain, oracle, gpu_ctx_factory):
    osc = oracle.OracleScene(terrain.tris, terrain.mats, terrain.lights, bvh2=build_bvh2(terrain.tris))
    o, d, t = oracle.camera_rays(terrain.views, 0, 0)
    ref = osc.trace_closest(o, d, t)
    with gpu_ctx_factory(terrain) as ctx:
        g = ctx.trace_closest(o, d, t)
        _assert_hits_equal(g, ref, osc, o, d, t)
        # secondary rays: from the primary hit points (offset along the ray
        # back-side) in random directions
        hit = ref["ref"] != 0xFFFFFFFF
        P = o[hit] + (ref["t"][hit, None] * 0.9999) * d[hit]
        rng = np.random.default_rng(2)
        d2 = rng.normal(size=P.shape).astype(np.float32)
        d2[:, 1] = np.abs(d2[:, 1])
        d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
        t2 = np.full(len(P), FMAX, np.float32)
        _assert_hits_equal(ctx.trace_closest(P, d2, t2), osc.trace_closest(P, d2, t2), osc, P, d2, t2)


def test_any_hit_parity(terrain, oracle, gpu_ctx_factory):
    rng = np.random.default_rng(5)
    o, d = _rays(rng, 200_000, 0, 100, 0, 20)
    tmax = rng.uniform(0, 50, len(o)).astype(np.float32)
    osc = oracle.OracleScene(terrain.tris)
    ref = osc.trace_any(o, d, tmax)
    with gpu_ctx_factory(terrain) as ctx:
        occ = ctx.trace_any(o, d, tmax)
        assert np.array_equal(occ, ref)
        h = ctx.trace_closest(o, d, np.full(len(o), FMAX, np.float32))
        assert np.array_equal(occ.astype(bool), (h["ref"] != 0xFFFFFFFF) & (h["t"] <= tmax))


def test_edge_cases(c1, oracle, gpu_ctx_factory):
    osc = oracle.OracleScene(c1.tris)
    with gpu_ctx_factory(c1) as ctx:
        assert len(ctx.trace_closest(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0))) == 0
        # axis-parallel directions (zero components -> clamped reciprocal)
        o = np.array([[0.5, 0.5, 0.5]] * 6, np.float32)
        d = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
        t = np.full(6, FMAX, np.float32)
        g = ctx.trace_closest(o, d, t)
        assert (g == osc.trace_closest(o, d, t)).all()
        assert (g["ref"][[0, 1, 2, 3, 4]] != 0xFFFFFFFF).all() and g["ref"][5] == 0xFFFFFFFF  # front open
        # max_t <= 0: nothing can be hit
        g = ctx.trace_closest(o, d, np.array([0, -1, 0, -5, 0, 0], np.float32))
        assert (g["ref"] == 0xFFFFFFFF).all() and (g["t"] == FMAX).all()
        assert not ctx.trace_any(o, d, np.zeros(6, np.float32)).any()
        # sizes around the wave width
        rng = np.random.default_rng(1)
        for n in (1, 63, 64, 65, 1000, 1_000_003):
            o, d = _rays(rng, n, -0.2, 1.2)
            t = np.full(n, FMAX, np.float32)
            _assert_hits_equal(ctx.trace_closest(o, d, t), osc.brute_closest(o, d, t), osc, o, d, t)
            _assert_hits_equal(ctx.trace_closest(o, d, t), osc.trace_closest(o, d, t), osc, o, d, t)


def test_call_order_and_argument_errors(c1, gpu_ctx_factory):
    with gpu_ctx_factory() as ctx:
        with pytest.raises(GiError) as e:
            ctx.trace_closest(np.zeros((1, 3)), np.ones((1, 3)), np.ones(1))
        assert e.value.status == 4
    with gpu_ctx_factory() as ctx:
        ctx.upload(c1)
        bad_nodes = np.zeros(8, np.float32)
        bad_nodes.view(np.int32)[0] = 3  # inner node pointing at child 1 of a 1-node tree
        with pytest.raises(GiError) as e:
            ctx.upload(c1, bvh2=(bad_nodes, np.zeros(0, np.uint32)))
        assert e.value.status == 1


# ---- device-built BVH (linear BVH or PLOC, 4-wide collapse) ----------------
# Closest hit is the minimum (t, id) over all triangles, so the hits do not
# depend on the tree: the device-built tree must give the oracle's answers.

@pytest.mark.parametrize("builder,max_leaf,radius", [("device", 1, 16), ("device", 4, 16), ("device", 15, 16),
                                                     ("ploc", 1, 16), ("ploc", 4, 16), ("ploc", 15, 16),
                                                     ("ploc", 4, 1), ("ploc", 4, 64),
                                                     ("sah", 1, 0), ("sah", 4, 0), ("sah", 15, 0)])
def test_device_bvh_closest_and_any(builder, max_leaf, radius, terrain, oracle, gpu_ctx_factory):
    sc = terrain
    rng = np.random.default_rng(5 + max_leaf)
    o, d = _rays(rng, 200_000, -5, 105, 0, 20)
    tmax = np.where(rng.random(len(o)) < 0.3, rng.uniform(0, 30, len(o)), FMAX).astype(np.float32)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights)
    ref = osc.trace_closest(o, d, tmax)
    occ = osc.trace_any(o, d, tmax)
    with gpu_ctx_factory(sc, builder=builder, max_leaf=max_leaf, radius=radius) as ctx:
        _assert_hits_equal(ctx.trace_closest(o, d, tmax), ref, osc, o, d, tmax)
        assert np.array_equal(ctx.trace_any(o, d, tmax), occ)


@pytest.mark.parametrize("builder", ["sah", "ploc", "device"])
def test_device_bvh_many_equal_centroids(builder, oracle, gpu_ctx_factory):
    """6000 coplanar triangles scaled about one centroid: no axis has a
    centroid extent, so every binned-SAH node above the leaves is an
    object-median split (the root's segment spans several binning chunks),
    and rays hit many triangles at the same t (ties go to the lower id)."""
    rng = np.random.default_rng(3)
    sc = scenes.cornell_boxes(w=8, h=8, spp=1)
    n = 6000
    tris = np.zeros(n, sc.tris.dtype)
    tris[:] = sc.tris[0]
    c = np.array([0.5, 0.5, 0.5], np.float32)
    base = np.array([[-0.3, -0.2, 0.0], [0.35, -0.25, 0.0], [-0.05, 0.45, 0.0]], np.float32)
    s_ = rng.uniform(0.05, 1.0, n).astype(np.float32)
    for k, name in enumerate("abc"):
        tris[name] = c + s_[:, None] * base[k]
    o = np.stack([rng.uniform(0.1, 0.9, 20_000), rng.uniform(0.1, 0.9, 20_000), np.full(20_000, -1.0)],
                 1).astype(np.float32)
    d = np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (len(o), 1))
    d[:, :2] = rng.normal(0, 0.05, (len(o), 2))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = np.full(len(o), FMAX, np.float32)
    osc = oracle.OracleScene(tris, sc.mats, sc.lights)
    ref = osc.trace_closest(o, d, t)
    assert (ref["ref"] != 0xFFFFFFFF).mean() > 0.3
    one = scenes.Scene(tris, sc.mats, sc.lights, sc.views, sc.opts)
    with gpu_ctx_factory(one, builder=builder) as ctx:
        _assert_hits_equal(ctx.trace_closest(o, d, t), ref, osc, o, d, t)
        assert np.array_equal(ctx.trace_any(o, d, t), osc.trace_any(o, d, t))


def test_device_sah_is_the_host_tree(terrain, gpu_ctx_factory):
    """The device binned-SAH builder takes the host builder's split decisions,
    so its tree makes a closest-hit traversal visit the same nodes and test
    the same triangles (the counts of a primitive call on a fixed ray array
    repeat exactly; the collapse's costs are float on the device, double on
    the host, so a near tie may still go the other way).  The radix tree
    and clustering trees differ from it by tens of percent."""
    rng = np.random.default_rng(11)
    o, d = _rays(rng, 100_000, -5, 105, 0, 20)
    t = np.full(len(o), FMAX, np.float32)
    counts = {}
    for builder in ("host", "sah", "ploc"):
        with gpu_ctx_factory(terrain, builder=builder) as ctx:
            ctx.traversal_stats(True)
            ctx.trace_closest(o, d, t)
            s = ctx.traversal_stats()
        counts[builder] = np.array([s["closest_nodes"], s["closest_triangles"]], np.float64)
    print({k: v.tolist() for k, v in counts.items()})
    assert np.all(np.abs(counts["sah"] / counts["host"] - 1) <= 2e-3)
    assert np.any(np.abs(counts["ploc"] / counts["host"] - 1) > 0.05)


def test_ploc_argument_errors(c1, gpu_ctx_factory):
    with gpu_ctx_factory(c1) as ctx:
        for method, radius in (("ploc", 65), ("ploc", -1)):
            with pytest.raises(GiError) as e:
                ctx.build_bvh_device(4, method, radius)
            assert e.value.status == 1
        with pytest.raises(GiError) as e:
            ctx._ck(ctx.lib.gi_build_bvh_device_ex(ctx.h, 4, 3, 16))  # 0..2 are the builders
        assert e.value.status == 1
        with pytest.raises(ValueError):
            ctx.build_bvh_device(4, "host")


@pytest.mark.parametrize("builder", ["device", "ploc", "sah"])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 17])
def test_device_bvh_tiny_and_degenerate(n, builder, oracle, gpu_ctx_factory):
    rng = np.random.default_rng(n)
    sc = scenes.cornell_boxes(w=8, h=8, spp=1)
    tris = sc.tris[:n].copy() if n <= len(sc.tris) else np.concatenate([sc.tris] * 2)[:n].copy()
    if n == 17:
        # many triangles with one centroid: equal Morton keys, ties broken by index
        base = tris[0].copy()
        for i in range(n):
            tris[i] = base
            rot = rng.permutation(3)
            for k, src in zip("abc", rot):
                tris[i][k] = base["abc"[src]]
    o, d = _rays(rng, 20_000, -0.2, 1.2)
    t = np.full(len(o), FMAX, np.float32)
    osc = oracle.OracleScene(tris, sc.mats, sc.lights)
    ref = osc.trace_closest(o, d, t)
    one = scenes.Scene(tris, sc.mats, sc.lights, sc.views, sc.opts)
    with gpu_ctx_factory(one, builder=builder) as ctx:
        _assert_hits_equal(ctx.trace_closest(o, d, t), ref, osc, o, d, t)


@pytest.mark.parametrize("builder", ["host", "device", "ploc", "sah"])
def test_half_nodes_enclose_wide_boxes(builder, terrain, gpu_ctx_factory):
    """The traversal reads each node as f16 plane offsets from an f32 origin
    (HalfNode): every child box it decodes encloses the f32 box widened by
    the node's margins, and the child references are the same."""
    with gpu_ctx_factory(terrain, builder=builder) as ctx:
        assert ctx.check_half_nodes() == 0


@pytest.mark.parametrize("scale,offset", [(1e-5, 0.0), (1e-4, 30.0), (250.0, -2e6), (3e4, 0.0)])
def test_extreme_coordinates(scale, offset, oracle, gpu_ctx_factory):
    """Scenes far from the origin or of very small or large size: the f16
    node planes (scale and offsets relative to each node's origin) stay
    conservative, so hits still match the oracle bit for bit."""
    sc = scenes.terrain_spheres(cells=(48, 48), n_spheres=60, w=32, h=18, spp=1)
    tris = sc.tris.copy()
    for k in "abc":
        tris[k] = (tris[k].astype(np.float64) * scale + offset).astype(np.float32)
    one = scenes.Scene(tris, sc.mats, sc.lights, sc.views, sc.opts)
    rng = np.random.default_rng(17)
    o, d = _rays(rng, 100_000, -5, 105, 0, 20)
    o = (o.astype(np.float64) * scale + offset).astype(np.float32)
    t = np.full(len(o), FMAX, np.float32)
    osc = oracle.OracleScene(tris, sc.mats, sc.lights)
    ref = osc.trace_closest(o, d, t)
    assert (ref["ref"] != 0xFFFFFFFF).mean() > 0.2
    with gpu_ctx_factory(one) as ctx:
        assert ctx.check_half_nodes() == 0
        _assert_hits_equal(ctx.trace_closest(o, d, t), ref, osc, o, d, t)
        assert np.array_equal(ctx.trace_any(o, d, t), osc.trace_any(o, d, t))


def test_triangle_reciprocal_exhaustive():
    """The triangle test's reciprocal (hardware rcp + one Newton step) equals
    the IEEE division 1/x bit for bit for all 2^32 floats, so t, u and v keep
    the oracle's bits (globillum_amd/csrc/rcp_check.hip)."""
    import os
    import subprocess

    import globillum_amd

    exe = os.path.join(os.path.dirname(globillum_amd.__file__), "gi_rcp_check")
    assert os.path.exists(exe), "gi_rcp_check not built (python -m globillum_amd.build)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "total mismatches 0" in r.stdout
