"""Context state across calls, through the C-ABI: camera-ray generation
between a render and its readback / adjoint, a failed render, the kept
hipGraph relaunched by later renders, the autograd function's backward
after another forward, and the device-pointer entry points
(gi_get_image_device, gi_trace_closest_device, gi_trace_any_device)."""
import ctypes as C

import numpy as np
import pytest

from globillum_amd import scenes
from globillum_amd._lib import FLAG_BATCHED, FLAG_GRAPH, HIT, GiError, ptr

pytestmark = pytest.mark.gpu


def _small_diff(spp=8, seed=4):
    rng = np.random.default_rng(seed)
    sc = scenes.diff_scene(seed=seed, cells=(16, 16), n_spheres=10, n_mats=4, n_lights=2, w=48, h=32, spp=spp,
                           radius=(3.0, 8.0))
    sc.mats["color"] = rng.uniform(0.2, 0.9, (len(sc.mats), 3)).astype(np.float32)
    return sc


def test_camera_rays_leave_the_render_state_alone(gpu_ctx_factory):
    sc = _small_diff()
    dL = np.random.default_rng(1).normal(size=(48 * 32, 3)).astype(np.float32)
    with gpu_ctx_factory(sc) as ctx:
        img0 = ctx.render().image(0)
        gm0, ge0 = ctx.adjoint(dL)
        # a larger view, another seed: the render's views, images and Halton
        # points must be untouched
        big = scenes.make_view((10.0, 20.0, -30.0), (0.3, -0.5, 1.0), (0, 1, 0), 50.0, 200, 150, 4)
        o, d, t = ctx.camera_rays(big, 2, seed=99)
        assert o.shape == (200 * 150, 3) and np.isfinite(d).all()
        img1 = ctx.image(0)
        gm1, ge1 = ctx.adjoint(dL)
    assert img1.shape == img0.shape
    assert np.array_equal(img0, img1)
    # the same replay: equal up to the float atomics' summation order
    assert np.allclose(gm0, gm1, rtol=1e-4, atol=1e-5 * np.abs(gm0).max())
    assert np.allclose(ge0, ge1, rtol=1e-4, atol=1e-5 * np.abs(ge0).max())


def test_progressive_sequence_survives_camera_rays(gpu_ctx_factory):
    sc = scenes.cornell_boxes(w=40, h=32, spp=3)
    with gpu_ctx_factory(sc) as ctx:
        one = ctx.render().image(0)
        ctx.progress_begin()
        ctx.progress_step()
        ctx.camera_rays(scenes.make_view((0.5, 0.5, -1.3), (0, 0, 1), (0, 1, 0), 40.0, 64, 64, 2), 1, seed=7)
        while not ctx.progress_step():
            pass
        assert np.array_equal(ctx.image(0), one)


def test_failed_render_leaves_no_render_state(gpu_ctx_factory):
    sc = _small_diff()
    with gpu_ctx_factory(sc) as ctx:
        ctx.render()
        bad = sc.views.copy()
        bad["up"] = bad["dir"]  # up parallel to dir
        with pytest.raises(GiError) as e:
            ctx.render(bad)
        assert e.value.status == 1
        lib, h = ctx.lib, ctx.h
        dm = np.zeros(len(sc.mats) * 5, np.float32)
        de = np.zeros(len(sc.lights) * 3, np.float32)
        dL = np.zeros(48 * 32 * 3, np.float32)
        assert lib.gi_adjoint(h, ptr(dL), ptr(dm), ptr(de)) == 4  # GI_ERR_STATE
        done = C.c_int32(0)
        assert lib.gi_progress_step(h, C.byref(done)) == 4
        out = np.zeros(48 * 32 * 3, np.float32)
        assert lib.gi_get_image(h, 0, ptr(out)) == 1  # no view 0 any more
        # a progressive sequence that fails to start leaves none behind
        ctx.progress_begin()
        assert lib.gi_progress_begin(h, ptr(bad), 1, ptr(np.array(sc.opts))) == 1
        assert lib.gi_progress_step(h, C.byref(done)) == 4
        # and the context renders again
        img = ctx.render().image(0)
        assert np.isfinite(img).all() and img.max() > 0


def test_graph_is_kept_and_relaunched(gpu_ctx_factory):
    sc = scenes.cornell_boxes(w=96, h=64, spp=6)
    flags = FLAG_BATCHED | FLAG_GRAPH
    with gpu_ctx_factory(sc) as ctx:
        plain = ctx.render().image(0)
        first = ctx.render(flags=flags).image(0)
        c1 = ctx.counters()
        assert c1["ms_graph_build"] > 0
        # replays with eager renders in between, each checked against the eager image
        for _ in range(3):
            again = ctx.render(flags=flags).image(0)
            c2 = ctx.counters()
            assert c2["ms_graph_build"] == 0.0, "the kept graph was not relaunched"
            assert c2["launches_closest"] == c1["launches_closest"]
            assert np.array_equal(again, plain)
            assert np.array_equal(ctx.render().image(0), plain)
        assert np.array_equal(first, plain)
        # new material values: same launches, the graph is relaunched and
        # reads the new values
        m2 = sc.mats.copy()
        m2["diffuse_color"] *= 0.5
        ctx.set_materials(m2)
        eager = ctx.render().image(0)
        got = ctx.render(flags=flags).image(0)
        assert ctx.counters()["ms_graph_build"] == 0.0
        assert np.array_equal(got, eager) and not np.array_equal(got, plain)
        # a new camera pose with the same sizes: relaunched too
        v2 = sc.views.copy()
        v2["pos"][0, 0] += 0.05
        eager = ctx.render(v2).image(0)
        got = ctx.render(v2, flags=flags).image(0)
        assert ctx.counters()["ms_graph_build"] == 0.0
        assert np.array_equal(got, eager)
        # another size: captured again
        v3 = v2.copy()
        v3["w"] = 80
        eager = ctx.render(v3).image(0)
        got = ctx.render(v3, flags=flags).image(0)
        assert ctx.counters()["ms_graph_build"] > 0
        assert np.array_equal(got, eager)


def test_autograd_backward_after_a_second_forward():
    torch = pytest.importorskip("torch")
    from globillum_amd import diff
    from globillum_amd.context import Context

    sc = _small_diff()
    p0 = torch.tensor(diff.material_params(sc), requires_grad=True)
    e0 = torch.tensor(sc.lights["col"].astype(np.float32), requires_grad=True)
    w = torch.from_numpy(np.random.default_rng(3).normal(size=(48 * 32, 3)).astype(np.float32))
    with Context(0) as ctx:
        ctx.upload(sc)
        img = diff.render(ctx, p0, e0)
        (img * w).sum().backward()
        g_single = (p0.grad.clone(), e0.grad.clone())
        p0.grad = None
        e0.grad = None
        img_a = diff.render(ctx, p0, e0)
        p1 = (p0.detach() * 0.7).requires_grad_(True)
        e1 = (e0.detach() * 1.3).requires_grad_(True)
        img_b = diff.render(ctx, p1, e1)  # another forward before the first backward
        assert not torch.equal(img_a, img_b)
        (img_a * w).sum().backward()
        # the first forward's gradient (up to the atomics' summation order),
        # not the second's
        for g, r in ((p0.grad, g_single[0]), (e0.grad, g_single[1])):
            assert torch.allclose(g, r, rtol=1e-4, atol=1e-5 * float(r.abs().max()))
        (img_b * w).sum().backward()
        assert p1.grad is not None
        assert not torch.allclose(p1.grad, g_single[0], rtol=1e-2, atol=1e-3 * float(g_single[0].abs().max()))


class _Dev:
    """Device buffers through the HIP runtime libgi_hip.so itself links
    (loaded by its soname once the library is loaded), so one runtime
    serves the allocation and the call."""

    def __init__(self):
        from globillum_amd import _lib
        _lib.load()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(nbytes, 256)) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, nbytes):
        out = np.zeros(nbytes, np.uint8)
        assert self.hip.hipMemcpy(out.ctypes.data, p, nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def test_get_image_device_is_the_image(gpu_ctx_factory):
    sc = scenes.cornell_boxes(w=40, h=24, spp=2)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render().image(0).reshape(-1, 3)
        p = C.c_void_p()
        assert ctx.lib.gi_get_image_device(ctx.h, 0, C.byref(p)) == 0 and p.value
        dev = _Dev()
        got = dev.get(p, 40 * 24 * 16).view(np.float32).reshape(-1, 4)
        assert np.array_equal(got[:, :3], img)
        assert ctx.lib.gi_get_image_device(ctx.h, 1, C.byref(p)) == 1


def test_trace_device_entry_points_on_device_buffers(oracle, gpu_ctx_factory):
    sc = scenes.terrain_spheres(cells=(32, 32), n_spheres=60, w=8, h=8, spp=1)
    rng = np.random.default_rng(11)
    n = 20000
    lo = sc.tris["a"].min(0)
    hi = sc.tris["a"].max(0)
    o = rng.uniform(lo - 5, hi + 5, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = np.where(rng.uniform(size=n) < 0.5, np.float32(3.4e38), rng.uniform(1, 60, n)).astype(np.float32)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights)
    ref = osc.trace_closest(o, d, t)
    ref_any = osc.trace_any(o, d, t)
    dev = _Dev()
    try:
        with gpu_ctx_factory(sc) as ctx:
            args = [dev.put(a) for a in (o, d, t)]
            hits = dev.alloc(n * 16)
            occ = dev.alloc(n)
            assert ctx.lib.gi_trace_closest_device(ctx.h, *args, n, hits) == 0
            assert ctx.lib.gi_trace_any_device(ctx.h, *args, n, occ) == 0
            got = dev.get(hits, n * 16).view(HIT)
            got_any = dev.get(occ, n)
            # the host-pointer wrappers agree with the device entry points
            assert np.array_equal(ctx.trace_closest(o, d, t).view(np.uint8), got.view(np.uint8))
    finally:
        dev.free()
    from test_gpu_trace import _assert_hits_equal
    _assert_hits_equal(got, ref, osc, o, d, t)
    assert np.array_equal(got_any, ref_any)
