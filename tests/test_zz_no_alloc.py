"""gi_trace_closest_device_noalloc through the C-ABI: the same hits as
gi_trace_closest_device on the same device buffers, on two calls in a row
(the second with fewer rays, so stale scratch contents would show), on a
flat and a two-level (instanced) scene; sizes beyond its limit are refused
before any GPU work."""
import ctypes as C

import numpy as np
import pytest

from globillum_amd import scenes
from globillum_amd._lib import HIT

pytestmark = pytest.mark.gpu
FMAX = np.float32(3.4e38)


class _Dev:
    """Device buffers through the HIP runtime libgi_hip.so links."""

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
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def get(self, p, nbytes):
        out = np.zeros(nbytes, np.uint8)
        assert self.hip.hipMemcpy(out.ctypes.data, p, nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _rays(rng, n, lo, hi):
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = np.where(rng.uniform(size=n) < 0.3, rng.uniform(0.5, 40.0, n), FMAX).astype(np.float32)
    return o, d, t


def _both(ctx, dev, o, d, t):
    n = len(t)
    args = [dev.put(a) for a in (o, d, t)]
    h_ref = dev.alloc(n * 16)
    h_new = dev.alloc(n * 16)
    # poison the output so a ray the new entry point skips would show
    dev.hip.hipMemcpy(h_new, np.full(n * 4, np.nan, np.float32).ctypes.data, n * 16, 1)
    assert ctx.lib.gi_trace_closest_device(ctx.h, *args, n, h_ref) == 0
    rc = ctx.lib.gi_trace_closest_device_noalloc(ctx.h, *args, n, h_new)
    assert rc == 0, ctx.lib.gi_last_error(ctx.h)
    return dev.get(h_ref, n * 16), dev.get(h_new, n * 16)


@pytest.mark.parametrize("which", ["terrain", "instanced"])
def test_noalloc_matches_trace_closest_device(which, gpu_ctx_factory):
    if which == "terrain":
        sc = scenes.terrain_spheres(cells=(64, 64), n_spheres=200, w=8, h=8, spp=1)
    else:
        sc = scenes.instanced_terrain(cells=(32, 32), n_instances=120, w=8, h=8, spp=1)
    lo = sc.tris["a"].min(0) - 5
    hi = sc.tris["a"].max(0) + 5
    rng = np.random.default_rng(5)
    dev = _Dev()
    try:
        with gpu_ctx_factory(sc) as ctx:
            print(f"[noalloc] {which}: call 1, 300000 rays", flush=True)
            o, d, t = _rays(rng, 300_000, lo, hi)
            ref, new = _both(ctx, dev, o, d, t)
            assert np.array_equal(ref, new)
            hits = ref.view(HIT)
            assert (hits["ref"] != 0xFFFFFFFF).sum() > 1000
            print(f"[noalloc] {which}: call 2, 4099 rays", flush=True)
            o, d, t = _rays(rng, 4099, lo, hi)
            ref, new = _both(ctx, dev, o, d, t)
            assert np.array_equal(ref, new)
    finally:
        dev.free()


def test_noalloc_refuses_more_rays_than_its_scratch(gpu_ctx_factory):
    sc = scenes.cornell_boxes(w=8, h=8, spp=1)
    with gpu_ctx_factory(sc) as ctx:
        # the size check comes before any pointer is touched
        p = C.c_void_p(16)
        rc = ctx.lib.gi_trace_closest_device_noalloc(ctx.h, p, p, p, (1 << 19) + 1, p)
        assert rc == 6  # GI_ERR_LIMIT
        assert b"524288" in ctx.lib.gi_last_error(ctx.h)
        assert ctx.lib.gi_trace_closest_device_noalloc(ctx.h, p, p, p, 0, p) == 0
