"""Python face of the C-ABI: one :class:`Context` per device owning the
scene, the BVH and the render buffers (a gi_ctx)."""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from ._lib import (COUNTERS, HIT, INSTANCE, LIGHT, MATERIAL, OPTS, TRAVERSAL_STATS, TRI, TUNING, VIEW, GiError,
                   ptr)

# device builders of gi_build_bvh_device_ex (GI_BVH_LBVH, GI_BVH_PLOC, GI_BVH_SAH)
_DEVICE_METHODS = {"device": 0, "ploc": 1, "sah": 2}


def _check(lib, ctx, rc):
    if rc != _lib.GI_OK:
        msg = lib.gi_last_error(ctx).decode() if ctx else ""
        raise GiError(rc, msg)


class Context:
    """A gi_ctx: scene upload, tracing primitives, renderer, telemetry."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.gi_create(C.byref(h), device)
        if rc != _lib.GI_OK:
            raise GiError(rc, f"gi_create(device={device})")
        self.h = h
        self.scene = None
        # bumped by every call that changes what the next gi_adjoint would
        # differentiate (materials, lights, views, options, geometry)
        self.generation = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.gi_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc):
        _check(self.lib, self.h, rc)

    # ---- scene ---------------------------------------------------------
    def upload(self, scene, max_leaf: int = 4, bvh2: tuple | None = None, builder: str = "host",
               radius: int = 16):
        """Triangles, materials and lights of `scene`, and its BVH: the
        host binned-SAH builder (builder="host"), the device LBVH builder
        (builder="device"), the device PLOC builder with the given search
        radius (builder="ploc"), the device binned-SAH builder
        (builder="sah"), or an external tree in the reference's 2 x float4
        layout (bvh2=(nodes, indices))."""
        tris = np.ascontiguousarray(scene.tris, TRI)
        self.generation += 1
        self._ck(self.lib.gi_upload_triangles(self.h, ptr(tris), len(tris)))
        for t in getattr(scene, "textures", None) or []:
            self.upload_texture(t)
        if scene.mats is not None and len(scene.mats):
            mats = np.ascontiguousarray(scene.mats, MATERIAL)
            self._ck(self.lib.gi_upload_materials(self.h, ptr(mats), len(mats)))
        if scene.lights is not None and len(scene.lights):
            lights = np.ascontiguousarray(scene.lights, LIGHT)
            self._ck(self.lib.gi_upload_lights(self.h, ptr(lights), len(lights)))
        sky = getattr(scene, "sky", None)
        if sky is not None:
            m = np.ascontiguousarray(sky, np.float32)
            self._ck(self.lib.gi_upload_sky(self.h, ptr(m), m.shape[1], m.shape[0]))
        if bvh2 is None and builder in _DEVICE_METHODS:
            self._ck(self.lib.gi_build_bvh_device_ex(self.h, max_leaf, _DEVICE_METHODS[builder], radius))
        elif bvh2 is None:
            if builder != "host":
                raise ValueError(f"builder must be 'host', 'device', 'ploc' or 'sah', not {builder!r}")
            self._ck(self.lib.gi_build_bvh(self.h, max_leaf))
        else:
            nodes, idx = bvh2
            nodes = np.ascontiguousarray(nodes, np.float32)
            idx = np.ascontiguousarray(idx, np.uint32)
            self._ck(self.lib.gi_upload_bvh2(self.h, ptr(nodes), len(nodes) // 8, ptr(idx), len(idx)))
        if getattr(scene, "instances", None) is not None and len(scene.instances):
            self.upload_instances(scene.meshes, scene.instances)
        # own copy: set_materials / set_lights must not rewrite the caller's scene
        self.scene = dataclasses.replace(scene, mats=None if scene.mats is None else scene.mats.copy(),
                                         lights=None if scene.lights is None else scene.lights.copy())
        return self

    def build_bvh_device(self, max_leaf: int = 4, method: str = "device", radius: int = 16):
        """Rebuild the BVH of the uploaded triangles on the device, as a
        linear BVH (method="device"), by PLOC clustering (method="ploc") or
        by the host builder's binned SAH (method="sah").  A new world tree
        drops the instances; the scene's are placed again."""
        if method not in _DEVICE_METHODS:
            raise ValueError(f"method must be 'device', 'ploc' or 'sah', not {method!r}")
        self._ck(self.lib.gi_build_bvh_device_ex(self.h, max_leaf, _DEVICE_METHODS[method], radius))
        sc = self.scene
        if sc is not None and getattr(sc, "instances", None) is not None and len(sc.instances):
            self._ck(self.lib.gi_upload_instances(self.h, ptr(self._inst), len(self._inst)))

    def upload_texture(self, rgba) -> int:
        """An RGBA8 texture ((h, w, 4) uint8, row 0 at t = 1); its mip chain
        is built on the device.  Returns its id (gi_material.texture)."""
        t = np.ascontiguousarray(rgba, np.uint8)
        if t.ndim != 3 or t.shape[2] != 4:
            raise ValueError("texture must be (h, w, 4) uint8")
        tid = C.c_int32()
        self._ck(self.lib.gi_upload_texture(self.h, ptr(t), t.shape[1], t.shape[0], C.byref(tid)))
        return tid.value

    def texture_chain(self, tex: int = 0):
        """(levels, flat uint8 RGBA chain: level 0, then each halved level)."""
        lv, n = C.c_int32(), C.c_int64()
        self._ck(self.lib.gi_texture_info(self.h, tex, C.byref(lv), C.byref(n)))
        out = np.zeros(4 * n.value, np.uint8)
        self._ck(self.lib.gi_download_texture(self.h, tex, ptr(out)))
        return lv.value, out

    def texture_grad(self, tex: int = 0) -> np.ndarray:
        """(h, w, 3) gradient of the last adjoint w.r.t. the texture's
        level-0 texel values (byte / 255)."""
        t = self.scene.textures[tex]
        out = np.zeros((t.shape[0], t.shape[1], 3), np.float32)
        self._ck(self.lib.gi_texture_grad(self.h, tex, ptr(out)))
        return out

    def upload_instances(self, meshes, instances):
        """Meshes (object-space TRI arrays) and INSTANCE records placing
        them over the world geometry (two-level BVH).  Call after the world
        BVH; meshes uploaded before are kept, so the ids in `instances`
        count from the first mesh of this scene."""
        base = None
        for m in meshes:
            m = np.ascontiguousarray(m, TRI)
            mid = C.c_int32()
            self._ck(self.lib.gi_upload_mesh(self.h, ptr(m), len(m), C.byref(mid)))
            base = mid.value if base is None else base
        inst = np.array(instances, INSTANCE).reshape(-1)
        if base:
            inst["mesh"] += base
        self._inst = inst
        self._ck(self.lib.gi_upload_instances(self.h, ptr(inst), len(inst)))

    # ---- tracing primitives ---------------------------------------------
    def trace_closest(self, orig, dirs, max_t) -> np.ndarray:
        o = np.ascontiguousarray(orig, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(max_t, np.float32).reshape(-1)
        out = np.zeros(len(t), HIT)
        self._ck(self.lib.gi_trace_closest(self.h, ptr(o), ptr(d), ptr(t), len(t), ptr(out)))
        return out

    def trace_any(self, orig, dirs, max_t) -> np.ndarray:
        o = np.ascontiguousarray(orig, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(max_t, np.float32).reshape(-1)
        out = np.zeros(len(t), np.uint8)
        self._ck(self.lib.gi_trace_any(self.h, ptr(o), ptr(d), ptr(t), len(t), ptr(out)))
        return out

    def camera_rays(self, view, sample: int, seed: int = 0):
        v = np.ascontiguousarray(view, VIEW).reshape(1)
        n = int(v["w"][0]) * int(v["h"][0])
        o = np.zeros((n, 3), np.float32)
        d = np.zeros((n, 3), np.float32)
        t = np.zeros(n, np.float32)
        self._ck(self.lib.gi_generate_camera_rays(self.h, ptr(v), sample, seed, ptr(o), ptr(d), ptr(t)))
        return o, d, t

    # ---- renderer -------------------------------------------------------
    def render(self, views=None, opts=None, **kw):
        views = np.ascontiguousarray(self.scene.views if views is None else views, VIEW).reshape(-1)
        o = np.array(self.scene.opts if opts is None else opts, OPTS).reshape(1).copy()
        for k, v in kw.items():
            o[k] = v
        self.generation += 1
        self._views = None
        self._ck(self.lib.gi_render(self.h, ptr(views), len(views), ptr(o)))
        self._views = views
        self._opts = o
        return self

    def image(self, view: int = 0) -> np.ndarray:
        if getattr(self, "_views", None) is None:
            raise GiError(_lib.GI_ERR_STATE, "no render: render() or progress_begin() first")
        v = self._views[view]
        out = np.zeros((int(v["h"]), int(v["w"]), 3), np.float32)
        self._ck(self.lib.gi_get_image(self.h, view, ptr(out)))
        return out

    def images(self):
        return [self.image(i) for i in range(len(self._views))]

    def progress_begin(self, views=None, opts=None):
        views = np.ascontiguousarray(self.scene.views if views is None else views, VIEW).reshape(-1)
        o = np.array(self.scene.opts if opts is None else opts, OPTS).reshape(1).copy()
        self.generation += 1
        self._views = None
        self._ck(self.lib.gi_progress_begin(self.h, ptr(views), len(views), ptr(o)))
        self._views = views
        self._opts = o

    def progress_step(self) -> bool:
        done = C.c_int32(0)
        self._ck(self.lib.gi_progress_step(self.h, C.byref(done)))
        return bool(done.value)

    def adjoint(self, dL_dimage):
        """Gradient of sum(dL_dimage * image) over the views of the last
        render (concatenated, 3 floats per pixel): (d_mats (n_mats, 5) as
        color rgb / roughness / metallic, d_emission (n_lights, 3))."""
        from ._lib import MATGRAD
        dL = np.ascontiguousarray(dL_dimage, np.float32).reshape(-1)
        nm = len(self.scene.mats)
        nl = len(self.scene.lights)
        dm = np.zeros(nm, MATGRAD)
        de = np.zeros((nl, 3), np.float32)
        self._ck(self.lib.gi_adjoint(self.h, ptr(dL), ptr(dm), ptr(de)))
        g = np.concatenate([dm["d_color"], dm["d_roughness"][:, None], dm["d_metallic"][:, None]], 1)
        return g, de

    def set_materials(self, mats):
        mats = np.ascontiguousarray(mats, MATERIAL)
        self.generation += 1
        self._ck(self.lib.gi_upload_materials(self.h, ptr(mats), len(mats)))
        self.scene.mats = mats

    def set_lights(self, lights):
        lights = np.ascontiguousarray(lights, LIGHT)
        self.generation += 1
        self._ck(self.lib.gi_upload_lights(self.h, ptr(lights), len(lights)))
        self.scene.lights = lights

    def counters(self) -> dict:
        c = np.zeros(1, COUNTERS)
        self._ck(self.lib.gi_get_counters(self.h, ptr(c)))
        return {k: c[k][0].item() for k in COUNTERS.names}

    def check_half_nodes(self) -> int:
        """Child slots of the current tree whose half-precision traversal box
        fails to enclose the f32 box widened by its margins (0 expected)."""
        n = C.c_int64(0)
        self._ck(self.lib.gi_check_half_nodes(self.h, C.byref(n)))
        return n.value

    def traversal_stats(self, enable: bool | None = None) -> dict | None:
        """enable=True zeroes the traversal statistics and counts every
        later trace; enable=False stops counting; with no argument returns
        the totals (rays, interior nodes visited, triangles tested,
        instance entries per kind of trace)."""
        if enable is not None:
            self._ck(self.lib.gi_set_traversal_stats(self.h, 1 if enable else 0))
            return None
        t = np.zeros(1, TRAVERSAL_STATS)
        self._ck(self.lib.gi_get_traversal_stats(self.h, ptr(t)))
        return {k: t[k][0].item() for k in TRAVERSAL_STATS.names}

    def set_tuning(self, **kw):
        t = np.zeros(1, TUNING)
        self._ck(self.lib.gi_get_tuning(self.h, ptr(t)))
        for k, v in kw.items():
            t[k] = v
        self._ck(self.lib.gi_set_tuning(self.h, ptr(t)))

    def tuning(self) -> dict:
        t = np.zeros(1, TUNING)
        self._ck(self.lib.gi_get_tuning(self.h, ptr(t)))
        return {k: int(t[k][0]) for k in TUNING.names}


def build_bvh2(tris: np.ndarray, max_leaf: int = 4):
    """Host binned-SAH BVH (no device needed): (nodes float32[8*n], indices uint32[m])
    in the reference's 2 x float4 layout (src/tracers.cu:26-36)."""
    lib = _lib.load()
    tris = np.ascontiguousarray(tris, TRI)
    h = C.c_void_p()
    rc = lib.gi_bvh_build(ptr(tris), len(tris), max_leaf, C.byref(h))
    if rc != _lib.GI_OK:
        raise GiError(rc, "gi_bvh_build")
    try:
        nn, ni = C.c_int64(), C.c_int64()
        lib.gi_bvh_sizes(h, C.byref(nn), C.byref(ni))
        nodes = np.zeros(8 * nn.value, np.float32)
        idx = np.zeros(max(1, ni.value), np.uint32)
        rc = lib.gi_bvh_export2(h, ptr(nodes), ptr(idx))
        if rc != _lib.GI_OK:
            raise GiError(rc, "gi_bvh_export2")
        return nodes, idx[: ni.value]
    finally:
        lib.gi_bvh_free(h)
