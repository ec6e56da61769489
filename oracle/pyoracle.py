"""ctypes wrapper of the oracle (oracle.h) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline import
this module; the product (globillum_amd/) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# layouts of include/gi_hip.h (duplicated so the oracle does not import the product)
TRI = np.dtype([("a", "<f4", 3), ("b", "<f4", 3), ("c", "<f4", 3), ("na", "<f4", 3), ("nb", "<f4", 3),
                ("nc", "<f4", 3), ("ta", "<f4", 2), ("tb", "<f4", 2), ("tc", "<f4", 2),
                ("material_index", "<i4")])
INSTANCE = np.dtype([("xform", "<f4", 12), ("mesh", "<i4"), ("material", "<i4")])
HIT = np.dtype([("t", "<f4"), ("beta", "<f4"), ("gamma", "<f4"), ("ref", "<u4")])
STATS = np.dtype([(k, "<u8") for k in ("rays_closest", "rays_shadow", "vertices", "samples", "box_closest",
                                       "tri_closest", "box_shadow", "tri_shadow")])

_lib = None


def build() -> str:
    r = subprocess.run(["make", "-C", HERE, "-B", "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB_PATH


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        P, I32, I64, U32, U64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64
        sig = {
            "or_scene_create": (P, [P, I64, P, I32, P, I32, P, I64, P, I64, I32]),
            "or_scene_free": (None, [P]),
            "or_scene_add_mesh": (I32, [P, P, I64, I32]),
            "or_scene_add_texture": (I32, [P, P, I32, I32]),
            "or_texture_levels": (I32, [P, I32]),
            "or_texture_chain": (I64, [P, I32, P]),
            "or_tex_sample": (None, [P, I32, C.c_float, C.c_float, C.c_float, P]),
            "or_tex_level": (I32, [C.c_float, I32, I32, I32]),
            "or_pow22": (C.c_float, [C.c_float]),
            "or_scene_set_instances": (C.c_int, [P, P, I32]),
            "or_scene_bvh_nodes": (I64, [P]),
            "or_scene_export_bvh": (None, [P, P, P]),
            "or_scene_diagonal": (C.c_float, [P]),
            "or_trace_closest": (None, [P, P, P, P, I64, P, P, P]),
            "or_trace_any": (None, [P, P, P, P, I64, P]),
            "or_brute_closest": (None, [P, P, P, P, I64, P]),
            "or_brute_closest_f64": (None, [P, P, P, P, I64, P, P, P, P]),
            "or_intersect_tri": (C.c_int, [P, P, P, P, P, P, P, P]),
            "or_intersect_tri_f64": (C.c_int, [P, P, P, P, P, P, P, P]),
            "or_radical_inverse": (C.c_float, [U64, I32]),
            "or_halton_sample": (C.c_float, [U64, I32, U32]),
            "or_prime": (I32, [I32]),
            "or_halton_perm": (None, [I32, U32, P]),
            "or_camera_rays": (None, [P, I32, U32, U64, U32, P, P, P]),
            "or_tangent_frame": (None, [P, P, P]),
            "or_nee": (None, [P, P, P, P, P, P, P, P]),
            "or_brdf_sample": (C.c_int, [P, P, P, P, P]),
            "or_brdf_eval": (None, [P, C.c_int, P, P, P, P]),
            "or_brdf_sample_f64": (C.c_int, [P, P, P, P, P]),
            "or_brdf_eval_f64": (None, [P, C.c_int, P, P, P, P]),
            "or_render": (None, [P, P, I32, P, P, P, I32, I64, I64]),
            "or_render_f64": (None, [P, P, I32, P, P, I32]),
            "or_adjoint": (C.c_int, [P, P, I32, P, P, P, P, I32, I32]),
            "or_adjoint_tex": (C.c_int, [P, P, I32, P, P, P, P, P, I32, I32]),
            "or_scene_set_sky": (C.c_int, [P, P, I32, I32]),
            "or_sky_eval": (None, [P, P, P]),
            "or_sky_sample": (None, [P, P, P, P, P]),
            "or_atan2_turns": (C.c_float, [C.c_float, C.c_float]),
            "or_cdf_sample": (C.c_float, [P, P, I32, C.c_float, P]),
            "or_gtr2_half_f64": (None, [C.c_double, C.c_double, C.c_double, C.c_double, P]),
            "or_walter_weight_f64": (C.c_double, [P, P, P, C.c_double, C.c_double]),
            "or_gtr1_f64": (C.c_double, [C.c_double, C.c_double]),
            "or_camera_basis": (None, [P, P, P, P, P]),
            "or_phong_pow100": (C.c_float, [C.c_float]),
            "or_phong_root101": (C.c_float, [C.c_float]),
            "or_phong_vertex": (C.c_int, [P, P, P, P, P, P, P, P, P]),
        }
        for k, (res, args) in sig.items():
            f = getattr(lib, k)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


class OracleScene:
    """The oracle's scene: its own object-median BVH unless bvh2 is given."""

    def __init__(self, tris, mats=None, lights=None, bvh2=None, max_leaf=4, sky=None, meshes=None, instances=None,
                 textures=None):
        self.lib = load()
        self._keep = [np.ascontiguousarray(tris)]
        mats = None if mats is None else np.ascontiguousarray(mats)
        lights = None if lights is None else np.ascontiguousarray(lights)
        self._keep += [mats, lights]
        nodes = idx = None
        nn = ni = 0
        if bvh2 is not None:
            nodes = np.ascontiguousarray(bvh2[0], np.float32)
            idx = np.ascontiguousarray(bvh2[1], np.uint32)
            nn, ni = len(nodes) // 8, len(idx)
            self._keep += [nodes, idx]
        self.h = self.lib.or_scene_create(_p(self._keep[0]), len(tris), _p(mats), 0 if mats is None else len(mats),
                                          _p(lights), 0 if lights is None else len(lights), _p(nodes), nn, _p(idx),
                                          ni, max_leaf)
        if sky is not None:
            m = np.ascontiguousarray(sky, np.float32)
            self._keep.append(m)
            if self.lib.or_scene_set_sky(self.h, _p(m), m.shape[1], m.shape[0]) != 0:
                raise ValueError("or_scene_set_sky: the scene needs exactly one sky light")
        self.tex_shapes = []
        for t in textures or []:
            t = np.ascontiguousarray(t, np.uint8)
            self.tex_shapes.append(t.shape[:2])
            if self.lib.or_scene_add_texture(self.h, _p(t), t.shape[1], t.shape[0]) < 0:
                raise ValueError("or_scene_add_texture")
        for m in meshes or []:
            m = np.ascontiguousarray(m)
            self._keep.append(m)
            if self.lib.or_scene_add_mesh(self.h, _p(m), len(m), max_leaf) < 0:
                raise ValueError("or_scene_add_mesh")
        if instances is not None:
            inst = np.ascontiguousarray(instances, INSTANCE)
            if self.lib.or_scene_set_instances(self.h, _p(inst), len(inst)) != 0:
                raise ValueError("or_scene_set_instances: bad mesh id, material or singular transform")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_scene_free(self.h)
            self.h = None

    def texture_chain(self, tex: int = 0):
        """(levels, flat uint8 RGBA chain: level 0 then each halved level)."""
        n = self.lib.or_texture_chain(self.h, tex, None)
        out = np.zeros(n, np.uint8)
        self.lib.or_texture_chain(self.h, tex, _p(out))
        return self.lib.or_texture_levels(self.h, tex), out

    def tex_sample(self, tex: int, u: float, v: float, diff: float):
        out = np.zeros(3, np.float32)
        self.lib.or_tex_sample(self.h, tex, u, v, diff, _p(out))
        return out

    def diagonal(self) -> float:
        return float(self.lib.or_scene_diagonal(self.h))

    def bvh2(self):
        n = self.lib.or_scene_bvh_nodes(self.h)
        nodes = np.zeros(8 * n, np.float32)
        idx = np.zeros(len(self._keep[0]), np.uint32)
        self.lib.or_scene_export_bvh(self.h, _p(nodes), _p(idx))
        return nodes, idx

    @staticmethod
    def _rays(o, d, t):
        return (np.ascontiguousarray(o, np.float32).reshape(-1, 3), np.ascontiguousarray(d, np.float32).reshape(-1, 3),
                np.ascontiguousarray(t, np.float32).reshape(-1))

    def trace_closest(self, o, d, t, with_counts=False):
        o, d, t = self._rays(o, d, t)
        out = np.zeros(len(t), HIT)
        nb, nt = C.c_uint64(), C.c_uint64()
        self.lib.or_trace_closest(self.h, _p(o), _p(d), _p(t), len(t), _p(out), C.byref(nb), C.byref(nt))
        return (out, nb.value, nt.value) if with_counts else out

    def trace_any(self, o, d, t):
        o, d, t = self._rays(o, d, t)
        out = np.zeros(len(t), np.uint8)
        self.lib.or_trace_any(self.h, _p(o), _p(d), _p(t), len(t), _p(out))
        return out

    def brute_closest(self, o, d, t):
        o, d, t = self._rays(o, d, t)
        out = np.zeros(len(t), HIT)
        self.lib.or_brute_closest(self.h, _p(o), _p(d), _p(t), len(t), _p(out))
        return out

    def brute_closest_f64(self, o, d, t):
        o, d, t = self._rays(o, d, t)
        n = len(t)
        tt, bb, gg = np.zeros(n), np.zeros(n), np.zeros(n)
        rr = np.zeros(n, np.uint32)
        self.lib.or_brute_closest_f64(self.h, _p(o), _p(d), _p(t), n, _p(tt), _p(bb), _p(gg), _p(rr))
        return tt, bb, gg, rr

    def render(self, views, opts, threads=0, p_begin=0, p_end=-1):
        views = np.ascontiguousarray(views).reshape(-1)
        opts = np.ascontiguousarray(opts).reshape(1)
        npix = int(sum(int(v["w"]) * int(v["h"]) for v in views))
        img = np.zeros((npix, 3), np.float32)
        st = np.zeros(1, STATS)
        self.lib.or_render(self.h, _p(views), len(views), _p(opts), _p(img), _p(st), threads, p_begin, p_end)
        return img, {k: int(st[k][0]) for k in STATS.names}

    def render_f64(self, views, opts, threads=0):
        views = np.ascontiguousarray(views).reshape(-1)
        opts = np.ascontiguousarray(opts).reshape(1)
        npix = int(sum(int(v["w"]) * int(v["h"]) for v in views))
        img = np.zeros((npix, 3), np.float64)
        self.lib.or_render_f64(self.h, _p(views), len(views), _p(opts), _p(img), threads)
        return img


def adjoint(osc: OracleScene, views, opts, dL, threads=0, f64=True, textures=False):
    """(d_mats (n_mats, 5): color rgb, roughness, metallic; d_emission (n_lights, 3))
    and with textures=True also the list of (h, w, 3) level-0 texel gradients."""
    views = np.ascontiguousarray(views).reshape(-1)
    opts = np.ascontiguousarray(opts).reshape(1)
    dL = np.ascontiguousarray(dL, np.float32).reshape(-1, 3)
    nm = len(osc._keep[1]) if osc._keep[1] is not None else 0
    nl = len(osc._keep[2]) if osc._keep[2] is not None else 0
    dm = np.zeros((nm, 5))
    de = np.zeros((nl, 3))
    dt = np.zeros(sum(3 * h * w for h, w in osc.tex_shapes))
    rc = osc.lib.or_adjoint_tex(osc.h, _p(views), len(views), _p(opts), _p(dL), _p(dm), _p(de),
                                _p(dt) if len(dt) else None, threads, int(f64))
    if rc != 0:
        raise ValueError("or_adjoint: unsupported options")
    if not textures:
        return dm, de
    out, k = [], 0
    for h, w in osc.tex_shapes:
        out.append(dt[k:k + 3 * h * w].reshape(h, w, 3))
        k += 3 * h * w
    return dm, de, out


def sky_eval(osc: OracleScene, d):
    d = np.ascontiguousarray(d, np.float32)
    c = np.zeros(3, np.float32)
    osc.lib.or_sky_eval(osc.h, _p(d), _p(c))
    return c


def sky_sample(osc: OracleScene, xi, N):
    xi, N = np.ascontiguousarray(xi, np.float32), np.ascontiguousarray(N, np.float32)
    d, c = np.zeros(3, np.float32), np.zeros(3, np.float32)
    osc.lib.or_sky_sample(osc.h, _p(xi), _p(N), _p(d), _p(c))
    return d, c


def atan2_turns(y, x):
    return float(load().or_atan2_turns(y, x))


def cdf_sample(cdf, pdf, u):
    cdf, pdf = np.ascontiguousarray(cdf, np.float32), np.ascontiguousarray(pdf, np.float32)
    p = C.c_float()
    x = load().or_cdf_sample(_p(cdf), _p(pdf), len(pdf), u, C.byref(p))
    return float(x), p.value


def radical_inverse(n, base):
    return float(load().or_radical_inverse(n, base))


def halton_sample(n, dim, seed):
    return float(load().or_halton_sample(n, dim, seed))


def prime(dim):
    return int(load().or_prime(dim))


def halton_perm(dim, seed):
    p = np.zeros(prime(dim), np.uint16)
    load().or_halton_perm(dim, seed, _p(p))
    return p


def camera_rays(view, sample, seed, index_base=0, flags=0):
    v = np.ascontiguousarray(view).reshape(1)
    n = int(v["w"][0]) * int(v["h"][0])
    o = np.zeros((n, 3), np.float32)
    d = np.zeros((n, 3), np.float32)
    t = np.zeros(n, np.float32)
    load().or_camera_rays(_p(v), sample, seed, index_base, flags, _p(o), _p(d), _p(t))
    return o, d, t


def camera_basis(view):
    """(U, V, W, (width, height)) of the view's camera."""
    v = np.ascontiguousarray(view).reshape(1)
    U, V, W, wh = (np.zeros(k, np.float32) for k in (3, 3, 3, 2))
    load().or_camera_basis(_p(v), _p(U), _p(V), _p(W), _p(wh))
    return U, V, W, wh


def gtr2_half(u, v, ax, ay):
    h = np.zeros(3)
    load().or_gtr2_half_f64(u, v, ax, ay, _p(h))
    return h


def walter_weight(L, V, m, ax, ay):
    L, V, m = (np.ascontiguousarray(x, np.float64) for x in (L, V, m))
    return float(load().or_walter_weight_f64(_p(L), _p(V), _p(m), ax, ay))


def phong_vertex(kd, ks, N, org_dir, sd, xp):
    """(alive, NEE weight, new direction, throughput factor) of one Phong-mode vertex (FP32)."""
    args = [np.ascontiguousarray(x, np.float32) for x in (kd, ks, N, org_dir, sd, xp)]
    out = [np.zeros(3, np.float32) for _ in range(3)]
    alive = load().or_phong_vertex(*[_p(x) for x in args], *[_p(x) for x in out])
    return bool(alive), out[0], out[1], out[2]


def gtr1(ndh, a):
    return float(load().or_gtr1_f64(ndh, a))


def tangent_frame(n):
    n = np.ascontiguousarray(n, np.float32)
    t = np.zeros(3, np.float32)
    b = np.zeros(3, np.float32)
    load().or_tangent_frame(_p(n), _p(t), _p(b))
    return t, b


def intersect_tri(a, b, c, o, d, f64=False):
    dt = np.float64 if f64 else np.float32
    arrs = [np.ascontiguousarray(x, dt) for x in (a, b, c, o, d)]
    t, u, v = (C.c_double() if f64 else C.c_float() for _ in range(3))
    fn = load().or_intersect_tri_f64 if f64 else load().or_intersect_tri
    hit = fn(*[_p(x) for x in arrs], C.byref(t), C.byref(u), C.byref(v))
    return bool(hit), t.value, u.value, v.value


def nee(scene: OracleScene, P, N, xi):
    P, N, xi = (np.ascontiguousarray(x, np.float32) for x in (P, N, xi))
    d = np.zeros(3, np.float32)
    c = np.zeros(3, np.float32)
    ln = C.c_float()
    li = C.c_int32()
    scene.lib.or_nee(scene.h, _p(P), _p(N), _p(xi), _p(d), C.byref(ln), _p(c), C.byref(li))
    return d, ln.value, c, li.value


def brdf_sample(mat, wo_ts, xi, f64=False):
    dt = np.float64 if f64 else np.float32
    m = np.ascontiguousarray(mat).reshape(1)
    wo, x = np.ascontiguousarray(wo_ts, dt), np.ascontiguousarray(xi, dt)
    wi = np.zeros(3, dt)
    pdf = C.c_double() if f64 else C.c_float()
    fn = load().or_brdf_sample_f64 if f64 else load().or_brdf_sample
    typ = fn(_p(m), _p(wo), _p(x), _p(wi), C.byref(pdf))
    return typ, wi, pdf.value


def brdf_eval(mat, typ, wo, wi, n, f64=False):
    dt = np.float64 if f64 else np.float32
    m = np.ascontiguousarray(mat).reshape(1)
    wo, wi, n = (np.ascontiguousarray(x, dt) for x in (wo, wi, n))
    f = np.zeros(3, dt)
    fn = load().or_brdf_eval_f64 if f64 else load().or_brdf_eval
    fn(_p(m), typ, _p(wo), _p(wi), _p(n), _p(f))
    return f


def pow22(x: float) -> float:
    return float(load().or_pow22(x))


def tex_level(diff: float, w: int, h: int, levels: int) -> int:
    return int(load().or_tex_level(diff, w, h, levels))
