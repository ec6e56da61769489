"""ctypes binding of libgi_hip.so (include/gi_hip.h).

Record layouts are numpy structured dtypes with exactly the C layout, so
arrays of triangles / materials / lights / views pass straight through.
The library is built in-tree by globillum_amd.build; loading fails loudly
when it is missing -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libgi_hip.so")

TRI = np.dtype([("a", "<f4", 3), ("b", "<f4", 3), ("c", "<f4", 3), ("na", "<f4", 3), ("nb", "<f4", 3),
                ("nc", "<f4", 3), ("ta", "<f4", 2), ("tb", "<f4", 2), ("tc", "<f4", 2),
                ("material_index", "<i4")])
MATERIAL = np.dtype([("diffuse_color", "<f4", 3), ("specular_color", "<f4", 3), ("alpha", "<f4"),
                     ("principled", "<i4"), ("metallic", "<f4"), ("subsurface", "<f4"), ("specular", "<f4"),
                     ("roughness", "<f4"), ("specular_tint", "<f4"), ("anisotropic", "<f4"), ("sheen", "<f4"),
                     ("sheen_tint", "<f4"), ("clearcoat", "<f4"), ("clearcoat_gloss", "<f4"), ("opacity", "<f4"),
                     ("transmission", "<f4"), ("color", "<f4", 3), ("texture", "<i4")])
LIGHT = np.dtype([("center", "<f4", 3), ("dir", "<f4", 3), ("up", "<f4", 3), ("col", "<f4", 3),
                  ("wh", "<f4", 2), ("power", "<f4"), ("type", "<i4")])
VIEW = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("up", "<f4", 3), ("fovy", "<f4"), ("w", "<i4"),
                 ("h", "<i4"), ("spp", "<i4"), ("aperture", "<f4"), ("focus_distance", "<f4")])
OPTS = np.dtype([("max_depth", "<i4"), ("rr_start", "<i4"), ("seed", "<u4"), ("flags", "<u4"),
                 ("offset_bounce", "<f4"), ("offset_shadow", "<f4"), ("background", "<f4", 3),
                 ("samples_per_pass", "<i4")])
HIT = np.dtype([("t", "<f4"), ("beta", "<f4"), ("gamma", "<f4"), ("ref", "<u4")])
COUNTERS = np.dtype([("rays_closest", "<u8"), ("rays_shadow", "<u8"), ("samples", "<u8"),
                     ("path_vertices", "<u8"), ("ms_total", "<f8"), ("ms_trace_closest", "<f8"),
                     ("ms_trace_any", "<f8"), ("ms_shade", "<f8"), ("ms_graph_build", "<f8"),
                     ("launches_closest", "<i4"), ("passes", "<i4")])
TUNING = np.dtype([("trace_block", "<i4"), ("trace_blocks_per_cu", "<i4"), ("shade_block", "<i4"),
                   ("samples_per_pass", "<i4"), ("stack_lds", "<i4"), ("target_slots", "<i4"),
                   ("node_width", "<i4"), ("refill_lanes", "<i4")])
INSTANCE = np.dtype([("xform", "<f4", 12), ("mesh", "<i4"), ("material", "<i4")])
TRAVERSAL_STATS = np.dtype([(f"{kind}_{what}", "<u8") for kind in ("closest", "any")
                            for what in ("rays", "nodes", "triangles", "instances")])
MATGRAD = np.dtype([("d_color", "<f4", 3), ("d_roughness", "<f4"), ("d_metallic", "<f4")])

assert TRI.itemsize == 100 and MATERIAL.itemsize == 96 and LIGHT.itemsize == 64
assert VIEW.itemsize == 60 and OPTS.itemsize == 40 and HIT.itemsize == 16 and COUNTERS.itemsize == 80

GI_OK = 0
GI_INVALID_REF = 0xFFFFFFFF
FLAG_BATCHED = 0x1
FLAG_GRAPH = 0x2
FLAG_PHONG = 0x4
FLAG_PIXEL_CENTRE = 0x8
STATUS = {0: "ok", 1: "invalid argument", 2: "no HIP device", 3: "HIP runtime error", 4: "invalid call order",
          5: "out of device memory", 6: "size beyond a documented limit", 7: "unknown name"}

# Every symbol include/gi_hip.h declares, with (restype, argtypes).
_P = C.c_void_p
_I = C.c_int
_I32 = C.c_int32
_I64 = C.c_int64
_U32 = C.c_uint32
_U64 = C.c_uint64
_F = C.c_float
SIGNATURES = {
    "gi_abi_version": (_I, []),
    "gi_status_string": (C.c_char_p, [_I]),
    "gi_create": (_I, [C.POINTER(_P), _I]),
    "gi_destroy": (None, [_P]),
    "gi_last_error": (C.c_char_p, [_P]),
    "gi_bvh_build": (_I, [_P, _I64, _I32, C.POINTER(_P)]),
    "gi_bvh_free": (None, [_P]),
    "gi_bvh_sizes": (_I, [_P, C.POINTER(_I64), C.POINTER(_I64)]),
    "gi_bvh_export2": (_I, [_P, _P, _P]),
    "gi_upload_triangles": (_I, [_P, _P, _I64]),
    "gi_upload_materials": (_I, [_P, _P, _I32]),
    "gi_upload_lights": (_I, [_P, _P, _I32]),
    "gi_build_bvh": (_I, [_P, _I32]),
    "gi_upload_bvh2": (_I, [_P, _P, _I64, _P, _I64]),
    "gi_trace_closest": (_I, [_P, _P, _P, _P, _I64, _P]),
    "gi_trace_any": (_I, [_P, _P, _P, _P, _I64, _P]),
    "gi_trace_closest_device": (_I, [_P, _P, _P, _P, _I64, _P]),
    "gi_trace_any_device": (_I, [_P, _P, _P, _P, _I64, _P]),
    "gi_trace_closest_device_noalloc": (_I, [_P, _P, _P, _P, _I64, _P]),
    "gi_generate_camera_rays": (_I, [_P, _P, _I32, _U32, _P, _P, _P]),
    "gi_render": (_I, [_P, _P, _I32, _P]),
    "gi_get_image": (_I, [_P, _I32, _P]),
    "gi_get_image_device": (_I, [_P, _I32, C.POINTER(_P)]),
    "gi_get_counters": (_I, [_P, _P]),
    "gi_set_tuning": (_I, [_P, _P]),
    "gi_get_tuning": (_I, [_P, _P]),
    "gi_progress_begin": (_I, [_P, _P, _I32, _P]),
    "gi_progress_step": (_I, [_P, C.POINTER(_I32)]),
    "gi_adjoint": (_I, [_P, _P, _P, _P]),
    "gi_halton_sample": (_F, [_U64, _I32, _U32]),
    "gi_radical_inverse": (_F, [_U64, _I32]),
    "gi_halton_prime": (_I32, [_I32]),
    "gi_halton_permutation": (_I, [_I32, _U32, _P]),
    "gi_rect_light_power": (_F, [_P]),
    "gi_sky_light_power": (_F, [_F, _F]),
    "gi_upload_sky": (_I, [_P, _P, _I32, _I32]),
    "gi_build_bvh_device": (_I, [_P, _I32]),
    "gi_build_bvh_device_ex": (_I, [_P, _I32, _I32, _I32]),
    "gi_set_traversal_stats": (_I, [_P, _I32]),
    "gi_get_traversal_stats": (_I, [_P, _P]),
    "gi_check_half_nodes": (_I, [_P, C.POINTER(_I64)]),
    "gi_upload_mesh": (_I, [_P, _P, _I64, C.POINTER(C.c_int32)]),
    "gi_upload_texture": (_I, [_P, _P, _I32, _I32, C.POINTER(C.c_int32)]),
    "gi_texture_info": (_I, [_P, _I32, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "gi_download_texture": (_I, [_P, _I32, _P]),
    "gi_texture_grad": (_I, [_P, _I32, _P]),
    "gi_image_size": (_I, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "gi_image_read": (_I, [C.c_char_p, _P]),
    "gi_obj_texture_count": (_I32, [_P]),
    "gi_obj_texture_path": (C.c_char_p, [_P, _I32]),
    "gi_upload_instances": (_I, [_P, _P, _I32]),
    "gi_obj_open": (_I, [C.POINTER(_P)]),
    "gi_obj_append": (_I, [_P, C.c_char_p, _P]),
    "gi_obj_apply_pbrdf": (_I, [_P, C.c_char_p]),
    "gi_obj_sizes": (_I, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "gi_obj_export": (_I, [_P, _P, _P]),
    "gi_obj_material_name": (C.c_char_p, [_P, _I32]),
    "gi_obj_error": (C.c_char_p, [_P]),
    "gi_obj_free": (None, [_P]),
    "gi_pbrdf_read": (_I, [C.c_char_p, _P]),
    "gi_write_png": (_I, [C.c_char_p, _P, _I32, _I32]),
    "gi_write_exr": (_I, [C.c_char_p, _P, _I32, _I32]),
}

_lib = None


def load() -> C.CDLL:
    """Load libgi_hip.so (built in-tree); raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m globillum_amd.build` first")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class GiError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the C-ABI must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)
