"""Scene ingestion through the C-ABI: Wavefront OBJ / MTL files into the
triangle and material tables (add_objfile_to_flat_tri_list,
src/hl-main.cpp:277-350) and .pbrdf principled parameters
(PrincipledBRDFParameters::readFrom, src/principledParameters.h:64-125;
convert_and_upload_materials, src/material-conversion.cpp:59-151)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import MATERIAL, TRI, GiError, ptr
from .scenes import Scene


def load_obj(paths, trafo=None, pbrdf_dir: str | None = None, with_textures: bool = False):
    """Read one or more OBJ files into one table.  Returns (tris, mats,
    names): gi_tri records, gi_material records (index 0 = the fallback
    material) and the material names; with_textures also the list of map_Kd
    files that gi_material.texture indexes.  trafo: optional 4x4
    (row-major) applied to every file.  pbrdf_dir: directory searched for
    <material name>.pbrdf files."""
    if isinstance(paths, (str, os.PathLike)):
        paths = [paths]
    lib = _lib.load()
    h = C.c_void_p()
    if lib.gi_obj_open(C.byref(h)) != _lib.GI_OK:
        raise GiError(1, "gi_obj_open")
    try:
        m = None
        if trafo is not None:
            m = np.ascontiguousarray(np.asarray(trafo, np.float32).reshape(16))
        for p in paths:
            rc = lib.gi_obj_append(h, os.fsencode(p), None if m is None else ptr(m))
            if rc != _lib.GI_OK:
                raise GiError(rc, lib.gi_obj_error(h).decode())
        if pbrdf_dir is not None:
            rc = lib.gi_obj_apply_pbrdf(h, os.fsencode(pbrdf_dir))
            if rc != _lib.GI_OK:
                raise GiError(rc, lib.gi_obj_error(h).decode())
        nt, nm = C.c_int64(), C.c_int32()
        lib.gi_obj_sizes(h, C.byref(nt), C.byref(nm))
        tris = np.zeros(nt.value, TRI)
        mats = np.zeros(nm.value, MATERIAL)
        lib.gi_obj_export(h, ptr(tris) if len(tris) else None, ptr(mats) if len(mats) else None)
        names = [lib.gi_obj_material_name(h, i).decode() for i in range(nm.value)]
        if with_textures:
            tex = [lib.gi_obj_texture_path(h, i).decode() for i in range(lib.gi_obj_texture_count(h))]
            return tris, mats, names, tex
        return tris, mats, names
    finally:
        lib.gi_obj_free(h)


def read_pbrdf(path: str, base=None) -> np.ndarray:
    """One .pbrdf file applied to a material record (a Lambert record with
    diffuse colour (1,1,1) by default)."""
    lib = _lib.load()
    if base is None:
        base = np.zeros(1, MATERIAL)
        base["diffuse_color"] = 1.0
        base["texture"] = -1
    m = np.array(base, MATERIAL).reshape(1).copy()
    rc = lib.gi_pbrdf_read(os.fsencode(path), ptr(m))
    if rc != _lib.GI_OK:
        raise GiError(rc, f"gi_pbrdf_read({path})")
    return m


def obj_scene(paths, views, lights, opts, trafo=None, pbrdf_dir: str | None = None, name: str = "",
              textures: str = "load") -> Scene:
    """A renderable Scene from OBJ files plus the camera(s), lights and
    options a scene script would give.  textures="load" decodes the map_Kd
    images (PNG / PPM; a missing or unreadable file is an error),
    "ignore" renders those materials untextured."""
    from .image_io import read_image

    tris, mats, names, tex = load_obj(paths, trafo, pbrdf_dir, with_textures=True)
    if textures == "ignore":
        mats["texture"] = -1
        images = None
    elif textures == "load":
        images = [read_image(p) for p in tex] or None
    else:
        raise ValueError(f"textures must be 'load' or 'ignore', not {textures!r}")
    sc = Scene(tris, mats, np.array(lights).reshape(-1), np.array(views).reshape(-1), np.array(opts).reshape(1),
               name=name or os.path.basename(str(paths if isinstance(paths, (str, os.PathLike)) else paths[0])),
               textures=images)
    sc.meta["material_names"] = names
    sc.meta["texture_files"] = tex
    return sc
