"""Image output through the C-ABI (save_image, src/util.cpp:111-150): PNG
(8-bit, clamp(int(255*v))) and EXR (RGBA half), both axes flipped as the
reference writes them."""
from __future__ import annotations

import os

import numpy as np

from . import _lib
from ._lib import GiError, ptr


def _write(fn, path, rgb):
    img = np.ascontiguousarray(rgb, np.float32)
    if img.ndim != 3 or img.shape[2] != 3:
        raise ValueError("image must be (h, w, 3)")
    h, w = img.shape[:2]
    rc = fn(os.fsencode(path), ptr(img), w, h)
    if rc != _lib.GI_OK:
        raise GiError(rc, f"cannot write {path}")


def save_png(path, rgb):
    _write(_lib.load().gi_write_png, path, rgb)


def save_exr(path, rgb):
    _write(_lib.load().gi_write_exr, path, rgb)


def save(path, rgb):
    """By extension: .png, .exr, or .pfm (float RGB, rows as given)."""
    ext = os.path.splitext(str(path))[1].lower()
    if ext == ".png":
        save_png(path, rgb)
    elif ext == ".exr":
        save_exr(path, rgb)
    elif ext == ".pfm":
        img = np.ascontiguousarray(rgb, np.float32)
        with open(path, "wb") as f:
            f.write(f"PF\n{img.shape[1]} {img.shape[0]}\n-1.0\n".encode())
            f.write(img.tobytes())
    else:
        raise ValueError(f"unknown image extension {ext!r}")


def read_image(path) -> np.ndarray:
    """A texture image (PNG or binary PPM) as (h, w, 4) uint8 RGBA, row 0 the
    file's top row -- the layout gi_upload_texture takes."""
    lib = _lib.load()
    import ctypes as C
    w, h = C.c_int32(), C.c_int32()
    p = os.fsencode(str(path))
    rc = lib.gi_image_size(p, C.byref(w), C.byref(h))
    if rc != _lib.GI_OK:
        raise GiError(rc, f"cannot read image {path}")
    out = np.zeros((h.value, w.value, 4), np.uint8)
    rc = lib.gi_image_read(p, ptr(out))
    if rc != _lib.GI_OK:
        raise GiError(rc, f"cannot decode image {path}")
    return out
