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
