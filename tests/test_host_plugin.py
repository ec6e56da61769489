"""The C++ host side: the gi_algorithm registry semantics (CPU, no device
call) and the mi355x_pt plugin driven by the headless driver (GPU), whose
image must equal the oracle's bit for bit."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from globillum_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "globillum_amd")


def test_registry_semantics():
    r = subprocess.run([os.path.join(PKG, "registry_test")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "registry test ok" in r.stdout


def test_scene_dump_layout(tmp_path):
    sc = scenes.cornell_boxes(w=8, h=8, spp=1)
    p = tmp_path / "c1.gisc"
    scenes.save_scene(sc, str(p))
    assert p.stat().st_size == 4 + 4 + 8 + 12 + 34 * 100 + 7 * 96 + 64 + 60 + 40


def _read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        f.readline()
        return np.fromfile(f, "<f4").reshape(h, w, 3)


@pytest.mark.gpu
def test_driver_progressive_render_matches_oracle(tmp_path, oracle):
    sc = scenes.cornell_boxes(w=80, h=60, spp=6, depth=4)
    path = tmp_path / "c1.gisc"
    scenes.save_scene(sc, str(path))
    out = tmp_path / "c1.pfm"
    r = subprocess.run([os.path.join(PKG, "gi_driver"), str(path), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["steps"] == 6 and info["samples"] == 80 * 60 * 6
    img = _read_pfm(out).reshape(-1, 3)
    ref, st = oracle.OracleScene(sc.tris, sc.mats, sc.lights).render(sc.views, sc.opts)
    assert np.array_equal(img, ref)
    assert info["rays_closest"] == st["rays_closest"] and info["rays_shadow"] == st["rays_shadow"]


@pytest.mark.gpu
def test_driver_writes_png_and_exr(tmp_path, oracle):
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_image_io import _read_exr, _read_png

    sc = scenes.cornell_boxes(w=40, h=30, spp=2, depth=3)
    path = tmp_path / "c1.gisc"
    scenes.save_scene(sc, str(path))
    ref, _ = oracle.OracleScene(sc.tris, sc.mats, sc.lights).render(sc.views, sc.opts)
    ref = ref.reshape(30, 40, 3)
    for ext in ("png", "exr"):
        out = tmp_path / f"c1.{ext}"
        r = subprocess.run([os.path.join(PKG, "gi_driver"), str(path), str(out)], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        if ext == "png":
            want = np.where(ref > 0, np.minimum(np.trunc(255.0 * ref), 255), 0).astype(np.uint8)[::-1, ::-1]
            assert np.array_equal(_read_png(out), want)
        else:
            got = _read_exr(out)
            assert np.array_equal(got[..., 3], ref.astype(np.float16)[::-1, ::-1, 0])
