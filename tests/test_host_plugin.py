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


def _combined(w=48, h=32, spp=3):
    """Instanced geometry, textures and a sky light in one scene."""
    import dataclasses

    sc = scenes.textured_scene(w=w, h=h, spp=spp, depth=3)
    return dataclasses.replace(sc, lights=np.concatenate([sc.lights, scenes.make_sky_light(0.5, 40.0)]),
                               sky=scenes.sky_map(32, 16))


def test_scene_dump_v2_sections(tmp_path):
    sc = _combined()
    p = tmp_path / "x.gisc"
    scenes.save_scene(sc, str(p))
    b = p.read_bytes()
    assert b[:4] == b"GISC" and int.from_bytes(b[4:8], "little") == 2 and b.endswith(b"END ")
    base = 4 + 4 + 8 + 12 + len(sc.tris) * 100 + len(sc.mats) * 96 + len(sc.lights) * 64 + 60 + 40
    sky = 4 + 8 + 32 * 16 * 12
    tex = sum(4 + 8 + t.size for t in sc.textures)
    mesh = sum(4 + 8 + len(m) * 100 for m in sc.meshes)
    inst = 4 + 4 + len(sc.instances) * 56
    assert len(b) == base + sky + tex + mesh + inst + 4


@pytest.mark.gpu
def test_driver_renders_instances_textures_and_sky(tmp_path, oracle):
    sc = _combined()
    path = tmp_path / "x.gisc"
    scenes.save_scene(sc, str(path))
    out = tmp_path / "x.pfm"
    r = subprocess.run([os.path.join(PKG, "gi_driver"), str(path), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    img = _read_pfm(out).reshape(-1, 3)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights, sky=sc.sky, textures=sc.textures, meshes=sc.meshes,
                             instances=sc.instances)
    ref, st = osc.render(sc.views, sc.opts)
    assert img.max() > 0 and np.array_equal(img, ref)
