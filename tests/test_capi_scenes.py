"""The C-ABI library loads and exports every symbol include/gi_hip.h
declares; host-only entry points behave; scene generators are
deterministic and have the configured sizes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from globillum_amd import _lib, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "gi_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gi_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    names = _declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_read_benchmark_library_loads():
    """bench.py's memory read-bandwidth kernel: built beside the library,
    exports its one entry point, and refuses bad sizes before touching a
    device."""
    lib = C.CDLL(os.path.join(ROOT, "globillum_amd", "libgi_membench.so"))
    f = lib.gi_membench_read
    f.restype = C.c_int
    f.argtypes = [C.c_int32, C.c_int64, C.c_int32, C.POINTER(C.c_double)]
    g = C.c_double(0.0)
    for nbytes, passes in ((1024, 1), (16 * 1024 + 8, 1), (1 << 20, 0)):
        assert f(0, nbytes, passes, C.byref(g)) != 0
    assert f(0, 1 << 20, 1, None) != 0


def test_version_and_status_strings():
    lib = _lib.load()
    assert lib.gi_abi_version() == 1
    assert lib.gi_status_string(0) == b"ok"
    assert lib.gi_status_string(4) == b"invalid call order"
    assert lib.gi_status_string(99) == b"unknown status"


def test_create_without_device_reports_status():
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.gi_create(None, 0) == 1
    rc = lib.gi_create(C.byref(h), 0)
    if rc == 0:
        lib.gi_destroy(h)  # a GPU is present
    else:
        assert rc == 2 and not h.value


def test_null_context_calls_fail_cleanly():
    lib = _lib.load()
    assert lib.gi_upload_triangles(None, None, 0) == 1
    assert lib.gi_render(None, None, 0, None) != 0
    assert lib.gi_get_counters(None, None) == 1
    assert lib.gi_set_traversal_stats(None, 1) == 1 and lib.gi_get_traversal_stats(None, None) == 1
    assert lib.gi_last_error(None) == b"null context"
    lib.gi_destroy(None)


def test_rect_light_power_matches_reference_formula():
    lib = _lib.load()
    l = scenes.make_light((0, 1, 0), (0, -1, 0), (10, 20, 30), 0.5, 0.25)
    # add-rectlight: w*h*(r+g+b)*.333 (src/lights.cpp:92)
    assert lib.gi_rect_light_power(_lib.ptr(l)) == pytest.approx(0.5 * 0.25 * 60 * 0.333, rel=1e-6)
    assert l["power"][0] == pytest.approx(0.5 * 0.25 * 60 * 0.333, rel=1e-6)


def test_config_scenes():
    c1 = scenes.cornell_boxes()
    assert len(c1.tris) == 34 and len(c1.mats) == 7 and len(c1.lights) == 1
    assert (c1.mats["diffuse_color"] >= 0.2).all() and (c1.mats["diffuse_color"] <= 0.8).all()
    assert np.array_equal(c1.tris, scenes.cornell_boxes().tris)
    assert not np.array_equal(c1.tris, scenes.cornell_boxes(seed=2).tris)
    c2 = scenes.config_scene(2)
    assert len(c2.tris) == 256 * 256 * 2 + 1000 * 160
    assert len(c2.mats) == 32 and len(c2.lights) == 4
    assert int(c2.views["w"][0]) == 1280 and int(c2.views["spp"][0]) == 64 and int(c2.opts["max_depth"][0]) == 8
    m = c2.mats
    assert set(np.unique(m["metallic"]).tolist()) <= {0.0, 1.0}
    assert (m["roughness"] >= 0.05).all() and (m["roughness"] <= 1).all()
    c4 = scenes.config_scene(4)
    assert len(c4.tris) == 200 * 250 * 2 + 50 * 160 and len(c4.mats) == 64


def test_sphere_normals_are_radial():
    s = scenes.uv_sphere((1.0, 2.0, 3.0), 0.5)
    assert len(s) == 160
    for k in ("a", "b", "c"):
        r = s[k] - np.array([1.0, 2.0, 3.0], np.float32)
        assert np.allclose(np.linalg.norm(r, axis=1), 0.5, atol=1e-6)
        assert np.allclose(s["n" + k], r / 0.5, atol=1e-5)


def test_context_does_not_alias_the_callers_scene():
    # set_materials / set_lights on a context must leave the uploaded Scene
    # object untouched (checked without a device: the copy is made host-side)
    import dataclasses
    from globillum_amd.context import Context
    sc = scenes.cornell_boxes(w=4, h=4, spp=1)
    ctx = Context.__new__(Context)
    ctx.scene = None
    copy = dataclasses.replace(sc, mats=sc.mats.copy(), lights=sc.lights.copy())
    assert copy.mats is not sc.mats and np.array_equal(copy.mats, sc.mats)


def test_integration_record_checks(tmp_path):
    """The compile-time record checks of the integration binding against two
    stand-ins for rta's records (tests/csrc/layout_check.cpp): 4-byte-aligned
    texture coordinates pass straight through, an 8-byte-aligned float2
    (simple_triangle padded to 104 bytes) is copied field by field; the
    uploaded gi_tri records hold the same values either way."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path / "layout_check")
    r = subprocess.run([cxx, "-std=c++17", "-Wall", "-Werror", "-o", exe,
                        os.path.join(ROOT, "tests", "csrc", "layout_check.cpp")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    assert r.stdout.split("\n")[:2] == ["vec2f: sizeof 100, cast, ok", "float2: sizeof 104, copy, ok"]
