"""Scene ingestion (OBJ / MTL / .pbrdf): the product's C++ reader against the
plain-Python restatement in oracle/obj_oracle.py, closed-form geometry
checks, and the reference's own asset files when they are present."""
import os

import numpy as np
import pytest

from globillum_amd import _lib, scene_io

import obj_oracle

REF = "/root/reference"


def _compare(path_list, trafo=None, pbrdf_dir=None):
    tris, mats, names, tex = scene_io.load_obj(path_list, trafo, pbrdf_dir, with_textures=True)
    o = obj_oracle.ObjTable()
    for p in path_list:
        o.append(p, trafo)
    if pbrdf_dir:
        o.apply_pbrdf(pbrdf_dir)
    assert names == o.names
    assert tex == o.textures
    assert len(tris) == len(o.tris)
    for k in ("a", "b", "c", "na", "nb", "nc", "ta", "tb", "tc"):
        want = np.array([t[k] for t in o.tris], np.float32).reshape(len(o.tris), -1)
        assert np.array_equal(tris[k], want), k
    assert np.array_equal(tris["material_index"], np.array([t["mat"] for t in o.tris], np.int32))
    for i, m in enumerate(o.mats):
        assert np.array_equal(mats[i]["diffuse_color"], np.float32(m["kd"]))
        assert np.array_equal(mats[i]["specular_color"], np.float32(m["ks"]))
        assert mats[i]["alpha"] == np.float32(m["alpha"])
        assert mats[i]["texture"] == m["texture"]
        if "pbrdf" in m:
            assert mats[i]["principled"] == 1
            for f, v in m["pbrdf"].items():
                assert np.array_equal(mats[i][f], np.float32(v)), (names[i], f)
        else:
            assert mats[i]["principled"] == 0
    return tris, mats, names


@pytest.fixture
def objdir(tmp_path):
    (tmp_path / "a.mtl").write_text(
        "# materials\nnewmtl red\nKd 0.8 0.1 0.1\nKs 0.2 0.2 0.2\nd 0.5\n\n"
        "newmtl tex\nKd 0.3 0.6 0.9\nTr 0.25\nmap_Kd some/file.png\n\n"
        "newmtl shiny metal\nKd 0.9 0.9 0.9\n")
    (tmp_path / "a.obj").write_text(
        "mtllib a.mtl\n"
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0 0 1\n"
        "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
        "vn 0 0 1\nvn 0 1 0\n"
        "f 1 2 3\n"                      # no material yet, no normals
        "usemtl red\n"
        "f 1/1/1 2/2/1 3/3/1 4/4/1\n"    # quad -> fan of 2
        "usemtl tex\n"
        "f 1/1 2/2 5/3\n"                # texture coordinates kept, geometric normal
        "f -5//2 -4//2 -1//2\n"          # negative (relative) references
        "usemtl undefined_name\n"
        "f 2 3 5\n"                      # undefined -> fallback
        "usemtl shiny metal\n"
        "f 1 3 5 4 2\n")                 # pentagon -> fan of 3
    (tmp_path / "b.obj").write_text(
        "mtllib a.mtl\nv 2 0 0\nv 3 0 0\nv 2 1 0\nusemtl red\nf 1 2 3\nusemtl tex\nf 3 2 1\n")
    (tmp_path / "mats").mkdir()
    (tmp_path / "mats" / "red.pbrdf").write_text(
        "color 0.5 0.7 0.8\nroughness 0.2\nmetallic 1.0\nopacity 1.0\nsubsurface 0.0\nspecular 0.2\n"
        "tintSpecular 0.1\nanisotropic 0.0\nsheen 0.3\ntintSheen 0.4\nclearcoat 0.2\nglossClearcoat 0.6\n"
        "transmission0.0\n")
    (tmp_path / "mats" / "shiny metal.pbrdf").write_text("metallic 0.75 roughness 0.4\n")
    return tmp_path


def test_obj_matches_oracle(objdir):
    tris, mats, names = _compare([str(objdir / "a.obj")])
    assert names == ["gi/fallback", "red", "tex", "shiny metal"]
    # 1 + 2 + 1 + 1 + 1 + 3 triangles
    assert len(tris) == 9
    assert list(tris["material_index"]) == [0, 1, 1, 2, 2, 0, 3, 3, 3]
    # fallback material is red/red, alpha from d and from Tr
    assert np.array_equal(mats[0]["diffuse_color"], [1, 0, 0]) and np.array_equal(mats[0]["specular_color"], [1, 0, 0])
    assert mats[1]["alpha"] == np.float32(0.5) and mats[2]["alpha"] == np.float32(0.75)
    # map_Kd: the texture list holds the file (relative to the MTL's directory)
    _, mats, _, tex = scene_io.load_obj(str(objdir / "a.obj"), with_textures=True)
    assert tex == [str(objdir) + "/some/file.png"] and list(mats["texture"]) == [-1, -1, 0, -1]


def test_obj_closed_forms(objdir):
    tris, _, _ = scene_io.load_obj(str(objdir / "a.obj"))
    # fan of the quad 1 2 3 4: (1,2,3), (1,3,4)
    assert np.array_equal(tris["a"][1], [0, 0, 0]) and np.array_equal(tris["c"][1], [1, 1, 0])
    assert np.array_equal(tris["b"][2], [1, 1, 0]) and np.array_equal(tris["c"][2], [0, 1, 0])
    # geometric normal of (0,0,0),(1,0,0),(1,1,0) is +z (right-handed)
    assert np.array_equal(tris["na"][0], [0, 0, 1]) and np.array_equal(tris["nc"][0], [0, 0, 1])
    # texture coordinates only where the material has a diffuse texture
    assert not tris["ta"][1].any() and not tris["tc"][2].any()
    assert np.array_equal(tris["tb"][3], [1, 0]) and np.array_equal(tris["tc"][3], [1, 1])
    # relative references -5, -4, -1 of 5 vertices are 1, 2, 5
    assert np.array_equal(tris["c"][4], [0, 0, 1]) and np.array_equal(tris["na"][4], [0, 1, 0])
    # the unit-square fan covers area 1
    cr = np.cross(tris["b"][1:3] - tris["a"][1:3], tris["c"][1:3] - tris["a"][1:3])
    assert np.isclose(0.5 * np.linalg.norm(cr, axis=1).sum(), 1.0)


def test_multiple_files_share_materials(objdir):
    tris, mats, names = _compare([str(objdir / "a.obj"), str(objdir / "b.obj")])
    assert names == ["gi/fallback", "red", "tex", "shiny metal"]
    assert list(tris["material_index"][-2:]) == [1, 2]
    assert len(tris) == 11


def test_transform(objdir):
    M = np.array([[0, -2, 0, 1], [2, 0, 0, 2], [0, 0, 2, 3], [0, 0, 0, 1]], np.float32)
    tris, _, _ = _compare([str(objdir / "a.obj")], trafo=M)
    # vertex (1,0,0) -> (1,4,3); normal (0,1,0) -> rotated by +90 degrees about z: (-1,0,0)
    assert np.array_equal(tris["b"][0], [1, 4, 3])
    assert np.allclose(tris["na"][4], [-1, 0, 0], atol=1e-7)
    assert np.allclose(np.linalg.norm(tris["na"], axis=1), 1.0, atol=1e-6)


def test_pbrdf(objdir):
    tris, mats, names = _compare([str(objdir / "a.obj")], pbrdf_dir=str(objdir / "mats"))
    red = mats[names.index("red")]
    assert red["principled"] == 1 and red["metallic"] == 1.0 and red["roughness"] == np.float32(0.2)
    assert np.array_equal(red["color"], np.float32([0.5, 0.7, 0.8]))
    # "transmission0.0" is one unknown token: transmission keeps its default 1
    assert red["transmission"] == 1.0 and red["opacity"] == 1.0
    sm = mats[names.index("shiny metal")]
    # unset color defaults to the material's diffuse colour
    assert sm["metallic"] == 0.75 and np.array_equal(sm["color"], np.float32([0.9, 0.9, 0.9]))
    assert mats[names.index("tex")]["principled"] == 0


def test_errors(objdir, tmp_path):
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nf 1 2 7\n")
    with pytest.raises(_lib.GiError, match="bad face reference 7"):
        scene_io.load_obj(str(bad))
    bad.write_text("mtllib missing.mtl\n")
    with pytest.raises(_lib.GiError, match="cannot open material library"):
        scene_io.load_obj(str(bad))
    with pytest.raises(_lib.GiError, match="cannot open"):
        scene_io.load_obj(str(tmp_path / "nope.obj"))
    p = tmp_path / "bad.pbrdf"
    p.write_text("color 1 x 1\n")
    with pytest.raises(_lib.GiError):
        scene_io.read_pbrdf(str(p))


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference asset files not present")
def test_reference_assets():
    tris, mats, names = _compare([os.path.join(REF, "plainSphere.obj")], pbrdf_dir=os.path.join(REF, "materials"))
    # 720 triangles on the unit sphere; usemtl "default" is not in its MTL -> fallback
    assert len(tris) == 720 and names == ["gi/fallback"] and (tris["material_index"] == 0).all()
    for k in ("a", "b", "c"):
        assert np.allclose(np.linalg.norm(tris[k], axis=1), 1.0, atol=1e-5)
        assert np.allclose(np.linalg.norm(tris["n" + k], axis=1), 1.0, atol=1e-3)
    tris, mats, names = _compare([os.path.join(REF, "boxes.obj")], pbrdf_dir=os.path.join(REF, "materials"))
    assert len(tris) == 22752 and names == ["gi/fallback", "None"]
    assert np.array_equal(mats[1]["diffuse_color"], np.float32([0.7, 0.0, 0.0]))
    # every .pbrdf file the reference ships parses as the oracle reads it
    d = os.path.join(REF, "materials")
    for f in sorted(os.listdir(d)):
        if not f.endswith(".pbrdf"):
            continue
        m = scene_io.read_pbrdf(os.path.join(d, f))[0]
        want = obj_oracle.read_pbrdf(os.path.join(d, f))
        for k, v in want.items():
            assert np.array_equal(m[k], np.float32(v)), (f, k)


def test_reference_material_fixture_is_current(tmp_path):
    """tests/golden/reference_materials.json holds what the reader makes of
    the reference's material assets (checked where the reference is)."""
    import json
    import shutil
    root = "/root/reference"
    if not os.path.isdir(os.path.join(root, "materials")):
        pytest.skip("reference tree not present")
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    committed = json.load(open(os.path.join(here, "reference_materials.json")))
    shutil.copy(os.path.join(here, "make_reference_materials.py"), tmp_path)
    import importlib.util
    spec = importlib.util.spec_from_file_location("mkref", os.path.join(here, "make_reference_materials.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.HERE = str(tmp_path)
    mod.main(root)
    assert json.load(open(tmp_path / "reference_materials.json")) == committed
