"""GPU renderer parity through the C-ABI: gi_render images against the
oracle's CPU path tracer on the same scene, camera and Halton indices.
The device code uses only correctly rounded IEEE operations in the
oracle's order, so the bar is bit-exact images and equal ray counts; the
1%-RMSE bar of the specification is asserted as well.  At full
configuration size the oracle renders a pixel range of the same view."""
import numpy as np
import pytest

from globillum_amd import scenes
from globillum_amd._lib import FLAG_BATCHED, FLAG_GRAPH, GiError

pytestmark = pytest.mark.gpu


def _same(a, b):
    a = np.asarray(a).reshape(-1, 3)
    b = np.asarray(b).reshape(-1, 3)
    bad = np.nonzero((a != b).any(1))[0]
    assert len(bad) == 0, (f"{len(bad)} of {len(a)} pixels differ, first {bad[:5].tolist()}, "
                           f"max |d| {np.abs(a - b).max():.3g}")


def _rel_rmse(a, b):
    return float(np.sqrt(((a - b) ** 2).mean()) / max(np.sqrt((b ** 2).mean()), 1e-12))


def _oracle(oracle, sc, views=None, opts=None, **kw):
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights)
    return osc.render(sc.views if views is None else views, sc.opts if opts is None else opts, **kw)


def test_c1_full_config_bit_exact(oracle, gpu_ctx_factory):
    sc = scenes.cornell_boxes()  # 256x256, 16 spp, depth 4
    ref, st = _oracle(oracle, sc)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render().image(0).reshape(-1, 3)
        c = ctx.counters()
    assert _rel_rmse(img, ref) <= 0.01
    _same(img, ref)
    assert c["rays_closest"] == st["rays_closest"] and c["rays_shadow"] == st["rays_shadow"]
    assert c["samples"] == 256 * 256 * 16


def test_depth_32_bit_exact(oracle, gpu_ctx_factory):
    # the deepest render samples dimensions up to 227 (primes up to 1439);
    # each shading launch copies its vertex's seven chunk tables into LDS,
    # the last window (dimensions 221..227) holding 9959 entries.  No
    # Russian roulette, so some paths of the open box reach the last vertices
    sc = scenes.cornell_boxes(w=48, h=48, spp=16, depth=32)
    ref, st = _oracle(oracle, sc)
    _, st24 = _oracle(oracle, sc, opts=scenes.make_opts(24))
    assert st["rays_closest"] > st24["rays_closest"]  # paths went past vertex 24
    for flags in (0, FLAG_BATCHED):
        with gpu_ctx_factory(sc) as ctx:
            img = ctx.render(flags=flags).image(0).reshape(-1, 3)
            c = ctx.counters()
        _same(img, ref)
        assert c["rays_closest"] == st["rays_closest"] and c["rays_shadow"] == st["rays_shadow"]


@pytest.mark.parametrize("flags", [FLAG_BATCHED, FLAG_BATCHED | FLAG_GRAPH])
def test_plain_batched_graph_identical(flags, gpu_ctx_factory):
    sc = scenes.cornell_boxes(w=128, h=96, spp=8)
    with gpu_ctx_factory(sc) as ctx:
        a = ctx.render().image(0)
        b = ctx.render(flags=flags).image(0)
        c2 = ctx.render(flags=flags, samples_per_pass=3).image(0)
    _same(a, b)
    _same(a, c2)


def test_progressive_equals_one_shot(gpu_ctx_factory):
    sc = scenes.cornell_boxes(w=64, h=64, spp=5)
    with gpu_ctx_factory(sc) as ctx:
        one = ctx.render().image(0)
        ctx.progress_begin()
        steps = 0
        while not ctx.progress_step():
            steps += 1
        assert steps == 4
        assert np.array_equal(ctx.image(0), one)
        with pytest.raises(GiError):
            ctx.progress_step()


def test_russian_roulette_bit_exact(oracle, gpu_ctx_factory):
    sc = scenes.cornell_boxes(w=96, h=96, spp=16, depth=8)
    sc.opts["rr_start"] = 3
    ref, st = _oracle(oracle, sc)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render().image(0).reshape(-1, 3)
        c = ctx.counters()
    _same(img, ref)
    assert c["rays_closest"] == st["rays_closest"]
    assert st["rays_closest"] < 0.9 * 8 * 96 * 96 * 16


def test_principled_scene_bit_exact(oracle, gpu_ctx_factory):
    sc = scenes.terrain_spheres(cells=(64, 64), n_spheres=200, w=160, h=90, spp=8, depth=8)
    ref, st = _oracle(oracle, sc)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render().image(0).reshape(-1, 3)
        c = ctx.counters()
    assert _rel_rmse(img, ref) <= 0.01
    _same(img, ref)
    assert c["rays_closest"] == st["rays_closest"] and c["rays_shadow"] == st["rays_shadow"]


def test_multi_view_batched_matches_per_view_oracle(oracle, gpu_ctx_factory):
    sc = scenes.cornell_boxes()
    rng = np.random.default_rng(7)
    views = np.concatenate([
        scenes.make_view((rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7), -1.2), (0.5, 0.4, 0.5), (0, 1, 0),
                         40.0, w, h, spp) for (w, h, spp) in ((64, 48, 4), (40, 40, 16), (96, 54, 1))])
    views[1]["aperture"] = 0.05
    views[1]["focus_distance"] = 1.6
    opts = scenes.make_opts(4, flags=FLAG_BATCHED | FLAG_GRAPH)
    ref, _ = _oracle(oracle, sc, views, opts)
    with gpu_ctx_factory(sc) as ctx:
        ctx.render(views, opts)
        got = np.concatenate([im.reshape(-1, 3) for im in ctx.images()])
    _same(got, ref)


def test_thin_lens_camera_rays_bit_exact(oracle, gpu_ctx_factory):
    sc = scenes.cornell_boxes()
    v = scenes.make_view((0.5, 0.5, -1.3), (0.5, 0.5, 0.0), (0, 1, 0), 40.0, 33, 17, 4, aperture=0.1, focus=1.5)
    with gpu_ctx_factory(sc) as ctx:
        for s in (0, 3):
            go, gd, _ = ctx.camera_rays(v, s, 5)
            oo, od, _ = oracle.camera_rays(v, s, 5)
            assert np.array_equal(go, oo) and np.array_equal(gd, od), s


def test_converged_1024spp_within_one_percent(oracle, gpu_ctx_factory):
    sc = scenes.cornell_boxes(w=64, h=64, spp=1024)
    ref, _ = _oracle(oracle, sc)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render().image(0).reshape(-1, 3)
    assert _rel_rmse(img, ref) <= 0.01
    _same(img, ref)


def test_full_size_c2_properties_and_crop_parity(oracle, gpu_ctx_factory):
    sc = scenes.config_scene(2)  # 1280x720, 64 spp, depth 8, ~291k triangles
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render(flags=FLAG_BATCHED).image(0).reshape(-1, 3)
        c = ctx.counters()
    n = 1280 * 720 * 64
    assert np.isfinite(img).all() and (img >= 0).all()
    assert c["samples"] == n and n <= c["rays_closest"] <= 8 * n and c["rays_shadow"] <= c["rays_closest"]
    # oracle on a pixel range of the same view (same Halton indices)
    p0, p1 = 360 * 1280 + 400, 360 * 1280 + 1400
    ref, _ = _oracle(oracle, sc, p_begin=p0, p_end=p1)
    _same(img[p0:p1], ref[p0:p1])


def test_full_size_large_configuration_properties_and_crop_parity(oracle, gpu_ctx_factory):
    """One view of the large configuration at full size (3840x2160, 16 spp,
    depth 6: 4.2M-triangle world, 2000 instanced spheres, 1024x1024 albedo
    texture with ray-differential mip levels, 64 lights), and the oracle on
    pixel ranges of the same view."""
    sc = scenes.config_scene(5, n_views=1)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render(flags=FLAG_BATCHED).image(0).reshape(-1, 3)
        c = ctx.counters()
    n = 3840 * 2160 * 16
    assert np.isfinite(img).all() and (img >= 0).all() and img.mean() > 0
    assert c["samples"] == n and n <= c["rays_closest"] <= 6 * n and c["rays_shadow"] <= c["rays_closest"]
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights, textures=sc.textures, meshes=sc.meshes,
                             instances=sc.instances)
    for row, x0 in ((1400, 1500), (800, 2600)):
        p0 = row * 3840 + x0
        ref, _ = osc.render(sc.views, sc.opts, p_begin=p0, p_end=p0 + 300)
        _same(img[p0:p0 + 300], ref[p0:p0 + 300])


def test_render_argument_errors(gpu_ctx_factory):
    sc = scenes.cornell_boxes(w=8, h=8, spp=1)
    with gpu_ctx_factory(sc) as ctx:
        for bad in (dict(max_depth=0), dict(max_depth=33)):
            with pytest.raises(GiError) as e:
                ctx.render(**bad)
            assert e.value.status == 1
        v = sc.views.copy()
        v["spp"] = 0
        with pytest.raises(GiError):
            ctx.render(v)


def _obj_room(path):
    """A room of quads (open front), a box and a pyramid with its own MTL;
    some faces without normals."""
    path.joinpath("room.mtl").write_text(
        "newmtl wall\nKd 0.7 0.7 0.7\nnewmtl redwall\nKd 0.7 0.1 0.1\nnewmtl gold\nKd 0.8 0.6 0.2\n")
    v = ["0 0 0", "4 0 0", "4 3 0", "0 3 0", "0 0 4", "4 0 4", "4 3 4", "0 3 4",   # room 1-8
         "1 0 2", "2 0 2", "2 1 2", "1 1 2", "1 0 3", "2 0 3", "2 1 3", "1 1 3",   # box 9-16
         "3 0 1", "3.8 0 1", "3.4 0 1.8", "3.4 1.5 1.4"]                           # pyramid 17-20
    # windings give the geometric normals inside the room and outside the solids
    f = ["usemtl wall", "f 1 5 6 2", "f 4 3 7 8", "f 5 8 7 6",
         "usemtl redwall", "f 1 4 8 5",
         "usemtl wall", "f 2 6 7 3",
         "usemtl gold", "f 9 12 11 10", "f 13 14 15 16", "f 9 13 16 12", "f 10 11 15 14", "f 12 16 15 11",
         "usemtl nosuchmaterial", "f 17 20 18", "f 18 20 19", "f 19 20 17"]
    path.joinpath("room.obj").write_text("mtllib room.mtl\n" + "".join(f"v {x}\n" for x in v) + "\n".join(f) + "\n")
    return str(path / "room.obj")


def test_obj_scene_bit_exact(oracle, gpu_ctx_factory, tmp_path):
    from globillum_amd import scene_io

    mats_dir = tmp_path / "materials"
    mats_dir.mkdir()
    (mats_dir / "gold.pbrdf").write_text("color 0.9 0.7 0.3 roughness 0.3 metallic 1.0 specular 0.5\n")
    view = scenes.make_view((2.0, 1.5, -3.0), (2.0, 1.2, 2.0), (0, 1, 0), 50.0, 64, 48, 4)
    light = scenes.make_light((2.0, 2.95, 2.0), (0, -1, 0), (8, 8, 8), 1.0, 1.0)
    opts = scenes.make_opts(4)
    sc = scene_io.obj_scene(_obj_room(tmp_path), view, light, opts, pbrdf_dir=str(mats_dir))
    assert sc.mats["principled"].tolist() == [0, 0, 0, 1]
    ref, st = _oracle(oracle, sc)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render().image(0).reshape(-1, 3)
        c = ctx.counters()
    assert img.max() > 0
    _same(img, ref)
    assert c["rays_closest"] == st["rays_closest"] and c["rays_shadow"] == st["rays_shadow"]


def test_obj_scene_with_map_kd_texture_bit_exact(oracle, gpu_ctx_factory, tmp_path):
    """The room with map_Kd PNG textures on two materials (one Lambert, one
    principled through its .pbrdf): obj_scene decodes the images, and the
    render matches the oracle given the same decoded textures."""
    from globillum_amd import image_io, scene_io

    rng = np.random.default_rng(0)
    image_io.save_png(tmp_path / "bricks.png", rng.uniform(0.2, 1.0, (32, 48, 3)).astype(np.float32))
    image_io.save_png(tmp_path / "gold.png", rng.uniform(0.5, 1.0, (16, 16, 3)).astype(np.float32))
    (tmp_path / "room.mtl").write_text(
        "newmtl wall\nKd 0.7 0.7 0.7\nnewmtl redwall\nKd 0.9 0.3 0.3\nmap_Kd bricks.png\n"
        "newmtl gold\nKd 0.8 0.6 0.2\nmap_Kd -bm 1.0 gold.png\n")
    v = ["0 0 0", "4 0 0", "4 3 0", "0 3 0", "0 0 4", "4 0 4", "4 3 4", "0 3 4",
         "1 0 2", "2 0 2", "2 1 2", "1 1 2", "1 0 3", "2 0 3", "2 1 3", "1 1 3"]
    vt = ["0 0", "3 0", "3 2", "0 2"]
    f = ["usemtl wall", "f 1 5 6 2", "f 4 3 7 8", "f 5 8 7 6",
         "usemtl redwall", "f 1/1 4/4 8/3 5/2",
         "usemtl wall", "f 2 6 7 3",
         "usemtl gold", "f 9/1 12/4 11/3 10/2", "f 13/1 14/2 15/3 16/4", "f 9/1 13/2 16/3 12/4",
         "f 10/1 11/4 15/3 14/2", "f 12/1 16/2 15/3 11/4"]
    (tmp_path / "room.obj").write_text("mtllib room.mtl\n" + "".join(f"v {x}\n" for x in v)
                                       + "".join(f"vt {x}\n" for x in vt) + "\n".join(f) + "\n")
    mats_dir = tmp_path / "materials"
    mats_dir.mkdir()
    (mats_dir / "gold.pbrdf").write_text("color 0.9 0.7 0.3 roughness 0.3 metallic 0.4 specular 0.5\n")
    view = scenes.make_view((2.0, 1.5, -3.0), (1.5, 1.2, 2.0), (0, 1, 0), 50.0, 64, 48, 4)
    light = scenes.make_light((2.0, 2.95, 2.0), (0, -1, 0), (8, 8, 8), 1.0, 1.0)
    sc = scene_io.obj_scene(str(tmp_path / "room.obj"), view, light, scenes.make_opts(4), pbrdf_dir=str(mats_dir))
    assert sc.meta["texture_files"] == [str(tmp_path) + "/bricks.png", str(tmp_path) + "/gold.png"]
    assert sc.mats["texture"].tolist() == [-1, -1, 0, 1] and [t.shape for t in sc.textures] == [(32, 48, 4), (16, 16, 4)]
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights, textures=sc.textures)
    ref, st = osc.render(sc.views, sc.opts)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render().image(0).reshape(-1, 3)
    assert img.max() > 0
    _same(img, ref)
    plain = scene_io.obj_scene(str(tmp_path / "room.obj"), view, light, scenes.make_opts(4),
                               pbrdf_dir=str(mats_dir), textures="ignore")
    with gpu_ctx_factory(plain) as ctx:
        assert np.abs(ctx.render().image(0).reshape(-1, 3) - img).max() > 0.01


@pytest.mark.parametrize("builder", ["device", "ploc", "sah"])
def test_device_bvh_render_bit_exact(builder, oracle, gpu_ctx_factory):
    sc = scenes.terrain_spheres(cells=(64, 64), n_spheres=100, w=80, h=45, spp=4)
    ref, st = _oracle(oracle, sc)
    with gpu_ctx_factory(sc, builder=builder) as ctx:
        img = ctx.render().image(0).reshape(-1, 3)
        c = ctx.counters()
    _same(img, ref)
    assert c["rays_closest"] == st["rays_closest"] and c["rays_shadow"] == st["rays_shadow"]


@pytest.mark.parametrize("with_rect", [True, False])
def test_sky_light_bit_exact(with_rect, oracle, gpu_ctx_factory):
    sc = scenes.sky_scene(w=96, h=64, spp=4, depth=4, with_rect=with_rect)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights, sky=sc.sky)
    ref, st = osc.render(sc.views, sc.opts)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render(flags=FLAG_BATCHED).image(0).reshape(-1, 3)
        c = ctx.counters()
    _same(img, ref)
    assert c["rays_closest"] == st["rays_closest"] and c["rays_shadow"] == st["rays_shadow"]


def test_sky_light_errors(gpu_ctx_factory):
    import dataclasses

    sc = scenes.sky_scene(w=16, h=16, spp=1)
    with gpu_ctx_factory(dataclasses.replace(sc, sky=None)) as ctx:
        with pytest.raises(GiError, match="sky light without a map"):
            ctx.render()
    two = dataclasses.replace(sc, lights=np.concatenate([sc.lights, scenes.make_sky_light(1.0, 1.0)]))
    with pytest.raises(GiError, match="at most one sky light"):
        gpu_ctx_factory(two)
    with gpu_ctx_factory(sc) as ctx:
        ctx.render()
        with pytest.raises(GiError, match="sky"):
            ctx.adjoint(np.ones((16, 16, 3), np.float32))


def _reference_materials():
    """The reference's material assets (materials/*.pbrdf, from the
    fixture): clearcoat, anisotropic GTR2, specular tint, sheen and glass
    (opacity 0) lobes; plus a subsurface material, which no asset sets."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_materials.json")
    d = json.load(open(path))["materials"]
    keys = ("subsurface", "specular", "specular_tint", "anisotropic", "sheen", "sheen_tint", "clearcoat",
            "clearcoat_gloss", "opacity", "transmission")
    mats = [scenes.principled(r["color"], r["roughness"], r["metallic"], **{k: r[k] for k in keys})
            for _, r in sorted(d.items())]
    mats.append(scenes.principled((0.8, 0.5, 0.3), 0.6, 0.1, subsurface=1.0, sheen=0.3, sheen_tint=0.5))
    return np.concatenate(mats)


@pytest.mark.parametrize("spp,depth", [(8, 6), (2, 12)])
def test_reference_material_assets_bit_exact(spp, depth, oracle, gpu_ctx_factory):
    mats = _reference_materials()
    assert (mats["clearcoat"] > 0).sum() >= 8 and (mats["anisotropic"] > 0).sum() >= 3
    assert (mats["opacity"] < 1).sum() >= 1 and (mats["sheen"] > 0).sum() >= 1 and (mats["subsurface"] > 0).any()
    sc = scenes.terrain_spheres(cells=(48, 48), n_spheres=160, n_mats=len(mats), w=128, h=72, spp=spp, depth=depth)
    sc.mats = mats
    ref, st = _oracle(oracle, sc)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render().image(0).reshape(-1, 3)
        c = ctx.counters()
        b = ctx.render(flags=FLAG_BATCHED | FLAG_GRAPH).image(0).reshape(-1, 3)
    _same(img, ref)
    _same(b, ref)
    assert c["rays_closest"] == st["rays_closest"] and c["rays_shadow"] == st["rays_shadow"]


def _phong_mats(n, seed):
    """Materials for gpu_pt's model: diffuse and specular colours (the
    reference's .mtl assets give several materials both), some with colour
    sums above 1 (divided by the sum), some without a specular part."""
    rng = np.random.default_rng(seed)
    mats = np.concatenate([scenes.lambert(rng.uniform(0.05, 0.7, 3)) for _ in range(n)])
    mats["specular_color"] = rng.uniform(0.0, 0.45, (n, 3))
    mats["specular_color"][0] = 0.0
    mats["diffuse_color"][1] = 0.0
    return mats


@pytest.mark.parametrize("flags,rr", [(0, 0), (FLAG_BATCHED | FLAG_GRAPH, 0), (FLAG_BATCHED, 3)])
def test_phong_mode_bit_exact(flags, rr, oracle, gpu_ctx_factory):
    from globillum_amd._lib import FLAG_PHONG
    sc = scenes.terrain_spheres(cells=(48, 48), n_spheres=120, n_mats=8, w=128, h=72, spp=8, depth=6)
    sc.mats = _phong_mats(8, 5)
    assert ((sc.mats["diffuse_color"].sum(1) + sc.mats["specular_color"].sum(1)) > 1).any()
    opts = sc.opts.copy()
    opts["flags"] = flags | FLAG_PHONG
    opts["rr_start"] = rr
    ref, st = _oracle(oracle, sc, opts=opts)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render(sc.views, opts).image(0).reshape(-1, 3)
        c = ctx.counters()
        plain = ctx.render().image(0).reshape(-1, 3)
    _same(img, ref)
    assert c["rays_closest"] == st["rays_closest"] and c["rays_shadow"] == st["rays_shadow"]
    assert not np.array_equal(img, plain)  # the flag does change the material model


def test_phong_mode_textured_bit_exact(oracle, gpu_ctx_factory):
    from globillum_amd._lib import FLAG_PHONG
    sc = scenes.textured_scene(w=64, h=48, spp=4, depth=4)
    rng = np.random.default_rng(2)
    sc.mats["specular_color"] = rng.uniform(0.0, 0.3, (len(sc.mats), 3))
    opts = sc.opts.copy()
    opts["flags"] = FLAG_BATCHED | FLAG_PHONG
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights, textures=sc.textures, meshes=sc.meshes,
                             instances=sc.instances, sky=sc.sky)
    ref, st = osc.render(sc.views, opts)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render(sc.views, opts).image(0)
        c = ctx.counters()
        with pytest.raises(GiError) as e:
            ctx.adjoint(np.ones(64 * 48 * 3, np.float32))
        assert e.value.status == 1
    _same(img, ref)
    assert c["rays_closest"] == st["rays_closest"]


@pytest.mark.parametrize("aperture", [0.0, 0.06])
def test_pixel_centre_generator_bit_exact(aperture, oracle, gpu_ctx_factory):
    from globillum_amd._lib import FLAG_PIXEL_CENTRE
    sc = scenes.cornell_boxes(w=80, h=64, spp=6, depth=4)
    sc.views["aperture"] = aperture
    sc.views["focus_distance"] = 1.6
    opts = sc.opts.copy()
    opts["flags"] = FLAG_BATCHED | FLAG_PIXEL_CENTRE
    opts["background"] = 0.0
    ref, st = _oracle(oracle, sc, opts=opts)
    dL = np.random.default_rng(4).normal(size=(80 * 64, 3)).astype(np.float32)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render(sc.views, opts).image(0).reshape(-1, 3)
        c = ctx.counters()
        gm, ge = ctx.adjoint(dL)
        jittered = ctx.render(flags=FLAG_BATCHED).image(0).reshape(-1, 3)
    _same(img, ref)
    assert c["rays_closest"] == st["rays_closest"] and c["rays_shadow"] == st["rays_shadow"]
    assert not np.array_equal(img, jittered)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights)
    om, oe = oracle.adjoint(osc, sc.views, opts, dL, f64=False)
    assert np.allclose(gm, om, rtol=2e-3, atol=2e-4 * np.abs(om).max())
    assert np.allclose(ge, oe, rtol=2e-3, atol=2e-4 * np.abs(oe).max())


def test_many_lights_selection_bit_exact(oracle, gpu_ctx_factory):
    """37 lights of uneven power (light selection by bisection over the
    running power sums must pick what the oracle's linear scan picks)."""
    sc = scenes.terrain_spheres(cells=(32, 32), n_spheres=60, n_lights=37, w=96, h=64, spp=8, depth=3)
    rng = np.random.default_rng(12)
    sc.lights["col"] = rng.uniform(0.5, 30.0, (37, 3)).astype(np.float32)
    sc.lights["col"][5] = sc.lights["col"][6]  # equal powers next to each other
    sc.lights["power"] = [scenes.rect_light_power(l["wh"][0], l["wh"][1], l["col"]) for l in sc.lights]
    ref, st = _oracle(oracle, sc)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render(flags=FLAG_BATCHED).image(0).reshape(-1, 3)
        c = ctx.counters()
    _same(img, ref)
    assert c["rays_shadow"] == st["rays_shadow"]
