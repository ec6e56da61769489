"""Textures in the oracle, pinned without the product: pow 2.2 against
numpy, the box-filtered mip chain against a numpy restatement and closed
forms (constant texture, odd sizes), the level choice at powers of two,
sample_bilin_lod at level 0 against plain bilinear interpolation with
repeat addressing, and renders: a white texture leaves a Lambert image
bit-identical, a grey texture scales it, the FP32 render against FP64,
and a one-texel checkerboard that turns into its mean grey with distance
(the ray-differential level choice)."""
import dataclasses

import numpy as np
import pytest

from globillum_amd import scenes


def test_pow22_against_numpy(oracle):
    xs = np.concatenate([np.float32([0.0, 1.0, 0.5, 2.0, 1e-12, 7.0]),
                         np.random.default_rng(0).uniform(0, 1.5, 20000).astype(np.float32),
                         np.geomspace(1e-14, 1e6, 5000).astype(np.float32)])
    got = np.array([oracle.pow22(float(x)) for x in xs], np.float64)
    want = xs.astype(np.float64) ** 2.2
    ok = xs >= 2.0 ** -50
    assert np.all(np.abs(got[ok] / want[ok] - 1) < 3e-7)
    assert np.all(got[~ok] == 0.0)
    assert oracle.pow22(1.0) == 1.0 and oracle.pow22(-1.0) == 0.0 and oracle.pow22(float("nan")) == 0.0


def _np_chain(img):
    """numpy restatement of the mip chain: truncating mean of the existing
    texels of each 2x2 block."""
    h, w = img.shape[:2]
    levels = int(np.floor(np.log2(min(w, h))))
    out = [img.astype(np.int64)]
    cur = out[0]
    for _ in range(levels):
        H, W = cur.shape[:2]
        dh, dw = (H + 1) // 2, (W + 1) // 2
        pad = np.zeros((2 * dh, 2 * dw, 4), np.int64)
        cnt = np.zeros((2 * dh, 2 * dw, 1), np.int64)
        pad[:H, :W] = cur
        cnt[:H, :W] = 1
        s = pad.reshape(dh, 2, dw, 2, 4).sum((1, 3))
        c = cnt.reshape(dh, 2, dw, 2, 1).sum((1, 3))
        cur = s // c
        out.append(cur)
    return levels, np.concatenate([o.reshape(-1) for o in out]).astype(np.uint8)


@pytest.mark.parametrize("shape", [(1, 1), (2, 2), (4, 4), (5, 3), (3, 7), (16, 9), (64, 64), (33, 65)])
def test_mip_chain(oracle, shape):
    h, w = shape
    img = np.random.default_rng(h * 100 + w).integers(0, 256, (h, w, 4), dtype=np.uint8)
    osc = oracle.OracleScene(scenes.cornell_boxes().tris, textures=[img])
    lv, chain = osc.texture_chain(0)
    want_lv, want = _np_chain(img)
    assert lv == want_lv and np.array_equal(chain, want)
    # closed form: a constant texture stays constant on every level
    const = np.full((h, w, 4), 77, np.uint8)
    osc = oracle.OracleScene(scenes.cornell_boxes().tris, textures=[const])
    assert (osc.texture_chain(0)[1] == 77).all()


def test_level_choice(oracle):
    # levels counted as i < levels with 2^i < diff * max(w, h)
    for x, want in ((0.0, 0), (0.5, 0), (1.0, 0), (1.0001, 1), (2.0, 1), (2.001, 2), (4.0, 2), (1e9, 6)):
        assert oracle.tex_level(x / 64.0, 64, 32, 6) == want, x
    assert oracle.tex_level(float("nan"), 64, 64, 6) == 6
    assert oracle.tex_level(float("inf"), 64, 64, 6) == 6


def test_bilinear_at_level_0(oracle):
    rng = np.random.default_rng(1)
    h, w = 13, 21
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    osc = oracle.OracleScene(scenes.cornell_boxes().tris, textures=[img])
    f = img[..., :3].astype(np.float64) / 255.0
    for s, t in rng.uniform(-2.5, 2.5, (400, 2)):
        x = (s % 1.0) * w
        y = (1.0 - t % 1.0) * h
        # texel centres at i + 0.5, repeat addressing
        fx, fy = x - 0.5, y - 0.5
        i0, j0 = int(np.floor(fx)), int(np.floor(fy))
        ax, ay = fx - i0, fy - j0
        c = lambda i, j: f[j % h, i % w]
        want = ((1 - ay) * ((1 - ax) * c(i0, j0) + ax * c(i0 + 1, j0))
                + ay * ((1 - ax) * c(i0, j0 + 1) + ax * c(i0 + 1, j0 + 1)))
        got = osc.tex_sample(0, float(s), float(t), 0.0)
        assert np.allclose(got, want, atol=2e-6), (s, t)
    # texel centres give the texel
    got = osc.tex_sample(0, (3 + 0.5) / w, 1.0 - (5 + 0.5) / h, 0.0)
    assert np.array_equal(got, np.float32(img[5, 3, :3]) / np.float32(255.0))


def _lambert_plane(tex=None, albedo=0.8, w=48, h=32, spp=2):
    ground = np.array([[(-10, 0, -10), (-10, 0, 30), (10, 0, 30), (10, 0, -10)]], np.float32)
    tris = scenes._quads_to_tris(ground, np.array([(0, 1, 0)], np.float32), np.zeros(1, np.int32))
    tris["ta"], tris["tb"], tris["tc"] = tris["a"][:, [0, 2]] / 4, tris["b"][:, [0, 2]] / 4, tris["c"][:, [0, 2]] / 4
    mats = scenes.lambert((albedo, albedo, albedo))
    if tex is not None:
        mats["texture"] = 0
    light = scenes.make_light((0, 5, 5), (0, -1, 0), (20, 20, 20), 2.0, 2.0)
    view = scenes.make_view((0, 2, -8), (0, 0, 10), (0, 1, 0), 60.0, w, h, spp)
    return scenes.Scene(tris, mats, light, view, scenes.make_opts(3), textures=None if tex is None else [tex])


def _render(oracle, sc, **kw):
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights, textures=sc.textures, meshes=sc.meshes,
                             instances=sc.instances)
    return osc.render(sc.views, sc.opts, threads=4, **kw)


def test_white_texture_is_identity_and_grey_scales(oracle):
    plain, _ = _render(oracle, _lambert_plane())
    white, _ = _render(oracle, _lambert_plane(np.full((8, 8, 4), 255, np.uint8)))
    assert plain.sum() > 0 and np.array_equal(plain, white)
    # grey 51/255 = 0.2 at every level: the Lambert albedo 0.8 becomes 0.16
    grey, _ = _render(oracle, _lambert_plane(np.full((8, 8, 4), 51, np.uint8)))
    ref, _ = _render(oracle, _lambert_plane(albedo=0.8 * 0.2))
    assert np.allclose(grey, ref, rtol=1e-5, atol=1e-7)


def test_checkerboard_greys_with_distance(oracle):
    """One-texel checkerboard of 0 and 255: level 0 alternates, every coarser
    level is (255 + 255 + 0 + 0) // 4 = 127.  Far rows of the plane (a
    footprint of several texels) render exactly as a constant 127 texture
    does; near rows keep the pattern (the ray-differential level choice)."""
    def render(tex):
        sc = _lambert_plane(tex, albedo=1.0, w=160, h=120, spp=1)
        sc.opts["max_depth"] = 1
        sc.lights = scenes.make_light((0, 30, 10), (0, -1, 0), (5e5, 5e5, 5e5), 0.02, 0.02)
        return _render(oracle, sc)[0].reshape(120, 160, 3)

    checker = render(scenes.checker_texture(64, 64, cell=1, c0=(255, 255, 255), c1=(0, 0, 0)))
    grey = render(np.full((64, 64, 4), 127, np.uint8))
    contrast = lambda img, r: float(np.abs(np.diff(img[r, 40:120, 0])).mean() / img[r, 40:120, 0].mean())
    # image row 0 is the bottom of the view: the nearest ground
    assert min(contrast(checker, r) for r in range(0, 12)) > 0.1 > 50 * contrast(grey, 6)
    assert np.array_equal(checker[30:60], grey[30:60]) and grey[30:60].min() > 0


def test_textured_render_fp32_vs_fp64(oracle):
    sc = scenes.textured_scene(w=64, h=48, spp=2, depth=3)
    img, st = _render(oracle, sc)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights, textures=sc.textures, meshes=sc.meshes,
                             instances=sc.instances)
    img64 = osc.render_f64(sc.views, sc.opts, threads=4)
    assert st["rays_shadow"] > 0 and np.isfinite(img).all()
    assert np.sqrt(((img - img64) ** 2).mean()) / np.sqrt((img64 ** 2).mean()) < 2e-3
    # the textures change the image
    plain = dataclasses.replace(sc, mats=sc.mats.copy())
    plain.mats["texture"] = -1
    img0, _ = _render(oracle, plain)
    assert np.abs(img - img0).mean() > 0.01 * np.abs(img0).mean()


def _topdown(tex, principled=False, depth=1, w=24, h=24, spp=2):
    """A textured plane seen from straight above, close enough that every
    lookup is on level 0 (a footprint well under one texel)."""
    sc = _lambert_plane(tex, albedo=0.9, w=w, h=h, spp=spp)
    if principled:
        sc.mats = scenes.principled((0.9, 0.8, 0.7), 0.6, 0.0)
        sc.mats["texture"] = 0
    sc.views = scenes.make_view((0.3, 3.0, 5.2), (0.3, 0.0, 5.0), (0, 0, 1), 30.0, w, h, spp)
    sc.opts["max_depth"] = depth
    return sc


def _loss_grads(oracle, sc, dL, f64=True):
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights, textures=sc.textures)
    return oracle.adjoint(osc, sc.views, sc.opts, dL, threads=4, f64=f64, textures=True)


def test_texel_gradient_exact_for_a_linear_image(oracle):
    """Depth 1, Lambert, level-0 lookups: the image is linear in every texel
    value, so 40 bytes more on texel k change sum(dL * I) by 40 g_k / 255."""
    rng = np.random.default_rng(3)
    tex = rng.integers(60, 200, (16, 16, 4), dtype=np.uint8)
    sc = _topdown(tex)
    dL = rng.normal(size=(24 * 24, 3)).astype(np.float32)
    _, _, (g,) = _loss_grads(oracle, sc, dL)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights, textures=[tex])
    base = (osc.render_f64(sc.views, sc.opts, threads=4).reshape(-1, 3) * dL).sum()
    hot = np.argwhere(np.abs(g).sum(-1) > 0)
    assert len(hot) > 10
    for (y, x) in hot[rng.choice(len(hot), 6, replace=False)]:
        for c in range(3):
            t2 = tex.copy()
            t2[y, x, c] += 40
            o2 = oracle.OracleScene(sc.tris, sc.mats, sc.lights, textures=[t2])
            l2 = (o2.render_f64(sc.views, sc.opts, threads=4).reshape(-1, 3) * dL).sum()
            # texel values are float32(byte / 255): 40 bytes are 40/255 up to float rounding
            assert (l2 - base) * 255.0 / 40.0 == pytest.approx(g[y, x, c], rel=2e-5, abs=1e-9), (y, x, c)


@pytest.mark.parametrize("principled", [False, True])
def test_texel_gradient_sum_through_the_mip_chain(oracle, principled):
    """A constant texture is constant on every level, so scaling all its
    texels moves every lookup, whatever its level: the sum of the level-0
    gradients (coarser levels passed down) is the derivative along that
    direction, checked by central differences of the FP64 render."""
    c0 = 140
    sc = _lambert_plane(np.full((32, 32, 4), c0, np.uint8), albedo=0.9, w=40, h=30, spp=2)
    sc.opts["max_depth"] = 2
    if principled:
        sc.mats = scenes.principled((0.9, 0.8, 0.7), 0.6, 0.0)
        sc.mats["texture"] = 0
    dL = np.random.default_rng(4).normal(size=(40 * 30, 3)).astype(np.float32)
    _, _, (g,) = _loss_grads(oracle, sc, dL)

    def loss(c):
        t = np.full((32, 32, 4), c, np.uint8)
        t[..., 3] = 255
        o = oracle.OracleScene(sc.tris, sc.mats, sc.lights, textures=[t])
        return (o.render_f64(sc.views, sc.opts, threads=4).reshape(-1, 3) * dL).sum()

    fd = (loss(c0 + 4) - loss(c0 - 4)) / (8 / 255.0)
    assert g.sum() == pytest.approx(fd, rel=2e-3)


def test_texel_gradients_fp32_vs_fp64_and_white_texture(oracle):
    sc = scenes.textured_scene(w=32, h=24, spp=2, depth=3)
    dL = np.random.default_rng(5).normal(size=(32 * 24, 3)).astype(np.float32)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights, textures=sc.textures, meshes=sc.meshes,
                             instances=sc.instances)
    m64, e64, t64 = oracle.adjoint(osc, sc.views, sc.opts, dL, threads=4, f64=True, textures=True)
    m32, e32, t32 = oracle.adjoint(osc, sc.views, sc.opts, dL, threads=4, f64=False, textures=True)
    for a, b in zip(t32, t64):
        assert np.abs(b).max() > 0
        assert np.allclose(a, b, rtol=5e-2, atol=5e-3 * np.abs(b).max())
    assert np.allclose(m32, m64, rtol=5e-2, atol=5e-3 * np.abs(m64).max())
    # a white texture: material gradients equal the untextured ones (Lambert)
    tex = _lambert_plane(np.full((8, 8, 4), 255, np.uint8))
    plain = _lambert_plane()
    dL = np.random.default_rng(6).normal(size=(32 * 48, 3)).astype(np.float32)
    gt, _, _ = _loss_grads(oracle, tex, dL)
    gp = oracle.adjoint(oracle.OracleScene(plain.tris, plain.mats, plain.lights), plain.views, plain.opts, dL,
                        threads=4)[0]
    assert np.allclose(gt, gp, rtol=1e-12, atol=0)


def _rn32(fr):
    """The float32 nearest to the rational fr (ties to even)."""
    from fractions import Fraction
    c = np.float32(float(fr))
    best = None
    for cand in (np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))):
        d = abs(Fraction(float(cand)) - fr)
        if best is None or d < best[0] or (d == best[0] and int(np.float32(cand).view(np.uint32)) & 1 == 0):
            best = (d, np.float32(cand))
    return best[1]


def test_fma_corrected_division_by_255_is_the_ieee_quotient():
    # the kernels take a texel channel c / 255 as q0 = c * RN(1/255) and
    # fma(fma(-255, q0, c), RN(1/255), q0); with each fma rounded once
    # (exact rationals here) that equals the IEEE quotient for every byte
    from fractions import Fraction
    r = np.float32(1) / np.float32(255)
    for c in range(256):
        x = np.float32(c)
        q0 = np.float32(x * r)
        e = _rn32(Fraction(-255) * Fraction(float(q0)) + Fraction(float(x)))
        assert Fraction(float(e)) == Fraction(-255) * Fraction(float(q0)) + Fraction(float(x))  # exact residual
        q = _rn32(Fraction(float(e)) * Fraction(float(r)) + Fraction(float(q0)))
        assert q == x / np.float32(255), c


def test_texel_index_wrap_by_selects():
    # the lookup's indices lie in [-1, n + 1] (x = s n with s in [0, 1), and
    # one neighbour either side); three selects give the modulo wrap there
    def wrap(i, n):
        i = i + n if i < 0 else i
        i = i - n if i >= n else i
        return i - n if i >= n else i
    for n in range(1, 70):
        for i in range(-1, n + 2):
            assert wrap(i, n) == ((i % n) + n) % n == i % n
    # x = s * W rounds to at most W for the largest s below 1, and 1 - t to
    # at most 1 for t in [0, 1)
    s_max = np.nextafter(np.float32(1), np.float32(0))
    for w in list(range(1, 5000)) + [8191, 8192, 16383, 16384, 1 << 20]:
        x = np.float32(s_max * np.float32(w))
        assert 0 <= int(x) <= w
    t = np.random.default_rng(3).random(10000).astype(np.float32)
    assert np.all((np.float32(1) - t) <= 1) and np.all((np.float32(1) - t) > 0)
