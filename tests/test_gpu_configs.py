"""The benchmark configurations at their full sizes on the GPU, through the
C-ABI, against the oracle on pixel ranges (the oracle renders / replays
only those pixels: zero dL elsewhere) and through size-independent
properties:

- batching (16 views of the cornell scene, 256x256 .. 1920x1080, 4/16/64
  spp, depth 8, Russian roulette from 3): the batched + hipGraph frame with
  tuned launch parameters equals the per-view renders (the reference's
  schedule, src/gpu-pt.h:102-226, and per-view full passes) bit for bit,
  and every view's pixel range equals the oracle;
- differentiation (108k triangles, 64 principled materials, 4 lights,
  512x512, 16 spp, depth 4): forward pixel ranges bit-exact; the adjoint of
  an L2 loss restricted to a pixel range against the FP32 oracle's replay
  of the same paths (summation order) and the FP64 gradient; the emission
  gradients exactly the per-light images (the image is linear in each
  light's emission with the selection weights held fixed);
- the large combined configuration (4 views of 3840x2160, 16 spp, depth 6,
  4.2M-triangle world, 2000 instances, a 1024x1024 texture): forward pixel
  ranges of every view bit-exact; the adjoint restricted to pixel ranges of
  every view against the FP32 oracle; and over the whole frame, texture and
  base colour enter only as their product, so sum_t T_t dL/dT_t equals
  sum_m C_m dL/dC_m per channel (up to the mip chain's integer rounding)."""
import numpy as np
import pytest

from globillum_amd import scenes
from globillum_amd._lib import FLAG_BATCHED, FLAG_GRAPH

pytestmark = pytest.mark.gpu


def _same(a, b, what=""):
    a = np.asarray(a).reshape(-1, 3)
    b = np.asarray(b).reshape(-1, 3)
    bad = np.nonzero((a != b).any(1))[0]
    assert len(bad) == 0, (f"{what}: {len(bad)} of {len(a)} pixels differ, first {bad[:5].tolist()}, "
                           f"max |d| {np.abs(a - b).max():.3g}")


def _offsets(views):
    n = [int(v["w"]) * int(v["h"]) for v in views]
    return np.concatenate([[0], np.cumsum(n)]).astype(np.int64)


def test_c3_batched_graph_equals_per_view_and_oracle(oracle, gpu_ctx_factory):
    sc = scenes.config_scene(3)
    sizes = [(int(v["w"]), int(v["h"])) for v in sc.views]
    assert len(sc.views) == 16 and (1920, 1080) in sizes and (256, 256) in sizes and sc.opts["rr_start"][0] == 3
    batched = sc.opts.copy()
    batched["flags"] = FLAG_BATCHED | FLAG_GRAPH
    ref_sched = sc.opts.copy()
    ref_sched["flags"] = 0
    ref_sched["samples_per_pass"] = 1
    full = sc.opts.copy()
    full["flags"] = 0
    with gpu_ctx_factory(sc) as ctx:
        ctx.set_tuning(trace_blocks_per_cu=6, refill_lanes=16)
        print("[c3] batched + graph", flush=True)
        b = [im.copy() for im in ctx.render(sc.views, batched).images()]
        cb = ctx.counters()
        b2 = [im.copy() for im in ctx.render(sc.views, batched).images()]  # the kept graph, relaunched
        assert ctx.counters()["ms_graph_build"] == 0.0
        print("[c3] reference schedule", flush=True)
        r = ctx.render(sc.views, ref_sched).images()
        cr = ctx.counters()
        print("[c3] per-view full passes", flush=True)
        f = ctx.render(sc.views, full).images()
    for v in range(16):
        _same(b[v], r[v], f"view {v} batched vs reference schedule")
        _same(b[v], f[v], f"view {v} batched vs per-view full passes")
        _same(b[v], b2[v], f"view {v} graph relaunch")
    assert cb["rays_closest"] == cr["rays_closest"] and cb["rays_shadow"] == cr["rays_shadow"]
    # Russian roulette shortens paths: fewer than depth closest rays per sample
    assert cb["rays_closest"] < 0.8 * 8 * cb["samples"]
    # the oracle on pixel ranges of every view
    print("[c3] oracle pixel ranges", flush=True)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights)
    for p0 in (0, 40000):
        ref, _ = osc.render(sc.views, sc.opts, p_begin=p0, p_end=p0 + 600)
        off = _offsets(sc.views)
        for v in range(16):
            n = off[v + 1] - off[v]
            if p0 >= n:
                continue
            e = min(p0 + 600, n)
            _same(b[v].reshape(-1, 3)[p0:e], ref[off[v] + p0:off[v] + e], f"view {v} pixels {p0}..{e}")


@pytest.fixture(scope="module")
def c4():
    sc = scenes.config_scene(4)
    assert len(sc.tris) >= 100_000 and len(sc.mats) == 64 and len(sc.lights) == 4
    return sc


def test_c4_full_size_forward_and_adjoint(c4, oracle, gpu_ctx_factory):
    sc = c4
    tgt = scenes.perturbed(sc, seed=0)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights)
    with gpu_ctx_factory(tgt) as ctx:
        timg = ctx.render().image(0).reshape(-1, 3)
    with gpu_ctx_factory(sc) as ctx:
        img = ctx.render().image(0).reshape(-1, 3)
        for p0 in (70000, 190000):
            ref, _ = osc.render(sc.views, sc.opts, p_begin=p0, p_end=p0 + 1500)
            _same(img[p0:p0 + 1500], ref[p0:p0 + 1500], f"pixels {p0}..")
        dL = (2.0 * (img - timg) / img.size).astype(np.float32)
        # the adjoint restricted to a block of rows, against the oracle
        mask = np.zeros_like(dL)
        mask[150 * 512:158 * 512] = 1.0
        dLm = dL * mask
        gm, ge = ctx.adjoint(dLm)
        # the whole frame's emission gradients, and the per-light images
        gm_all, ge_all = ctx.adjoint(dL)
        per_light = []
        for k in range(len(sc.lights)):
            lk = sc.lights.copy()
            for j in range(len(lk)):
                if j != k:
                    lk["col"][j] = 0.0
            ctx.set_lights(lk)
            per_light.append(ctx.render().image(0).reshape(-1, 3).astype(np.float64))
        ctx.set_lights(sc.lights)
    om32, oe32 = oracle.adjoint(osc, sc.views, sc.opts, dLm, f64=False)
    om64, oe64 = oracle.adjoint(osc, sc.views, sc.opts, dLm, f64=True)
    sm, se = np.abs(om64).max(), np.abs(oe64).max()
    assert sm > 0 and se > 0
    assert np.allclose(gm, om32, rtol=2e-3, atol=2e-4 * sm)
    assert np.allclose(ge, oe32, rtol=2e-3, atol=2e-4 * se)
    assert np.allclose(gm, om64, rtol=3e-2, atol=3e-3 * sm)
    assert np.allclose(ge, oe64, rtol=1e-2, atol=1e-3 * se)
    # emission: d/dE_kc sum(dL * I) = sum(dL_c * I_k,c) / E_kc
    col = sc.lights["col"].astype(np.float64)
    exact = np.array([(dL.astype(np.float64) * per_light[k]).sum(0) / col[k] for k in range(len(sc.lights))])
    scale = np.array([(np.abs(dL) * per_light[k]).sum(0) / col[k] for k in range(len(sc.lights))])
    assert np.allclose(ge_all, exact, rtol=1e-3, atol=1e-4 * scale.max())
    assert np.isfinite(gm_all).all() and np.abs(gm_all).max() > 0


def test_c5_four_views_forward_and_adjoint(oracle, gpu_ctx_factory):
    sc = scenes.config_scene(5)
    assert len(sc.views) == 4 and len(sc.instances) == 2000 and (sc.mats["texture"] >= 0).all()
    W, H = 3840, 2160
    npix = W * H
    tgt = scenes.perturbed(sc, seed=1)
    with gpu_ctx_factory(sc) as ctx:
        print("[c5] target frame", flush=True)
        ctx.set_materials(tgt.mats)
        ctx.set_lights(tgt.lights)
        timgs = [im.reshape(-1, 3).copy() for im in ctx.render(flags=FLAG_BATCHED).images()]
        ctx.set_materials(sc.mats)
        ctx.set_lights(sc.lights)
        print("[c5] frame", flush=True)
        imgs = [im.reshape(-1, 3).copy() for im in ctx.render(flags=FLAG_BATCHED).images()]
        c = ctx.counters()
        img = np.concatenate(imgs)
        dL = (2.0 * (img - np.concatenate(timgs)) / img.size).astype(np.float32)
        print("[c5] adjoint, whole frame", flush=True)
        gm, ge = ctx.adjoint(dL)
        gt = ctx.texture_grad(0)
        # restricted to 48 pixels of each view, for the oracle
        rows = (1200, 700, 1500, 1000)
        dLm = np.zeros_like(dL)
        for v, r in enumerate(rows):
            p = v * npix + r * W + 1800
            dLm[p:p + 48] = dL[p:p + 48] * 1e3
        print("[c5] adjoint, pixel ranges", flush=True)
        gmm, gem = ctx.adjoint(dLm)
        gtm = ctx.texture_grad(0)
    n = 4 * npix * 16
    assert c["samples"] == n and n <= c["rays_closest"] <= 6 * n
    assert all(np.isfinite(x).all() for x in (img, gm, ge, gt, gmm, gem, gtm))
    # forward: a pixel range of every view against the oracle
    print("[c5] oracle pixel ranges", flush=True)
    osc = oracle.OracleScene(sc.tris, sc.mats, sc.lights, textures=sc.textures, meshes=sc.meshes,
                             instances=sc.instances)
    p0 = 1100 * W + 1700
    ref, _ = osc.render(sc.views, sc.opts, p_begin=p0, p_end=p0 + 160)
    for v in range(4):
        a = v * npix + p0
        _same(img[a:a + 160], ref[a:a + 160], f"view {v}")
    # adjoint on the pixel ranges: the FP32 oracle replays the same paths
    print("[c5] oracle adjoint", flush=True)
    om, oe, ot = oracle.adjoint(osc, sc.views, sc.opts, dLm, f64=False, textures=True)
    sm, se, st = np.abs(om).max(), np.abs(oe).max(), np.abs(ot[0]).max()
    assert sm > 0 and se > 0 and st > 0
    assert np.allclose(gmm, om, rtol=2e-3, atol=2e-4 * sm)
    assert np.allclose(gem, oe, rtol=2e-3, atol=2e-4 * se)
    assert np.allclose(gtm, ot[0], rtol=2e-3, atol=2e-4 * st)
    # whole frame: texture and base colour enter as their product b = C * T,
    # so scaling every texel by s and every colour by 1/s leaves the image
    # as it is -- sum_t T_t g_t = sum_m C_m g_m per channel (the coarse mip
    # levels' integer rounding is not differentiated: about 1/255 apart)
    T = sc.textures[0][..., :3].astype(np.float64) / 255.0
    lhs = (T * gt.astype(np.float64)).reshape(-1, 3).sum(0)
    cg = sc.mats["color"].astype(np.float64) * gm[:, :3].astype(np.float64)
    rhs = cg.sum(0)
    scale = np.abs(cg).sum(0)
    assert (np.abs(lhs - rhs) <= 2e-2 * scale).all(), (lhs, rhs, scale)
    assert (np.abs(gt).sum(-1) > 0).mean() > 0.99
