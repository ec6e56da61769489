"""Benchmark: path-tracing throughput on one MI355X.

Workload (one step): a full render of the ~300k-triangle configuration --
256x256-cell fBm heightfield (131,072 triangles) + 1000 UV spheres of 160
triangles, 32 principled materials, 4 rect lights, 1280x720, 64 spp, depth
8, FP32 -- with scene, BVH and camera already resident in HBM.

value = (closest-hit + shadow rays traced) / wall time of the K timed
renders, in Mrays/s.  roofline: the closest-hit traversal kernel's
algorithmic bytes (per ray 32 B in + 16 B out + 32 B per child box + 40 B
per triangle test, with the box / triangle test counts of the reference's
binary-BVH traversal rules measured by the oracle on the same BVH) over its
HIP-event time, against 8 TB/s HBM, and beside it against the read
bandwidth of L2, Infinity Cache and HBM measured on the same GPU by a
streaming-read kernel (libgi_membench.so).  cpu_baseline: the oracle (a C++
restatement of the reference's CPU path tracer, OpenMP) on a bounded sample
of the same workload.  Further legs: the small 34-triangle configuration
(with RMSE against its 1024-spp image), 16-view batching, the adjoint, BVH
construction and the large instanced configuration.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s incl. shadow rays + adjoint grads/s at fixed RMSE, 1 MI355X"
HBM_PEAK_GBS = 8000.0
# Memory traffic of one k_trace_closest launch of this workload at the
# default tuning (one pass, 8 launches per frame, 18.7 M rays per launch),
# from two rocprofv3 counter runs of `bench.py --steps 1 --warmup 0 --no-cpu
# --no-batching --no-adjoint --no-bvh --no-instancing --no-small
# --no-membench --copy-mib 512`: FETCH_SIZE 2,914,794 KiB and WRITE_SIZE
# 510,834 KiB per dispatch, each scaled by what the counter reads for the
# 512 MiB device copy (FETCH 262,168 KiB -> x2.000, WRITE 524,288 KiB ->
# x1.000).  FETCH counts L2 misses to the fabric, so Infinity Cache hits
# are included: an upper bound of the HBM bytes.
TRACE_CLOSEST_FETCH_BYTES = 2914793.7 * 1.99982 * 1024
TRACE_CLOSEST_TRAFFIC_BYTES = TRACE_CLOSEST_FETCH_BYTES + 510833.6 * 1.0 * 1024


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--graph", action="store_true", help="replay the render as one hipGraph")
    p.add_argument("--no-batching", action="store_true", help="skip the multi-view batching leg")
    p.add_argument("--no-tune", action="store_true", help="skip the autotuner in the batching and small legs")
    p.add_argument("--no-adjoint", action="store_true", help="skip the differentiation leg")
    p.add_argument("--no-bvh", action="store_true", help="skip the BVH construction leg")
    p.add_argument("--no-instancing", action="store_true", help="skip the two-level instancing leg")
    p.add_argument("--no-small", action="store_true", help="skip the small (34-triangle) configuration leg")
    p.add_argument("--no-membench", action="store_true", help="skip the memory read-bandwidth measurement")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="first run one device-to-device copy of this many MiB (untimed): the known-size "
                        "kernel that scales FETCH_SIZE / WRITE_SIZE counter runs")
    return p.parse_args()


def read_bandwidth():
    """Measured read bandwidth (GB/s) of each level of the memory hierarchy:
    the streaming-read kernel of libgi_membench.so over 16 MiB (2 MiB per
    XCD: L2), 128 MiB (16 MiB per XCD: Infinity Cache) and 4 GiB (HBM)."""
    import ctypes

    path = os.path.join(ROOT, "globillum_amd", "libgi_membench.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (python -m globillum_amd.build)")
    lib = ctypes.CDLL(path)
    lib.gi_membench_read.restype = ctypes.c_int
    lib.gi_membench_read.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    out = {}
    for name, mib, passes in (("l2", 16, 5), ("infinity_cache", 128, 5), ("hbm", 4096, 3)):
        g = ctypes.c_double(0.0)
        rc = lib.gi_membench_read(0, mib << 20, passes, ctypes.byref(g))
        if rc != 0:
            raise RuntimeError(f"gi_membench_read({mib} MiB) failed with HIP error {rc}")
        out[name] = round(g.value, 1)
    print(f"[bench] measured read bandwidth GB/s: {out}", file=sys.stderr, flush=True)
    return out


def small_leg(steps: int, with_cpu: bool, bw: dict | None, tune: bool):
    """The small plain configuration (cornell_boxes: 5 walls and 2 boxes, 34
    triangles, 1 rect light of emission 10, Lambert albedo U[0.2, 0.8],
    256x256, 16 spp, depth 4), rendered as one hipGraph with autotuned
    launch parameters: Mrays/s,
    samples/s, the relative RMSE of its image against the same scene at
    1024 spp, and the closest-hit kernel's algorithmic bytes (box and
    triangle test counts from the oracle on the same BVH, which runs the
    whole configuration as the CPU baseline) against the measured L2 read
    bandwidth -- the scene's BVH and triangles take a few KB -- and
    against the HBM peak."""
    import numpy as np

    from globillum_amd import autotune, scenes
    from globillum_amd._lib import FLAG_BATCHED, FLAG_GRAPH
    from globillum_amd.context import Context, build_bvh2

    sc = scenes.config_scene(1)
    bvh2 = build_bvh2(sc.tris)
    flags = FLAG_BATCHED | FLAG_GRAPH
    tot = {"rays_closest": 0, "rays_shadow": 0, "samples": 0}
    kern = {"rays_closest": 0, "ms_trace_closest": 0.0, "launches_closest": 0}
    tuned = None
    with Context(0) as ctx:
        ctx.upload(sc, bvh2=bvh2)
        if tune:
            o = sc.opts.copy()
            o["flags"] = flags
            # launch shape only: the tree stays the exported one the oracle counts on
            space = {k: v for k, v in autotune.SPACE.items() if k != "node_width"}
            tuned = autotune.tune(ctx, sc.views, o, space=space, repeats=3)["params"]
        for _ in range(3):
            ctx.render(flags=flags)
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.render(flags=flags)
            c = ctx.counters()
            for k in tot:
                tot[k] += c[k]
        dt = time.perf_counter() - t0
        # kernel times come from renders outside a graph (events per launch)
        for _ in range(steps):
            ctx.render(flags=FLAG_BATCHED)
            c = ctx.counters()
            for k in kern:
                kern[k] += c[k]
        img = ctx.image(0).astype(np.float64)
        v = sc.views.copy()
        v["spp"] = 1024
        ctx.render(v, sc.opts, flags=FLAG_BATCHED)
        ref = ctx.image(0).astype(np.float64)
    rays = tot["rays_closest"] + tot["rays_shadow"]
    out = {"workload": "cornell_boxes: 5 walls + 2 boxes (34 triangles), 1 rect light, Lambert, 256x256, 16 spp, "
                       "depth 4 (the small plain configuration), batched + hipGraph",
           "mrays_per_s": round(rays / dt / 1e6, 2), "samples_per_s": round(tot["samples"] / dt, 1),
           "ms_per_frame": round(dt / steps * 1e3, 3),
           "rmse_rel_vs_1024spp": round(float(np.sqrt(((img - ref) ** 2).mean()) / np.sqrt((ref ** 2).mean())), 5),
           "tuned": tuned}
    if with_cpu:
        base, per_ray = cpu_baseline(sc, bvh2, full=True)
        out["cpu_baseline"] = base
        if kern["launches_closest"] > 0 and kern["ms_trace_closest"] > 0:
            b_ray = 32 + 16 + 32 * per_ray["box_closest"] + 40 * per_ray["tri_closest"]
            avg_s = kern["ms_trace_closest"] / kern["launches_closest"] / 1e3
            achieved = kern["rays_closest"] / kern["launches_closest"] * b_ray / avg_s / 1e9
            roof = {"kernel": "k_trace_closest", "achieved": round(achieved, 1), "unit": "GB/s",
                    "algorithmic_bytes_per_ray": round(b_ray, 1), "avg_launch_ms": round(avg_s * 1e3, 4),
                    "frac_of_hbm_peak": round(achieved / HBM_PEAK_GBS, 4)}
            if bw:
                roof["measured_l2_read"] = bw["l2"]
                roof["frac_of_measured_l2"] = round(achieved / bw["l2"], 4)
            out["roofline"] = roof
    return out


def batching_leg(steps: int, tune: bool):
    """16 views of the cornell scene (resolutions from 256x256 to 1920x1080,
    4/16/64 spp, depth 8, Russian roulette from 3): the per-view wavefront
    scheduled as the reference schedules it (one view after the other, one
    sample per pixel per wavefront, the host reading each bounce's queue
    length; src/gpu-pt.h:102-226) against the batched renderer (all views
    and samples in shared queues, device-side lengths, the whole frame
    replayed as one hipGraph, autotuned launch parameters).  Both run the
    same kernels and produce the same image bits."""
    import numpy as np

    from globillum_amd import autotune, scenes
    from globillum_amd._lib import FLAG_BATCHED, FLAG_GRAPH
    from globillum_amd.context import Context

    sc = scenes.config_scene(3)
    plain = sc.opts.copy()
    plain["flags"] = 0
    plain["samples_per_pass"] = 1
    batched = sc.opts.copy()
    batched["flags"] = FLAG_BATCHED | FLAG_GRAPH
    out = {"workload": "16 views of the cornell scene, 256x256..1920x1080, spp 4/16/64, depth 8, RR from 3; "
                       "plain = per view, 1 sample per pixel per wavefront, host-synchronised bounces",
           "samples": int(sum(int(v["w"]) * int(v["h"]) * int(v["spp"]) for v in sc.views))}
    with Context(0) as ctx:
        ctx.upload(sc)
        res = {}
        for name, o in (("plain", plain), ("batched", batched)):
            if name == "batched" and tune:
                t = autotune.tune(ctx, sc.views, o, repeats=1,
                                  log=lambda m: print(m, file=sys.stderr, flush=True))
                out["tuned"] = t["params"]
            ctx.render(sc.views, o)
            rays = 0
            t0 = time.perf_counter()
            for _ in range(steps):
                ctx.render(sc.views, o)
                c = ctx.counters()
                rays += c["rays_closest"] + c["rays_shadow"]
            dt = time.perf_counter() - t0
            res[name] = {"mrays_per_s": round(rays / dt / 1e6, 2), "ms_per_frame": round(dt / steps * 1e3, 2),
                         "image": [im.copy() for im in ctx.images()[:3]]}
            print(f"[bench] batching leg {name}: {res[name]['mrays_per_s']} Mrays/s", file=sys.stderr, flush=True)
        same = all(np.array_equal(a, b) for a, b in zip(res["plain"]["image"], res["batched"]["image"]))
    for k in res:
        res[k].pop("image")
    out.update(res)
    out["speedup"] = round(res["batched"]["mrays_per_s"] / res["plain"]["mrays_per_s"], 3)
    out["images_identical"] = bool(same)
    return out


def adjoint_leg(steps: int):
    """Differentiation workload: heightfield + spheres (108k triangles), 64
    principled materials, 4 lights, 512x512, 16 spp, depth 4; loss = image L2
    against a target rendered with parameters perturbed by N(0, 0.1^2).  One
    gradient = the full vector (64 x (rgb, roughness, metallic) + 4 x rgb
    emission) of one loss evaluation."""
    import numpy as np

    from globillum_amd import scenes
    from globillum_amd.context import Context

    sc = scenes.config_scene(4)
    tgt = scenes.perturbed(sc, seed=0)
    with Context(0) as ctx:
        ctx.upload(tgt)
        timg = ctx.render().image(0).reshape(-1, 3)
        ctx.set_materials(sc.mats)
        ctx.set_lights(sc.lights)
        img = ctx.render().image(0).reshape(-1, 3)
        dL = (2.0 * (img - timg) / img.size).astype(np.float32)
        ctx.adjoint(dL)
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.render()
        t_fwd = (time.perf_counter() - t0) / steps
        t0 = time.perf_counter()
        for _ in range(steps):
            gm, ge = ctx.adjoint(dL)
        t_adj = (time.perf_counter() - t0) / steps
    n_params = gm.size + ge.size
    out = {"workload": "diff_scene 108k triangles, 64 principled materials, 4 lights, 512x512, 16 spp, depth 4",
           "params_per_gradient": int(n_params), "ms_forward": round(t_fwd * 1e3, 2),
           "ms_adjoint": round(t_adj * 1e3, 2), "adjoint_over_forward": round(t_adj / t_fwd, 3),
           "grads_per_s": round(1.0 / t_adj, 2), "loss": float(((img - timg) ** 2).mean()),
           "grad_finite": bool(np.isfinite(gm).all() and np.isfinite(ge).all())}
    print(f"[bench] adjoint leg: fwd {out['ms_forward']} ms, adj {out['ms_adjoint']} ms", file=sys.stderr,
          flush=True)
    return out


def bvh_leg(steps: int):
    """BVH construction: the host binned-SAH builder against the device
    linear-BVH and PLOC (radius 16) builders on the headline scene (291k
    triangles) and on a 2048x1024-cell heightfield (4.19M triangles, the
    large configuration's terrain), plus the headline frame time
    traversing each tree."""
    import numpy as np

    from globillum_amd import scenes
    from globillum_amd._lib import FLAG_BATCHED
    from globillum_amd.context import Context

    out = {}
    sc = scenes.config_scene(2)
    with Context(0) as ctx:
        t0 = time.perf_counter()
        ctx.upload(sc)  # host SAH (includes the upload of the tables)
        host_ms = (time.perf_counter() - t0) * 1e3
        frames, dev_ms = {}, {}
        for name, method in (("sah", None), ("lbvh", "device"), ("ploc", "ploc")):
            if method is not None:
                ctx.build_bvh_device(4, method)  # warm (allocations, code objects)
                t0 = time.perf_counter()
                ctx.build_bvh_device(4, method)
                dev_ms[name] = (time.perf_counter() - t0) * 1e3
            ctx.render(flags=FLAG_BATCHED)
            t0 = time.perf_counter()
            for _ in range(steps):
                ctx.render(flags=FLAG_BATCHED)
            frames[name] = (time.perf_counter() - t0) / steps * 1e3
    out["terrain_spheres_291k"] = {"host_sah_ms_incl_upload": round(host_ms, 1),
                                   "device_lbvh_ms": round(dev_ms["lbvh"], 2), "device_ploc_ms": round(dev_ms["ploc"], 2),
                                   "frame_ms_sah": round(frames["sah"], 2), "frame_ms_lbvh": round(frames["lbvh"], 2),
                                   "frame_ms_ploc": round(frames["ploc"], 2)}
    rng = np.random.default_rng(5)
    hf, _ = scenes.heightfield(rng, 2048, 1024, 200.0, 100.0, 10.0, 8)
    big = scenes.Scene(hf, sc.mats, sc.lights, sc.views, sc.opts)
    with Context(0) as ctx:
        t0 = time.perf_counter()
        ctx.upload(big)
        host_ms = (time.perf_counter() - t0) * 1e3
        dev_ms = {}
        for name, method in (("lbvh", "device"), ("ploc", "ploc")):
            ctx.build_bvh_device(4, method)
            t0 = time.perf_counter()
            ctx.build_bvh_device(4, method)
            dev_ms[name] = (time.perf_counter() - t0) * 1e3
    out["heightfield_4194k"] = {"host_sah_ms_incl_upload": round(host_ms, 1), "device_lbvh_ms": round(dev_ms["lbvh"], 2),
                                "device_ploc_ms": round(dev_ms["ploc"], 2)}
    return out


def instancing_leg(steps: int):
    """The large combined configuration's forward pass (2048x1024-cell
    heightfield world + 2000 instanced 160-triangle spheres, 256 principled
    materials with a 1024x1024 albedo texture, 64 rect lights, 4 views of
    3840x2160, 16 spp, depth 6) rendered with the two-level BVH and with the
    same spheres flattened into the world BVH (4.51M triangles), and the
    instanced scene without its texture."""
    import numpy as np

    from globillum_amd import scenes
    from globillum_amd._lib import FLAG_BATCHED
    from globillum_amd.context import Context

    sc = scenes.config_scene(5)
    out = {"workload": "large_combined: 2048x1024-cell heightfield (4194304 triangles) + 2000 instances of a "
                       "160-triangle sphere, 256 principled materials with a 1024x1024 albedo texture, 64 rect "
                       "lights, 4 views 3840x2160, 16 spp, depth 6"}
    means = {}
    untextured = dataclasses.replace(sc, mats=sc.mats.copy(), textures=None)
    untextured.mats["texture"] = -1
    for name, s in (("instanced", sc), ("flattened", scenes.flatten(sc)), ("instanced_untextured", untextured)):
        with Context(0) as ctx:
            t0 = time.perf_counter()
            ctx.upload(s)
            up_ms = (time.perf_counter() - t0) * 1e3
            ctx.render(flags=FLAG_BATCHED)
            rays, kms = 0, {"ms_trace_closest": 0.0, "ms_trace_any": 0.0, "ms_shade": 0.0}
            t0 = time.perf_counter()
            for _ in range(steps):
                ctx.render(flags=FLAG_BATCHED)
                c = ctx.counters()
                rays += c["rays_closest"] + c["rays_shadow"]
                for k in kms:
                    kms[k] += c[k]
            dt = time.perf_counter() - t0
            means[name] = np.array([ctx.image(v).mean(0).mean(0) for v in range(len(s.views))])
            out[name] = {"mrays_per_s": round(rays / dt / 1e6, 1), "ms_per_frame": round(dt / steps * 1e3, 1),
                         "upload_and_bvh_ms": round(up_ms, 1),
                         "kernel_ms_per_frame": {k[3:]: round(v / steps, 2) for k, v in kms.items()}}
        print(f"[bench] instancing leg: {name} {out[name]}", file=sys.stderr, flush=True)
    # the configuration's adjoint: material, light and per-texel gradients
    # of an L2 loss against a perturbed target (same geometry and texture)
    tgt = dataclasses.replace(scenes.perturbed(sc, seed=1), meshes=sc.meshes, instances=sc.instances,
                              textures=sc.textures)
    with Context(0) as ctx:
        ctx.upload(tgt)
        ctx.render(flags=FLAG_BATCHED)
        timgs = [ctx.image(v) for v in range(len(sc.views))]
    with Context(0) as ctx:
        ctx.upload(sc)
        ctx.render(flags=FLAG_BATCHED)
        imgs = [ctx.image(v) for v in range(len(sc.views))]
        dL = np.concatenate([(2.0 * (a - b) / a.size).reshape(-1, 3) for a, b in zip(imgs, timgs)]).astype(np.float32)
        t0 = time.perf_counter()
        ctx.render(flags=FLAG_BATCHED)
        fwd_ms = (time.perf_counter() - t0) * 1e3
        ctx.adjoint(dL)  # warm (gradient buffers)
        t0 = time.perf_counter()
        gm, ge = ctx.adjoint(dL)
        adj_ms = (time.perf_counter() - t0) * 1e3
        gt = ctx.texture_grad(0)
    n_par = gm.size + ge.size + gt.size
    out["adjoint"] = {"ms_forward": round(fwd_ms, 1), "ms_adjoint": round(adj_ms, 1),
                      "adjoint_over_forward": round(adj_ms / fwd_ms, 3), "params_per_gradient": int(n_par),
                      "texels_with_gradient": int((np.abs(gt).sum(-1) > 0).sum()),
                      "grad_finite": bool(np.isfinite(gm).all() and np.isfinite(ge).all() and np.isfinite(gt).all())}
    print(f"[bench] large configuration adjoint: {out['adjoint']}", file=sys.stderr, flush=True)
    out["instanced_over_flattened"] = round(out["instanced"]["mrays_per_s"] / out["flattened"]["mrays_per_s"], 3)
    out["mean_radiance_rel_diff"] = float(np.abs(means["instanced"] - means["flattened"]).max()
                                          / np.abs(means["flattened"]).max())
    return out


def cpu_baseline(scene, bvh2, full: bool = False):
    """Oracle on 1/4 of the samples (same scene, camera, depth and BVH at
    640x360, 32 spp), or on the whole workload (full=True, for the small
    configuration).  Returns the baseline record and the per-ray traversal
    statistics (box / triangle tests per closest and shadow ray)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    threads = min(16, os.cpu_count() or 1)
    osc = pyoracle.OracleScene(scene.tris, scene.mats, scene.lights, bvh2=bvh2)
    v = scene.views.copy()
    if not full:
        v["w"], v["h"], v["spp"] = 640, 360, 32
    t0 = time.perf_counter()
    _, st = osc.render(v, scene.opts, threads=threads)
    dt = time.perf_counter() - t0
    rays = st["rays_closest"] + st["rays_shadow"]
    what = (f"the whole workload ({int(v['w'][0])}x{int(v['h'][0])}, {int(v['spp'][0])} spp)" if full else
            "same scene/camera/depth/BVH at 640x360, 32 spp (1/4 of the samples)")
    rec = {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
           "sample": f"{what}: {rays} rays in {dt:.2f} s on {threads} threads",
           "samples_per_s": round(st["samples"] / dt, 1)}
    per_ray = {"box_closest": st["box_closest"] / max(1, st["rays_closest"]),
               "tri_closest": st["tri_closest"] / max(1, st["rays_closest"]),
               "box_shadow": st["box_shadow"] / max(1, st["rays_shadow"]),
               "tri_shadow": st["tri_shadow"] / max(1, st["rays_shadow"])}
    return rec, per_ray


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the path shards by view / image "
              f"tile / spp range; a multi-GPU run is not implemented", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    from globillum_amd import scenes
    from globillum_amd._lib import FLAG_BATCHED, FLAG_GRAPH
    from globillum_amd.context import Context, build_bvh2

    if a.copy_mib > 0:
        import torch
        src = torch.empty(a.copy_mib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        print(f"[bench] copied {a.copy_mib} MiB device to device", file=sys.stderr, flush=True)
    scene = scenes.config_scene(2)
    bvh2 = build_bvh2(scene.tris)
    flags = FLAG_BATCHED | (FLAG_GRAPH if a.graph else 0)
    ctx = Context(0)
    ctx.upload(scene, bvh2=bvh2)
    print(f"[bench] scene {len(scene.tris)} triangles uploaded; warmup {a.warmup}", file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        ctx.render(flags=flags)
    tot = {"rays_closest": 0, "rays_shadow": 0, "samples": 0, "ms_trace_closest": 0.0, "ms_trace_any": 0.0,
           "ms_shade": 0.0, "launches_closest": 0}
    t0 = time.perf_counter()
    for i in range(a.steps):
        ctx.render(flags=flags)  # returns after the stream synchronised
        c = ctx.counters()
        for k in tot:
            tot[k] += c[k]
    dt = time.perf_counter() - t0
    print(f"[bench] {a.steps} steps in {dt:.3f} s", file=sys.stderr, flush=True)
    img = ctx.image(0)
    assert np.isfinite(img).all()
    rays = tot["rays_closest"] + tot["rays_shadow"]
    value = rays / dt / 1e6
    # one more (untimed) frame through the counting build of the traversal
    # kernels: node visits and triangle tests of the wide BVH per ray
    ctx.traversal_stats(True)
    ctx.render(flags=flags)
    ts = ctx.traversal_stats()
    ctx.traversal_stats(False)

    base, per_ray = (None, None) if a.no_cpu else cpu_baseline(scene, bvh2)
    bw = None if a.no_membench else read_bandwidth()
    roof = None
    if per_ray is not None and tot["launches_closest"] > 0 and tot["ms_trace_closest"] > 0:
        b_ray = 32 + 16 + 32 * per_ray["box_closest"] + 40 * per_ray["tri_closest"]
        bytes_per_launch = tot["rays_closest"] * b_ray / tot["launches_closest"]
        avg_launch_s = tot["ms_trace_closest"] / tot["launches_closest"] / 1e3
        achieved = bytes_per_launch / avg_launch_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(TRACE_CLOSEST_TRAFFIC_BYTES) if tot["launches_closest"] == 8 * a.steps else None,
                "kernel": "k_trace_closest", "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                "algorithmic_bytes_per_ray": round(b_ray, 1)}
        # bytes the wide traversal requests per closest ray: ray in, hit
        # out, one 128-byte node per interior visit, 48 bytes per triangle
        req = 32 + 16 + (128 * ts["closest_nodes"] + 48 * ts["closest_triangles"]) / max(1, ts["closest_rays"])
        roof["wide_traversal"] = {"nodes_per_ray": round(ts["closest_nodes"] / max(1, ts["closest_rays"]), 3),
                                  "triangles_per_ray": round(ts["closest_triangles"] / max(1, ts["closest_rays"]), 3),
                                  "requested_bytes_per_ray": round(req, 1),
                                  "requested_gbs": round(bytes_per_launch / b_ray * req / avg_launch_s / 1e9, 1)}
        if bw:
            roof["wide_traversal"]["requested_frac_of_measured_l2"] = round(
                roof["wide_traversal"]["requested_gbs"] / bw["l2"], 4)
        if bw:
            # the BVH and triangles (~27 MB) live in the Infinity Cache
            roof["measured_read_gbs"] = bw
            roof["frac_of_measured_infinity_cache"] = round(achieved / bw["infinity_cache"], 4)
            roof["frac_of_measured_l2"] = round(achieved / bw["l2"], 4)
            if roof["traffic"] is not None:
                # time bound of the measured hierarchy: the bytes fetched past
                # L2 at the Infinity Cache read rate, the rest at the L2 rate
                t_bound = ((bytes_per_launch - TRACE_CLOSEST_FETCH_BYTES) / (bw["l2"] * 1e9) +
                           TRACE_CLOSEST_FETCH_BYTES / (bw["infinity_cache"] * 1e9))
                roof["frac_of_measured_hierarchy"] = round(t_bound / avg_launch_s, 4)
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "terrain_spheres: 256x256-cell fBm heightfield + 1000 UV spheres "
                               "(291072 triangles), 32 principled materials, 4 rect lights, 1280x720, "
                               "64 spp, depth 8 (the ~300k-triangle plain configuration)",
                   "mode": "batched wavefront" + (" + hipGraph" if a.graph else ""),
                   "samples_per_step": tot["samples"] // max(1, a.steps)},
        "samples_per_s": round(tot["samples"] / dt, 1),
        "rays_per_step": {"closest": tot["rays_closest"] // a.steps, "shadow": tot["rays_shadow"] // a.steps},
        "kernel_ms_per_step": {"trace_closest": round(tot["ms_trace_closest"] / a.steps, 3),
                               "trace_shadow": round(tot["ms_trace_any"] / a.steps, 3),
                               "shade": round(tot["ms_shade"] / a.steps, 3)},
        "roofline": roof,
        "cpu_baseline": base,
    }
    ctx.close()
    if not a.no_small:
        out["small"] = small_leg(20, not a.no_cpu, bw, not a.no_tune)
    if not a.no_batching:
        out["batching"] = batching_leg(max(1, min(a.steps, 3)), not a.no_tune)
    if not a.no_adjoint:
        out["adjoint"] = adjoint_leg(max(1, min(a.steps, 3)))
    if not a.no_bvh:
        out["bvh_build"] = bvh_leg(max(1, min(a.steps, 3)))
    if not a.no_instancing:
        out["instancing"] = instancing_leg(max(1, min(a.steps, 2)))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
