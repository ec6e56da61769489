"""Benchmark: path-tracing throughput on one MI355X.

Headline workload (one step): one frame of the large combined configuration
-- 2048x1024-cell fBm heightfield (4,194,304 triangles) + 2000 instances of
a 160-triangle sphere (two-level BVH), 256 principled materials with a
1024x1024 RGB albedo texture (ray-differential mip levels), 64 rect lights,
4 views of 3840x2160 at 16 spp, depth 6, batched, FP32 -- with scene, BVH
and cameras resident in HBM; then the configuration's path-replay adjoint
(material, light and per-texel gradients of an image-L2 loss).

value = (closest-hit + shadow rays traced) / wall time of the K timed
frames, in Mrays/s; adjoint_grads_per_s = full gradient vectors per second.
roofline: per kernel (closest hit, any hit, shading), algorithmic bytes per
launch -- per traced ray 32 B in + 16 B out + 32 B per child box + 40 B per
triangle test, with the box / triangle test counts of the reference's
binary-BVH traversal rules measured by the oracle on the same rays' sample;
per shaded path vertex 308 B, plus 112 B with a texture in use (ray
differentials in and out, four texels) -- over the launch's HIP-event time.
The traversal kernels' reads are served mostly by the caches (their
measured HBM traffic is a fifth of the algorithmic bytes), so they are held
against the L2 read bandwidth measured on the same GPU, with the HBM
figures beside; the shading kernel streams path state and is held against
the 8 TB/s HBM peak.  traffic: the HBM bytes one launch moves, from two
rocprofv3 counter runs of this file in its --traffic-probe mode (FETCH_SIZE,
then WRITE_SIZE, per dispatch), each scaled by what the counter reads for a
copy kernel of known size in the same run.  cpu_baseline: the oracle (a C++
restatement of the reference's CPU path tracer, OpenMP over pixels, on
every CPU this process may use) on a bounded sample of the same workload.

Further legs: the ~300k-triangle plain configuration (its roofline against
the measured L2 read bandwidth -- its scene lives in the caches), the small
34-triangle configuration, 16-view batching against the per-view wavefront,
the differentiation configuration, BVH construction, and the large
configuration flattened / untextured.
"""
from __future__ import annotations

import argparse
import csv
import dataclasses
import glob
import json
import os
import re
import shutil
import signal
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s incl. shadow rays + adjoint grads/s at fixed RMSE, 1 MI355X"
HBM_PEAK_GBS = 8000.0
B_VERTEX = 308  # algorithmic bytes per shaded path vertex (hit, triangle, material, path state in and out)
# ... with a texture in use: also the path's two ray differentials (origin
# and direction each, 48 B in and 48 B out) and the four blended texels
B_VERTEX_TEX = B_VERTEX + 96 + 16
COPY_BYTES = 1 << 30  # the known-size copy of the traffic probe


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    p.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM-traffic counter runs")
    p.add_argument("--no-plain", action="store_true", help="skip the ~300k-triangle plain configuration leg")
    p.add_argument("--no-batching", action="store_true", help="skip the multi-view batching leg")
    p.add_argument("--no-tune", action="store_true", help="skip the autotuner in the batching and small legs")
    p.add_argument("--no-adjoint", action="store_true", help="skip the differentiation legs")
    p.add_argument("--no-bvh", action="store_true", help="skip the BVH construction leg")
    p.add_argument("--no-instancing", action="store_true",
                   help="skip the large configuration's flattened / untextured variants")
    p.add_argument("--no-small", action="store_true", help="skip the small (34-triangle) configuration leg")
    p.add_argument("--no-membench", action="store_true", help="skip the memory read-bandwidth measurement")
    p.add_argument("--tuning-cache", default=os.path.join(os.path.expanduser("~"), ".cache", "globillum_amd",
                                                           "tuning.json"),
                   help="autotuner results per workload (JSON); reused when present")
    p.add_argument("--traffic-probe", action="store_true",
                   help="(internal) one known-size copy, one frame of the plain and of the large configuration; "
                        "run under rocprofv3 --pmc by the main benchmark")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="first run one device-to-device copy of this many MiB (untimed): the known-size "
                        "kernel that scales FETCH_SIZE / WRITE_SIZE counter runs of the main benchmark")
    return p.parse_args()


# ------------------------------------------------------------------- host --
def host_cpus():
    """CPUs this process may use (affinity, capped by a cgroup CPU quota),
    and what the host reports."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except OSError:
        aff = nproc
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) / int(period)))
    except (OSError, ValueError):
        pass
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    threads = min(aff, quota) if quota else aff
    return threads, {"nproc": nproc, "affinity": aff, "cgroup_cpus": quota, "model": model}


def membench():
    import ctypes

    path = os.path.join(ROOT, "globillum_amd", "libgi_membench.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (python -m globillum_amd.build)")
    lib = ctypes.CDLL(path)
    lib.gi_membench_read.restype = ctypes.c_int
    lib.gi_membench_read.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    lib.gi_membench_copy.restype = ctypes.c_int
    lib.gi_membench_copy.argtypes = [ctypes.c_int32, ctypes.c_int64]
    return lib


def read_bandwidth():
    """Measured read bandwidth (GB/s) of each level of the memory hierarchy:
    the streaming-read kernel of libgi_membench.so over 16 MiB (2 MiB per
    XCD: L2), 128 MiB (16 MiB per XCD: Infinity Cache) and 4 GiB (HBM)."""
    import ctypes

    lib = membench()
    out = {}
    for name, mib, passes in (("l2", 16, 5), ("infinity_cache", 128, 5), ("hbm", 4096, 3)):
        g = ctypes.c_double(0.0)
        rc = lib.gi_membench_read(0, mib << 20, passes, ctypes.byref(g))
        if rc != 0:
            raise RuntimeError(f"gi_membench_read({mib} MiB) failed with HIP error {rc}")
        out[name] = round(g.value, 1)
    log(f"measured read bandwidth GB/s: {out}")
    return out


# ---------------------------------------------------------------- traffic --
# kernel families of the two workloads the probe renders: (regex on the
# demangled kernel name, label).  Template arguments tell the flat scene's
# traversal kernels from the two-level ones.
PROBE_KERNELS = {
    "large": {"trace_closest": r"k_trace_closest<false, true,", "trace_any": r"k_trace_any<true, true,",
              "shade": r"k_shade_tex\("},
    "plain": {"trace_closest": r"k_trace_closest<false, false,", "trace_any": r"k_trace_any<true, false,",
              "shade": r"k_shade\("},
}


def traffic_probe() -> None:
    """The counter runs' workload: a 1 GiB device copy, then one frame of the
    plain configuration and one of the large configuration (batched)."""
    from globillum_amd import scenes
    from globillum_amd._lib import FLAG_BATCHED
    from globillum_amd.context import Context, build_bvh2

    rc = membench().gi_membench_copy(0, COPY_BYTES)
    if rc != 0:
        raise RuntimeError(f"gi_membench_copy failed with HIP error {rc}")
    sc = scenes.config_scene(2)
    with Context(0) as ctx:
        ctx.upload(sc, bvh2=build_bvh2(sc.tris))
        ctx.render(flags=FLAG_BATCHED)
    sc = scenes.config_scene(5)
    with Context(0) as ctx:
        ctx.upload(sc)
        ctx.render(flags=FLAG_BATCHED)
    print("probe done", flush=True)


def _counter_rows(outdir: str):
    rows = []
    for f in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(f, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def run_traffic_probes() -> dict | None:
    """Two rocprofv3 --pmc runs of the probe (FETCH_SIZE, then WRITE_SIZE):
    HBM bytes per launch of each kernel family, each counter scaled by its
    reading for the known-size copy.  None when rocprofv3 is missing or a
    run fails; each run is bounded in time and ended by process group."""
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        log("rocprofv3 not found: traffic not measured")
        return None
    per = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        outdir = tempfile.mkdtemp(prefix="gi_traffic_")
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", outdir, "-o", "probe", "--",
               sys.executable, os.path.abspath(__file__), "--traffic-probe"]
        log(f"traffic probe: {counter}")
        env = dict(os.environ, TMPDIR="/tmp")
        p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, start_new_session=True, env=env)
        try:
            out, _ = p.communicate(timeout=240)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.communicate()
            log(f"traffic probe {counter}: timed out")
            return None
        text = out.decode(errors="replace")
        if p.returncode != 0 and "probe done" not in text:
            log(f"traffic probe {counter}: exit {p.returncode}: {text[-600:]}")
            return None
        rows = _counter_rows(outdir)
        shutil.rmtree(outdir, ignore_errors=True)
        vals = {}
        for r in rows:
            if r.get("Counter_Name") != counter:
                continue
            vals.setdefault(r.get("Kernel_Name", ""), []).append(float(r["Counter_Value"]))
        copy = [v for k, vs in vals.items() if "k_copy" in k for v in vs]
        if not copy or copy[0] <= 0:
            log(f"traffic probe {counter}: no reading for the copy kernel")
            return None
        scale = (COPY_BYTES / 1024.0) / copy[0]
        res = {"copy_kib": copy[0], "scale": round(scale, 4)}
        for wl, fams in PROBE_KERNELS.items():
            for fam, pat in fams.items():
                xs = [v for k, vs in vals.items() if re.search(pat, k) for v in vs]
                if xs:
                    res[f"{wl}.{fam}"] = sum(xs) / len(xs) * 1024.0 * scale  # bytes per launch
        per[counter] = res
    out = {"fetch_scale": per["FETCH_SIZE"]["scale"], "write_scale": per["WRITE_SIZE"]["scale"],
           "copy_fetch_kib": per["FETCH_SIZE"]["copy_kib"], "copy_write_kib": per["WRITE_SIZE"]["copy_kib"]}
    for wl, fams in PROBE_KERNELS.items():
        for fam in fams:
            k = f"{wl}.{fam}"
            if k in per["FETCH_SIZE"] and k in per["WRITE_SIZE"]:
                out[k] = {"fetch": round(per["FETCH_SIZE"][k]), "write": round(per["WRITE_SIZE"][k]),
                          "bytes": round(per["FETCH_SIZE"][k] + per["WRITE_SIZE"][k])}
    log(f"traffic per launch: {out}")
    return out


# ------------------------------------------------------------ cpu oracle --
def oracle_sample(scene, bvh2, views, opts, threads, p_begin=0, p_end=-1):
    """Oracle render of pixels [p_begin, p_end) of every view: (images,
    seconds, stats, per-ray box / triangle tests)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    osc = pyoracle.OracleScene(scene.tris, scene.mats, scene.lights, bvh2=bvh2,
                               textures=getattr(scene, "textures", None))
    t0 = time.perf_counter()
    img, st = osc.render(views, opts, threads=threads, p_begin=p_begin, p_end=p_end)
    dt = time.perf_counter() - t0
    per_ray = {"box_closest": st["box_closest"] / max(1, st["rays_closest"]),
               "tri_closest": st["tri_closest"] / max(1, st["rays_closest"]),
               "box_shadow": st["box_shadow"] / max(1, st["rays_shadow"]),
               "tri_shadow": st["tri_shadow"] / max(1, st["rays_shadow"])}
    return img, dt, st, per_ray


def b_ray(per_ray: dict, kind: str) -> float:
    return 32 + 16 + 32 * per_ray[f"box_{kind}"] + 40 * per_ray[f"tri_{kind}"]


def roofline_entry(kernel: str, bytes_per_launch: float, ms_total: float, launches: int, bw: dict | None,
                   cached: bool, traffic: dict | None = None, unit_bytes: dict | None = None) -> dict:
    """One kernel's roofline record.  cached: the kernel's reads are mostly
    served by the caches (the traversal kernels: the measured HBM traffic is
    a fraction of the algorithmic bytes), so its bound is the measured L2
    read bandwidth; otherwise (and without that measurement) the 8 TB/s HBM
    peak.  The HBM figures stay beside: the algorithmic bytes over the HBM
    peak and the measured traffic's rate over the HBM peak."""
    avg_s = ms_total / max(1, launches) / 1e3
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    if cached and bw:
        bound, peak = "l2", bw["l2"]
    else:
        bound, peak = "hbm", HBM_PEAK_GBS
    e = {"bound": bound, "achieved": round(achieved, 1), "peak": peak, "unit": "GB/s",
         "frac": round(achieved / peak, 4), "traffic": traffic["bytes"] if traffic else None, "kernel": kernel,
         "avg_launch_ms": round(avg_s * 1e3, 4), "launches": launches,
         "algorithmic_bytes_per_launch": round(bytes_per_launch)}
    if bound != "hbm":
        e["frac_of_hbm_peak_algorithmic"] = round(achieved / HBM_PEAK_GBS, 4)
    if unit_bytes:
        e.update(unit_bytes)
    if traffic:
        e["traffic_fetch"] = traffic["fetch"]
        e["traffic_write"] = traffic["write"]
        e["traffic_over_algorithmic"] = round(traffic["bytes"] / bytes_per_launch, 4)
        e["traffic_gbs"] = round(traffic["bytes"] / avg_s / 1e9, 1)
        e["traffic_frac_of_hbm_peak"] = round(traffic["bytes"] / avg_s / 1e9 / HBM_PEAK_GBS, 4)
    return e


# --------------------------------------------------------------- headline --
def headline(a, threads, host, traffic, bw):
    """The large combined configuration: K timed forward frames (4 views,
    batched), its adjoint, the CPU baseline and the three kernels' rooflines."""
    import numpy as np

    from globillum_amd import scenes
    from globillum_amd._lib import FLAG_BATCHED
    from globillum_amd.context import Context, build_bvh2

    sc = scenes.config_scene(5)
    flags = FLAG_BATCHED
    ctx = Context(0)
    ctx.upload(sc)
    log(f"large configuration uploaded ({len(sc.tris)} world triangles, {len(sc.instances)} instances); "
        f"warmup {a.warmup}")
    for _ in range(a.warmup):
        ctx.render(flags=flags)
    keys = ("rays_closest", "rays_shadow", "samples", "ms_trace_closest", "ms_trace_any", "ms_shade",
            "launches_closest")
    tot = dict.fromkeys(keys, 0)
    t0 = time.perf_counter()
    for i in range(a.steps):
        ctx.render(flags=flags)  # returns after the stream synchronised
        c = ctx.counters()
        for k in keys:
            tot[k] += c[k]
        log(f"step {i + 1}/{a.steps}")
    dt = time.perf_counter() - t0
    log(f"{a.steps} steps in {dt:.3f} s")
    imgs = [ctx.image(v) for v in range(len(sc.views))]
    assert all(np.isfinite(im).all() for im in imgs)
    # one untimed frame through the counting build of the traversal kernels
    ctx.traversal_stats(True)
    ctx.render(flags=flags)
    ts = ctx.traversal_stats()
    ctx.traversal_stats(False)
    # the adjoint: gradients of an L2 loss against a target rendered with
    # perturbed parameters (same geometry, texture and views)
    adj = None
    if not a.no_adjoint:
        tgt = scenes.perturbed(sc, seed=1)
        ctx.set_materials(tgt.mats)
        ctx.set_lights(tgt.lights)
        ctx.render(flags=flags)
        timgs = [ctx.image(v) for v in range(len(sc.views))]
        ctx.set_materials(sc.mats)
        ctx.set_lights(sc.lights)
        ctx.render(flags=flags)
        dL = np.concatenate([(2.0 * (x - y) / x.size).reshape(-1, 3) for x, y in zip(imgs, timgs)]).astype(np.float32)
        ctx.adjoint(dL)  # warm (gradient buffers)
        n_adj = max(1, min(a.steps, 3))
        t1 = time.perf_counter()
        for _ in range(n_adj):
            gm, ge = ctx.adjoint(dL)
        t_adj = (time.perf_counter() - t1) / n_adj
        gt = ctx.texture_grad(0)
        n_par = gm.size + ge.size + gt.size
        adj = {"ms_adjoint": round(t_adj * 1e3, 1), "ms_forward": round(dt / a.steps * 1e3, 1),
               "adjoint_over_forward": round(t_adj / (dt / a.steps), 3), "grads_per_s": round(1.0 / t_adj, 3),
               "params_per_gradient": int(n_par), "texels_with_gradient": int((np.abs(gt).sum(-1) > 0).sum()),
               "loss": float(sum(((x - y) ** 2).sum() for x, y in zip(imgs, timgs)) /
                             sum(x.size for x in imgs)),
               "grad_finite": bool(np.isfinite(gm).all() and np.isfinite(ge).all() and np.isfinite(gt).all())}
        log(f"adjoint: {adj}")
    ctx.close()

    # CPU baseline and per-ray test counts: the oracle on the flattened scene
    # (the reference has no instancing; one BVH over all 4.51M triangles,
    # identical triangle ids), one band of pixels in the middle of every view,
    # sized to about 12 s of CPU work
    base, per_ray, rmse = None, None, None
    if not a.no_cpu:
        flat = scenes.flatten(sc)
        bvh2 = build_bvh2(flat.tris)
        W, H = int(sc.views["w"][0]), int(sc.views["h"][0])
        mid = (H // 2) * W
        cal = 4 * W
        _, t_cal, st_cal, _ = oracle_sample(flat, bvh2, sc.views, sc.opts, threads, mid, mid + cal)
        rate = (st_cal["rays_closest"] + st_cal["rays_shadow"]) / max(t_cal, 1e-3)
        px = int(min(H * W // 4, max(cal, 12.0 * rate / max(1.0, (st_cal["rays_closest"] + st_cal["rays_shadow"]) / cal))))
        px = max(W, px // W * W)
        p0 = mid - px // 2 // W * W
        img, t_cpu, st, per_ray = oracle_sample(flat, bvh2, sc.views, sc.opts, threads, p0, p0 + px)
        rays = st["rays_closest"] + st["rays_shadow"]
        base = {"value": round(rays / t_cpu / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                "sample": f"the oracle on rows {p0 // W}..{(p0 + px) // W - 1} of each of the 4 views (16 spp, depth 6; "
                          f"the spheres flattened into one BVH, as the reference has no instancing): "
                          f"{rays} rays in {t_cpu:.2f} s on {threads} threads",
                "samples_per_s": round(st["samples"] / t_cpu, 1), "host": host}
        # the GPU image on the same pixels: the instanced scene against the
        # oracle's flattened one (transforms applied in float64 there)
        ref = np.asarray(img).reshape(-1, 3)
        got = np.concatenate([im.reshape(-1, 3) for im in imgs])
        npix = W * H
        sel = np.concatenate([np.arange(v * npix + p0, v * npix + p0 + px) for v in range(len(sc.views))])
        rmse = float(np.sqrt(((got[sel] - ref[sel]) ** 2).mean()) / np.sqrt((ref[sel] ** 2).mean()))
        log(f"cpu baseline: {base['value']} Mrays/s; GPU vs oracle on the sample: relative RMSE {rmse:.3g}")

    L = max(1, tot["launches_closest"])
    roof = None
    if per_ray is not None:
        rc = tot["rays_closest"] / L
        rs = tot["rays_shadow"] / L
        tr = traffic or {}
        roof = roofline_entry("k_trace_closest (two-level)", rc * b_ray(per_ray, "closest"), tot["ms_trace_closest"],
                              L, bw, True, traffic=tr.get("large.trace_closest"),
                              unit_bytes={"algorithmic_bytes_per_ray": round(b_ray(per_ray, "closest"), 1),
                                          "rays_per_launch": round(rc)})
        roof["others"] = {
            "k_trace_any": roofline_entry("k_trace_any (two-level)", rs * b_ray(per_ray, "shadow"),
                                          tot["ms_trace_any"], L, bw, True, traffic=tr.get("large.trace_any"),
                                          unit_bytes={"algorithmic_bytes_per_ray": round(b_ray(per_ray, "shadow"), 1),
                                                      "rays_per_launch": round(rs)}),
            "k_shade_tex": roofline_entry("k_shade_tex", rc * B_VERTEX_TEX, tot["ms_shade"], L, bw, False,
                                          traffic=tr.get("large.shade"),
                                          unit_bytes={"algorithmic_bytes_per_vertex": B_VERTEX_TEX,
                                                      "vertices_per_launch": round(rc)})}
        req = 32 + 16 + (128 * ts["closest_nodes"] + 48 * ts["closest_triangles"]) / max(1, ts["closest_rays"])
        roof["wide_traversal"] = {"nodes_per_ray": round(ts["closest_nodes"] / max(1, ts["closest_rays"]), 3),
                                  "triangles_per_ray": round(ts["closest_triangles"] / max(1, ts["closest_rays"]), 3),
                                  "instances_per_ray": round(ts["closest_instances"] / max(1, ts["closest_rays"]), 3),
                                  "requested_bytes_per_ray": round(req, 1)}
        if bw:
            roof["measured_read_gbs"] = bw
    rays = tot["rays_closest"] + tot["rays_shadow"]
    out = {
        "metric": METRIC,
        "value": round(rays / dt / 1e6, 2),
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
        "config": {"workload": "large_combined: 2048x1024-cell fBm heightfield (4194304 triangles) + 2000 instances "
                               "of a 160-triangle sphere (two-level BVH), 256 principled materials with a 1024x1024 "
                               "RGB albedo texture, 64 rect lights, 4 views 3840x2160, 16 spp, depth 6 (the large "
                               "combined configuration); one step = one forward frame of the 4 views",
                   "mode": "batched wavefront",
                   "samples_per_step": tot["samples"] // max(1, a.steps)},
        "samples_per_s": round(tot["samples"] / dt, 1),
        "adjoint_grads_per_s": adj["grads_per_s"] if adj else None,
        "adjoint": adj,
        "rmse_rel_vs_oracle_sample": round(rmse, 8) if rmse is not None else None,
        "rays_per_step": {"closest": tot["rays_closest"] // a.steps, "shadow": tot["rays_shadow"] // a.steps},
        "kernel_ms_per_step": {"trace_closest": round(tot["ms_trace_closest"] / a.steps, 3),
                               "trace_shadow": round(tot["ms_trace_any"] / a.steps, 3),
                               "shade": round(tot["ms_shade"] / a.steps, 3)},
        "roofline": roof,
        "cpu_baseline": base,
        "traffic_probe": traffic,
    }
    return out


# ------------------------------------------------------------------- legs --
def plain_leg(steps: int, threads: int, host, traffic, bw, with_cpu: bool):
    """The ~300k-triangle plain configuration (256x256-cell heightfield + 1000
    UV spheres, 291,072 triangles, 32 principled materials, 4 rect lights,
    1280x720, 64 spp, depth 8), batched.  Its BVH and triangles (~27 MB) sit
    in the caches, so the closest-hit kernel's algorithmic bytes are held
    against the measured L2 read bandwidth (and, beside it, the HBM peak)."""
    from globillum_amd import scenes
    from globillum_amd._lib import FLAG_BATCHED
    from globillum_amd.context import Context, build_bvh2

    sc = scenes.config_scene(2)
    bvh2 = build_bvh2(sc.tris)
    keys = ("rays_closest", "rays_shadow", "samples", "ms_trace_closest", "ms_trace_any", "ms_shade",
            "launches_closest")
    tot = dict.fromkeys(keys, 0)
    with Context(0) as ctx:
        ctx.upload(sc, bvh2=bvh2)
        ctx.render(flags=FLAG_BATCHED)
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.render(flags=FLAG_BATCHED)
            c = ctx.counters()
            for k in keys:
                tot[k] += c[k]
        dt = time.perf_counter() - t0
        ctx.traversal_stats(True)
        ctx.render(flags=FLAG_BATCHED)
        ts = ctx.traversal_stats()
        ctx.traversal_stats(False)
    rays = tot["rays_closest"] + tot["rays_shadow"]
    out = {"workload": "terrain_spheres: 256x256-cell fBm heightfield + 1000 UV spheres (291072 triangles), 32 "
                       "principled materials, 4 rect lights, 1280x720, 64 spp, depth 8, batched",
           "mrays_per_s": round(rays / dt / 1e6, 2), "ms_per_frame": round(dt / steps * 1e3, 3),
           "samples_per_s": round(tot["samples"] / dt, 1),
           "kernel_ms_per_frame": {"trace_closest": round(tot["ms_trace_closest"] / steps, 3),
                                   "trace_shadow": round(tot["ms_trace_any"] / steps, 3),
                                   "shade": round(tot["ms_shade"] / steps, 3)}}
    if with_cpu:
        v = sc.views.copy()
        v["w"], v["h"], v["spp"] = 640, 360, 32
        _, t_cpu, st, per_ray = oracle_sample(sc, bvh2, v, sc.opts, threads)
        r = st["rays_closest"] + st["rays_shadow"]
        out["cpu_baseline"] = {"value": round(r / t_cpu / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                               "sample": f"same scene/camera/depth/BVH at 640x360, 32 spp (1/4 of the samples): "
                                         f"{r} rays in {t_cpu:.2f} s on {threads} threads", "host": host}
        L = max(1, tot["launches_closest"])
        rc, rs = tot["rays_closest"] / L, tot["rays_shadow"] / L
        tr = traffic or {}
        roof = {}
        for name, nbytes, ms, tk, ub in (
                ("k_trace_closest", rc * b_ray(per_ray, "closest"), tot["ms_trace_closest"], "plain.trace_closest",
                 {"algorithmic_bytes_per_ray": round(b_ray(per_ray, "closest"), 1)}),
                ("k_trace_any", rs * b_ray(per_ray, "shadow"), tot["ms_trace_any"], "plain.trace_any",
                 {"algorithmic_bytes_per_ray": round(b_ray(per_ray, "shadow"), 1)}),
                ("k_shade", rc * B_VERTEX, tot["ms_shade"], "plain.shade", {"algorithmic_bytes_per_vertex": B_VERTEX})):
            roof[name] = roofline_entry(name, nbytes, ms, L, bw, name != "k_shade", traffic=tr.get(tk), unit_bytes=ub)
        req = 32 + 16 + (128 * ts["closest_nodes"] + 48 * ts["closest_triangles"]) / max(1, ts["closest_rays"])
        roof["k_trace_closest"]["wide_traversal"] = {
            "nodes_per_ray": round(ts["closest_nodes"] / max(1, ts["closest_rays"]), 3),
            "triangles_per_ray": round(ts["closest_triangles"] / max(1, ts["closest_rays"]), 3),
            "requested_bytes_per_ray": round(req, 1)}
        out["roofline"] = roof
    log(f"plain leg: {out['mrays_per_s']} Mrays/s")
    return out


def small_leg(steps: int, threads: int, host, bw: dict | None, tune: bool, with_cpu: bool, cache: str):
    """The small plain configuration (cornell_boxes: 5 walls and 2 boxes, 34
    triangles, 1 rect light of emission 10, Lambert albedo U[0.2, 0.8],
    256x256, 16 spp, depth 4), rendered as one hipGraph with autotuned
    launch parameters: Mrays/s, samples/s, the relative RMSE of its image
    against the oracle's 1024-spp image, and the closest-hit kernel's
    algorithmic bytes against the measured L2 read bandwidth (its BVH and
    triangles take a few KB)."""
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
            tuned = autotune.tune(ctx, sc.views, o, space=space, repeats=3, cache=cache,
                                  key=autotune.workload_key(sc.views, o, "cornell_boxes"))["params"]
    with Context(0) as ctx:
        # a fresh context: the first graph render captures and instantiates,
        # the later ones relaunch the kept graph
        ctx.upload(sc, bvh2=bvh2)
        if tuned:
            ctx.set_tuning(**tuned)
        ctx.render(flags=flags)
        graph_ms = ctx.counters()["ms_graph_build"]
        for _ in range(2):
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
    rays = tot["rays_closest"] + tot["rays_shadow"]
    out = {"workload": "cornell_boxes: 5 walls + 2 boxes (34 triangles), 1 rect light, Lambert, 256x256, 16 spp, "
                       "depth 4 (the small plain configuration), batched + hipGraph (kept and relaunched)",
           "mrays_per_s": round(rays / dt / 1e6, 2), "samples_per_s": round(tot["samples"] / dt, 1),
           "ms_per_frame": round(dt / steps * 1e3, 3), "ms_graph_capture_once": round(graph_ms, 3), "tuned": tuned}
    if with_cpu:
        _, t_cpu, st, per_ray = oracle_sample(sc, bvh2, sc.views, sc.opts, threads)
        r = st["rays_closest"] + st["rays_shadow"]
        out["cpu_baseline"] = {"value": round(r / t_cpu / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                               "sample": f"the whole workload (256x256, 16 spp): {r} rays in {t_cpu:.2f} s on "
                                         f"{threads} threads", "host": host}
        v = sc.views.copy()
        v["spp"] = 1024
        ref, _, _, _ = oracle_sample(sc, bvh2, v, sc.opts, threads)
        ref = np.asarray(ref, np.float64).reshape(img.shape)
        out["rmse_rel_vs_oracle_1024spp"] = round(float(np.sqrt(((img - ref) ** 2).mean()) /
                                                        np.sqrt((ref ** 2).mean())), 5)
        if kern["launches_closest"] > 0 and kern["ms_trace_closest"] > 0:
            L = kern["launches_closest"]
            out["roofline"] = roofline_entry("k_trace_closest", kern["rays_closest"] / L * b_ray(per_ray, "closest"),
                                             kern["ms_trace_closest"], L, bw, True,
                                             unit_bytes={"algorithmic_bytes_per_ray": round(b_ray(per_ray, "closest"), 1)})
    log(f"small leg: {out['mrays_per_s']} Mrays/s")
    return out


def batching_leg(steps: int, tune: bool, cache: str):
    """16 views of the cornell scene (resolutions from 256x256 to 1920x1080,
    4/16/64 spp, depth 8, Russian roulette from 3).  Three schedules of the
    same kernels, with identical image bits: the reference's (one view after
    the other, one sample per pixel per wavefront, the host reading each
    bounce's queue length; src/gpu-pt.h:102-226); per view with full passes
    (all of a view's samples in flight, host-read queue lengths); and the
    batched renderer (all views and samples in shared queues, device-side
    lengths, the frame as one kept hipGraph, autotuned)."""
    import numpy as np

    from globillum_amd import autotune, scenes
    from globillum_amd._lib import FLAG_BATCHED, FLAG_GRAPH
    from globillum_amd.context import Context

    sc = scenes.config_scene(3)
    ref_sched = sc.opts.copy()
    ref_sched["flags"] = 0
    ref_sched["samples_per_pass"] = 1
    full = sc.opts.copy()
    full["flags"] = 0
    full["samples_per_pass"] = 0
    batched = sc.opts.copy()
    batched["flags"] = FLAG_BATCHED | FLAG_GRAPH
    out = {"workload": "16 views of the cornell scene, 256x256..1920x1080, spp 4/16/64, depth 8, RR from 3",
           "samples": int(sum(int(v["w"]) * int(v["h"]) * int(v["spp"]) for v in sc.views))}
    with Context(0) as ctx:
        ctx.upload(sc)
        res = {}
        for name, o in (("reference_schedule", ref_sched), ("per_view_full_passes", full), ("batched", batched)):
            if name == "batched" and tune:
                t = autotune.tune(ctx, sc.views, o, repeats=3, cache=cache,
                                  key=autotune.workload_key(sc.views, o, "cornell_batch"),
                                  log=lambda m: print(m, file=sys.stderr, flush=True))
                out["tuned"] = t["params"]
                out["tuned_from_cache"] = t["cached"]
            ctx.render(sc.views, o)
            rays, times = 0, []
            for _ in range(steps):
                t0 = time.perf_counter()
                ctx.render(sc.views, o)
                times.append(time.perf_counter() - t0)
                c = ctx.counters()
                rays += c["rays_closest"] + c["rays_shadow"]
            dt = sum(times)
            res[name] = {"mrays_per_s": round(rays / dt / 1e6, 2), "ms_per_frame": round(dt / steps * 1e3, 2),
                         "ms_per_frame_min": round(min(times) * 1e3, 2), "image": [im.copy() for im in ctx.images()]}
            log(f"batching leg {name}: {res[name]['mrays_per_s']} Mrays/s")
        same = all(all(np.array_equal(a, b) for a, b in zip(res[k]["image"], res["batched"]["image"]))
                   for k in ("reference_schedule", "per_view_full_passes"))
    for k in res:
        res[k].pop("image")
    out.update(res)
    out["speedup_vs_reference_schedule"] = round(res["batched"]["mrays_per_s"] / res["reference_schedule"]["mrays_per_s"], 3)
    out["speedup_vs_per_view_full_passes"] = round(res["batched"]["mrays_per_s"] /
                                                   res["per_view_full_passes"]["mrays_per_s"], 3)
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
    log(f"adjoint leg: fwd {out['ms_forward']} ms, adj {out['ms_adjoint']} ms")
    return out


def bvh_leg(steps: int):
    """BVH construction: the host binned-SAH builder against the device
    linear-BVH, PLOC (radius 16) and binned-SAH builders on the ~300k-triangle
    scene and on a 2048x1024-cell heightfield (4.19M triangles, the large
    configuration's terrain), plus the ~300k-triangle frame time traversing
    each tree."""
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
        for name, method in (("sah", None), ("lbvh", "device"), ("ploc", "ploc"), ("device_sah", "sah")):
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
                                   "device_sah_ms": round(dev_ms["device_sah"], 2),
                                   "frame_ms_sah": round(frames["sah"], 2), "frame_ms_lbvh": round(frames["lbvh"], 2),
                                   "frame_ms_ploc": round(frames["ploc"], 2),
                                   "frame_ms_device_sah": round(frames["device_sah"], 2)}
    rng = np.random.default_rng(5)
    hf, _ = scenes.heightfield(rng, 2048, 1024, 200.0, 100.0, 10.0, 8)
    big = scenes.Scene(hf, sc.mats, sc.lights, sc.views, sc.opts)
    with Context(0) as ctx:
        t0 = time.perf_counter()
        ctx.upload(big)
        host_ms = (time.perf_counter() - t0) * 1e3
        dev_ms = {}
        for name, method in (("lbvh", "device"), ("ploc", "ploc"), ("device_sah", "sah")):
            ctx.build_bvh_device(4, method)
            t0 = time.perf_counter()
            ctx.build_bvh_device(4, method)
            dev_ms[name] = (time.perf_counter() - t0) * 1e3
    out["heightfield_4194k"] = {"host_sah_ms_incl_upload": round(host_ms, 1), "device_lbvh_ms": round(dev_ms["lbvh"], 2),
                                "device_ploc_ms": round(dev_ms["ploc"], 2),
                                "device_sah_ms": round(dev_ms["device_sah"], 2)}
    log(f"bvh leg: {out}")
    return out


def variants_leg(steps: int):
    """The large configuration's forward frame with the spheres flattened
    into the world BVH (4.51M triangles), and instanced without its
    texture: Mrays/s and kernel times beside the headline's two-level,
    textured frame."""
    from globillum_amd import scenes
    from globillum_amd._lib import FLAG_BATCHED
    from globillum_amd.context import Context

    sc = scenes.config_scene(5)
    untextured = dataclasses.replace(sc, mats=sc.mats.copy(), textures=None)
    untextured.mats["texture"] = -1
    out = {}
    for name, s in (("flattened", scenes.flatten(sc)), ("instanced_untextured", untextured)):
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
            out[name] = {"mrays_per_s": round(rays / dt / 1e6, 1), "ms_per_frame": round(dt / steps * 1e3, 1),
                         "upload_and_bvh_ms": round(up_ms, 1),
                         "kernel_ms_per_frame": {k[3:]: round(v / steps, 2) for k, v in kms.items()}}
        log(f"variants leg: {name} {out[name]}")
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the path shards by view / image "
              f"tile / spp range; a multi-GPU run is not implemented", file=sys.stderr)
        sys.exit(2)
    if a.traffic_probe:
        traffic_probe()
        return
    if a.copy_mib > 0:
        rc = membench().gi_membench_copy(0, a.copy_mib << 20)
        if rc != 0:
            raise RuntimeError(f"gi_membench_copy failed with HIP error {rc}")
        log(f"copied {a.copy_mib} MiB device to device")
    threads, host = host_cpus()
    log(f"host: {host}; CPU baselines on {threads} threads")
    # counter runs first, in child processes, before this process uses the GPU
    traffic = None if a.no_traffic else run_traffic_probes()
    bw = None if a.no_membench else read_bandwidth()
    out = headline(a, threads, host, traffic, bw)
    legs = max(1, min(a.steps, 3))
    if not a.no_plain:
        out["plain"] = plain_leg(legs, threads, host, traffic, bw, not a.no_cpu)
    if not a.no_small:
        out["small"] = small_leg(20, threads, host, bw, not a.no_tune, not a.no_cpu, a.tuning_cache)
    if not a.no_batching:
        out["batching"] = batching_leg(legs, not a.no_tune, a.tuning_cache)
    if not a.no_adjoint:
        out["adjoint_diff_config"] = adjoint_leg(legs)
    if not a.no_bvh:
        out["bvh_build"] = bvh_leg(legs)
    if not a.no_instancing:
        out["large_variants"] = variants_leg(max(1, min(a.steps, 2)))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
