"""Procedural scenes of the benchmark configurations (seeded, numpy).

There is no scene file for these workloads; each generator returns a
:class:`Scene` -- triangle, material and light records in the C-ABI layout
(the simple_triangle field set of src/hl-main.cpp:332-348, material_t /
PrincipledBRDFParameters, gi::light rect lights) plus its views.

  cornell_boxes      closed box of 5 walls + 2 boxes, 1 rect light (34 tris)
  terrain_spheres    fBm heightfield + UV spheres, principled materials
  instanced_terrain  fBm heightfield + instanced meshes (two-level BVH)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ._lib import INSTANCE, LIGHT, MATERIAL, OPTS, TRI, VIEW


@dataclass
class Scene:
    tris: np.ndarray
    mats: np.ndarray
    lights: np.ndarray
    views: np.ndarray
    opts: np.ndarray
    name: str = ""
    meta: dict = field(default_factory=dict)
    sky: np.ndarray | None = None  # (h, w, 3) float32 map of the scene's sky light
    # two-level instancing: object-space meshes (TRI arrays) and INSTANCE
    # records placing them; tris stays the world geometry
    meshes: list | None = None
    instances: np.ndarray | None = None
    # RGBA8 textures, (h, w, 4) uint8, row 0 at t = 1; gi_material.texture
    # names one by its index
    textures: list | None = None


def rect_light_power(w: float, h: float, col) -> np.float32:
    """add-rectlight's selection weight (src/lights.cpp:92), in float32."""
    col = np.asarray(col, np.float32)
    return np.float32(np.float32(w) * np.float32(h) * (col[0] + col[1] + col[2]) * np.float32(0.333))


def make_light(center, direction, col, w, h) -> np.ndarray:
    l = np.zeros(1, LIGHT)
    l["center"] = center
    l["dir"] = direction
    l["up"] = (0.0, 0.0, 1.0)
    l["col"] = col
    l["wh"] = (w, h)
    l["power"] = rect_light_power(w, h, col)
    l["type"] = 0
    return l


def sky_light_power(scale: float, diameter: float) -> np.float32:
    """add-skylight's selection weight s*4*pi*(d*.5)^2 (src/lights.cpp:104),
    evaluated in double as there."""
    s4 = np.float32(np.float32(scale) * np.float32(4.0))
    d = float(np.float32(diameter))
    return np.float32(float(s4) * np.pi * (d * 0.5) * (d * 0.5))


def make_sky_light(scale: float, diameter: float) -> np.ndarray:
    """A GI_LIGHT_SKY record: scale in col[0]; its map goes in Scene.sky."""
    l = np.zeros(1, LIGHT)
    l["col"] = (scale, 0.0, 0.0)
    l["power"] = sky_light_power(scale, diameter)
    l["type"] = 1
    return l


def lambert(rgb) -> np.ndarray:
    m = np.zeros(1, MATERIAL)
    m["diffuse_color"] = rgb
    m["alpha"] = 1.0
    m["principled"] = 0
    m["opacity"] = 1.0
    m["transmission"] = 1.0
    m["color"] = rgb
    m["texture"] = -1
    return m


def principled(color, roughness, metallic, **kw) -> np.ndarray:
    """PrincipledBRDFParameters defaults (src/principledParameters.h:23-24)
    with the given base colour, roughness and metallic."""
    m = np.zeros(1, MATERIAL)
    m["diffuse_color"] = color
    m["alpha"] = 1.0
    m["principled"] = 1
    m["color"] = color
    m["roughness"] = roughness
    m["metallic"] = metallic
    m["opacity"] = 1.0
    m["transmission"] = 1.0
    m["texture"] = -1
    for k, v in kw.items():
        m[k] = v
    return m


def make_view(pos, look_at, up, fovy, w, h, spp, aperture=0.0, focus=1.0) -> np.ndarray:
    v = np.zeros(1, VIEW)
    pos = np.asarray(pos, np.float32)
    d = np.asarray(look_at, np.float32) - pos
    v["pos"] = pos
    v["dir"] = d / np.linalg.norm(d)
    v["up"] = up
    v["fovy"] = fovy
    v["w"], v["h"], v["spp"] = w, h, spp
    v["aperture"] = aperture
    v["focus_distance"] = focus
    return v


def make_opts(depth, rr_start=0, seed=0, flags=0, offset_bounce=0.0, offset_shadow=0.0,
              background=(1.0, 1.0, 1.0), samples_per_pass=0) -> np.ndarray:
    o = np.zeros(1, OPTS)
    o["max_depth"] = depth
    o["rr_start"] = rr_start
    o["seed"] = seed
    o["flags"] = flags
    o["offset_bounce"] = offset_bounce
    o["offset_shadow"] = offset_shadow
    o["background"] = background
    o["samples_per_pass"] = samples_per_pass
    return o


def _quads_to_tris(quads: np.ndarray, normals: np.ndarray, mat: np.ndarray) -> np.ndarray:
    """quads (n,4,3) counter-clockwise, normals (n,3), mat (n,) -> 2n tris."""
    n = len(quads)
    t = np.zeros(2 * n, TRI)
    t["a"][0::2], t["b"][0::2], t["c"][0::2] = quads[:, 0], quads[:, 1], quads[:, 2]
    t["a"][1::2], t["b"][1::2], t["c"][1::2] = quads[:, 0], quads[:, 2], quads[:, 3]
    for k in ("na", "nb", "nc"):
        t[k][0::2] = normals
        t[k][1::2] = normals
    t["material_index"][0::2] = mat
    t["material_index"][1::2] = mat
    return t


def _box_quads(lo, hi, inward: bool):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [  # (quad, outward normal)
        ([(x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)], (0, -1, 0)),
        ([(x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0)], (0, 1, 0)),
        ([(x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0)], (0, 0, -1)),
        ([(x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)], (0, 0, 1)),
        ([(x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)], (-1, 0, 0)),
        ([(x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1)], (1, 0, 0)),
    ]
    q = np.array([f[0] for f in faces], np.float32)
    n = np.array([f[1] for f in faces], np.float32)
    return q, (-n if inward else n)


def cornell_boxes(seed: int = 1, w: int = 256, h: int = 256, spp: int = 16, depth: int = 4) -> Scene:
    """Closed box [0,1]^3 of 5 walls (open towards -z, 10 triangles), two
    axis-aligned boxes of seeded size U[0.1,0.35]^3 resting on the floor (12
    triangles each), one 0.25 x 0.25 rect light of emission 10 just below the
    ceiling, Lambert albedo U[0.2,0.8]^3 per object; pinhole camera at
    (0.5, 0.5, -1.3) looking +z, fovy 40."""
    rng = np.random.default_rng(seed)
    q, n = _box_quads((0, 0, 0), (1, 1, 1), inward=True)
    walls = [0, 1, 3, 4, 5]  # floor, ceiling, back, left, right (no front)
    parts = [_quads_to_tris(q[walls], n[walls], np.arange(5, dtype=np.int32))]
    mats = [lambert(rng.uniform(0.2, 0.8, 3).astype(np.float32)) for _ in range(5)]
    for b in range(2):
        ext = rng.uniform(0.1, 0.35, 3).astype(np.float32)
        x = rng.uniform(0.05, 0.95 - ext[0])
        z = rng.uniform(0.05, 0.95 - ext[2])
        lo = np.array([x, 0.0, z], np.float32)
        bq, bn = _box_quads(lo, lo + ext, inward=False)
        parts.append(_quads_to_tris(bq, bn, np.full(6, 5 + b, np.int32)))
        mats.append(lambert(rng.uniform(0.2, 0.8, 3).astype(np.float32)))
    tris = np.concatenate(parts)
    lights = make_light((0.5, 0.999, 0.5), (0.0, -1.0, 0.0), (10.0, 10.0, 10.0), 0.25, 0.25)
    view = make_view((0.5, 0.5, -1.3), (0.5, 0.5, 0.0), (0, 1, 0), 40.0, w, h, spp)
    return Scene(tris, np.concatenate(mats), lights, view, make_opts(depth), name="cornell_boxes")


# ------------------------------------------------------------- terrain ----
def _value_noise(rng, res: int):
    return rng.uniform(-1.0, 1.0, (res + 1, res + 1))


def _fbm(x, z, rng, octaves=6, gain=0.5, lacunarity=2.0, base_res=4):
    """Fractal sum of bilinearly interpolated value noise on [0,1]^2."""
    h = np.zeros_like(x, dtype=np.float64)
    amp, res = 1.0, base_res
    for _ in range(octaves):
        g = _value_noise(rng, res)
        fx, fz = x * res, z * res
        ix = np.clip(np.floor(fx).astype(int), 0, res - 1)
        iz = np.clip(np.floor(fz).astype(int), 0, res - 1)
        tx, tz = fx - ix, fz - iz
        tx = tx * tx * (3 - 2 * tx)
        tz = tz * tz * (3 - 2 * tz)
        v = (g[ix, iz] * (1 - tx) * (1 - tz) + g[ix + 1, iz] * tx * (1 - tz) + g[ix, iz + 1] * (1 - tx) * tz
             + g[ix + 1, iz + 1] * tx * tz)
        h += amp * v
        amp *= gain
        res = int(res * lacunarity)
    return h


def heightfield(rng, nx: int, nz: int, size_x: float, size_z: float, amplitude: float, n_patch_mats: int,
                mat_offset: int = 0, uv_repeat: float = 1.0):
    """nx x nz cells -> 2*nx*nz triangles with smooth vertex normals and
    texture coordinates (x, z) / size * uv_repeat; the material index cycles
    over 8x8 patches."""
    xs = np.linspace(0.0, 1.0, nx + 1)
    zs = np.linspace(0.0, 1.0, nz + 1)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    H = amplitude * _fbm(X, Z, rng)
    P = np.stack([X * size_x, H, Z * size_z], -1).astype(np.float64)
    dx = np.gradient(H, size_x / nx, axis=0)
    dz = np.gradient(H, size_z / nz, axis=1)
    N = np.stack([-dx, np.ones_like(dx), -dz], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    i, k = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    i, k = i.ravel(), k.ravel()
    v00, v10, v11, v01 = (i, k), (i + 1, k), (i + 1, k + 1), (i, k + 1)
    t = np.zeros(2 * len(i), TRI)
    t["a"][0::2], t["b"][0::2], t["c"][0::2] = P[v00], P[v01], P[v11]
    t["na"][0::2], t["nb"][0::2], t["nc"][0::2] = N[v00], N[v01], N[v11]
    t["a"][1::2], t["b"][1::2], t["c"][1::2] = P[v00], P[v11], P[v10]
    t["na"][1::2], t["nb"][1::2], t["nc"][1::2] = N[v00], N[v11], N[v10]
    UV = np.stack([X, Z], -1) * uv_repeat  # texture coordinates: (x, z) over the extent
    t["ta"][0::2], t["tb"][0::2], t["tc"][0::2] = UV[v00], UV[v01], UV[v11]
    t["ta"][1::2], t["tb"][1::2], t["tc"][1::2] = UV[v00], UV[v11], UV[v10]
    patch = ((i * 8) // nx) * 8 + (k * 8) // nz
    mi = (mat_offset + patch % n_patch_mats).astype(np.int32)
    t["material_index"][0::2] = mi
    t["material_index"][1::2] = mi
    return t, H


def uv_sphere(center, radius, segments=10, rings=9, material=0) -> np.ndarray:
    """UV sphere of 2*segments*(rings-1) triangles (160 for 10 x 9) with
    smooth normals."""
    th = np.linspace(0.0, np.pi, rings + 1)
    ph = np.linspace(0.0, 2 * np.pi, segments + 1)
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    N = np.stack([np.sin(T) * np.cos(Pp), np.cos(T), np.sin(T) * np.sin(Pp)], -1)
    V = np.asarray(center)[None, None, :] + radius * N
    tris = []
    for r in range(rings):
        for s in range(segments):
            a, b, c, d = (r, s), (r + 1, s), (r + 1, s + 1), (r, s + 1)
            if r > 0:
                tris.append((a, b, d))
            if r < rings - 1:
                tris.append((b, c, d))
    tris = tris
    out = np.zeros(len(tris), TRI)
    UV = np.stack([Pp / (2 * np.pi), 1.0 - T / np.pi], -1)  # texture coordinates (phi, theta)
    for j, key in enumerate(("a", "b", "c")):
        idx = np.array([tt[j] for tt in tris])
        out[key] = V[idx[:, 0], idx[:, 1]]
        out["n" + key] = N[idx[:, 0], idx[:, 1]]
        out["t" + key] = UV[idx[:, 0], idx[:, 1]]
    out["material_index"] = material
    return out


def terrain_spheres(seed: int = 2, cells=(256, 256), n_spheres: int = 1000, n_mats: int = 32, n_lights: int = 4,
                    w: int = 1280, h: int = 720, spp: int = 64, depth: int = 8, size: float = 100.0,
                    radius=(0.1, 2.0)) -> Scene:
    """fBm heightfield (2*cells triangles, 6 octaves, gain 0.5, lacunarity 2,
    amplitude 10) over [0,size]^2 plus n_spheres 160-triangle UV spheres of
    radius U[radius], principled materials (baseColor U[0,1]^3, roughness
    U[0.05,1], metallic in {0,1}), n_lights rect lights above the terrain."""
    rng = np.random.default_rng(seed)
    nx, nz = cells
    ter, H = heightfield(rng, nx, nz, size, size, 10.0, n_mats)
    parts = [ter]
    for _ in range(n_spheres):
        cx, cz = rng.uniform(0.0, size, 2)
        r = rng.uniform(*radius)
        hx = min(int(cx / size * nx), nx)
        hz = min(int(cz / size * nz), nz)
        cy = H[hx, hz] + 0.5 * r
        parts.append(uv_sphere((cx, cy, cz), r, material=int(rng.integers(0, n_mats))))
    tris = np.concatenate(parts)
    mats = np.concatenate([
        principled(rng.uniform(0.0, 1.0, 3).astype(np.float32), rng.uniform(0.05, 1.0),
                   float(rng.random() < 0.5)) for _ in range(n_mats)])
    g = int(np.ceil(np.sqrt(n_lights)))
    lights = []
    for li in range(n_lights):
        lx = (li % g + 0.5) / g * size
        lz = (li // g + 0.5) / g * size
        lights.append(make_light((lx, 40.0, lz), (0.0, -1.0, 0.0), (40.0, 40.0, 40.0), 8.0, 8.0))
    view = make_view((0.5 * size, 40.0, 0.0), (0.5 * size, 0.0, 0.5 * size), (0, 1, 0), 50.0, w, h, spp)
    return Scene(tris, mats, np.concatenate(lights), view, make_opts(depth), name="terrain_spheres")


def sky_map(w: int = 64, h: int = 32, seed: int = 3) -> np.ndarray:
    """A procedural sky: bright zenith-to-horizon gradient, a small sun and
    some noise, (h, w, 3) float32, row 0 at theta = 0."""
    rng = np.random.default_rng(seed)
    th = (np.arange(h) + 0.5) / h * np.pi
    ph = (np.arange(w) + 0.5) / w * 2 * np.pi
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    base = np.stack([0.4 + 0.6 * np.cos(T / 2) ** 2, 0.6 + 0.4 * np.cos(T / 2) ** 2, np.ones_like(T)], -1)
    sun = np.exp(-(((T - 0.6) / 0.08) ** 2 + ((Pp - 1.0) / 0.15) ** 2))[..., None] * np.array([40.0, 36.0, 30.0])
    return (base * (0.8 + 0.4 * rng.random((h, w, 1))) + sun).astype(np.float32)


def sky_scene(w=96, h=64, spp=4, depth=4, with_rect=True, seed=4) -> Scene:
    """Ground plane, a few boxes and a UV sphere under a sky light (and
    optionally a rect light)."""
    rng = np.random.default_rng(seed)
    ground = np.array([[(-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10)]], np.float32)
    tris = [_quads_to_tris(ground, np.array([(0, 1, 0)], np.float32), np.zeros(1, np.int32))]
    for i in range(4):
        lo = np.array([rng.uniform(-4, 3), 0.0, rng.uniform(-2, 5)], np.float32)
        ext = rng.uniform(0.5, 2.0, 3).astype(np.float32)
        bq, bn = _box_quads(lo, lo + ext, inward=False)
        tris.append(_quads_to_tris(bq, bn, np.full(6, 1 + i % 3, np.int32)))
    tris.append(uv_sphere((1.0, 1.2, 0.0), 1.2, material=3))
    mats = np.concatenate([lambert((0.6, 0.6, 0.6)), lambert((0.7, 0.3, 0.2)), lambert((0.2, 0.5, 0.7)),
                           principled((0.8, 0.7, 0.5), roughness=0.4, metallic=1.0)])
    lights = [make_sky_light(0.5, 30.0)]
    if with_rect:
        lights.append(make_light((0.0, 6.0, 2.0), (0, -1, 0), (30, 30, 30), 1.0, 1.0))
    view = make_view((0.0, 3.0, -9.0), (0.0, 1.0, 2.0), (0, 1, 0), 55.0, w, h, spp)
    return Scene(np.concatenate(tris), mats, np.concatenate(lights), view, make_opts(depth),
                 name="sky_scene", sky=sky_map())


def rotation(axis, angle: float) -> np.ndarray:
    """3x3 rotation about `axis` by `angle` radians (Rodrigues)."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def make_instance(A, t, mesh: int, material: int = -1) -> np.ndarray:
    """An INSTANCE record: object-to-world x -> A x + t (rows of (A | t))."""
    r = np.zeros(1, INSTANCE)
    r["xform"] = np.concatenate([np.asarray(A, np.float64), np.asarray(t, np.float64).reshape(3, 1)], 1).ravel()
    r["mesh"] = mesh
    r["material"] = material
    return r


def unit_box(material=0) -> np.ndarray:
    """[-1,1]^3 as 12 outward-facing triangles."""
    q, n = _box_quads(np.float32([-1, -1, -1]), np.float32([1, 1, 1]), inward=False)
    return _quads_to_tris(q, n, np.full(6, material, np.int32))


def instanced_terrain(seed: int = 5, cells=(64, 64), n_instances: int = 200, n_mats: int = 16, n_lights: int = 4,
                      w: int = 160, h: int = 96, spp: int = 4, depth: int = 4, size: float = 40.0,
                      sphere=(10, 9)) -> Scene:
    """fBm heightfield world plus n_instances placed copies of three meshes
    (a UV sphere and a box of radius 1, a coarse sphere): random rotations,
    uniform and non-uniform scales, translations above the terrain; about a
    third of the instances override the mesh's material."""
    rng = np.random.default_rng(seed)
    nx, nz = cells
    ter, H = heightfield(rng, nx, nz, size, size, 4.0, n_mats)
    meshes = [uv_sphere((0.0, 0.0, 0.0), 1.0, *sphere, material=1), unit_box(material=2),
              uv_sphere((0.0, 0.0, 0.0), 1.0, 6, 5, material=3)]
    inst = []
    for _ in range(n_instances):
        cx, cz = rng.uniform(0.0, size, 2)
        hx = min(int(cx / size * nx), nx)
        hz = min(int(cz / size * nz), nz)
        s = rng.uniform(0.2, 1.0)
        S = np.diag(rng.uniform(0.5, 1.5, 3) * s) if rng.random() < 0.4 else np.eye(3) * s
        A = rotation(rng.normal(size=3), rng.uniform(0.0, 2 * np.pi)) @ S
        m = int(rng.integers(0, len(meshes)))
        mat = int(rng.integers(0, n_mats)) if rng.random() < 0.35 else -1
        inst.append(make_instance(A, (cx, H[hx, hz] + 0.6 * s, cz), m, mat))
    mats = np.concatenate([
        principled(rng.uniform(0.0, 1.0, 3).astype(np.float32), rng.uniform(0.05, 1.0),
                   float(rng.random() < 0.5)) for _ in range(n_mats)])
    g = int(np.ceil(np.sqrt(n_lights)))
    lights = []
    for li in range(n_lights):
        lx = (li % g + 0.5) / g * size
        lz = (li // g + 0.5) / g * size
        lights.append(make_light((lx, 20.0, lz), (0.0, -1.0, 0.0), (40.0, 40.0, 40.0), 4.0, 4.0))
    view = make_view((0.5 * size, 14.0, -2.0), (0.5 * size, 0.0, 0.5 * size), (0, 1, 0), 55.0, w, h, spp)
    return Scene(ter, mats, np.concatenate(lights), view, make_opts(depth), name="instanced_terrain",
                 meshes=meshes, instances=np.concatenate(inst))


def large_combined(seed: int = 6, cells=(2048, 1024), n_spheres: int = 2000, n_mats: int = 256,
                   n_lights: int = 64, w: int = 3840, h: int = 2160, spp: int = 16, depth: int = 6,
                   n_views: int = 4, size=(200.0, 100.0), radius=(0.2, 1.5), texture=(1024, 1024)) -> Scene:
    """The large combined configuration: a 2048x1024-cell fBm heightfield
    (4,194,304 triangles) as the world, 2000 instances of one 160-triangle
    unit UV sphere (uniform scale U[radius], resting on the terrain, each
    with its own principled material), 256 principled materials sharing a
    1024x1024 RGBA8 albedo texture (texture=None: untextured), 64 rect
    lights on an 8x8 grid, n_views views of 3840x2160 at 16 spp, depth 6."""
    rng = np.random.default_rng(seed)
    nx, nz = cells
    sx, sz = size
    ter, H = heightfield(rng, nx, nz, sx, sz, 10.0, n_mats, uv_repeat=4.0)
    sphere = uv_sphere((0.0, 0.0, 0.0), 1.0, material=0)
    inst = []
    for _ in range(n_spheres):
        cx, cz = rng.uniform(0.0, sx), rng.uniform(0.0, sz)
        r = rng.uniform(*radius)
        hx = min(int(cx / sx * nx), nx)
        hz = min(int(cz / sz * nz), nz)
        inst.append(make_instance(np.eye(3) * r, (cx, H[hx, hz] + 0.5 * r, cz), 0, int(rng.integers(0, n_mats))))
    mats = np.concatenate([
        principled(rng.uniform(0.0, 1.0, 3).astype(np.float32), rng.uniform(0.05, 1.0),
                   float(rng.random() < 0.5)) for _ in range(n_mats)])
    g = int(np.ceil(np.sqrt(n_lights)))
    lights = []
    for li in range(n_lights):
        lx = (li % g + 0.5) / g * sx
        lz = (li // g + 0.5) / g * sz
        lights.append(make_light((lx, 40.0, lz), (0.0, -1.0, 0.0), (20.0, 20.0, 20.0), 6.0, 6.0))
    views = []
    for v in range(n_views):
        x = (v + 0.5) / n_views * sx
        views.append(make_view((x, 35.0, -5.0), (x, 0.0, 0.5 * sz), (0, 1, 0), 55.0, w, h, spp))
    if texture:
        mats["texture"] = 0
    return Scene(ter, mats, np.concatenate(lights), np.concatenate(views), make_opts(depth), name="large_combined",
                 meshes=[sphere], instances=np.concatenate(inst),
                 textures=[noise_texture(*texture, seed=seed)] if texture else None)


def checker_texture(w: int = 64, h: int = 64, cell: int = 1, c0=(230, 230, 230), c1=(20, 40, 160)) -> np.ndarray:
    """RGBA8 checkerboard of cell x cell texel squares."""
    y, x = np.mgrid[0:h, 0:w]
    on = ((x // cell + y // cell) % 2).astype(bool)
    t = np.empty((h, w, 4), np.uint8)
    t[...] = (*c0, 255)
    t[on] = (*c1, 255)
    return t


def noise_texture(w: int = 1024, h: int = 1024, seed: int = 7) -> np.ndarray:
    """RGBA8 albedo: fBm value noise tinted per channel plus texel noise."""
    rng = np.random.default_rng(seed)
    ys, xs = np.meshgrid(np.linspace(0.0, 1.0, h), np.linspace(0.0, 1.0, w), indexing="ij")
    base = 0.5 + 0.25 * _fbm(xs, ys, rng, octaves=7, base_res=8)
    rgb = np.stack([base * 0.9, base * 0.8 + 0.1, base * 0.6], -1) + 0.08 * rng.random((h, w, 3))
    out = np.empty((h, w, 4), np.uint8)
    out[..., :3] = np.clip(rgb * 255.0, 0, 255).astype(np.uint8)
    out[..., 3] = 255
    return out


def textured_scene(seed: int = 8, cells=(32, 32), w: int = 96, h: int = 64, spp: int = 4, depth: int = 4,
                   n_instances: int = 40, tex_size=(64, 64)) -> Scene:
    """A heightfield under a grazing camera with a 1-texel checkerboard
    (Lambert and principled patches) and instanced textured spheres: every
    mip level of the texture is reached."""
    rng = np.random.default_rng(seed)
    size = 30.0
    ter, H = heightfield(rng, *cells, size, size, 1.5, 4, uv_repeat=4.0)
    mats = np.concatenate([
        lambert((0.9, 0.9, 0.9)), principled((0.8, 0.7, 0.6), 0.5, 0.0), lambert((0.7, 0.8, 0.9)),
        principled((0.9, 0.9, 0.9), 0.3, 1.0), lambert((0.6, 0.6, 0.6)), principled((0.5, 0.8, 0.5), 0.8, 0.0)])
    mats["texture"] = [0, 0, 1, 1, -1, 0]
    sphere = uv_sphere((0.0, 0.0, 0.0), 1.0, 12, 10, material=5)
    inst = []
    for _ in range(n_instances):
        cx, cz = rng.uniform(2.0, size - 2.0, 2)
        hx = min(int(cx / size * cells[0]), cells[0])
        hz = min(int(cz / size * cells[1]), cells[1])
        r = rng.uniform(0.3, 1.2)
        A = rotation(rng.normal(size=3), rng.uniform(0.0, 2 * np.pi)) * r
        inst.append(make_instance(A, (cx, H[hx, hz] + r, cz), 0, int(rng.integers(-1, 6))))
    lights = np.concatenate([make_light((10.0, 12.0, 10.0), (0, -1, 0), (60, 60, 60), 3.0, 3.0),
                             make_light((22.0, 10.0, 20.0), (0, -1, 0), (40, 40, 40), 2.0, 2.0)])
    view = make_view((0.3 * size, 3.0, -2.0), (0.5 * size, 0.0, size), (0, 1, 0), 60.0, w, h, spp)
    return Scene(ter, mats, lights, view, make_opts(depth), name="textured_scene", meshes=[sphere],
                 instances=np.concatenate(inst),
                 textures=[checker_texture(*tex_size, cell=1), noise_texture(37, 23, seed=seed)])


def flatten(scene: Scene) -> Scene:
    """The same scene without instancing: every instance's mesh mapped to
    world space in float64 (positions by the transform, normals by its
    inverse transpose) and appended after the world triangles, so triangle
    ids equal the instanced scene's virtual ids."""
    if scene.instances is None or len(scene.instances) == 0:
        return scene
    parts = [scene.tris]
    for r in scene.instances:
        M = r["xform"].astype(np.float64).reshape(3, 4)
        A, t = M[:, :3], M[:, 3]
        Nt = np.linalg.inv(A).T
        m = scene.meshes[int(r["mesh"])].copy()
        for k in ("a", "b", "c"):
            m[k] = m[k].astype(np.float64) @ A.T + t
            m["n" + k] = m["n" + k].astype(np.float64) @ Nt.T
        if r["material"] >= 0:
            m["material_index"] = r["material"]
        parts.append(m)
    return Scene(np.concatenate(parts), scene.mats, scene.lights, scene.views, scene.opts, name=scene.name + "_flat",
                 sky=scene.sky, textures=scene.textures)


def save_scene(scene: Scene, path: str) -> None:
    """Binary dump read by the C++ driver (host/gi_driver.cpp): "GISC", u32
    version, u64 n_tris, u32 n_mats, n_lights, n_views, then the records;
    version 2 (a sky map, textures or instances) appends tagged sections
    ("SKY ", "TEX ", "MESH", "INST") and "END "."""
    extra = scene.sky is not None or scene.textures or (scene.instances is not None and len(scene.instances))
    with open(path, "wb") as f:
        f.write(b"GISC")
        np.array([2 if extra else 1], "<u4").tofile(f)
        np.array([len(scene.tris)], "<u8").tofile(f)
        np.array([len(scene.mats), len(scene.lights), len(scene.views)], "<u4").tofile(f)
        for a, dt in ((scene.tris, TRI), (scene.mats, MATERIAL), (scene.lights, LIGHT), (scene.views, VIEW),
                      (scene.opts, OPTS)):
            np.ascontiguousarray(a, dt).tofile(f)
        if not extra:
            return
        if scene.sky is not None:
            m = np.ascontiguousarray(scene.sky, "<f4")
            f.write(b"SKY ")
            np.array([m.shape[1], m.shape[0]], "<u4").tofile(f)
            m.tofile(f)
        for t in scene.textures or []:
            t = np.ascontiguousarray(t, np.uint8)
            f.write(b"TEX ")
            np.array([t.shape[1], t.shape[0]], "<u4").tofile(f)
            t.tofile(f)
        if scene.instances is not None and len(scene.instances):
            for m in scene.meshes:
                f.write(b"MESH")
                np.array([len(m)], "<u8").tofile(f)
                np.ascontiguousarray(m, TRI).tofile(f)
            f.write(b"INST")
            np.array([len(scene.instances)], "<u4").tofile(f)
            np.ascontiguousarray(scene.instances, INSTANCE).tofile(f)
        f.write(b"END ")


def cornell_batch(seed: int = 3, n_views: int = 16, depth: int = 8, rr_start: int = 3,
                  resolutions=((256, 256), (640, 480), (1280, 720), (1920, 1080)), spps=(4, 16, 64)) -> Scene:
    """The cornell_boxes scene seen by n_views random cameras (position in front
    of the opening, look-at inside the box, fovy U[30,50]), each with a
    resolution and spp drawn from the given sets; depth 8 with Russian
    roulette from 3 segments on."""
    base = cornell_boxes()
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(n_views):
        w, h = resolutions[int(rng.integers(0, len(resolutions)))]
        spp = int(spps[int(rng.integers(0, len(spps)))])
        pos = (rng.uniform(0.25, 0.75), rng.uniform(0.25, 0.75), rng.uniform(-1.6, -0.9))
        at = (rng.uniform(0.35, 0.65), rng.uniform(0.25, 0.6), rng.uniform(0.4, 0.8))
        views.append(make_view(pos, at, (0, 1, 0), rng.uniform(30.0, 50.0), w, h, spp))
    return Scene(base.tris, base.mats, base.lights, np.concatenate(views), make_opts(depth, rr_start=rr_start),
                 name="cornell_batch")


def diff_scene(seed: int = 4, cells=(200, 250), n_spheres: int = 50, n_mats: int = 64, n_lights: int = 4,
               w: int = 512, h: int = 512, spp: int = 16, depth: int = 4, radius=(0.5, 4.0)) -> Scene:
    """The differentiation workload: heightfield + spheres (200x250 cells + 50
    spheres = 108k triangles), 64 principled materials with metallic and
    roughness drawn continuously (U[0.1,0.9], U[0.2,0.9]) so both have
    gradients, 4 lights, 512x512, 16 spp, depth 4, black background (the
    image is then smooth in every parameter except at visibility changes)."""
    sc = terrain_spheres(seed=seed, cells=cells, n_spheres=n_spheres, n_mats=n_mats, n_lights=n_lights, w=w, h=h,
                         spp=spp, depth=depth, radius=radius)
    rng = np.random.default_rng(seed + 100)
    sc.mats["metallic"] = rng.uniform(0.1, 0.9, n_mats).astype(np.float32)
    sc.mats["roughness"] = rng.uniform(0.2, 0.9, n_mats).astype(np.float32)
    sc.opts["background"] = (0.0, 0.0, 0.0)
    sc.name = "diff_scene"
    return sc


def perturbed(scene: Scene, sigma: float = 0.1, seed: int = 0) -> Scene:
    """Copy of the scene with baseColor, roughness, metallic and emission
    perturbed by N(0, sigma^2) (relative for emission) and clamped."""
    rng = np.random.default_rng(seed)
    m = scene.mats.copy()
    m["color"] = np.clip(m["color"] + rng.normal(0, sigma, m["color"].shape), 0.0, 1.0)
    lam = m["principled"] == 0
    m["diffuse_color"][lam] = m["color"][lam]
    m["roughness"] = np.clip(m["roughness"] + rng.normal(0, sigma, len(m)), 0.05, 1.0)
    m["metallic"] = np.clip(m["metallic"] + rng.normal(0, sigma, len(m)), 0.0, 1.0)
    l = scene.lights.copy()
    l["col"] = np.clip(l["col"] * (1 + rng.normal(0, sigma, l["col"].shape)), 0.0, None)
    return Scene(scene.tris, m, l, scene.views, scene.opts, name=scene.name + "_target")


def config_scene(idx: int, **kw) -> Scene:
    """Scenes of the benchmark configurations by number (1-based)."""
    if idx == 1:
        return cornell_boxes(**kw)
    if idx == 2:
        return terrain_spheres(**kw)
    if idx == 3:
        return cornell_batch(**kw)
    if idx == 4:
        return diff_scene(**kw)
    if idx == 5:
        return large_combined(**kw)
    raise ValueError(f"no scene generator for configuration {idx}")
