/*
 * ORACLE -- test infrastructure only (see oracle.h).  CPU restatement of
 * globillum's CPU path tracer, templated on the arithmetic type so the same
 * code gives the FP32 checker and the FP64 reference.  Compiled with
 * -ffp-contract=off: every a*b+c rounds twice, as in the device code.
 *
 * Reference files restated (all under src/ of kiselgra/globillum):
 *   halton                 util.cpp:22-32
 *   camera rays            raygen.cu:10-118, raygen.h:122-181
 *   BVH node layout        tracers.cu:26-36
 *   closest-hit traversal  tracers.cu:42-138
 *   NEE                    arealight-sampler-cpu.cpp:25-143
 *   shading + bounce       hybrid-pt.cpp:196-206, 330-371, 421-448, 450-577
 *   BRDFs                  simpleMaterial.h:82-113, principledMaterial.h:22-394,
 *                          materialBase.h:16-40, material-wrapper.h:15-51
 *   accumulate             util-kernels.cu:17-25
 *   state machine          hybrid-pt.h:115-249, gpu-pt.h:102-226
 */
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class R> struct V3 {
  R x, y, z;
};
template <class R> inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> mul(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> divs(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline V3<R> normalize(V3<R> a) { return a * (R(1) / std::sqrt(dot(a, a))); }
template <class R> inline V3<R> ld(const float *p) { return {R(p[0]), R(p[1]), R(p[2])}; }
template <class R> inline R maxc(V3<R> a) { return std::max(a.x, std::max(a.y, a.z)); }

template <class R> struct K {
  static constexpr R pi = R(3.14159265358979323846);
  static constexpr R inv_pi = R(0.31830988618379067154);
  static constexpr R two_pi = R(6.28318530717958647692);
};

/* ---------------------------------------------------------------- Halton -- */
struct Primes {
  std::vector<int32_t> p;
  Primes() {
    for (int32_t c = 2; (int)p.size() < 1024; ++c) {
      bool ok = true;
      for (int32_t q : p) {
        if (q * q > c) break;
        if (c % q == 0) { ok = false; break; }
      }
      if (ok) p.push_back(c);
    }
  }
};
const Primes &primes() {
  static Primes P;
  return P;
}

/* PCG32 (O'Neill 2014, pcg32_srandom_r / pcg32_random_r /
 * pcg32_boundedrand_r), used to draw the per-dimension digit permutations. */
struct Pcg32 {
  uint64_t state = 0, inc = 1;
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  void seed(uint64_t initstate, uint64_t initseq) {
    state = 0;
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
  }
  uint32_t bounded(uint32_t bound) {
    uint32_t threshold = (0u - bound) % bound;
    for (;;) {
      uint32_t r = next();
      if (r >= threshold) return r % bound;
    }
  }
};

void halton_perm(int32_t dim, uint32_t seed, uint16_t *perm) {
  int32_t p = primes().p[dim];
  for (int32_t i = 0; i < p; ++i) perm[i] = (uint16_t)i;
  Pcg32 r;
  r.seed(((uint64_t)seed << 32) ^ (0x9E3779B97F4A7C15ULL * (uint64_t)(dim + 1)), (uint64_t)(dim + 1));
  for (int32_t i = p - 1; i >= 1; --i) {
    uint32_t j = r.bounded((uint32_t)(i + 1));
    std::swap(perm[i], perm[j]);
  }
}

/* Scrambled radical inverse: digits d_k of n (K of them) map to perm[d_k]
 * and the infinitely many leading zero digits to perm[0], i.e. the exact
 * rational  A / B  with  A = rev*(p-1) + perm[0],  B = p^K * (p-1),
 * rev = sum_k perm[d_k] p^(K-1-k).  A and B are integers below 2^53, so the
 * value is one correctly rounded FP64 division, then one rounding to FP32.
 * (Grouping digits differently scales A and B by the same power of p and
 * leaves the value unchanged.) */
float scrambled_ri(uint64_t n, int32_t p, const uint16_t *perm) {
  uint64_t rev = 0, pk = 1;
  while (n > 0) {
    uint64_t q = n / (uint64_t)p;
    uint64_t d = n - q * (uint64_t)p;
    rev = rev * (uint64_t)p + perm[d];
    pk *= (uint64_t)p;
    n = q;
  }
  const uint64_t A = rev * (uint64_t)(p - 1) + perm[0];
  const uint64_t B = pk * (uint64_t)(p - 1);
  float f = (float)((double)A / (double)B);
  if (f >= 1.0f) f = 0x1.fffffep-1f;
  return f;
}

struct HaltonTable {
  uint32_t seed;
  int32_t ndims;
  std::vector<uint32_t> off;
  std::vector<uint16_t> perm;
  HaltonTable(uint32_t s, int32_t nd) : seed(s), ndims(nd) {
    off.resize(nd + 1);
    off[0] = 0;
    for (int32_t d = 0; d < nd; ++d) off[d + 1] = off[d] + primes().p[d];
    perm.resize(off[nd]);
    for (int32_t d = 0; d < nd; ++d) halton_perm(d, s, perm.data() + off[d]);
  }
  float sample(uint64_t n, int32_t dim) const {
    return scrambled_ri(n, primes().p[dim], perm.data() + off[dim]);
  }
};

/* Dimension map of one sample: (jitter x, y), (lens x, y), then 7 per path
 * vertex b: NEE x, y, light choice; BSDF x, y, lobe; Russian roulette. */
enum { DIM_JX = 0, DIM_JY = 1, DIM_LX = 2, DIM_LY = 3, DIM_VTX0 = 4, DIMS_PER_VTX = 7 };
inline int32_t dims_for_depth(int32_t depth) { return DIM_VTX0 + DIMS_PER_VTX * depth; }

/* -------------------------------------------------------- tangent frames -- */
/* Orthonormal basis around a unit vector n: Duff et al., "Building an
 * Orthonormal Basis, Revisited", JCGT 6(1), 2017.  Stands in for libhyb's
 * make_tangent_frame (used at src/hybrid-pt.cpp:526); T and B are normalised
 * afterwards as the reference does (src/hybrid-pt.cpp:527-528). */
template <class R> void tangent_frame(V3<R> n, V3<R> &t, V3<R> &b) {
  R sign = std::copysign(R(1), n.z);
  R a = R(-1) / (sign + n.z);
  R bb = n.x * n.y * a;
  t = mk<R>(R(1) + sign * n.x * n.x * a, sign * bb, -sign * n.x);
  b = mk<R>(bb, sign + n.y * n.y * a, -n.y);
  t = normalize(t);
  b = normalize(b);
}
/* libhyb make_tangential(v, n): v given in the frame (T, B, n) to world. */
template <class R> V3<R> tangential(V3<R> v, V3<R> n) {
  V3<R> t, b;
  tangent_frame(n, t, b);
  return (t * v.x + b * v.y) + n * v.z;
}

/* ------------------------------------------------------------ intersection */
/* Moller & Trumbore 1997, no back-face culling; hit iff det != 0,
 * u,v >= 0, u+v <= 1, t > 0.  Stands in for rta's intersect_tri_opt
 * (src/tracers.cu:112).  beta = u weights vertex b, gamma = v weights c. */
template <class R>
inline bool intersect_tri(V3<R> a, V3<R> e1, V3<R> e2, V3<R> o, V3<R> d, R &t, R &u, R &v) {
  V3<R> pvec = cross(d, e2);
  R det = dot(e1, pvec);
  if (det == R(0)) return false;
  R inv_det = R(1) / det;
  V3<R> tvec = o - a;
  u = dot(tvec, pvec) * inv_det;
  if (!(u >= R(0)) || u > R(1)) return false;
  V3<R> qvec = cross(tvec, e1);
  v = dot(d, qvec) * inv_det;
  if (!(v >= R(0)) || u + v > R(1)) return false;
  t = dot(e2, qvec) * inv_det;
  return t > R(0);
}

/* Slab test after Aila & Laine 2009 with the idir clamp of
 * src/tracers.cu:48-51; returns the entry distance. */
template <class R>
inline bool slab(const float *bmin, const float *bmax, V3<R> o, V3<R> idir, R tfar, R &dist) {
  R x0 = (R(bmin[0]) - o.x) * idir.x, x1 = (R(bmax[0]) - o.x) * idir.x;
  R y0 = (R(bmin[1]) - o.y) * idir.y, y1 = (R(bmax[1]) - o.y) * idir.y;
  R z0 = (R(bmin[2]) - o.z) * idir.z, z1 = (R(bmax[2]) - o.z) * idir.z;
  R tmin = std::max(std::max(std::min(x0, x1), std::min(y0, y1)), std::max(std::min(z0, z1), R(0)));
  R tmax = std::min(std::min(std::max(x0, x1), std::max(y0, y1)), std::min(std::max(z0, z1), tfar));
  dist = tmin;
  return tmin <= tmax;
}

template <class R> inline R clamp_idir(R d) {
  const R ooeps = R(8.2718061255302767e-25); /* 2^-80 */
  return R(1) / (std::fabs(d) > ooeps ? d : std::copysign(ooeps, d));
}

}  // namespace

/* ------------------------------------------------------------------ scene -- */
struct or_scene {
  std::vector<gi_tri> tris;
  std::vector<gi_material> mats;
  std::vector<gi_light> lights;
  std::vector<float> nodes; /* 8 floats per node, reference 2f4 layout */
  std::vector<uint32_t> indices;
  float lo[3], hi[3];
  float power_total = 0.f;
  /* sky map of the scene's GI_LIGHT_SKY light (gi::sky_light, src/lights.h:99-142):
   * rgb w*h, and the tables of its Distribution2D (row CDF/PDF, per-row CDF/PDF) */
  int32_t sky_w = 0, sky_h = 0, sky_light = -1;
  std::vector<float> sky_rgb, sky_ycdf, sky_ypdf, sky_xcdf, sky_xpdf;
  /* instancing: meshes (object space, each with its own median BVH) and the
   * instances placing them; instance i's triangle k has the virtual id
   * tris.size() + inst[i].vbase + k */
  std::vector<or_scene *> meshes;
  struct Instance {
    float fwd[12], inv[12]; /* object-to-world rows, world-to-object rows */
    int32_t mesh, material;
    uint32_t vbase;
  };
  std::vector<Instance> inst;
  /* RGBA8 textures with their mip chains (level 0, then each halved level) */
  struct Texture {
    int32_t w, h, levels;
    std::vector<uint8_t> chain;
  };
  std::vector<Texture> textures;
  bool textured = false; /* some material has a diffuse texture */
  ~or_scene() {
    for (or_scene *m : meshes) delete m;
  }
};

namespace {

inline int32_t fbits(float f) { int32_t i; std::memcpy(&i, &f, 4); return i; }
inline float bitsf(int32_t i) { float f; std::memcpy(&f, &i, 4); return f; }

/* Object-median binary BVH (the oracle's own, deliberately unlike the
 * product's binned SAH): split the centroid bounds' longest axis at the
 * median.  Emits the reference 2f4 layout of src/tracers.cu:26-36. */
struct MedianBuilder {
  const std::vector<gi_tri> &tris;
  std::vector<uint32_t> &idx;
  std::vector<float> &nodes;
  int32_t max_leaf;
  void bounds(uint32_t b, uint32_t e, float *lo, float *hi) {
    for (int k = 0; k < 3; ++k) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
    for (uint32_t i = b; i < e; ++i) {
      const gi_tri &t = tris[idx[i]];
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], std::min(t.a[k], std::min(t.b[k], t.c[k])));
        hi[k] = std::max(hi[k], std::max(t.a[k], std::max(t.b[k], t.c[k])));
      }
    }
  }
  /* Boxes grow by 2^-20 of their largest coordinate magnitude / extent so
   * the float slab test stays conservative for points on the triangles
   * (rays grazing an axis-aligned wall's zero-thickness box). */
  void set_box(int64_t n, const float *lo, const float *hi) {
    float m = 0.f;
    for (int k = 0; k < 3; ++k) m = std::max(m, std::max(std::fabs(lo[k]), std::fabs(hi[k])));
    for (int k = 0; k < 3; ++k) m = std::max(m, hi[k] - lo[k]);
    const float pad = m * 0x1p-20f + 1e-30f;
    float *p = &nodes[8 * n];
    p[2] = lo[0] - pad; p[3] = lo[1] - pad; p[4] = lo[2] - pad;
    p[5] = hi[0] + pad; p[6] = hi[1] + pad; p[7] = hi[2] + pad;
  }
  void build(int64_t node, uint32_t b, uint32_t e) {
    float lo[3], hi[3];
    bounds(b, e, lo, hi);
    set_box(node, lo, hi);
    if (e - b <= (uint32_t)max_leaf) {
      nodes[8 * node + 0] = bitsf((int32_t)((e - b) << 1));
      nodes[8 * node + 1] = bitsf((int32_t)b);
      return;
    }
    float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    auto cen = [&](uint32_t id, int k) {
      const gi_tri &t = tris[id];
      return (t.a[k] + t.b[k] + t.c[k]) * (1.0f / 3.0f);
    };
    for (uint32_t i = b; i < e; ++i)
      for (int k = 0; k < 3; ++k) {
        clo[k] = std::min(clo[k], cen(idx[i], k));
        chi[k] = std::max(chi[k], cen(idx[i], k));
      }
    int ax = 0;
    for (int k = 1; k < 3; ++k)
      if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
    uint32_t m = (b + e) / 2;
    std::nth_element(idx.begin() + b, idx.begin() + m, idx.begin() + e, [&](uint32_t x, uint32_t y) {
      float cx = cen(x, ax), cy = cen(y, ax);
      return cx < cy || (cx == cy && x < y);
    });
    int64_t left = (int64_t)nodes.size() / 8;
    nodes.resize(nodes.size() + 16);
    nodes[8 * node + 0] = bitsf((int32_t)((left << 1) | 1));
    nodes[8 * node + 1] = bitsf((int32_t)(left + 1));
    build(left, b, m);
    build(left + 1, m, e);
  }
};

/* --------------------------------------------------------------- traverse -- */
/* Ordered traversal of the binary 2f4 BVH after src/tracers.cu:42-138:
 * near child first, far child pushed, both children's boxes tested per inner
 * node.  Differences: the hit record starts invalid (t = FLT_MAX, see
 * src/tracers.cu:61-62), boxes are culled only when their entry distance is
 * strictly beyond the closest hit or t_max, and a triangle replaces the
 * closest hit when (t, id) is lexicographically smaller -- this makes the
 * result independent of traversal order and BVH topology. */
template <class R, bool ANY>
bool traverse(const or_scene &s, V3<R> o, V3<R> d, R tmax, R &bt, R &bu, R &bv, uint32_t &bid,
              uint64_t &nbox, uint64_t &ntri) {
  V3<R> idir = mk<R>(clamp_idir(d.x), clamp_idir(d.y), clamp_idir(d.z));
  bt = R(FLT_MAX);
  bid = GI_INVALID_REF;
  const float *N = s.nodes.data();
  if (s.nodes.empty()) return false;
  uint32_t stack[64];
  int sp = -1;
  int64_t cur = 0;
  for (;;) {
    const float *n = N + 8 * cur;
    int32_t x = fbits(n[0]), y = fbits(n[1]);
    if (x & 1) {
      int64_t l = x >> 1, r = y;
      R dl, dr;
      R lim = std::min(bt, tmax);
      bool hl = slab<R>(N + 8 * l + 2, N + 8 * l + 5, o, idir, lim, dl);
      bool hr = slab<R>(N + 8 * r + 2, N + 8 * r + 5, o, idir, lim, dr);
      nbox += 2;
      if (hl && hr) {
        if (dl <= dr) { stack[++sp] = (uint32_t)r; cur = l; }
        else { stack[++sp] = (uint32_t)l; cur = r; }
        continue;
      } else if (hl) { cur = l; continue; }
      else if (hr) { cur = r; continue; }
    } else {
      uint32_t cnt = (uint32_t)(x >> 1), off = (uint32_t)y;
      for (uint32_t i = 0; i < cnt; ++i) {
        uint32_t id = s.indices[off + i];
        const gi_tri &T = s.tris[id];
        V3<R> a = ld<R>(T.a), b = ld<R>(T.b), c = ld<R>(T.c);
        R t, u, v;
        ++ntri;
        if (intersect_tri<R>(a, b - a, c - a, o, d, t, u, v) && t <= tmax) {
          if (ANY) { bt = t; bu = u; bv = v; bid = id; return true; }
          if (t < bt || (t == bt && id < bid)) { bt = t; bu = u; bv = v; bid = id; }
        }
      }
    }
    if (sp < 0) break;
    cur = stack[sp--];
  }
  return bid != GI_INVALID_REF;
}

/* Affine maps by 3x4 row-major matrices, summed left to right:
 * ((m0 x + m1 y) + m2 z) + m3 per row.  Normals go by the transpose of the
 * inverse's 3x3 part (its columns). */
template <class R> inline V3<R> xf_point(const float *m, V3<R> p) {
  return mk<R>(((R(m[0]) * p.x + R(m[1]) * p.y) + R(m[2]) * p.z) + R(m[3]),
               ((R(m[4]) * p.x + R(m[5]) * p.y) + R(m[6]) * p.z) + R(m[7]),
               ((R(m[8]) * p.x + R(m[9]) * p.y) + R(m[10]) * p.z) + R(m[11]));
}
template <class R> inline V3<R> xf_dir(const float *m, V3<R> v) {
  return mk<R>((R(m[0]) * v.x + R(m[1]) * v.y) + R(m[2]) * v.z, (R(m[4]) * v.x + R(m[5]) * v.y) + R(m[6]) * v.z,
               (R(m[8]) * v.x + R(m[9]) * v.y) + R(m[10]) * v.z);
}
template <class R> inline V3<R> xf_normal(const float *inv, V3<R> n) {
  return mk<R>((R(inv[0]) * n.x + R(inv[4]) * n.y) + R(inv[8]) * n.z,
               (R(inv[1]) * n.x + R(inv[5]) * n.y) + R(inv[9]) * n.z,
               (R(inv[2]) * n.x + R(inv[6]) * n.y) + R(inv[10]) * n.z);
}

/* World geometry and every instance (no instance tree: each instance's mesh
 * is traversed with the ray mapped into its object space by the inverse
 * rows).  The object-space ray keeps the world ray's parameter, so t values
 * compare directly; the closest hit is the lexicographic minimum of (t,
 * virtual id) over everything. */
template <class R, bool ANY>
bool trace_scene(const or_scene &s, V3<R> o, V3<R> d, R tmax, R &bt, R &bu, R &bv, uint32_t &bid,
                 uint64_t &nbox, uint64_t &ntri) {
  bool hit = traverse<R, ANY>(s, o, d, tmax, bt, bu, bv, bid, nbox, ntri);
  if (ANY && hit) return true;
  const uint32_t nw = (uint32_t)s.tris.size();
  for (const or_scene::Instance &I : s.inst) {
    R t, u, v;
    uint32_t id;
    if (!traverse<R, ANY>(*s.meshes[(size_t)I.mesh], xf_point<R>(I.inv, o), xf_dir<R>(I.inv, d), tmax, t, u, v, id,
                          nbox, ntri))
      continue;
    const uint32_t vid = nw + I.vbase + id;
    if (ANY || t < bt || (t == bt && vid < bid)) { bt = t; bu = u; bv = v; bid = vid; }
    if (ANY) return true;
  }
  return bid != GI_INVALID_REF;
}

template <class R>
bool brute_tris(const or_scene &s, V3<R> o, V3<R> d, R tmax, uint32_t base, R &bt, R &bu, R &bv, uint32_t &bid) {
  for (uint32_t k = 0; k < s.tris.size(); ++k) {
    const gi_tri &T = s.tris[k];
    V3<R> a = ld<R>(T.a), b = ld<R>(T.b), c = ld<R>(T.c);
    R t, u, v;
    const uint32_t id = base + k;
    if (intersect_tri<R>(a, b - a, c - a, o, d, t, u, v) && t <= tmax &&
        (t < bt || (t == bt && id < bid))) {
      bt = t; bu = u; bv = v; bid = id;
    }
  }
  return bid != GI_INVALID_REF;
}

template <class R>
bool brute(const or_scene &s, V3<R> o, V3<R> d, R tmax, R &bt, R &bu, R &bv, uint32_t &bid) {
  bt = R(FLT_MAX);
  bid = GI_INVALID_REF;
  brute_tris<R>(s, o, d, tmax, 0u, bt, bu, bv, bid);
  const uint32_t nw = (uint32_t)s.tris.size();
  for (const or_scene::Instance &I : s.inst)
    brute_tris<R>(*s.meshes[(size_t)I.mesh], xf_point<R>(I.inv, o), xf_dir<R>(I.inv, d), tmax, nw + I.vbase, bt, bu,
                  bv, bid);
  return bid != GI_INVALID_REF;
}

/* The surface at hit (id, beta, gamma): position and unnormalised
 * interpolated normal, in world space, and the material.  An instance
 * triangle is interpolated in object space, then mapped (positions by the
 * forward rows, normals by the inverse transpose); the instance's material,
 * when set, replaces the triangle's. */
template <class R>
void surface(const or_scene &s, uint32_t id, R bu, R bv, V3<R> &P, V3<R> &Ni, int32_t &mat) {
  const or_scene::Instance *I = nullptr;
  const gi_tri *T;
  const uint32_t nw = (uint32_t)s.tris.size();
  if (id < nw) {
    T = &s.tris[id];
  } else {
    size_t lo = 0, hi = s.inst.size() - 1;
    while (lo < hi) {
      const size_t mid = (lo + hi + 1) / 2;
      if (nw + s.inst[mid].vbase <= id) lo = mid;
      else hi = mid - 1;
    }
    I = &s.inst[lo];
    T = &s.meshes[(size_t)I->mesh]->tris[id - nw - I->vbase];
  }
  V3<R> bc = mk<R>((R(1) - bu) - bv, bu, bv);
  P = (ld<R>(T->a) * bc.x + ld<R>(T->b) * bc.y) + ld<R>(T->c) * bc.z;
  Ni = (ld<R>(T->na) * bc.x + ld<R>(T->nb) * bc.y) + ld<R>(T->nc) * bc.z;
  mat = T->material_index;
  if (I) {
    P = xf_point<R>(I->fwd, P);
    Ni = xf_normal<R>(I->inv, Ni);
    if (I->material >= 0) mat = I->material;
  }
}

/* The hit triangle in world space (vertices mapped by the instance's
 * forward rows) and its texture coordinates. */
inline void world_tri(const or_scene &s, uint32_t id, V3<float> &A, V3<float> &B, V3<float> &C, const float *&ta,
                      const float *&tb, const float *&tc) {
  const uint32_t nw = (uint32_t)s.tris.size();
  const gi_tri *T;
  const or_scene::Instance *I = nullptr;
  if (id < nw) {
    T = &s.tris[id];
  } else {
    size_t lo = 0, hi = s.inst.size() - 1;
    while (lo < hi) {
      const size_t mid = (lo + hi + 1) / 2;
      if (nw + s.inst[mid].vbase <= id) lo = mid;
      else hi = mid - 1;
    }
    I = &s.inst[lo];
    T = &s.meshes[(size_t)I->mesh]->tris[id - nw - I->vbase];
  }
  A = ld<float>(T->a); B = ld<float>(T->b); C = ld<float>(T->c);
  if (I) { A = xf_point<float>(I->fwd, A); B = xf_point<float>(I->fwd, B); C = xf_point<float>(I->fwd, C); }
  ta = T->ta; tb = T->tb; tc = T->tc;
}

/* ------------------------------------------------------------- textures -- */
/* Box-filtered mip chain (k::mipmap + compute_mipmaps,
 * src/material.cu:235-274): floor(log2(min(w, h))) levels after level 0,
 * each ((W+1)/2) x ((H+1)/2); a texel is the truncating integer mean of the
 * (up to) 2x2 texels below it that exist.  (The reference zeroes the wrong
 * weight at an odd right edge and reads an unset value; here the missing
 * texel's weight is the one dropped.) */
void build_chain(or_scene::Texture &t, const uint8_t *rgba) {
  int32_t mn = std::min(t.w, t.h), levels = 0;
  while ((2 << levels) <= mn) ++levels;
  t.levels = levels;
  size_t total = 0;
  int32_t W = t.w, H = t.h;
  for (int l = 0; l <= levels; ++l) { total += (size_t)W * H; W = (W + 1) / 2; H = (H + 1) / 2; }
  t.chain.assign(4 * total, 0);
  std::memcpy(t.chain.data(), rgba, 4 * (size_t)t.w * t.h);
  size_t so = 0;
  W = t.w; H = t.h;
  for (int l = 0; l < levels; ++l) {
    const int32_t dw = (W + 1) / 2, dh = (H + 1) / 2;
    const size_t d0 = so + (size_t)W * H;
    for (int32_t y = 0; y < dh; ++y)
      for (int32_t x = 0; x < dw; ++x) {
        const int32_t X = 2 * x, Y = 2 * y;
        const bool rx = X + 1 < W, ry = Y + 1 < H;
        for (int c = 0; c < 4; ++c) {
          unsigned sum = t.chain[4 * (so + (size_t)Y * W + X) + c], cnt = 1;
          if (rx) { sum += t.chain[4 * (so + (size_t)Y * W + X + 1) + c]; ++cnt; }
          if (ry) { sum += t.chain[4 * (so + (size_t)(Y + 1) * W + X) + c]; ++cnt; }
          if (rx && ry) { sum += t.chain[4 * (so + (size_t)(Y + 1) * W + X + 1) + c]; ++cnt; }
          t.chain[4 * (d0 + (size_t)y * dw + x) + c] = (uint8_t)(sum / cnt);
        }
      }
    so = d0;
    W = dw; H = dh;
  }
}

/* Number of i < levels with 2^i < diff * max(w, h): the iterations of
 * sample_bilin_lod's `for (i = 0; i < lod; ++i)` (src/material.h:90-99), the
 * log2f compared exactly; NaN -> levels as fminf(NaN, max_mm) makes it. */
int tex_level(float diff, int32_t w, int32_t h, int32_t levels) {
  const float x = diff * (float)std::max(w, h);
  if (std::isnan(x)) return levels;
  int l = 0;
  for (double p = 1.0; l < levels && p < (double)x; p *= 2.0) ++l;
  return l;
}

/* sample_bilin_lod, src/material.h:88-137, with repeating addressing where
 * the reference leaves an index of W or H unwrapped (reading the next row
 * or level).  Coordinates that are NaN, infinite or >= 2^23 in magnitude
 * wrap to 0. */
struct TexFetch {
  size_t idx[4]; /* texel indices into the chain */
  float w[4];    /* bilinear weights: d value / d texel */
};

V3<float> tex_sample(const or_scene::Texture &T, float s, float t, float diff, TexFetch *tf = nullptr) {
  int32_t W = T.w, H = T.h;
  const int lv = tex_level(diff, T.w, T.h, T.levels);
  size_t off = 0;
  for (int i = 0; i < lv; ++i) { off += (size_t)W * H; W = (W + 1) / 2; H = (H + 1) / 2; }
  if (!(std::fabs(s) < 0x1p23f)) s = 0.0f;
  if (!(std::fabs(t) < 0x1p23f)) t = 0.0f;
  if (s < 0) s += -std::trunc(s) + 1;
  if (s >= 1) s -= std::trunc(s);
  if (t < 0) t += -std::trunc(t) + 1;
  if (t >= 1) t -= std::trunc(t);
  const float x = s * (float)W, y = (1.0f - t) * (float)H;
  int32_t nx = (int32_t)x, ny = (int32_t)y, ox = nx + 1, oy = ny + 1;
  if (x - std::floor(x) < .5) ox = nx - 1;
  if (y - std::floor(y) < .5) oy = ny - 1;
  const float wx = (float)std::fabs((double)ox + .5 - (double)x);
  const float wy = (float)std::fabs((double)oy + .5 - (double)y);
  nx %= W; ny %= H;
  ox = (ox % W + W) % W; oy = (oy % H + H) % H;
  auto texel = [&](int32_t xx, int32_t yy) {
    const uint8_t *q = &T.chain[4 * (off + (size_t)yy * W + xx)];
    return mk<float>(float(q[0]) / 255.0f, float(q[1]) / 255.0f, float(q[2]) / 255.0f);
  };
  const V3<float> a00 = texel(nx, ny), a01 = texel(ox, ny), a10 = texel(nx, oy), a11 = texel(ox, oy);
  if (tf) {
    tf->idx[0] = off + (size_t)ny * W + nx;
    tf->idx[1] = off + (size_t)ny * W + ox;
    tf->idx[2] = off + (size_t)oy * W + nx;
    tf->idx[3] = off + (size_t)oy * W + ox;
    tf->w[0] = wy * wx;
    tf->w[1] = wy * (1.0f - wx);
    tf->w[2] = (1.0f - wy) * wx;
    tf->w[3] = (1.0f - wy) * (1.0f - wx);
  }
  const V3<float> a0x = wx * a00 + (1.0f - wx) * a01, a1x = wx * a10 + (1.0f - wx) * a11;
  return wy * a0x + (1.0f - wy) * a1x;
}

/* The two ray differentials of a path (the "upper" and "right" rays of the
 * camera's neighbouring pixels, src/hybrid-pt.cpp:300-328). */
struct Diffs {
  V3<float> uo, ud, ro, rd;
};

/* reflect(d, N) = d - 2 (d.N) N */
inline V3<float> reflect_dir(V3<float> d, V3<float> N) {
  const float k = 2.0f * dot(d, N);
  return mk<float>(d.x - N.x * k, d.y - N.y * k, d.z - N.z * k);
}

/* Barycentrics of a ray's crossing with the triangle's plane, no inside
 * tests (intersect_tri_opt_nocheck): the arithmetic of intersect_tri. */
inline void plane_bary(V3<float> a, V3<float> e1, V3<float> e2, V3<float> o, V3<float> d, float &u, float &v) {
  V3<float> pvec = cross(d, e2);
  float det = dot(e1, pvec);
  float inv_det = 1.0f / det;
  V3<float> tvec = o - a;
  u = dot(tvec, pvec) * inv_det;
  V3<float> qvec = cross(tvec, e1);
  v = dot(d, qvec) * inv_det;
}

/* computeRayDiffs (src/hybrid-pt.cpp:300-328) at hit (id, bu, bv) with the
 * normalised shading normal N: the differentials' crossings with the hit
 * triangle's plane give their texture coordinates; the origins move there,
 * the upper direction is reflected about N (the right direction is kept, as
 * there).  Returns the texture footprint of the material's diffuse texture
 * lookup (material_t::diffuseColor, src/material.h:160-166) and the hit's
 * texture coordinates. */
float update_diffs(const or_scene &s, uint32_t id, float bu, float bv, V3<float> N, Diffs &df, float &tcx,
                   float &tcy) {
  V3<float> A, B, C;
  const float *ta, *tb, *tc;
  world_tri(s, id, A, B, C, ta, tb, tc);
  const V3<float> e1 = B - A, e2 = C - A;
  float ub, ug, rb, rg;
  plane_bary(A, e1, e2, df.uo, df.ud, ub, ug);
  plane_bary(A, e1, e2, df.ro, df.rd, rb, rg);
  const float h0 = (1.0f - bu) - bv, u0 = (1.0f - ub) - ug, r0 = (1.0f - rb) - rg;
  tcx = (ta[0] * h0 + tb[0] * bu) + tc[0] * bv;
  tcy = (ta[1] * h0 + tb[1] * bu) + tc[1] * bv;
  const float ux = (ta[0] * u0 + tb[0] * ub) + tc[0] * ug, uy = (ta[1] * u0 + tb[1] * ub) + tc[1] * ug;
  const float rx = (ta[0] * r0 + tb[0] * rb) + tc[0] * rg, ry = (ta[1] * r0 + tb[1] * rb) + tc[1] * rg;
  float dx = std::fabs(tcx - ux), dy = std::fabs(tcy - uy);
  dx = std::fmax(std::fabs(tcx - rx), dx);
  dy = std::fmax(std::fabs(tcy - ry), dy);
  df.uo = (A * u0 + B * ub) + C * ug;
  df.ro = (A * r0 + B * rb) + C * rg;
  df.ud = reflect_dir(df.ud, N);
  return std::fmax(dx, dy);
}

/* ------------------------------------------------------------------ BRDFs -- */
enum { T_DIFFUSE = 1, T_SPECULAR = 0, T_SPEC_REFL = 3, T_SPEC_TRANS = 4, T_DISNEY_REFL = 5 };

template <class R> inline R saturate(R a) { return std::min(std::max(a, R(0)), R(1)); }
template <class R> inline R schlick(R u) { R m = saturate(R(1) - u); R m2 = m * m; return m2 * m2 * m; }
template <class R> inline R mixf(R x, R y, R s) { return (R(1) - s) * x + s * y; }
template <class R> inline V3<R> mixv(V3<R> x, V3<R> y, R s) { return (R(1) - s) * x + s * y; }
template <class R> inline R sqr(R a) { return a * a; }
template <class R> inline R clamp01(R a) { return a < R(0) ? R(0) : a; }

/* sin(2 pi v), cos(2 pi v) for v in [0,1).  FP32: quadrant q = floor(4v)
 * and remainder f = 4v - q (both exact), Taylor series of sin(pi/2 f) to
 * degree 13 and cos(pi/2 f) to degree 14 (truncation error < 1e-9), rotated
 * into the quadrant -- basic operations only, so any IEEE machine gives the
 * same bits.  FP64: the C library. */
constexpr double kH = 1.5707963267948966;
inline void sincos2pi(float v, float &s, float &c) {
  const float S1 = (float)kH, S3 = (float)(-kH * kH * kH / 6.0), S5 = (float)(kH * kH * kH * kH * kH / 120.0);
  const float S7 = (float)(-kH * kH * kH * kH * kH * kH * kH / 5040.0);
  const float S9 = (float)(kH * kH * kH * kH * kH * kH * kH * kH * kH / 362880.0);
  const float S11 = (float)(-kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH / 39916800.0);
  const float S13 = (float)(kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH / 6227020800.0);
  const float C2 = (float)(-kH * kH / 2.0), C4 = (float)(kH * kH * kH * kH / 24.0);
  const float C6 = (float)(-kH * kH * kH * kH * kH * kH / 720.0);
  const float C8 = (float)(kH * kH * kH * kH * kH * kH * kH * kH / 40320.0);
  const float C10 = (float)(-kH * kH * kH * kH * kH * kH * kH * kH * kH * kH / 3628800.0);
  const float C12 = (float)(kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH / 479001600.0);
  const float C14 = (float)(-kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH / 87178291200.0);
  float x = v * 4.0f;
  float q = std::floor(x);
  float f = x - q;
  float f2 = f * f;
  float ps = f * (S1 + f2 * (S3 + f2 * (S5 + f2 * (S7 + f2 * (S9 + f2 * (S11 + f2 * S13))))));
  float pc = 1.0f + f2 * (C2 + f2 * (C4 + f2 * (C6 + f2 * (C8 + f2 * (C10 + f2 * (C12 + f2 * C14))))));
  int qi = (int)q & 3;
  if (qi == 0) { s = ps; c = pc; }
  else if (qi == 1) { s = pc; c = -ps; }
  else if (qi == 2) { s = -ps; c = -pc; }
  else { s = -pc; c = ps; }
}
inline void sincos2pi(double v, double &s, double &c) {
  s = std::sin(K<double>::two_pi * v);
  c = std::cos(K<double>::two_pi * v);
}

/* atan(t) / (2 pi) for t in [0, 1] as t * P(t^2), P of degree 8: a
 * least-squares fit (relative error < 2e-7) -- basic operations only, so the
 * device evaluates the same bits.  atan2 and acos below are built on it. */
inline float atan_turns01(float t) {
  const float c0 = 0.159154937f, c1 = -0.0530512258f, c2 = 0.0318191461f, c3 = -0.0226053596f,
              c4 = 0.0169349238f, c5 = -0.0119441384f, c6 = 0.00679757446f, c7 = -0.00256048539f,
              c8 = 0.000454627268f;
  const float x2 = t * t;
  float p = c8;
  p = p * x2 + c7;
  p = p * x2 + c6;
  p = p * x2 + c5;
  p = p * x2 + c4;
  p = p * x2 + c3;
  p = p * x2 + c2;
  p = p * x2 + c1;
  p = p * x2 + c0;
  return t * p;
}
inline double atan_turns01(double t) { return std::atan(t) / K<double>::two_pi; }
/* atan2(y, x) / (2 pi), in [-1/2, 1/2]; 0 for (0, 0). */
template <class R> inline R atan2_turns(R y, R x) {
  const R ax = std::fabs(x), ay = std::fabs(y);
  if (ax == R(0) && ay == R(0)) return R(0);
  R r = ay > ax ? R(0.25) - atan_turns01(ax / ay) : atan_turns01(ay / ax);
  if (x < R(0)) r = R(0.5) - r;
  return y < R(0) ? -r : r;
}

/* Distribution1D::sample (src/lights.h:50-57): upper_bound over CDF[0,size),
 * index clamped to [0,size-1], then the in-cell position as the reference
 * writes it, t = (CDF[i+1]-u)/(CDF[i+1]-CDF[i]), x = (1.0-t)*i + t*(i+1)
 * (the first product in double, the second in float). */
inline float cdf_sample(const float *cdf, const float *pdf, int size, float u, float &p) {
  int lo = 0, hi = size;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (cdf[mid] <= u) lo = mid + 1;
    else hi = mid;
  }
  const int idx = std::min(std::max(lo - 1, 0), size - 1);
  const float t = (cdf[idx + 1] - u) / (cdf[idx + 1] - cdf[idx]);
  p = pdf[idx];
  return (float)((1.0 - (double)t) * (double)idx + (double)(t * (float)(idx + 1)));
}

/* cosineSampleHemisphere, src/materialBase.h:16-24 */
template <class R> V3<R> cosine_sample(R u, R v) {
  R r = std::sqrt(u);
  R s, c;
  sincos2pi(v, s, c);
  return mk<R>(r * c, r * s, std::sqrt(std::max(R(0), R(1) - u)));
}

template <class R> V3<R> mon2lin(V3<R> p) {
  return mk<R>(std::pow(p.x, R(2.2)), std::pow(p.y, R(2.2)), std::pow(p.z, R(2.2)));
}

/* x^2.2 for a base colour that varies per hit (a texture): x = m 2^e,
 * m in [sqrt(1/2), sqrt(2)), 11 e = 5 k + j, x^2.2 = m^2.2 2^k 2^(j/5);
 * m^2.2 a degree-6 polynomial in m - 1 (relative least-squares fit, float
 * coefficients), 2^(j/5) a float table.  Basic operations only.  x < 2^-50
 * (or NaN) -> 0, x >= 2^50 -> infinity. */
inline float pow22(float x) {
  if (!(x >= 0x1p-50f)) return 0.0f;
  if (x >= 0x1p50f) return INFINITY;
  uint32_t bits;
  std::memcpy(&bits, &x, 4);
  int32_t e = (int32_t)(bits >> 23) - 127;
  uint32_t mb = (bits & 0x7fffffu) | 0x3f800000u;
  if (mb >= 0x3fb504f3u) { mb -= 0x00800000u; e += 1; }
  float m;
  std::memcpy(&m, &mb, 4);
  const float t = m - 1.0f;
  const float c6[7] = {1.0f, 2.200000286102295f, 1.3200013637542725f, 0.0879826545715332f,
                       -0.017642781138420105f, 0.006638647988438606f, -0.002734926762059331f};
  float p = c6[6];
  for (int i = 5; i >= 0; --i) p = p * t + c6[i];
  const int32_t q = 11 * e, j = ((q % 5) + 5) % 5, k = (q - j) / 5;
  const float tab[5] = {1.0f, 1.1486983549970351f, 1.3195079107728942f, 1.515716566510398f, 1.7411011265922482f};
  const uint32_t sb = (uint32_t)(k + 127) << 23;
  float scale;
  std::memcpy(&scale, &sb, 4);
  return (tab[j] * p) * scale;
}
inline double pow22(double x) { return std::pow(x, 2.2); }
template <class R> V3<R> mon2lin_poly(V3<R> p) { return mk<R>(pow22(p.x), pow22(p.y), pow22(p.z)); }

template <class R> R smith_ggx(R ndv, R ag) {
  R a = ag * ag, b = ndv * ndv;
  return R(1) / (ndv + std::sqrt(a + b - a * b));
}
template <class R> R gtr1(R ndh, R a) {
  if (a >= R(1)) return K<R>::inv_pi;
  R a2 = a * a;
  R t = R(1) + (a2 - R(1)) * ndh * ndh;
  return (a2 - R(1)) / (K<R>::pi * std::log(a2) * t);
}

template <class R> struct MatP {
  R metallic, subsurface, specular, roughness, spec_tint, aniso, sheen, sheen_tint, clearcoat,
      cc_gloss, opacity;
  V3<R> color, diffuse;
  bool principled;
  bool textured = false; /* colour scaled by a texel: mon2lin by pow22 */
  explicit MatP(const gi_material &m)
      : metallic(m.metallic), subsurface(m.subsurface), specular(m.specular), roughness(m.roughness),
        spec_tint(m.specular_tint), aniso(m.anisotropic), sheen(m.sheen), sheen_tint(m.sheen_tint),
        clearcoat(m.clearcoat), cc_gloss(m.clearcoat_gloss), opacity(m.opacity),
        color(ld<R>(m.color)), diffuse(ld<R>(m.diffuse_color)), principled(m.principled != 0) {}
  bool glass() const { return principled && opacity < R(1); }
  V3<R> cdlin() const { return textured ? mon2lin_poly(color) : mon2lin(color); }
  /* material_t::diffuseColor (src/material.h:155-168): the texel scales the
   * diffuse colour (Lambert) or the principled base colour */
  void apply_texel(V3<R> tx) {
    if (principled) { color = mul(color, tx); textured = true; }
    else diffuse = mul(diffuse, tx);
  }
};

/* evaluatePrincipledBRDF_specular, src/principledMaterial.h:210-252 */
template <class R> V3<R> eval_spec(const MatP<R> &p, V3<R> V, V3<R> N, V3<R> L) {
  R ndl = dot(N, L), ndv = dot(N, V);
  if (ndl < R(0) || ndv < R(0)) return mk<R>(0, 0, 0);
  V3<R> H = normalize(L + V);
  R ndh = dot(N, H), ldh = dot(L, H);
  V3<R> cd = p.cdlin();
  R lum = R(0.3) * cd.x + R(0.6) * cd.y + R(0.1) * cd.z;
  V3<R> one = mk<R>(1, 1, 1);
  V3<R> tint = lum > R(0) ? divs(cd, lum) : one;
  V3<R> cs0 = mixv((p.specular * R(0.08)) * mixv(one, tint, p.spec_tint), cd, p.metallic);
  R fh = schlick(ldh);
  V3<R> fs = mixv(cs0, one, fh);
  R cg = R(0);
  if (p.clearcoat > R(0)) {
    R dr = gtr1(ndh, mixf(R(0.1), R(0.001), p.cc_gloss));
    R fr = mixf(R(0.04), R(1), fh);
    R gr = smith_ggx(ndl, R(0.25)) * smith_ggx(ndv, R(0.25));
    cg = R(0.25) * p.clearcoat * gr * fr * dr;
  }
  V3<R> sum = fs + mk<R>(cg, cg, cg);
  return (sum * ndl) * p.metallic;
}

/* evaluatePrincipledBRDF_diffuse, src/principledMaterial.h:254-301 */
template <class R> V3<R> eval_diff(const MatP<R> &p, V3<R> V, V3<R> N, V3<R> L) {
  R ndl = dot(N, L), ndv = dot(N, V);
  if (ndl < R(0) || ndv < R(0)) return mk<R>(0, 0, 0);
  V3<R> H = normalize(L + V);
  R ldh = dot(L, H);
  V3<R> cd = p.cdlin();
  R lum = R(0.3) * cd.x + R(0.6) * cd.y + R(0.1) * cd.z;
  V3<R> one = mk<R>(1, 1, 1);
  V3<R> tint = lum > R(0) ? divs(cd, lum) : one;
  V3<R> csheen = mixv(one, tint, p.sheen_tint);
  R fl = schlick(ndl), fv = schlick(ndv);
  R fd90 = R(0.5) + R(2) * ldh * ldh * p.roughness;
  R fd = mixf(R(1), fd90, fl) * mixf(R(1), fd90, fv);
  R fh = schlick(ldh);
  V3<R> fsheen = (fh * p.sheen) * csheen;
  R fss90 = ldh * ldh * p.roughness;
  R fss = mixf(R(1), fss90, fl) * mixf(R(1), fss90, fv);
  R ss = R(1.25) * (fss * (R(1) / (ndl + ndv) - R(0.5)) + R(0.5));
  V3<R> dc = (K<R>::inv_pi * mixf(fd, ss, p.subsurface) * cd + fsheen) * (R(1) - p.metallic);
  return dc * ndl;
}

/* materialBRDF::evaluate, src/material-wrapper.h:34-38 with
 * LambertianMaterial::evaluate (src/simpleMaterial.h:99-101) and
 * PrincipledMaterial::evaluate (src/principledMaterial.h:336-352). */
template <class R> V3<R> brdf_eval(const MatP<R> &p, int type, V3<R> wo, V3<R> wi, V3<R> N) {
  if (!p.principled) return (p.diffuse * K<R>::inv_pi) * std::fabs(dot(wi, N));
  if (type == T_DISNEY_REFL) return eval_spec(p, wo, N, wi);
  if (type == T_SPEC_REFL) return mk<R>(1, 1, 1);
  if (type == T_SPEC_TRANS) return mk<R>(1, 1, 1) * (R(1) / std::fabs(dot(wi, N)));
  return eval_diff(p, wo, N, wi);
}

/* Halfway vector of sampleGTR2AnisoHalfwayVector (src/principledMaterial.h:
 * 46-61) in tangent space (T = x, B = y, N = z): the GGX half-vector sample
 * with slopes stretched by (ax, ay). */
template <class R> V3<R> gtr2_half(R u, R v, R ax, R ay) {
  R s, c;
  sincos2pi(u, s, c);
  R fac = std::sqrt(v / (R(1) - v));
  return normalize(mk<R>(fac * (ax * c), fac * (ay * s), R(1)));
}

/* The reference's specular "pdf": Walter et al. 2007's sample weight
 * (eq. 41) with the Smith G1 of eq. 34 (src/principledMaterial.h:76-84,
 * 157-175) -- as written there: the tangent in G1 is that of the angle
 * between the direction and m (not n), alpha_g = (ax*ay/2 + 1/2)^2, and the
 * weight is clamped at 0.01 from below. */
template <class R> R walter_weight(V3<R> L, V3<R> V, V3<R> m, R ax, R ay) {
  R roughg = sqr(ax * ay * R(0.5) + R(0.5));
  R dlm = dot(L, m), dvm = dot(V, m);
  R tl = std::sqrt(R(1) - dlm * dlm) / dlm;
  R tv = std::sqrt(R(1) - dvm * dvm) / dvm;
  R rg2 = roughg * roughg;
  auto g1 = [&](V3<R> w, R tn) {
    R xi = (dot(w, m) / w.z) > R(0) ? R(1) : R(0);
    return xi * R(2) / (R(1) + std::sqrt(R(1) + rg2 * tn * tn));
  };
  R gs = g1(L, tl) * g1(V, tv);
  R lw = (std::fabs(dot(L, m)) * gs) / (std::fabs(L.z) * std::fabs(m.z));
  if (lw < R(0.01)) lw = R(0.01);
  return lw;
}

/* sampleGTR2Aniso (src/principledMaterial.h:121-176): the halfway vector,
 * the reflection of V about it (flipped into the upper hemisphere), the
 * weight. */
template <class R> R sample_gtr2(R u, R v, R ax, R ay, V3<R> V, V3<R> &L) {
  V3<R> m = gtr2_half(u, v, ax, ay);
  L = V - (R(2) * dot(V, m)) * m; /* reflectR, src/materialBase.h:36-38 */
  if (L.z < R(0)) L = -L;
  return walter_weight(L, V, m, ax, ay);
}

/* sampleGlass, src/principledMaterial.h:86-118 (ior 1.9, :358). */
template <class R> int sample_glass(R u, V3<R> V, V3<R> &L) {
  const R ior = R(1.9);
  R c = V.z;
  R fr = schlick(c);
  if (u > fr) { L = -V; return T_SPEC_REFL; }
  R ti = R(1) + ior * (c * c - R(1));
  if (ti <= R(0)) {
    L = V - (R(2) * V.z) * mk<R>(0, 0, 1);
    return T_SPEC_REFL;
  }
  R sg = V.z > R(0) ? R(1) : R(-1);
  L = (ior * c - sg * std::sqrt(ti)) * mk<R>(0, 0, 1) - ior * V;
  return T_SPEC_TRANS;
}

/* materialBRDF::sample (src/material-wrapper.h:30-33):
 * LambertianMaterial::sample (src/simpleMaterial.h:103-106) or
 * PrincipledMaterial::sample (src/principledMaterial.h:354-390). */
template <class R> int brdf_sample(const MatP<R> &p, V3<R> wo, V3<R> xi, V3<R> &wi, R &pdf) {
  if (!p.principled) {
    wi = cosine_sample(xi.x, xi.y);
    pdf = std::fabs(wi.z) * K<R>::inv_pi;
    return T_DIFFUSE;
  }
  if (p.opacity < R(1)) {
    pdf = R(1);
    return sample_glass(xi.x, wo, wi);
  }
  if (xi.z < p.metallic) {
    R aspect = std::sqrt(R(1) - p.aniso * R(0.9));
    R ax = std::max(R(0.001), sqr(p.roughness) / aspect);
    R ay = std::max(R(0.001), sqr(p.roughness) * aspect);
    pdf = sample_gtr2(xi.x, xi.y, ax, ay, wo, wi);
    if (pdf == R(0) || std::isnan(pdf)) {
      wi = cosine_sample(xi.x, xi.y);
      pdf = clamp01(wi.z) / K<R>::pi;
      return T_DIFFUSE;
    }
    return T_DISNEY_REFL;
  }
  wi = cosine_sample(xi.x, xi.y);
  pdf = clamp01(wi.z) * K<R>::inv_pi;
  return T_DIFFUSE;
}

/* ------------------------------------------------------------------ camera */
struct Cam {
  float U[3], V[3], W[3], pos[3], vd[3];
  float width, height, fw, fh, aperture, focus;
};
/* Basis and image-plane extent of src/raygen.cu:17-39. */
Cam make_cam(const gi_view &v) {
  Cam c;
  float half = v.fovy / 2.0f;
  c.height = (float)std::tan(3.14159265358979323846 * (double)half / 180.0);
  c.width = ((float)v.w / (float)v.h) * c.height;
  V3<float> vd = ld<float>(v.dir), up = ld<float>(v.up);
  V3<float> W = divs(vd, length(vd));
  V3<float> txw = cross(up, W);
  V3<float> U = divs(txw, length(txw));
  V3<float> V = cross(W, U);
  std::memcpy(c.U, &U, 12); std::memcpy(c.V, &V, 12); std::memcpy(c.W, &W, 12);
  std::memcpy(c.pos, v.pos, 12); std::memcpy(c.vd, v.dir, 12);
  c.fw = (float)v.w; c.fh = (float)v.h;
  c.aperture = v.aperture; c.focus = v.focus_distance;
  return c;
}
/* One jittered camera ray: src/raygen.cu:22-58; thin lens :99-117 (the
 * lens jitter is uniform on the square [-1/2,1/2]^2 -- the reference's
 * rejection test against the unit disk never rejects such a point). */
void cam_ray(const Cam &c, int32_t x, int32_t y, float jx, float jy, float lx, float ly, V3<float> &o,
             V3<float> &d) {
  float u_s = ((float(x) + jx) / c.fw) * 2.0f - 1.0f;
  float v_s = ((float(y) + jy) / c.fh) * 2.0f - 1.0f;
  u_s = c.width * u_s;
  v_s = c.height * v_s;
  V3<float> U = ld<float>(c.U), V = ld<float>(c.V), W = ld<float>(c.W);
  d = normalize((U * u_s + V * v_s) + W);
  o = ld<float>(c.pos);
  if (c.aperture > 0.0f) {
    V3<float> vd = ld<float>(c.vd);
    V3<float> focal = o + d * ((1.0f / dot(d, vd)) * c.focus);
    float ox = lx - 0.5f, oy = ly - 0.5f;
    V3<float> jp = (o + U * (ox * c.aperture)) + V * (oy * c.aperture);
    d = normalize(focal - jp);
    o = jp;
  }
}

/* Ray differentials of pixel (x, y): pinhole rays from the camera position
 * through the centres of pixels (x, y+1) ("upper") and (x+1, y) ("right").
 * The reference takes them from rta's camera_ray_generator_shirley, an
 * external library (src/raygen.h:122-130), whose exact rays are not in the
 * reference; the lens generator keeps them pinhole. */
void cam_diffs(const Cam &c, int32_t x, int32_t y, Diffs &df) {
  Cam p = c;
  p.aperture = 0.0f;
  cam_ray(p, x, y + 1, 0.5f, 0.5f, 0.5f, 0.5f, df.uo, df.ud);
  cam_ray(p, x + 1, y, 0.5f, 0.5f, 0.5f, 0.5f, df.ro, df.rd);
}

/* ----------------------------------------------------------------- render -- */
template <class R> struct Tracer {
  const or_scene &s;
  bool closest(V3<R> o, V3<R> d, R tmax, R &t, R &u, R &v, uint32_t &id, uint64_t &nb, uint64_t &nt) const {
    return trace_scene<R, false>(s, o, d, tmax, t, u, v, id, nb, nt);
  }
  bool any(V3<R> o, V3<R> d, R tmax, uint64_t &nb, uint64_t &nt) const {
    R t, u, v;
    uint32_t id;
    return trace_scene<R, true>(s, o, d, tmax, t, u, v, id, nb, nt);
  }
};

struct Options {
  int32_t depth, rr_start;
  float off_b, off_s;
  float bg[3];
  bool phong; /* GI_FLAG_PHONG: gpu_pt's Lambert + Phong material model */
  bool centre; /* GI_FLAG_PIXEL_CENTRE: primary rays through the pixel centres */
};

Options resolve(const or_scene &s, const gi_render_opts &o) {
  Options r;
  r.depth = std::max(1, std::min(o.max_depth, 32));
  r.rr_start = o.rr_start;
  float diag = or_scene_diagonal(&s);
  r.off_b = o.offset_bounce > 0.f ? o.offset_bounce : 1e-4f * diag;
  r.off_s = o.offset_shadow > 0.f ? o.offset_shadow : 1e-4f * diag;
  for (int k = 0; k < 3; ++k) r.bg[k] = o.background[k];
  r.phong = (o.flags & GI_FLAG_PHONG) != 0;
  r.centre = (o.flags & GI_FLAG_PIXEL_CENTRE) != 0;
  return r;
}

/* ------------------------------------------------------ gpu_pt's Phong lobe
 * compute_path_contribution_and_bounce, src/gpu-pt-kernels.cu:89-199.  The
 * powers of the lobe (shininess 100): x^100 as a chain of squarings and the
 * sampler's x^(1/101) as 2^(log2(x)/101), log2 from the atanh series (m in
 * [sqrt(1/2), sqrt(2))), 2^f from its Taylor polynomial -- basic operations
 * in one order, in FP32; FP64 uses the C library's pow. */
inline float phong_pow100(float x) {
  const float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8, x32 = x16 * x16, x64 = x32 * x32;
  return (x64 * x32) * x4;
}
inline double phong_pow100(double x) { return std::pow(x, 100.0); }
inline float phong_root101(float x) {
  if (!(x > 0.0f)) return 0.0f;
  if (x >= 1.0f) return 1.0f;
  uint32_t bits;
  std::memcpy(&bits, &x, 4);
  if (bits < 0x00800000u) return 0.4214777f;
  int32_t e = (int32_t)(bits >> 23) - 127;
  uint32_t mb = (bits & 0x7fffffu) | 0x3f800000u;
  if (mb >= 0x3fb504f3u) { mb -= 0x00800000u; e += 1; }
  float m;
  std::memcpy(&m, &mb, 4);
  const float t = (m - 1.0f) / (m + 1.0f), t2 = t * t;
  const float ca[6] = {0.09090909090909091f, 0.1111111111111111f, 0.14285714285714285f, 0.2f,
                       0.3333333333333333f, 1.0f};
  float a = ca[0];
  for (int i = 1; i < 6; ++i) a = a * t2 + ca[i];
  const float l2 = (float)e + (a * t) * 2.8853900817779268f;
  const float y = l2 / 101.0f;
  const float k = std::floor(y), f = y - k;
  const float cp[10] = {1.0178086009239699e-07f, 1.3215486790144307e-06f, 1.5252733804059841e-05f,
                        0.00015403530393381609f, 0.0013333558146428443f, 0.009618129107628477f,
                        0.05550410866482158f, 0.2402265069591007f, 0.6931471805599453f, 1.0f};
  float p = cp[0];
  for (int i = 1; i < 10; ++i) p = p * f + cp[i];
  const uint32_t sb = (uint32_t)((int32_t)k + 127) << 23;
  float scale;
  std::memcpy(&scale, &sb, 4);
  return p * scale;
}
inline double phong_root101(double x) { return x > 0.0 ? std::pow(x, 1.0 / 101.0) : 0.0; }
template <class R> struct PhongK;
template <> struct PhongK<float> {
  static constexpr float norm = 16.07464925228143f, two_pi = 6.283185307179586f;
};
template <> struct PhongK<double> {
  static constexpr double norm = 101.0 / (2.0 * M_PI), two_pi = 2.0 * M_PI;
};

/* One Phong-mode vertex: the NEE weight fnee (no cosine: the NEE
 * contribution holds it) and the bounce -- false if the path is absorbed
 * (no lobe chosen, a specular direction below the surface, or a zero
 * density), else the new direction and the throughput factor.  kd is the
 * diffuse colour times the texel, ks the specular colour; N the
 * interpolated, normalised, unflipped shading normal. */
template <class R>
bool phong_vertex(V3<R> kd, V3<R> ks, V3<R> N, V3<R> org_dir, V3<R> sd, V3<R> xp, V3<R> &fnee, V3<R> &wi,
                  V3<R> &tpmul) {
  const R sum = ((((kd.x + kd.y) + kd.z) + ks.x) + ks.y) + ks.z;
  if (sum > R(1)) {
    kd = divs(kd, sum);
    ks = divs(ks, sum);
  }
  const R kr = R(2) * dot(org_dir, N);
  const V3<R> Rr = mk<R>(org_dir.x - N.x * kr, org_dir.y - N.y * kr, org_dir.z - N.z * kr);
  const V3<R> l = normalize(sd);
  const R cl = phong_pow100(std::max(dot(Rr, l), R(0)));
  fnee = kd * K<R>::inv_pi + ks * (PhongK<R>::norm * cl);
  R pd = (kd.x + kd.y) + kd.z;
  R ps = (ks.x + ks.y) + ks.z;
  if (pd + ps > R(1)) {
    pd = pd / (pd + ps);
    ps = ps / (pd + ps); /* the reference's order: with the renormalised pd */
  }
  const bool dif = xp.z <= pd;
  if (!dif && !(xp.z <= pd + ps)) return false;
  const V3<R> axis = dif ? N : Rr;
  V3<R> Tt, Bt;
  tangent_frame(axis, Tt, Bt);
  const R z = dif ? std::sqrt(xp.x) : phong_root101(xp.x);
  const R r = std::sqrt(R(1) - z * z);
  R sn, cs;
  sincos2pi(xp.y, sn, cs);
  const V3<R> dir = (Tt * (r * cs) + Bt * (r * sn)) + axis * z;
  const R dn = dot(dir, N);
  V3<R> f;
  R den;
  if (dif) {
    f = kd * K<R>::inv_pi;
    den = R(2) * z;
  } else {
    if (dn < R(0)) return false;
    f = ks * (PhongK<R>::norm * phong_pow100(std::max(dot(Rr, dir), R(0))));
    den = R(101) * phong_pow100(z);
  }
  if (!(den > R(0))) return false;
  const R w = (R(1) / (dif ? pd : ps)) * (PhongK<R>::two_pi / den);
  tpmul = (f * std::fabs(dn)) * w;
  wi = normalize(dir);
  return true;
}

/* Sky light sample for NEE (pixel_generate_arealight_sample_cpu with
 * USE_SKY_SAMPLING, src/arealight-sampler-cpu.cpp:87-96; sky_light::sample,
 * src/lights.h:121-139): (u, v) from the map's Distribution2D; direction
 * (sin t cos p, sin t sin p, cos t) with t = u pi / h, p = v 2 pi / w, turned
 * into the frame around the normal (make_tangential); pdf = pdf_uv w h /
 * (2 pi^2 sin t); contribution = scale * L / pdf * |dir . N| with the
 * unnormalised NEE normal, as for rect lights.  scale is col[0]. */
template <class R>
void sky_sample(const or_scene &s, R x0, R x1, V3<R> Ni, V3<R> &dir, V3<R> &contrib) {
  const int w = s.sky_w, h = s.sky_h;
  float yp, xp;
  const float u = cdf_sample(s.sky_ycdf.data(), s.sky_ypdf.data(), h, (float)x0, yp);
  const int row = std::min(std::max((int)u, 0), h - 1);
  const float v =
      cdf_sample(s.sky_xcdf.data() + (size_t)row * (w + 1), s.sky_xpdf.data() + (size_t)row * w, w, (float)x1, xp);
  const float pdf = yp * xp;
  R sth, cth, sph, cph;
  sincos2pi(R(u) / (R(2) * R(h)), sth, cth);
  sincos2pi(R(v) / R(w), sph, cph);
  const V3<R> dm = mk<R>(sth * cph, sth * sph, cth);
  const R outpdf = R((double)(pdf * (float)(w * h)) / (2.0 * M_PI * M_PI * (double)sth));
  const int ix = std::min(std::max((int)v, 0), w - 1);
  const float *L = &s.sky_rgb[3 * ((size_t)row * w + ix)];
  dir = tangential(dm, normalize(Ni));
  const R a = R(1) / outpdf;
  const R scale = R(s.lights[s.sky_light].col[0]);
  const R cosn = std::fabs(dot(dir, Ni));
  contrib = mk<R>(((scale * R(L[0])) * a) * cosn, ((scale * R(L[1])) * a) * cosn, ((scale * R(L[2])) * a) * cosn);
}

/* Sky radiance seen by a ray that leaves the scene (evaluateSkyLight,
 * src/hybrid-pt.cpp:175-189): theta = acos(d.y), phi = atan2(d.z, d.x) in
 * [0, 2 pi), texel (int(t h), int(s w)); a row index of h (d.y = -1) gives
 * black, as the reference's range check does; the column is clamped to w-1.
 * Unscaled, as there. */
template <class R> V3<R> sky_eval(const or_scene &s, V3<R> d) {
  if (s.sky_light < 0) return mk<R>(0, 0, 0);
  const int w = s.sky_w, h = s.sky_h;
  const R yc = std::min(std::max(d.y, R(-1)), R(1));
  const R st = std::sqrt((R(1) - yc) * (R(1) + yc));
  const R t = R(2) * atan2_turns(st, yc);
  R sx = atan2_turns(d.z, d.x);
  if (sx < R(0)) sx = sx + R(1);
  const int iy = (int)(t * R(h)), ix = std::min((int)(sx * R(w)), w - 1);
  if (iy < 0 || iy >= h || ix < 0) return mk<R>(0, 0, 0);
  const float *L = &s.sky_rgb[3 * ((size_t)iy * w + ix)];
  return mk<R>(R(L[0]), R(L[1]), R(L[2]));
}

/* One sample of one pixel: the hybrid_pt state machine (closest hit, NEE,
 * any hit, bounce; src/hybrid-pt.h:151-249) run to completion for a single
 * path.  Returns the path's accumulated colour (path_accum_color). */
template <class R>
V3<R> trace_sample(const or_scene &s, const Tracer<R> &tr, const HaltonTable &H, const Options &op,
                   uint64_t n, V3<R> o, V3<R> d, or_stats &st, Diffs *df = nullptr) {
  V3<R> L = mk<R>(0, 0, 0), TP = mk<R>(1, 1, 1);
  R tmax = R(FLT_MAX);
  const int nl = (int)s.lights.size();
  for (int32_t b = 0; b < op.depth; ++b) {
    const int32_t dim = DIM_VTX0 + DIMS_PER_VTX * b;
    R t, bu = R(0), bv = R(0);
    uint32_t id;
    ++st.rays_closest;
    bool hit = tr.closest(o, d, tmax, t, bu, bv, id, st.box_closest, st.tri_closest);
    V3<R> org_dir = normalize(d);
    if (!hit) {
      /* handle_no_hit with NO_BACKGROUND (src/hybrid-pt.cpp:196-206,567-574):
       * a path whose throughput is exactly white gets the background. */
      if (TP.x == R(1) && TP.y == R(1) && TP.z == R(1))
        L = L + mul(TP, mk<R>(R(op.bg[0]), R(op.bg[1]), R(op.bg[2])));
      else if (s.sky_light >= 0)
        L = L + mul(TP, sky_eval(s, org_dir));
      break;
    }
    ++st.vertices;
    V3<R> P, Ni;
    int32_t mat;
    surface<R>(s, id, bu, bv, P, Ni, mat);
    MatP<R> mp(s.mats[mat]);
    V3<R> texel = mk<R>(1, 1, 1);
    bool textured = false;
    if (df) {
      const V3<R> Nn = normalize(Ni);
      float tcx, tcy;
      const float diff = update_diffs(s, id, (float)bu, (float)bv, mk<float>((float)Nn.x, (float)Nn.y, (float)Nn.z),
                                      *df, tcx, tcy);
      const int32_t tex = s.mats[mat].texture;
      if (tex >= 0) {
        const V3<float> tx = tex_sample(s.textures[(size_t)tex], tcx, tcy, diff);
        texel = mk<R>(tx.x, tx.y, tx.z);
        textured = true;
        mp.apply_texel(texel);
      }
    }

    /* NEE: pixel_generate_arealight_sample_cpu, src/arealight-sampler-cpu.cpp:25-119.
     * The normal is the unnormalised interpolation, as there. */
    V3<R> xl = mk<R>(R(H.sample(n, dim + 0)), R(H.sample(n, dim + 1)), R(H.sample(n, dim + 2)));
    R choice = xl.z * R(s.power_total);
    R acc = R(0);
    int k = 0;
    while (k < nl && choice > acc + R(s.lights[k].power)) { acc += R(s.lights[k].power); ++k; }
    if (k >= nl) k = nl - 1;
    const gi_light &lt = s.lights[k];
    V3<R> sd, contrib;
    R len;
    if (lt.type == GI_LIGHT_SKY) {
      sky_sample(s, xl.x, xl.y, Ni, sd, contrib);
      len = R(FLT_MAX);
      sd = normalize(sd);
    } else {
      V3<R> ldir = ld<R>(lt.dir);
      V3<R> right = tangential(mk<R>(1, 0, 0), ldir), up = tangential(mk<R>(0, 1, 0), ldir);
      R offx = (xl.x - R(0.5)) * R(lt.wh[0]), offy = (xl.y - R(0.5)) * R(lt.wh[1]);
      V3<R> ls = (ld<R>(lt.center) + right * offx) + up * offy;
      sd = ls - P;
      len = length(sd);
      sd = divs(sd, len);
      R ndotl = std::max(dot(Ni, sd), R(0));
      R lcos = std::max(dot(ldir, -sd), R(0));
      R factor = (((R(lt.wh[0]) * R(lt.wh[1])) * ndotl) * lcos) / (len * len);
      contrib = ld<R>(lt.col) * factor;
      sd = normalize(sd);
    }
    V3<R> so = P + sd * R(op.off_s);
    V3<R> psc = contrib * (R(s.power_total) / R(lt.power));

    if (op.phong) {
      /* gpu_pt's material model: NEE with the whole lobe, then the bounce
       * (offset along the new direction, no throughput clamp) */
      V3<R> kd = ld<R>(s.mats[mat].diffuse_color);
      if (textured) kd = mul(kd, texel);
      const V3<R> ks = ld<R>(s.mats[mat].specular_color);
      const V3<R> xp = mk<R>(R(H.sample(n, dim + 3)), R(H.sample(n, dim + 4)), R(H.sample(n, dim + 5)));
      V3<R> fnee = mk<R>(0, 0, 0), wi = mk<R>(0, 0, 0), tpmul = mk<R>(0, 0, 0);
      const bool alive = phong_vertex(kd, ks, normalize(Ni), org_dir, sd, xp, fnee, wi, tpmul);
      const V3<R> C = mul(fnee, mul(TP, psc));
      if (C.x > R(0) || C.y > R(0) || C.z > R(0)) {
        ++st.rays_shadow;
        if (!tr.any(so, sd, len, st.box_shadow, st.tri_shadow)) L = L + C;
      }
      if (b + 1 >= op.depth || !alive) break;
      V3<R> tp = mul(TP, tpmul);
      if (op.rr_start > 0 && b + 1 >= op.rr_start) {
        R q = maxc(tp);
        R xr = R(H.sample(n, dim + 6));
        if (!(xr < q)) break;
        tp = divs(tp, q);
      }
      TP = tp;
      o = P + wi * R(op.off_b);
      d = wi;
      tmax = R(FLT_MAX);
      continue;
    }

    /* bounce: getHitDataTriangle + sample_material + evaluate_material,
     * src/hybrid-pt.cpp:330-371,421-448,519-563. */
    V3<R> N = normalize(Ni);
    V3<R> Tt, Bt;
    tangent_frame(N, Tt, Bt);
    if (dot(org_dir, N) > R(0)) N = -N;
    V3<R> inv = -org_dir;
    V3<R> wo_ts = mk<R>(dot(inv, Tt), dot(inv, Bt), dot(inv, N));
    V3<R> xp = mk<R>(R(H.sample(n, dim + 3)), R(H.sample(n, dim + 4)), R(H.sample(n, dim + 5)));
    V3<R> wi_ts;
    R pdf = R(1);
    int type = brdf_sample(mp, wo_ts, xp, wi_ts, pdf);
    V3<R> wi = normalize((Tt * wi_ts.x + Bt * wi_ts.y) + N * wi_ts.z);
    if (!mp.glass()) {
      V3<R> ldn = normalize(sd);
      V3<R> C = mul(brdf_eval(mp, type, inv, ldn, N), mul(TP, psc));
      if (C.x > R(0) || C.y > R(0) || C.z > R(0)) {
        ++st.rays_shadow;
        if (!tr.any(so, sd, len, st.box_shadow, st.tri_shadow)) L = L + C;
      }
    }
    if (b + 1 >= op.depth) break;
    V3<R> f = brdf_eval(mp, type, inv, wi, N);
    V3<R> np = mp.glass() ? P + org_dir * R(op.off_b) : P + N * R(op.off_b);
    V3<R> tp = mk<R>((TP.x * f.x) / pdf, (TP.y * f.y) / pdf, (TP.z * f.z) / pdf);
    if (tp.x > R(1)) tp.x = R(1);
    if (tp.y > R(1)) tp.y = R(1);
    if (tp.z > R(1)) tp.z = R(1);
    if (op.rr_start > 0 && b + 1 >= op.rr_start) {
      R q = maxc(tp);
      R xr = R(H.sample(n, dim + 6));
      if (!(xr < q)) break;
      tp = divs(tp, q);
    }
    TP = tp;
    o = np;
    d = wi;
    tmax = R(FLT_MAX);
  }
  return L;
}

template <class R, class Img>
void render_impl(const or_scene &s, const gi_view *views, int32_t nv, const gi_render_opts &ro, Img *img,
                 or_stats *st_out, int32_t threads, int64_t pb, int64_t pe) {
  Options op = resolve(s, ro);
  HaltonTable H(ro.seed, dims_for_depth(op.depth));
  Tracer<R> tr{s};
  or_stats total;
  std::memset(&total, 0, sizeof(total));
  uint64_t base = 0;
  int64_t img_off = 0;
#ifdef _OPENMP
  int nth = threads > 0 ? threads : omp_get_max_threads();
#else
  int nth = 1;
  (void)threads;
#endif
  for (int32_t vi = 0; vi < nv; ++vi) {
    const gi_view &v = views[vi];
    Cam cam = make_cam(v);
    const int64_t npix = (int64_t)v.w * v.h;
    int64_t b0 = std::max<int64_t>(0, pb), e0 = pe < 0 ? npix : std::min<int64_t>(npix, pe);
    std::vector<or_stats> per(nth);
    for (auto &q : per) std::memset(&q, 0, sizeof(q));
#pragma omp parallel for schedule(dynamic, 64) num_threads(nth)
    for (int64_t p = b0; p < e0; ++p) {
#ifdef _OPENMP
      or_stats &st = per[omp_get_thread_num()];
#else
      or_stats &st = per[0];
#endif
      int32_t x = (int32_t)(p % v.w), y = (int32_t)(p / v.w);
      V3<R> acc = mk<R>(0, 0, 0);
      for (int32_t sidx = 0; sidx < v.spp; ++sidx) {
        uint64_t n = base + (uint64_t)p * (uint64_t)v.spp + (uint64_t)sidx;
        V3<float> of, df;
        cam_ray(cam, x, y, op.centre ? 0.5f : H.sample(n, DIM_JX), op.centre ? 0.5f : H.sample(n, DIM_JY),
                H.sample(n, DIM_LX), H.sample(n, DIM_LY), of, df);
        Diffs dfs;
        if (s.textured) cam_diffs(cam, x, y, dfs);
        V3<R> Ls = trace_sample<R>(s, tr, H, op, n, mk<R>(of.x, of.y, of.z), mk<R>(df.x, df.y, df.z), st,
                                   s.textured ? &dfs : nullptr);
        ++st.samples;
        /* combine_color_samples, src/util-kernels.cu:17-25 */
        acc = divs(R(sidx) * acc + Ls, R(sidx + 1));
      }
      Img *dst = img + 3 * (img_off + p);
      dst[0] = (Img)acc.x; dst[1] = (Img)acc.y; dst[2] = (Img)acc.z;
    }
    for (auto &q : per) {
      total.rays_closest += q.rays_closest; total.rays_shadow += q.rays_shadow;
      total.vertices += q.vertices; total.samples += q.samples;
      total.box_closest += q.box_closest; total.tri_closest += q.tri_closest;
      total.box_shadow += q.box_shadow; total.tri_shadow += q.tri_shadow;
    }
    base += (uint64_t)npix * (uint64_t)v.spp;
    img_off += npix;
  }
  if (st_out) *st_out = total;
}


/* ---------------------------------------------------------------- adjoint --
 * Path-replay gradient of an image-weighted loss sum_p dL/dI_p I_p with
 * respect to (baseColor rgb, roughness, metallic) of every material and the
 * emission (rgb) of every light, for the estimator above with its samples
 * frozen: the lobe choice, the cosine direction or the GGX half-vector and
 * the light choice are fixed functions of the Halton point, so the path
 * does not move when a parameter changes.  Terms per vertex j:
 *   NEE      dC_j = V_j * d f_nee_j * T_j * psc_j
 *   chain    R_{j,c} * d log(f_j / pdf_j)_c, R = radiance after j up to the
 *            next throughput clamp in channel c (the clamp zeroes the chain)
 *   score    S_j * d log p(sample_j): lobe choice P(spec) = metallic, GGX
 *            half-vector density D(h) cos(theta_h) w.r.t. roughness; S_j =
 *            everything the path gathers from j on
 *   emission dC_j / d col = V_j * f_nee_j * T_j * factor * P_tot / P_k with
 *            the light-selection weight (gi_light.power) held fixed.
 * This restatement records the whole path and back-propagates afterwards;
 * the device replays the path instead.  Russian roulette is not supported. */
template <class R> struct MatD {
  V3<R> cd, dcd, tint;
  R lum, w[3];
  explicit MatD(const MatP<R> &p) {
    cd = p.cdlin();
    if (p.textured) /* d b^2.2 / d b = 2.2 b^2.2 / b for the textured base colour b */
      dcd = mk<R>(p.color.x > R(0) ? (R(2.2) * cd.x) / p.color.x : R(0),
                  p.color.y > R(0) ? (R(2.2) * cd.y) / p.color.y : R(0),
                  p.color.z > R(0) ? (R(2.2) * cd.z) / p.color.z : R(0));
    else
      dcd = mk<R>(p.color.x > R(0) ? R(2.2) * std::pow(p.color.x, R(1.2)) : R(0),
                  p.color.y > R(0) ? R(2.2) * std::pow(p.color.y, R(1.2)) : R(0),
                  p.color.z > R(0) ? R(2.2) * std::pow(p.color.z, R(1.2)) : R(0));
    lum = R(0.3) * cd.x + R(0.6) * cd.y + R(0.1) * cd.z;
    tint = lum > R(0) ? divs(cd, lum) : mk<R>(1, 1, 1);
    w[0] = R(0.3); w[1] = R(0.6); w[2] = R(0.1);
  }
  R c(const V3<R> &v, int i) const { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
  /* d tint_c / d Cd_i */
  R dtint(int c, int i) const {
    if (!(lum > R(0))) return R(0);
    return ((c == i ? R(1) : R(0)) - c3(tint, c) * w[i]) / lum;
  }
  static R c3(const V3<R> &v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
};
template <class R> inline R &at(V3<R> &v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
template <class R> inline R cat(const V3<R> &v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

/* d f / d (color r, g, b, roughness, metallic) for the lobe `type`. */
template <class R>
void brdf_grad(const MatP<R> &p, const MatD<R> &d, int type, V3<R> wo, V3<R> wi, V3<R> N, V3<R> df[5]) {
  for (int k = 0; k < 5; ++k) df[k] = mk<R>(0, 0, 0);
  if (!p.principled) {
    const R a = std::fabs(dot(wi, N)) * K<R>::inv_pi;
    df[0].x = a; df[1].y = a; df[2].z = a;
    return;
  }
  const V3<R> V = wo, L = wi;
  const R ndl = dot(N, L), ndv = dot(N, V);
  if (ndl < R(0) || ndv < R(0)) return;
  if (type == T_DISNEY_REFL) {
    V3<R> H = normalize(L + V);
    const R ldh = dot(L, H), ndh = dot(N, H);
    const R fh = schlick(ldh), m = p.metallic, sp08 = p.specular * R(0.08), st = p.spec_tint;
    R cg = R(0);
    if (p.clearcoat > R(0)) {
      R dr = gtr1(ndh, mixf(R(0.1), R(0.001), p.cc_gloss));
      R fr = mixf(R(0.04), R(1), fh);
      R gr = smith_ggx(ndl, R(0.25)) * smith_ggx(ndv, R(0.25));
      cg = R(0.25) * p.clearcoat * gr * fr * dr;
    }
    for (int c = 0; c < 3; ++c) {
      const R mt = (R(1) - st) + st * cat(d.tint, c);
      const R cs0 = (R(1) - m) * sp08 * mt + m * cat(d.cd, c);
      const R fs = (R(1) - fh) * cs0 + fh;
      at(df[4], c) = (fs + cg) * ndl + m * ndl * (R(1) - fh) * (cat(d.cd, c) - sp08 * mt);
      for (int i = 0; i < 3; ++i) {
        const R dcs0 = m * (c == i ? R(1) : R(0)) + (R(1) - m) * sp08 * st * d.dtint(c, i);
        at(df[i], c) = m * ndl * (R(1) - fh) * dcs0 * cat(d.dcd, i);
      }
    }
    return;
  }
  if (type != T_DIFFUSE) return;  /* glass lobes have no parameter */
  V3<R> H = normalize(L + V);
  const R ldh = dot(L, H);
  const R fl = schlick(ndl), fv = schlick(ndv), r = p.roughness, m = p.metallic, sub = p.subsurface;
  const R fd90 = R(0.5) + R(2) * ldh * ldh * r;
  const R fd = mixf(R(1), fd90, fl) * mixf(R(1), fd90, fv);
  const R fh = schlick(ldh);
  const R fss90 = ldh * ldh * r;
  const R fss = mixf(R(1), fss90, fl) * mixf(R(1), fss90, fv);
  const R q = R(1) / (ndl + ndv) - R(0.5);
  const R ss = R(1.25) * (fss * q + R(0.5));
  const R k = K<R>::inv_pi * mixf(fd, ss, sub);
  const R dfd = R(2) * ldh * ldh * (fl * mixf(R(1), fd90, fv) + mixf(R(1), fd90, fl) * fv);
  const R dss = R(1.25) * q * ldh * ldh * (fl * mixf(R(1), fss90, fv) + mixf(R(1), fss90, fl) * fv);
  const R dk = K<R>::inv_pi * ((R(1) - sub) * dfd + sub * dss);
  for (int c = 0; c < 3; ++c) {
    const R csh = (R(1) - p.sheen_tint) + p.sheen_tint * cat(d.tint, c);
    at(df[3], c) = dk * cat(d.cd, c) * (R(1) - m) * ndl;
    at(df[4], c) = -(k * cat(d.cd, c) + fh * p.sheen * csh) * ndl;
    for (int i = 0; i < 3; ++i)
      at(df[i], c) = (R(1) - m) * ndl * (k * (c == i ? R(1) : R(0)) + fh * p.sheen * p.sheen_tint * d.dtint(c, i)) *
                     cat(d.dcd, i);
  }
}

/* GGX roughness derivatives of the sampled specular lobe: d log(lweight)/dr
 * (the reference's specular "pdf", src/principledMaterial.h:158-175) and the
 * score d log D(h)/dr of the half-vector density. */
template <class R> void gtr2_grads(const MatP<R> &p, V3<R> V, V3<R> L, V3<R> h, R &dlogw, R &score) {
  const R aspect = std::sqrt(R(1) - p.aniso * R(0.9));
  const R r = p.roughness, r2 = r * r;
  const R ax = std::max(R(0.001), r2 / aspect), ay = std::max(R(0.001), r2 * aspect);
  const R dax = (r2 / aspect > R(0.001)) ? R(2) * r / aspect : R(0);
  const R day = (r2 * aspect > R(0.001)) ? R(2) * r * aspect : R(0);
  const R qq = ax * ay * R(0.5) + R(0.5);
  const R roughg = qq * qq, rg2 = roughg * roughg;
  const R drg2 = R(4) * qq * qq * qq * R(0.5) * (dax * ay + ax * day);
  const R dlm = dot(L, h), dvm = dot(V, h);
  const R tl = std::sqrt(R(1) - dlm * dlm) / dlm, tv = std::sqrt(R(1) - dvm * dvm) / dvm;
  const R xl = (dot(L, h) / L.z) > R(0) ? R(1) : R(0), xv = (dot(V, h) / V.z) > R(0) ? R(1) : R(0);
  const R sl = std::sqrt(R(1) + rg2 * tl * tl), sv = std::sqrt(R(1) + rg2 * tv * tv);
  const R g1l = xl * R(2) / (R(1) + sl), g1v = xv * R(2) / (R(1) + sv);
  const R lw = (std::fabs(dot(L, h)) * g1l * g1v) / (std::fabs(L.z) * std::fabs(h.z));
  dlogw = R(0);
  if (lw >= R(0.01) && g1l > R(0) && g1v > R(0)) {
    const R dg1l = -(R(1) / (R(1) + sl)) * (tl * tl / (R(2) * sl));
    const R dg1v = -(R(1) / (R(1) + sv)) * (tv * tv / (R(2) * sv));
    dlogw = (dg1l + dg1v) * drg2;
  }
  const R Q = h.x * h.x / (ax * ax) + h.y * h.y / (ay * ay) + h.z * h.z;
  const R dlx = -R(1) / ax + R(4) * h.x * h.x / (ax * ax * ax * Q);
  const R dly = -R(1) / ay + R(4) * h.y * h.y / (ay * ay * ay * Q);
  score = dlx * dax + dly * day;
}

template <class R> struct VRec {
  bool hit = false, bg = false, vis = false, chain = false;
  int mat = 0, light = 0, clamp = 0;
  V3<R> rec{0, 0, 0}, T{0, 0, 0}, psc{0, 0, 0}, fnee{0, 0, 0};
  V3<R> dfn[5] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  V3<R> dlr[5] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  R gk = 0, score[5] = {0, 0, 0, 0, 0};
  int tex = -1;              /* textured vertex: texture, lookup, texel value, untextured colour */
  TexFetch tf{};
  V3<R> tx{0, 0, 0}, col{0, 0, 0};
};

template <class R>
void adjoint_sample(const or_scene &s, const Tracer<R> &tr, const HaltonTable &H, const Options &op, uint64_t n,
                    V3<R> o, V3<R> d, V3<R> delta, std::vector<double> &gm, std::vector<double> &ge, Diffs *df,
                    std::vector<double> &gt, const std::vector<size_t> &toff) {
  VRec<R> vr[32];
  int nv = 0;
  V3<R> TP = mk<R>(1, 1, 1);
  R tmax = R(FLT_MAX);
  const int nl = (int)s.lights.size();
  uint64_t nb = 0, nt = 0;
  for (int32_t b = 0; b < op.depth; ++b) {
    const int32_t dim = DIM_VTX0 + DIMS_PER_VTX * b;
    R t, bu = R(0), bv = R(0);
    uint32_t id;
    bool hit = tr.closest(o, d, tmax, t, bu, bv, id, nb, nt);
    V3<R> org_dir = normalize(d);
    VRec<R> &v = vr[nv++];
    v.T = TP;
    if (!hit) {
      if (TP.x == R(1) && TP.y == R(1) && TP.z == R(1)) {
        v.rec = mul(TP, mk<R>(R(op.bg[0]), R(op.bg[1]), R(op.bg[2])));
        v.bg = true;
      }
      break;
    }
    v.hit = true;
    V3<R> P, Ni;
    int32_t mat;
    surface<R>(s, id, bu, bv, P, Ni, mat);
    MatP<R> mp(s.mats[mat]);
    if (df) {
      const V3<R> Nn = normalize(Ni);
      float tcx, tcy;
      const float diff = update_diffs(s, id, (float)bu, (float)bv, mk<float>((float)Nn.x, (float)Nn.y, (float)Nn.z),
                                      *df, tcx, tcy);
      const int32_t tex = s.mats[mat].texture;
      if (tex >= 0) {
        const V3<float> tx = tex_sample(s.textures[(size_t)tex], tcx, tcy, diff, &v.tf);
        v.tex = tex;
        v.tx = mk<R>(tx.x, tx.y, tx.z);
        v.col = mp.principled ? mp.color : mp.diffuse;
        mp.apply_texel(v.tx);
      }
    }
    MatD<R> md(mp);
    v.mat = mat;
    V3<R> xl = mk<R>(R(H.sample(n, dim + 0)), R(H.sample(n, dim + 1)), R(H.sample(n, dim + 2)));
    R choice = xl.z * R(s.power_total), acc = R(0);
    int k = 0;
    while (k < nl && choice > acc + R(s.lights[k].power)) { acc += R(s.lights[k].power); ++k; }
    if (k >= nl) k = nl - 1;
    const gi_light &lt = s.lights[k];
    v.light = k;
    V3<R> ldir = ld<R>(lt.dir);
    V3<R> right = tangential(mk<R>(1, 0, 0), ldir), up = tangential(mk<R>(0, 1, 0), ldir);
    R offx = (xl.x - R(0.5)) * R(lt.wh[0]), offy = (xl.y - R(0.5)) * R(lt.wh[1]);
    V3<R> ls = (ld<R>(lt.center) + right * offx) + up * offy;
    V3<R> sd = ls - P;
    R len = length(sd);
    sd = divs(sd, len);
    R ndotl = std::max(dot(Ni, sd), R(0)), lcos = std::max(dot(ldir, -sd), R(0));
    R factor = (((R(lt.wh[0]) * R(lt.wh[1])) * ndotl) * lcos) / (len * len);
    v.gk = factor * (R(s.power_total) / R(lt.power));
    V3<R> psc = ld<R>(lt.col) * factor * (R(s.power_total) / R(lt.power));
    v.psc = psc;
    sd = normalize(sd);
    V3<R> so = P + sd * R(op.off_s);
    V3<R> N = normalize(Ni), Tt, Bt;
    tangent_frame(N, Tt, Bt);
    if (dot(org_dir, N) > R(0)) N = -N;
    V3<R> inv = -org_dir;
    V3<R> wo_ts = mk<R>(dot(inv, Tt), dot(inv, Bt), dot(inv, N));
    V3<R> xp = mk<R>(R(H.sample(n, dim + 3)), R(H.sample(n, dim + 4)), R(H.sample(n, dim + 5)));
    V3<R> wi_ts;
    R pdf = R(1);
    int type = brdf_sample(mp, wo_ts, xp, wi_ts, pdf);
    V3<R> wi = normalize((Tt * wi_ts.x + Bt * wi_ts.y) + N * wi_ts.z);
    const bool glass = mp.glass();
    /* scores of the frozen sample */
    if (mp.principled && !glass) {
      const bool spec_chosen = xp.z < mp.metallic;
      v.score[4] = spec_chosen ? R(1) / mp.metallic : -R(1) / (R(1) - mp.metallic);
      if (type == T_DISNEY_REFL) {
        R sn, cs;
        sincos2pi(xp.x, sn, cs);
        R aspect = std::sqrt(R(1) - mp.aniso * R(0.9));
        R ax = std::max(R(0.001), sqr(mp.roughness) / aspect), ay = std::max(R(0.001), sqr(mp.roughness) * aspect);
        R fac = std::sqrt(xp.y / (R(1) - xp.y));
        V3<R> h = normalize(mk<R>(fac * (ax * cs), fac * (ay * sn), R(1)));
        R dlogw, sc;
        gtr2_grads(mp, wo_ts, wi_ts, h, dlogw, sc);
        v.score[3] = sc;
        v.dlr[3] = mk<R>(-dlogw, -dlogw, -dlogw);
      }
    }
    if (!glass) {
      V3<R> ldn = normalize(sd);
      V3<R> fn = brdf_eval(mp, type, inv, ldn, N);
      V3<R> C = mul(fn, mul(TP, psc));
      v.fnee = fn;
      brdf_grad(mp, md, type, inv, ldn, N, v.dfn);
      if (C.x > R(0) || C.y > R(0) || C.z > R(0)) {
        if (!tr.any(so, sd, len, nb, nt)) {
          v.vis = true;
          v.rec = C;
        }
      }
    }
    if (b + 1 >= op.depth) break;
    V3<R> f = brdf_eval(mp, type, inv, wi, N);
    V3<R> df[5];
    brdf_grad(mp, md, type, inv, wi, N, df);
    V3<R> np = glass ? P + org_dir * R(op.off_b) : P + N * R(op.off_b);
    V3<R> tp = mk<R>((TP.x * f.x) / pdf, (TP.y * f.y) / pdf, (TP.z * f.z) / pdf);
    for (int c = 0; c < 3; ++c)
      if (cat(tp, c) > R(1)) { at(tp, c) = R(1); v.clamp |= 1 << c; }
    v.chain = true;
    for (int q = 0; q < 5; ++q)
      for (int c = 0; c < 3; ++c)
        at(v.dlr[q], c) = cat(f, c) != R(0) ? cat(df[q], c) / cat(f, c) + cat(v.dlr[q], c) : R(0);
    TP = tp;
    o = np;
    d = wi;
    tmax = R(FLT_MAX);
  }
  for (int j = 0; j < nv; ++j) {
    const VRec<R> &v = vr[j];
    if (!v.hit) continue;
    V3<R> Rj = mk<R>(0, 0, 0), Sj = v.rec;
    for (int b = j + 1; b < nv; ++b) Sj = Sj + vr[b].rec;
    if (v.chain)
      for (int c = 0; c < 3; ++c)
        for (int b = j + 1; b < nv; ++b) {
          if (vr[b - 1].clamp & (1 << c)) break;
          if (!vr[b].bg) at(Rj, c) += cat(vr[b].rec, c);
        }
    const R sdot = dot(delta, Sj);
    for (int q = 0; q < 5; ++q) {
      R g = sdot * v.score[q];
      for (int c = 0; c < 3; ++c) {
        if (v.vis) g += cat(delta, c) * cat(v.dfn[q], c) * cat(v.T, c) * cat(v.psc, c);
        if (v.chain) g += cat(delta, c) * cat(Rj, c) * cat(v.dlr[q], c);
      }
      if (v.tex >= 0 && q < 3) {
        /* base colour b_q = colour_q * texel_q */
        const double gq = (double)(g * cat(v.col, q));
        for (int k = 0; k < 4; ++k) gt[3 * (toff[(size_t)v.tex] + v.tf.idx[k]) + q] += gq * (double)v.tf.w[k];
        g = g * cat(v.tx, q);
      }
      gm[5 * v.mat + q] += (double)g;
    }
    if (v.vis)
      for (int c = 0; c < 3; ++c)
        ge[3 * v.light + c] += (double)(cat(delta, c) * cat(v.fnee, c) * cat(v.T, c) * v.gk);
  }
}

template <class R>
void adjoint_impl(const or_scene &s, const gi_view *views, int32_t nv, const gi_render_opts &ro, const float *dL,
                  double *d_mats, double *d_em, double *d_tex, int32_t threads) {
  Options op = resolve(s, ro);
  HaltonTable H(ro.seed, dims_for_depth(op.depth));
  Tracer<R> tr{s};
#ifdef _OPENMP
  int nth = threads > 0 ? threads : omp_get_max_threads();
#else
  int nth = 1;
  (void)threads;
#endif
  const size_t nm = s.mats.size(), nl = s.lights.size();
  std::vector<std::vector<double>> gm(nth, std::vector<double>(5 * nm, 0.0)), ge(nth, std::vector<double>(3 * nl, 0.0));
  std::vector<size_t> toff(s.textures.size() + 1, 0);
  for (size_t t = 0; t < s.textures.size(); ++t) toff[t + 1] = toff[t] + s.textures[t].chain.size() / 4;
  std::vector<std::vector<double>> gt(nth, std::vector<double>(s.textured ? 3 * toff.back() : 0, 0.0));
  uint64_t base = 0;
  int64_t img_off = 0;
  for (int32_t vi = 0; vi < nv; ++vi) {
    const gi_view &v = views[vi];
    Cam cam = make_cam(v);
    const int64_t npix = (int64_t)v.w * v.h;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nth)
    for (int64_t p = 0; p < npix; ++p) {
#ifdef _OPENMP
      const int th = omp_get_thread_num();
#else
      const int th = 0;
#endif
      const float *w = dL + 3 * (img_off + p);
      V3<R> delta = mk<R>(R(w[0]) / R(v.spp), R(w[1]) / R(v.spp), R(w[2]) / R(v.spp));
      if (delta.x == R(0) && delta.y == R(0) && delta.z == R(0)) continue;
      int32_t x = (int32_t)(p % v.w), y = (int32_t)(p / v.w);
      for (int32_t si = 0; si < v.spp; ++si) {
        uint64_t n = base + (uint64_t)p * (uint64_t)v.spp + (uint64_t)si;
        V3<float> of, df;
        cam_ray(cam, x, y, op.centre ? 0.5f : H.sample(n, DIM_JX), op.centre ? 0.5f : H.sample(n, DIM_JY),
                H.sample(n, DIM_LX), H.sample(n, DIM_LY), of, df);
        Diffs dfs;
        if (s.textured) cam_diffs(cam, x, y, dfs);
        adjoint_sample<R>(s, tr, H, op, n, mk<R>(of.x, of.y, of.z), mk<R>(df.x, df.y, df.z), delta, gm[th], ge[th],
                          s.textured ? &dfs : nullptr, gt[th], toff);
      }
    }
    base += (uint64_t)npix * (uint64_t)v.spp;
    img_off += npix;
  }
  for (size_t i = 0; i < 5 * nm; ++i) {
    d_mats[i] = 0;
    for (int t = 0; t < nth; ++t) d_mats[i] += gm[t][i];
  }
  for (size_t i = 0; i < 3 * nl; ++i) {
    d_em[i] = 0;
    for (int t = 0; t < nth; ++t) d_em[i] += ge[t][i];
  }
  if (!d_tex || !s.textured) return;
  /* texel gradients: sum the threads, then pass each level's gradient down
   * to the (up to) 2x2 texels it averages, coarsest first; output level 0 */
  std::vector<double> g(3 * toff.back(), 0.0);
  for (int t = 0; t < nth; ++t)
    for (size_t i = 0; i < g.size(); ++i) g[i] += gt[t][i];
  size_t out = 0;
  for (size_t ti = 0; ti < s.textures.size(); ++ti) {
    const or_scene::Texture &T = s.textures[ti];
    std::vector<size_t> lo(T.levels + 1);
    std::vector<int32_t> lw(T.levels + 1), lh(T.levels + 1);
    size_t o = 0;
    int32_t W = T.w, Hh = T.h;
    for (int l = 0; l <= T.levels; ++l) { lo[l] = o; lw[l] = W; lh[l] = Hh; o += (size_t)W * Hh; W = (W + 1) / 2; Hh = (Hh + 1) / 2; }
    double *gt0 = g.data() + 3 * toff[ti];
    for (int l = T.levels - 1; l >= 0; --l)
      for (int32_t y = 0; y < lh[l]; ++y)
        for (int32_t x = 0; x < lw[l]; ++x) {
          const int32_t px = x / 2, py = y / 2;
          const double cnt = (double)((2 * px + 1 < lw[l] ? 2 : 1) * (2 * py + 1 < lh[l] ? 2 : 1));
          for (int c = 0; c < 3; ++c)
            gt0[3 * (lo[l] + (size_t)y * lw[l] + x) + c] += gt0[3 * (lo[l + 1] + (size_t)py * lw[l + 1] + px) + c] / cnt;
        }
    for (size_t i = 0; i < 3 * (size_t)T.w * T.h; ++i) d_tex[out + i] = gt0[i];
    out += 3 * (size_t)T.w * T.h;
  }
}

}  // namespace

/* ================================================================ C API === */
extern "C" {

or_scene *or_scene_create(const gi_tri *tris, int64_t n, const gi_material *mats, int32_t nm,
                          const gi_light *lights, int32_t nl, const float *nodes2f4, int64_t n_nodes,
                          const uint32_t *indices, int64_t n_indices, int32_t max_leaf) {
  or_scene *s = new or_scene;
  s->tris.assign(tris, tris + n);
  if (mats) s->mats.assign(mats, mats + nm);
  if (lights) s->lights.assign(lights, lights + nl);
  for (const gi_material &m : s->mats) s->textured = s->textured || m.texture >= 0;
  for (auto &l : s->lights) s->power_total += l.power;
  for (int k = 0; k < 3; ++k) { s->lo[k] = FLT_MAX; s->hi[k] = -FLT_MAX; }
  for (auto &t : s->tris)
    for (int k = 0; k < 3; ++k) {
      s->lo[k] = std::min(s->lo[k], std::min(t.a[k], std::min(t.b[k], t.c[k])));
      s->hi[k] = std::max(s->hi[k], std::max(t.a[k], std::max(t.b[k], t.c[k])));
    }
  if (nodes2f4) {
    s->nodes.assign(nodes2f4, nodes2f4 + 8 * n_nodes);
    s->indices.assign(indices, indices + n_indices);
  } else if (n > 0) {
    s->indices.resize(n);
    std::iota(s->indices.begin(), s->indices.end(), 0u);
    s->nodes.assign(8, 0.f);
    MedianBuilder mb{s->tris, s->indices, s->nodes, max_leaf > 0 ? max_leaf : 4};
    mb.build(0, 0, (uint32_t)n);
  }
  return s;
}

void or_scene_free(or_scene *s) { delete s; }

int32_t or_scene_add_texture(or_scene *s, const uint8_t *rgba, int32_t w, int32_t h) {
  if (!s || !rgba || w <= 0 || h <= 0) return -1;
  or_scene::Texture t;
  t.w = w;
  t.h = h;
  build_chain(t, rgba);
  s->textures.push_back(std::move(t));
  return (int32_t)s->textures.size() - 1;
}
int32_t or_texture_levels(const or_scene *s, int32_t id) { return s->textures[(size_t)id].levels; }
int64_t or_texture_chain(const or_scene *s, int32_t id, uint8_t *out) {
  const std::vector<uint8_t> &c = s->textures[(size_t)id].chain;
  if (out) std::memcpy(out, c.data(), c.size());
  return (int64_t)c.size();
}
void or_tex_sample(const or_scene *s, int32_t id, float u, float v, float diff, float *rgb) {
  const V3<float> c = tex_sample(s->textures[(size_t)id], u, v, diff);
  rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
int32_t or_tex_level(float diff, int32_t w, int32_t h, int32_t levels) { return tex_level(diff, w, h, levels); }
float or_pow22(float x) { return pow22(x); }

int32_t or_scene_add_mesh(or_scene *s, const gi_tri *tris, int64_t n, int32_t max_leaf) {
  if (!s || !tris || n <= 0) return -1;
  s->meshes.push_back(or_scene_create(tris, n, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, 0, max_leaf));
  return (int32_t)s->meshes.size() - 1;
}

/* Instance i: the rows of xform (object to world) and of the inverse map,
 * (A^-1 | -A^-1 t), computed in double by cofactors over the determinant and
 * rounded to float; vbase = sum of the mesh sizes of the instances before
 * i.  -1 for a bad mesh id or material, or a singular 3x3 part. */
int or_scene_set_instances(or_scene *s, const gi_instance *inst, int32_t n) {
  if (!s || n < 0 || (n > 0 && !inst)) return -1;
  std::vector<or_scene::Instance> out((size_t)n);
  uint32_t vbase = 0;
  for (int32_t i = 0; i < n; ++i) {
    const gi_instance &g = inst[i];
    or_scene::Instance &I = out[(size_t)i];
    if (g.mesh < 0 || g.mesh >= (int32_t)s->meshes.size() || g.material < -1) return -1;
    const float *m = g.xform;
    const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], gg = m[8], h = m[9], k = m[10];
    const double c00 = e * k - f * h, c01 = f * gg - d * k, c02 = d * h - e * gg;
    const double det = a * c00 + b * c01 + c * c02;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return -1;
    /* the adjugate (transposed cofactors) over det */
    const double r[3][3] = {{c00 / det, (c * h - b * k) / det, (b * f - c * e) / det},
                            {c01 / det, (a * k - c * gg) / det, (c * d - a * f) / det},
                            {c02 / det, (b * gg - a * h) / det, (a * e - b * d) / det}};
    for (int row = 0; row < 3; ++row) {
      for (int col = 0; col < 3; ++col) I.inv[4 * row + col] = (float)r[row][col];
      I.inv[4 * row + 3] = (float)-(r[row][0] * (double)m[3] + r[row][1] * (double)m[7] + r[row][2] * (double)m[11]);
    }
    std::memcpy(I.fwd, m, sizeof(I.fwd));
    I.mesh = g.mesh;
    I.material = g.material;
    I.vbase = vbase;
    vbase += (uint32_t)s->meshes[(size_t)g.mesh]->tris.size();
  }
  s->inst = std::move(out);
  return 0;
}
int64_t or_scene_bvh_nodes(const or_scene *s) { return (int64_t)s->nodes.size() / 8; }
void or_scene_export_bvh(const or_scene *s, float *nodes, uint32_t *indices) {
  std::memcpy(nodes, s->nodes.data(), s->nodes.size() * 4);
  std::memcpy(indices, s->indices.data(), s->indices.size() * 4);
}
float or_scene_diagonal(const or_scene *s) {
  float dx = s->hi[0] - s->lo[0], dy = s->hi[1] - s->lo[1], dz = s->hi[2] - s->lo[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

void or_trace_closest(const or_scene *s, const float *o, const float *d, const float *tmax, int64_t n,
                      gi_hit *out, uint64_t *n_box, uint64_t *n_tri) {
  uint64_t nb = 0, nt = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nb, nt)
  for (int64_t i = 0; i < n; ++i) {
    float t, u = 0.f, v = 0.f;
    uint32_t id;
    uint64_t b = 0, c = 0;
    trace_scene<float, false>(*s, ld<float>(o + 3 * i), ld<float>(d + 3 * i), tmax[i], t, u, v, id, b, c);
    nb += b; nt += c;
    out[i].t = t; out[i].beta = id == GI_INVALID_REF ? 0.f : u; out[i].gamma = id == GI_INVALID_REF ? 0.f : v;
    out[i].ref = id;
  }
  if (n_box) *n_box = nb;
  if (n_tri) *n_tri = nt;
}

void or_trace_any(const or_scene *s, const float *o, const float *d, const float *tmax, int64_t n, uint8_t *occ) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    uint64_t b = 0, c = 0;
    occ[i] = Tracer<float>{*s}.any(ld<float>(o + 3 * i), ld<float>(d + 3 * i), tmax[i], b, c) ? 1 : 0;
  }
}

void or_brute_closest(const or_scene *s, const float *o, const float *d, const float *tmax, int64_t n, gi_hit *out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    float t, u = 0.f, v = 0.f;
    uint32_t id;
    brute<float>(*s, ld<float>(o + 3 * i), ld<float>(d + 3 * i), tmax[i], t, u, v, id);
    out[i].t = t; out[i].beta = id == GI_INVALID_REF ? 0.f : u; out[i].gamma = id == GI_INVALID_REF ? 0.f : v;
    out[i].ref = id;
  }
}

void or_brute_closest_f64(const or_scene *s, const float *o, const float *d, const float *tmax, int64_t n,
                          double *t, double *beta, double *gamma, uint32_t *ref) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    double tt, u = 0, v = 0;
    uint32_t id;
    brute<double>(*s, ld<double>(o + 3 * i), ld<double>(d + 3 * i), (double)tmax[i], tt, u, v, id);
    t[i] = tt; beta[i] = u; gamma[i] = v; ref[i] = id;
  }
}

int or_intersect_tri(const float *a, const float *b, const float *c, const float *o, const float *d, float *t,
                     float *beta, float *gamma) {
  V3<float> A = ld<float>(a), B = ld<float>(b), C = ld<float>(c);
  return intersect_tri<float>(A, B - A, C - A, ld<float>(o), ld<float>(d), *t, *beta, *gamma) ? 1 : 0;
}
int or_intersect_tri_f64(const double *a, const double *b, const double *c, const double *o, const double *d,
                         double *t, double *beta, double *gamma) {
  V3<double> A{a[0], a[1], a[2]}, B{b[0], b[1], b[2]}, C{c[0], c[1], c[2]};
  V3<double> O{o[0], o[1], o[2]}, D{d[0], d[1], d[2]};
  return intersect_tri<double>(A, B - A, C - A, O, D, *t, *beta, *gamma) ? 1 : 0;
}

/* gi::halton, src/util.cpp:22-32 (f = f/base in float). */
float or_radical_inverse(uint64_t index, int32_t base) {
  float result = 0.0f;
  float f = 1.0f / float(base);
  uint64_t i = index;
  while (i > 0) {
    result += f * (float)(i % (uint64_t)base);
    i = i / (uint64_t)base;
    f = f / float(base);
  }
  return result;
}
float or_halton_sample(uint64_t n, int32_t dim, uint32_t seed) {
  std::vector<uint16_t> perm(primes().p[dim]);
  halton_perm(dim, seed, perm.data());
  return scrambled_ri(n, primes().p[dim], perm.data());
}
int32_t or_prime(int32_t dim) { return primes().p[dim]; }
void or_halton_perm(int32_t dim, uint32_t seed, uint16_t *perm) { halton_perm(dim, seed, perm); }

void or_camera_rays(const gi_view *v, int32_t sample, uint32_t seed, uint64_t index_base, uint32_t flags, float *o,
                    float *d, float *tmax) {
  const bool centre = (flags & GI_FLAG_PIXEL_CENTRE) != 0;
  Cam cam = make_cam(*v);
  HaltonTable H(seed, DIM_VTX0);
  int64_t npix = (int64_t)v->w * v->h;
  for (int64_t p = 0; p < npix; ++p) {
    uint64_t n = index_base + (uint64_t)p * (uint64_t)v->spp + (uint64_t)sample;
    V3<float> of, df;
    cam_ray(cam, (int32_t)(p % v->w), (int32_t)(p / v->w), centre ? 0.5f : H.sample(n, DIM_JX),
            centre ? 0.5f : H.sample(n, DIM_JY),
            H.sample(n, DIM_LX), H.sample(n, DIM_LY), of, df);
    std::memcpy(o + 3 * p, &of, 12);
    std::memcpy(d + 3 * p, &df, 12);
    tmax[p] = FLT_MAX;
  }
}

void or_tangent_frame(const float *n, float *t, float *b) {
  V3<float> T, B;
  tangent_frame(ld<float>(n), T, B);
  std::memcpy(t, &T, 12);
  std::memcpy(b, &B, 12);
}

void or_nee(const or_scene *s, const float *Pp, const float *Np, const float *xi, float *dir, float *len,
            float *contrib, int32_t *light) {
  const int nl = (int)s->lights.size();
  float choice = xi[2] * s->power_total, acc = 0.f;
  int k = 0;
  while (k < nl && choice > acc + s->lights[k].power) { acc += s->lights[k].power; ++k; }
  if (k >= nl) k = nl - 1;
  const gi_light &lt = s->lights[k];
  V3<float> ldir = ld<float>(lt.dir), P = ld<float>(Pp), N = ld<float>(Np);
  V3<float> right = tangential(mk<float>(1, 0, 0), ldir), up = tangential(mk<float>(0, 1, 0), ldir);
  float offx = (xi[0] - 0.5f) * lt.wh[0], offy = (xi[1] - 0.5f) * lt.wh[1];
  V3<float> ls = (ld<float>(lt.center) + right * offx) + up * offy;
  V3<float> sd = ls - P;
  float l = length(sd);
  sd = divs(sd, l);
  float ndotl = std::max(dot(N, sd), 0.f), lcos = std::max(dot(ldir, -sd), 0.f);
  float factor = (((lt.wh[0] * lt.wh[1]) * ndotl) * lcos) / (l * l);
  V3<float> c = ld<float>(lt.col) * factor;
  c = c * (s->power_total / lt.power);
  sd = normalize(sd);
  std::memcpy(dir, &sd, 12);
  *len = l;
  std::memcpy(contrib, &c, 12);
  *light = k;
}

int or_brdf_sample(const gi_material *m, const float *wo_ts, const float *xi, float *wi_ts, float *pdf) {
  MatP<float> p(*m);
  V3<float> wi;
  float pd = 1.f;
  int t = brdf_sample(p, ld<float>(wo_ts), ld<float>(xi), wi, pd);
  std::memcpy(wi_ts, &wi, 12);
  *pdf = pd;
  return t;
}
void or_brdf_eval(const gi_material *m, int type, const float *wo, const float *wi, const float *N, float *f) {
  MatP<float> p(*m);
  V3<float> r = brdf_eval(p, type, ld<float>(wo), ld<float>(wi), ld<float>(N));
  std::memcpy(f, &r, 12);
}
int or_brdf_sample_f64(const gi_material *m, const double *wo_ts, const double *xi, double *wi_ts, double *pdf) {
  MatP<double> p(*m);
  V3<double> wi, wo{wo_ts[0], wo_ts[1], wo_ts[2]}, x{xi[0], xi[1], xi[2]};
  double pd = 1.0;
  int t = brdf_sample(p, wo, x, wi, pd);
  wi_ts[0] = wi.x; wi_ts[1] = wi.y; wi_ts[2] = wi.z;
  *pdf = pd;
  return t;
}
void or_brdf_eval_f64(const gi_material *m, int type, const double *wo, const double *wi, const double *N, double *f) {
  MatP<double> p(*m);
  V3<double> r = brdf_eval(p, type, V3<double>{wo[0], wo[1], wo[2]}, V3<double>{wi[0], wi[1], wi[2]},
                           V3<double>{N[0], N[1], N[2]});
  f[0] = r.x; f[1] = r.y; f[2] = r.z;
}

void or_gtr2_half_f64(double u, double v, double ax, double ay, double *h) {
  V3<double> m = gtr2_half(u, v, ax, ay);
  h[0] = m.x; h[1] = m.y; h[2] = m.z;
}
double or_walter_weight_f64(const double *L, const double *V, const double *m, double ax, double ay) {
  return walter_weight(V3<double>{L[0], L[1], L[2]}, V3<double>{V[0], V[1], V[2]}, V3<double>{m[0], m[1], m[2]},
                       ax, ay);
}
double or_gtr1_f64(double ndh, double a) { return gtr1(ndh, a); }
float or_phong_pow100(float x) { return phong_pow100(x); }
float or_phong_root101(float x) { return phong_root101(x); }
int or_phong_vertex(const float *kd, const float *ks, const float *N, const float *org_dir, const float *sd,
                    const float *xp, float *fnee, float *wi, float *tpmul) {
  V3<float> f = mk<float>(0, 0, 0), w = mk<float>(0, 0, 0), t = mk<float>(0, 0, 0);
  const bool alive = phong_vertex(ld<float>(kd), ld<float>(ks), ld<float>(N), ld<float>(org_dir), ld<float>(sd),
                                  ld<float>(xp), f, w, t);
  std::memcpy(fnee, &f, 12);
  std::memcpy(wi, &w, 12);
  std::memcpy(tpmul, &t, 12);
  return alive ? 1 : 0;
}
void or_camera_basis(const gi_view *v, float *U, float *V, float *W, float *wh) {
  Cam c = make_cam(*v);
  std::memcpy(U, c.U, 12);
  std::memcpy(V, c.V, 12);
  std::memcpy(W, c.W, 12);
  wh[0] = c.width;
  wh[1] = c.height;
}

void or_render(const or_scene *s, const gi_view *views, int32_t n_views, const gi_render_opts *opts, float *images,
               or_stats *st, int32_t threads, int64_t p_begin, int64_t p_end) {
  render_impl<float, float>(*s, views, n_views, *opts, images, st, threads, p_begin, p_end);
}
void or_render_f64(const or_scene *s, const gi_view *views, int32_t n_views, const gi_render_opts *opts,
                   double *images, int32_t threads) {
  render_impl<double, double>(*s, views, n_views, *opts, images, nullptr, threads, 0, -1);
}

/* The sky map of the scene's (single) GI_LIGHT_SKY light, and its sampling
 * tables: gi::sky_light::initSkylight (src/lights.h:107-116) -- importance
 * sin(pi (y + 0.5) / h) * (r + g + b) -- and Distribution1D::init
 * (src/lights.h:27-44), in float, in order. */
int or_scene_set_sky(or_scene *s, const float *rgb, int32_t w, int32_t h) {
  int sky = -1, count = 0;
  for (size_t i = 0; i < s->lights.size(); ++i)
    if (s->lights[i].type == GI_LIGHT_SKY) { sky = (int)i; ++count; }
  if (count != 1 || !rgb || w <= 0 || h <= 0) return 1;
  s->sky_light = sky;
  s->sky_w = w;
  s->sky_h = h;
  s->sky_rgb.assign(rgb, rgb + 3 * (size_t)w * h);
  auto init = [](const float *data, int size, float *pdf, float *cdf) {
    float sum = 0.0f;
    for (int i = 0; i < size; ++i) sum += data[i];
    float scale = 0.0f;
    if (sum > 0.0f) scale = 1.0f / sum;
    cdf[0] = 0.0f;
    for (int i = 0; i < size; ++i) {
      pdf[i] = data[i] * scale;
      if (i > 0) cdf[i] = cdf[i - 1] + pdf[i - 1];
    }
    cdf[size] = 1.0f;
    return sum;
  };
  std::vector<float> imp((size_t)w * h), rows(h);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const float *c = &rgb[3 * ((size_t)y * w + x)];
      imp[(size_t)y * w + x] = std::sin((float)(M_PI * (y + 0.5f) / h)) * ((c[0] + c[1]) + c[2]);
    }
  s->sky_xpdf.assign((size_t)w * h, 0.f);
  s->sky_xcdf.assign((size_t)(w + 1) * h, 0.f);
  for (int y = 0; y < h; ++y)
    rows[y] = init(&imp[(size_t)y * w], w, &s->sky_xpdf[(size_t)y * w], &s->sky_xcdf[(size_t)y * (w + 1)]);
  s->sky_ypdf.assign(h, 0.f);
  s->sky_ycdf.assign(h + 1, 0.f);
  init(rows.data(), h, s->sky_ypdf.data(), s->sky_ycdf.data());
  return 0;
}

void or_sky_eval(const or_scene *s, const float *d, float *rgb) {
  V3<float> c = sky_eval(*s, ld<float>(d));
  std::memcpy(rgb, &c, 12);
}

void or_sky_sample(const or_scene *s, const float *xi, const float *N, float *dir, float *contrib) {
  V3<float> sd, c;
  sky_sample(*s, xi[0], xi[1], ld<float>(N), sd, c);
  std::memcpy(dir, &sd, 12);
  std::memcpy(contrib, &c, 12);
}

float or_atan2_turns(float y, float x) { return atan2_turns(y, x); }

float or_cdf_sample(const float *cdf, const float *pdf, int32_t size, float u, float *p) {
  return cdf_sample(cdf, pdf, size, u, *p);
}

int or_adjoint_tex(const or_scene *s, const gi_view *views, int32_t n_views, const gi_render_opts *opts,
                   const float *dL_dimage, double *d_mats, double *d_emission, double *d_tex, int32_t threads,
                   int32_t f64) {
  if (opts->rr_start > 0 || opts->max_depth > 32 || s->sky_light >= 0) return 1;
  if (f64) adjoint_impl<double>(*s, views, n_views, *opts, dL_dimage, d_mats, d_emission, d_tex, threads);
  else adjoint_impl<float>(*s, views, n_views, *opts, dL_dimage, d_mats, d_emission, d_tex, threads);
  return 0;
}

int or_adjoint(const or_scene *s, const gi_view *views, int32_t n_views, const gi_render_opts *opts,
               const float *dL_dimage, double *d_mats, double *d_emission, int32_t threads, int32_t f64) {
  return or_adjoint_tex(s, views, n_views, opts, dL_dimage, d_mats, d_emission, nullptr, threads, f64);
}

}  // extern "C"
