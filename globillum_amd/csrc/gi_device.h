// Device-side math of the gfx950 path-tracing kernels.  Every function that
// feeds a hit id, a sample direction or an image value uses only IEEE
// basic operations in a fixed order (the library is built with
// -ffp-contract=off), correctly rounded division and sqrt, and the
// polynomial sincos below -- so the results are reproducible bit-for-bit on
// any IEEE host.  Only the slab test uses the fast reciprocal: box tests
// just need to be conservative (boxes are padded by the builder).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gi_internal.h"

namespace gi {
namespace dev {

struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 scl(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 normalize(f3 a) { return scl(a, 1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
// max/min with the operand order of std::max / std::min
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ f3 ld3(const float *p) { return f3{p[0], p[1], p[2]}; }
__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;

// sin(2*pi*v), cos(2*pi*v) for v in [0,1): quadrant q = floor(4v), the
// remainder f = 4v - q (both exact), then the Taylor series of
// sin(pi/2 f) to degree 13 and cos(pi/2 f) to degree 14 (truncation error
// below 1e-9), rotated into the quadrant.
constexpr double kH = 1.5707963267948966;
constexpr float kS1 = (float)kH;
constexpr float kS3 = (float)(-kH * kH * kH / 6.0);
constexpr float kS5 = (float)(kH * kH * kH * kH * kH / 120.0);
constexpr float kS7 = (float)(-kH * kH * kH * kH * kH * kH * kH / 5040.0);
constexpr float kS9 = (float)(kH * kH * kH * kH * kH * kH * kH * kH * kH / 362880.0);
constexpr float kS11 = (float)(-kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH / 39916800.0);
constexpr float kS13 = (float)(kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH / 6227020800.0);
constexpr float kC2 = (float)(-kH * kH / 2.0);
constexpr float kC4 = (float)(kH * kH * kH * kH / 24.0);
constexpr float kC6 = (float)(-kH * kH * kH * kH * kH * kH / 720.0);
constexpr float kC8 = (float)(kH * kH * kH * kH * kH * kH * kH * kH / 40320.0);
constexpr float kC10 = (float)(-kH * kH * kH * kH * kH * kH * kH * kH * kH * kH / 3628800.0);
constexpr float kC12 = (float)(kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH / 479001600.0);
constexpr float kC14 =
    (float)(-kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH * kH / 87178291200.0);

__device__ __forceinline__ void sincos2pi(float v, float &s, float &c) {
  float x = v * 4.0f;
  float q = floorf(x);
  float f = x - q;
  float f2 = f * f;
  float ps = f * (kS1 + f2 * (kS3 + f2 * (kS5 + f2 * (kS7 + f2 * (kS9 + f2 * (kS11 + f2 * kS13))))));
  float pc = 1.0f + f2 * (kC2 + f2 * (kC4 + f2 * (kC6 + f2 * (kC8 + f2 * (kC10 + f2 * (kC12 + f2 * kC14))))));
  int qi = (int)q & 3;
  if (qi == 0) { s = ps; c = pc; }
  else if (qi == 1) { s = pc; c = -ps; }
  else if (qi == 2) { s = -ps; c = -pc; }
  else { s = -pc; c = ps; }
}

// ----------------------------------------------------------- atan2, acos --
// atan(t) / (2 pi) for t in [0, 1] as t * P(t^2), P of degree 8: a
// least-squares fit (relative error < 2e-7) evaluated with basic operations
// only, so the host restatement gives the same bits.
__host__ __device__ __forceinline__ float atan_turns01(float t) {
  const float x2 = t * t;
  float p = 0.000454627268f;
  p = p * x2 + -0.00256048539f;
  p = p * x2 + 0.00679757446f;
  p = p * x2 + -0.0119441384f;
  p = p * x2 + 0.0169349238f;
  p = p * x2 + -0.0226053596f;
  p = p * x2 + 0.0318191461f;
  p = p * x2 + -0.0530512258f;
  p = p * x2 + 0.159154937f;
  return t * p;
}
// atan2(y, x) / (2 pi) in [-1/2, 1/2]; 0 for (0, 0).
__host__ __device__ __forceinline__ float atan2_turns(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  if (ax == 0.0f && ay == 0.0f) return 0.0f;
  float r = ay > ax ? 0.25f - atan_turns01(ax / ay) : atan_turns01(ay / ax);
  if (x < 0.0f) r = 0.5f - r;
  return y < 0.0f ? -r : r;
}

// Distribution1D::sample (src/lights.h:50-57): upper_bound over cdf[0,size),
// index clamped to [0,size-1], in-cell position t = (cdf[i+1]-u) /
// (cdf[i+1]-cdf[i]), x = (1.0-t)*i + t*(i+1) (first product in double, the
// second in float, as the reference's expression evaluates).
__device__ __forceinline__ float cdf_sample(const float *__restrict__ cdf, const float *__restrict__ pdf, int size,
                                            float u, float &p) {
  int lo = 0, hi = size;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cdf[mid] <= u) lo = mid + 1;
    else hi = mid;
  }
  const int idx = min(max(lo - 1, 0), size - 1);
  const float t = (cdf[idx + 1] - u) / (cdf[idx + 1] - cdf[idx]);
  p = pdf[idx];
  return (float)((1.0 - (double)t) * (double)idx + (double)(t * (float)(idx + 1)));
}

// ---------------------------------------------------------------- Halton --
struct HaltonDev {
  const HaltonDim *dims;
  const uint16_t *chunks;
};
// Scrambled radical inverse of n in the prime base p of dimension dim: the
// digits d_k (K of them) map to perm[d_k] and the infinitely many leading
// zeros to perm[0], which is the exact rational
//   A / B,  A = rev*(p-1) + perm[0],  B = p^K (p-1),  rev = sum perm[d_k] p^(K-1-k).
// Digits are consumed m at a time through the chunk table (P = p^m <= 1024);
// a partial last chunk adds leading zeros, which scales A and B by the same
// power of p.  A, B < 2^53 for n < 2^32, so the value is one correctly
// rounded FP64 division and one rounding to FP32.
// Every dimension has P >= 37, so n < 2^32 has at most 7 chunks: all 7 are
// split off and looked up at once (independent loads), and the ones above
// the leading chunk are then dropped.
constexpr int kHaltonMaxChunks = 7;
__device__ __forceinline__ float halton(uint32_t n, int dim, HaltonDev H) {
  const HaltonDim hd = H.dims[dim];
  const uint16_t *chunk = H.chunks + hd.chunk_off;
  uint32_t c[kHaltonMaxChunks];
  uint32_t m = n, K = 0;
#pragma unroll
  for (int k = 0; k < kHaltonMaxChunks; ++k) {
    const uint32_t q = (uint32_t)__umul64hi((uint64_t)m, hd.magic);
    c[k] = m - q * hd.P;
    K += m > 0u ? 1u : 0u;
    m = q;
  }
  uint32_t v[kHaltonMaxChunks];
#pragma unroll
  for (int k = 0; k < kHaltonMaxChunks; ++k) v[k] = chunk[c[k]];
  uint64_t rev = 0, pk = 1;
#pragma unroll
  for (int k = 0; k < kHaltonMaxChunks; ++k) {
    if ((uint32_t)k < K) {
      rev = rev * hd.P + v[k];
      pk *= hd.P;
    }
  }
  const uint64_t A = rev * (uint64_t)(hd.prime - 1) + hd.perm0;
  const uint64_t B = pk * (uint64_t)(hd.prime - 1);
  const float f = (float)((double)A / (double)B);
  return f >= 1.0f ? 0x1.fffffep-1f : f;
}

// ---------------------------------------------------------- frames -------
// Orthonormal basis around unit n: Duff et al., JCGT 6(1) 2017; T and B
// normalised afterwards.
__device__ __forceinline__ void tangent_frame(f3 n, f3 &t, f3 &b) {
  float sign = copysignf(1.0f, n.z);
  float a = -1.0f / (sign + n.z);
  float bb = n.x * n.y * a;
  t = f3{1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x};
  b = f3{bb, sign + n.y * n.y * a, -n.y};
  t = normalize(t);
  b = normalize(b);
}
__device__ __forceinline__ f3 tangential(f3 v, f3 n) {
  f3 t, b;
  tangent_frame(n, t, b);
  return add(add(scl(t, v.x), scl(b, v.y)), scl(n, v.z));
}

// ---------------------------------------------------------- intersect ----
// Moller-Trumbore 1997, no culling: hit iff det != 0, u,v >= 0, u+v <= 1,
// t > 0.  e1 = b - a and e2 = c - a come precomputed (same float rounding).
__device__ __forceinline__ bool intersect_tri(f3 a, f3 e1, f3 e2, f3 o, f3 d, float &t, float &u,
                                              float &v) {
  // evaluated without early exits (one wave runs every lane's test anyway);
  // same operations and roundings as the early-exit form, so a hit has
  // bit-identical t, u, v
  const f3 pvec = cross(d, e2);
  const float det = dot(e1, pvec);
  const f3 tvec = sub(o, a);
  const f3 qvec = cross(tvec, e1);
  const float inv_det = 1.0f / det;
  u = dot(tvec, pvec) * inv_det;
  v = dot(d, qvec) * inv_det;
  t = dot(e2, qvec) * inv_det;
  return (det != 0.0f) & (u >= 0.0f) & (u <= 1.0f) & (v >= 0.0f) & (u + v <= 1.0f) & (t > 0.0f);
}

__device__ __forceinline__ float rcp_clamped(float d) {
  const float ooeps = 8.2718061255302767e-25f;  // 2^-80, as src/tracers.cu:48-51
  return __builtin_amdgcn_rcpf(fabsf(d) > ooeps ? d : copysignf(ooeps, d));
}

// ---------------------------------------------------------- materials ----
// Device material record: the material parameters plus the quantities that
// depend only on them (precomputed on the host with the C library's pow and
// log, so the device needs no transcendental besides sincos2pi).
struct alignas(16) DevMat {
  float diffuse[3];
  int32_t principled;
  float cdlin[3];   // mon2lin(color)
  float metallic;
  float cspec0[3];  // mix(specular*.08*mix(1,Ctint,specularTint), Cdlin, metallic)
  float roughness;
  float csheen[3];  // mix(1, Ctint, sheenTint)
  float sheen;
  float subsurface, clearcoat, cc_a, cc_log_a2;  // GTR1 alpha and log(alpha^2)
  float opacity, ax, ay, roughg2;
  float color[3];
  float specular;
};

enum { kTypeSpecular = 0, kTypeDiffuse = 1, kTypeSpecRefl = 3, kTypeSpecTrans = 4, kTypeDisneyRefl = 5 };

__device__ __forceinline__ float saturate(float a) { return smin(smax(a, 0.0f), 1.0f); }
__device__ __forceinline__ float schlick(float u) {
  float m = saturate(1.0f - u);
  float m2 = m * m;
  return m2 * m2 * m;
}
__device__ __forceinline__ float mixf(float x, float y, float s) { return (1.0f - s) * x + s * y; }
__device__ __forceinline__ f3 mixv(f3 x, f3 y, float s) {
  return add(f3{(1.0f - s) * x.x, (1.0f - s) * x.y, (1.0f - s) * x.z}, f3{s * y.x, s * y.y, s * y.z});
}
__device__ __forceinline__ float sqrf(float a) { return a * a; }

__device__ __forceinline__ f3 cosine_sample(float u, float v) {
  float r = sqrtf(u);
  float s, c;
  sincos2pi(v, s, c);
  return f3{r * c, r * s, sqrtf(smax(0.0f, 1.0f - u))};
}

__device__ __forceinline__ float smith_ggx(float ndv, float ag) {
  float a = ag * ag, b = ndv * ndv;
  return 1.0f / (ndv + sqrtf(a + b - a * b));
}
__device__ __forceinline__ float gtr1(float ndh, float a, float log_a2) {
  if (a >= 1.0f) return kInvPi;
  float a2 = a * a;
  float t = 1.0f + (a2 - 1.0f) * ndh * ndh;
  return (a2 - 1.0f) / (kPi * log_a2 * t);
}

// evaluatePrincipledBRDF_specular, src/principledMaterial.h:210-252
__device__ __forceinline__ f3 eval_spec(const DevMat &m, f3 V, f3 N, f3 L) {
  float ndl = dot(N, L), ndv = dot(N, V);
  if (ndl < 0.0f || ndv < 0.0f) return f3{0.f, 0.f, 0.f};
  f3 H = normalize(add(L, V));
  float ndh = dot(N, H), ldh = dot(L, H);
  float fh = schlick(ldh);
  f3 fs = mixv(ld3(m.cspec0), f3{1.f, 1.f, 1.f}, fh);
  float cg = 0.0f;
  if (m.clearcoat > 0.0f) {
    float dr = gtr1(ndh, m.cc_a, m.cc_log_a2);
    float fr = mixf(0.04f, 1.0f, fh);
    float gr = smith_ggx(ndl, 0.25f) * smith_ggx(ndv, 0.25f);
    cg = 0.25f * m.clearcoat * gr * fr * dr;
  }
  f3 sum = add(fs, f3{cg, cg, cg});
  return scl(scl(sum, ndl), m.metallic);
}

// evaluatePrincipledBRDF_diffuse, src/principledMaterial.h:254-301
__device__ __forceinline__ f3 eval_diff(const DevMat &m, f3 V, f3 N, f3 L) {
  float ndl = dot(N, L), ndv = dot(N, V);
  if (ndl < 0.0f || ndv < 0.0f) return f3{0.f, 0.f, 0.f};
  f3 H = normalize(add(L, V));
  float ldh = dot(L, H);
  float fl = schlick(ndl), fv = schlick(ndv);
  float fd90 = 0.5f + 2.0f * ldh * ldh * m.roughness;
  float fd = mixf(1.0f, fd90, fl) * mixf(1.0f, fd90, fv);
  float fh = schlick(ldh);
  f3 fsheen = scl(ld3(m.csheen), fh * m.sheen);
  float fss90 = ldh * ldh * m.roughness;
  float fss = mixf(1.0f, fss90, fl) * mixf(1.0f, fss90, fv);
  float ss = 1.25f * (fss * (1.0f / (ndl + ndv) - 0.5f) + 0.5f);
  float k = kInvPi * mixf(fd, ss, m.subsurface);
  f3 dc = scl(add(scl(ld3(m.cdlin), k), fsheen), 1.0f - m.metallic);
  return scl(dc, ndl);
}

// materialBRDF::evaluate (src/material-wrapper.h:34-38)
__device__ __forceinline__ f3 brdf_eval(const DevMat &m, int type, f3 wo, f3 wi, f3 N) {
  if (!m.principled) return scl(scl(ld3(m.diffuse), kInvPi), fabsf(dot(wi, N)));
  if (type == kTypeDisneyRefl) return eval_spec(m, wo, N, wi);
  if (type == kTypeSpecRefl) return f3{1.f, 1.f, 1.f};
  if (type == kTypeSpecTrans) return scl(f3{1.f, 1.f, 1.f}, 1.0f / fabsf(dot(wi, N)));
  return eval_diff(m, wo, N, wi);
}

// sampleGTR2Aniso, src/principledMaterial.h:121-176 (tangent space)
__device__ __forceinline__ float sample_gtr2(const DevMat &mt, float u, float v, f3 V, f3 &L) {
  float s, c;
  sincos2pi(u, s, c);
  float fac = sqrtf(v / (1.0f - v));
  f3 m = normalize(f3{fac * (mt.ax * c), fac * (mt.ay * s), 1.0f});
  L = sub(V, scl(m, 2.0f * dot(V, m)));
  if (L.z < 0.0f) L = neg(L);
  float dlm = dot(L, m), dvm = dot(V, m);
  float tl = sqrtf(1.0f - dlm * dlm) / dlm;
  float tv = sqrtf(1.0f - dvm * dvm) / dvm;
  float xl = (dot(L, m) / L.z) > 0.0f ? 1.0f : 0.0f;
  float xv = (dot(V, m) / V.z) > 0.0f ? 1.0f : 0.0f;
  float gl = xl * 2.0f / (1.0f + sqrtf(1.0f + mt.roughg2 * tl * tl));
  float gv = xv * 2.0f / (1.0f + sqrtf(1.0f + mt.roughg2 * tv * tv));
  float gs = gl * gv;
  float lw = (fabsf(dot(L, m)) * gs) / (fabsf(L.z) * fabsf(m.z));
  if (lw < 0.01f) lw = 0.01f;
  return lw;
}

// sampleGlass, src/principledMaterial.h:86-118 (ior 1.9)
__device__ __forceinline__ int sample_glass(float u, f3 V, f3 &L) {
  const float ior = 1.9f;
  float c = V.z;
  float fr = schlick(c);
  if (u > fr) {
    L = neg(V);
    return kTypeSpecRefl;
  }
  float ti = 1.0f + ior * (c * c - 1.0f);
  if (ti <= 0.0f) {
    L = sub(V, scl(f3{0.f, 0.f, 1.f}, 2.0f * V.z));
    return kTypeSpecRefl;
  }
  float sg = V.z > 0.0f ? 1.0f : -1.0f;
  L = sub(scl(f3{0.f, 0.f, 1.f}, ior * c - sg * sqrtf(ti)), scl(V, ior));
  return kTypeSpecTrans;
}

// materialBRDF::sample (src/material-wrapper.h:30-33)
__device__ __forceinline__ int brdf_sample(const DevMat &m, f3 wo, f3 xi, f3 &wi, float &pdf) {
  if (!m.principled) {
    wi = cosine_sample(xi.x, xi.y);
    pdf = fabsf(wi.z) * kInvPi;
    return kTypeDiffuse;
  }
  if (m.opacity < 1.0f) {
    pdf = 1.0f;
    return sample_glass(xi.x, wo, wi);
  }
  if (xi.z < m.metallic) {
    pdf = sample_gtr2(m, xi.x, xi.y, wo, wi);
    if (pdf == 0.0f || isnan(pdf)) {
      wi = cosine_sample(xi.x, xi.y);
      pdf = (wi.z < 0.0f ? 0.0f : wi.z) / kPi;
      return kTypeDiffuse;
    }
    return kTypeDisneyRefl;
  }
  wi = cosine_sample(xi.x, xi.y);
  pdf = (wi.z < 0.0f ? 0.0f : wi.z) * kInvPi;
  return kTypeDiffuse;
}

}  // namespace dev
}  // namespace gi
