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
constexpr float kTwoPi = 6.283185307179586f;
constexpr float kPhongNorm = 16.07464925228143f;  // (100 + 1) / (2 pi): the Phong lobe's normalisation

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
// chunks holds the chunk tables from entry chunk_base of the pool on (0: the
// whole pool in global memory; an LDS copy of some dimensions' tables starts
// at the first one's offset)
struct HaltonDev {
  const HaltonDim *dims;
  const uint16_t *chunks;
  uint32_t chunk_base;
};
// Scrambled radical inverse of n in the prime base p of dimension dim: the
// digits d_k (K of them) map to perm[d_k] and the infinitely many leading
// zeros to perm[0], which is the exact rational
//   A / B,  A = rev*(p-1) + perm[0],  B = p^K (p-1),  rev = sum perm[d_k] p^(K-1-k).
// Digits are consumed m at a time through the chunk table (P = p^m <= 1024).
// Leading zero digits may be counted as digits: each adds perm[0] to rev and
// a factor p to p^K, which multiplies A and B by p and leaves A / B as it
// is.  So every n is split into the same kmax chunks of its dimension (the
// least K with P^K >= 2^32), zero chunks above its leading one included:
// A, B <= 2^42 p stay exact in FP64 (every partial value is an integer below
// 2^53), and the value is one correctly rounded FP64 division and one
// rounding to FP32, as if only n's own digits were used.  kmax is the same
// for every lane, so the chunks are taken with no per-lane selection.
// floor(m / P) is the top 32 bits of the 96-bit m * magic (one 32-bit
// mulhi and one 32x32+32 multiply-add: exact as the 64-bit mulhi); no
// 64-bit integer multiplies, the quarter-rate part of this function.
constexpr int kHaltonMaxChunks = 7;  // kmax <= 7: every dimension has P >= 37
__device__ __forceinline__ float halton(uint32_t n, int dim, HaltonDev H) {
  const HaltonDim hd = H.dims[dim];
  const uint16_t *chunk = H.chunks + (hd.chunk_off - H.chunk_base);
  const uint32_t mlo = (uint32_t)hd.magic, mhi = (uint32_t)(hd.magic >> 32);
  uint32_t c[kHaltonMaxChunks];
  uint32_t m = n;
#pragma unroll
  for (int k = 0; k < kHaltonMaxChunks; ++k) {
    // (m * mhi + hi32(m * mlo)) >> 32 = hi64(m * magic): the low word of
    // m * mlo cannot carry into the top 32 bits of the 96-bit product
    const uint32_t q = (uint32_t)(((uint64_t)m * mhi + __umulhi(m, mlo)) >> 32);
    c[k] = m - q * hd.P;
    m = q;
  }
  uint32_t v[kHaltonMaxChunks];
#pragma unroll
  for (int k = 0; k < kHaltonMaxChunks; ++k) v[k] = chunk[c[k]];
  const double Pd = (double)hd.P;
  double rev = 0.0, pk = 1.0;
#pragma unroll
  for (int k = 0; k < kHaltonMaxChunks; ++k) {
    if ((uint32_t)k < hd.kmax) {
      rev = fma(rev, Pd, (double)v[k]);
      pk = pk * Pd;
    }
  }
  const double p1 = (double)(hd.prime - 1);
  const double A = fma(rev, p1, (double)hd.perm0), B = pk * p1;
  const float f = (float)(A / B);
  return f >= 1.0f ? 0x1.fffffep-1f : f;
}

// The chunk tables of any kDimsPerVertex consecutive dimensions of a render
// of depth <= 32 fit this many entries (the largest such window, dimensions
// 221..227 with P = p from 1399 to 1439, holds 9959); the host checks it.
constexpr uint32_t kHaltonLdsEntries = 10240;
// Copies the chunk tables of dimensions [d0, d0 + nd) into lds and returns
// the sampler that reads them there: the same values, with LDS reads in
// place of the scattered 16-bit global loads.  Called by every thread of
// the block, before any of them samples.  (The copy is addressed from its
// own start, chunk_base = the first table's offset: a pointer moved below
// the array would leave it, and in a flat address the borrow is lost.)
__device__ __forceinline__ HaltonDev halton_in_lds(HaltonDev H, int d0, int nd, uint16_t *lds) {
  const HaltonDim first = H.dims[d0], last = H.dims[d0 + nd - 1];
  const uint32_t off0 = first.chunk_off - H.chunk_base;
  const uint32_t n = min(last.chunk_off + last.P - first.chunk_off, kHaltonLdsEntries);
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lds[i] = H.chunks[off0 + i];
  __syncthreads();
  return HaltonDev{H.dims, lds, first.chunk_off};
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
// Correctly rounded 1/x.  For |x| in [2^-126, 2^126): the hardware
// reciprocal (within 1 ulp) refined by one Newton-Raphson step with fused
// multiply-adds (Markstein, IBM J. Res. Dev. 34(1), 1990): 4 instructions
// instead of the 11 of the IEEE division sequence.  rcp_check.hip compares
// it bit for bit with the IEEE division over all 2^32 floats.  Elsewhere
// (zero, subnormals, |x| >= 2^126, infinities, NaN) the IEEE division
// itself.
__device__ __forceinline__ float rcp_rn(float x) {
  const float a = fabsf(x);
  if (__builtin_expect(a >= 0x1p-126f && a < 0x1p126f, 1)) {
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r, 1.0f), r, r);
  }
  return 1.0f / x;
}

__device__ __forceinline__ bool intersect_tri(f3 a, f3 e1, f3 e2, f3 o, f3 d, float &t, float &u,
                                              float &v) {
  // evaluated without early exits (one wave runs every lane's test anyway);
  // same operations and roundings as the early-exit form, so a hit has
  // bit-identical t, u, v
  const f3 pvec = cross(d, e2);
  const float det = dot(e1, pvec);
  const f3 tvec = sub(o, a);
  const f3 qvec = cross(tvec, e1);
  const float inv_det = rcp_rn(det);  // = 1.0f / det, bit for bit
  u = dot(tvec, pvec) * inv_det;
  v = dot(d, qvec) * inv_det;
  t = dot(e2, qvec) * inv_det;
  return (det != 0.0f) & (u >= 0.0f) & (u <= 1.0f) & (v >= 0.0f) & (u + v <= 1.0f) & (t > 0.0f);
}

// Barycentrics of the ray's crossing with the triangle's plane, without
// the inside tests (intersect_tri_opt_nocheck, as computeRayDiffs uses it,
// src/hybrid-pt.cpp:300-328): the arithmetic of intersect_tri.
__device__ __forceinline__ void plane_bary(f3 a, f3 e1, f3 e2, f3 o, f3 d, float &u, float &v) {
  const f3 pvec = cross(d, e2);
  const float det = dot(e1, pvec);
  const f3 tvec = sub(o, a);
  const f3 qvec = cross(tvec, e1);
  const float inv_det = 1.0f / det;
  u = dot(tvec, pvec) * inv_det;
  v = dot(d, qvec) * inv_det;
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
  int32_t texture;  // diffuse texture, -1: none
  float spec_tint, sheen_tint, pad;
};

// --------------------------------------------------------------- pow 2.2 --
// x^2.2 (mon2lin) for a base colour that varies per hit.  x = m 2^e with m
// in [sqrt(1/2), sqrt(2)); x^2.2 = m^2.2 * 2^k * 2^(j/5) with 11 e = 5 k + j.
// m^2.2 is a degree-6 polynomial in m - 1 (relative least squares, float
// coefficients); 2^(j/5) a float table.  Relative error < 3e-7.  Basic
// operations only, so host and device agree bit for bit.  x <= 0 or NaN ->
// 0; x < 2^-50 -> 0; x >= 2^50 -> infinity.
__host__ __device__ __forceinline__ float pow22(float x) {
  if (!(x >= 0x1p-50f)) return 0.0f;
  if (x >= 0x1p50f) return __builtin_inff();
  uint32_t bits;
  __builtin_memcpy(&bits, &x, 4);
  int32_t e = (int32_t)(bits >> 23) - 127;
  uint32_t mb = (bits & 0x7fffffu) | 0x3f800000u;
  if (mb >= 0x3fb504f3u) {  // m >= sqrt(2): halve it
    mb -= 0x00800000u;
    e += 1;
  }
  float m;
  __builtin_memcpy(&m, &mb, 4);
  const float t = m - 1.0f;
  float p = -0.002734926762059331f;
  p = p * t + 0.006638647988438606f;
  p = p * t + -0.017642781138420105f;
  p = p * t + 0.0879826545715332f;
  p = p * t + 1.3200013637542725f;
  p = p * t + 2.200000286102295f;
  p = p * t + 1.0f;
  const int32_t q = 11 * e;
  const int32_t j = ((q % 5) + 5) % 5;
  const int32_t k = (q - j) / 5;
  const float c = j == 0 ? 1.0f
                         : j == 1 ? 1.1486983549970351f
                                  : j == 2 ? 1.3195079107728942f : j == 3 ? 1.515716566510398f : 1.7411011265922482f;
  const uint32_t sb = (uint32_t)(k + 127) << 23;
  float scale;
  __builtin_memcpy(&scale, &sb, 4);
  return (c * p) * scale;
}

// ----------------------------------------------------------------- Phong --
// The powers of gpu_pt's Phong lobe (shininess 100, src/gpu-pt-kernels.cu:
// 121-125,164-187), from basic operations only, so that host and device
// agree bit for bit: x^100 as a fixed chain of squarings, and the lobe
// sampler's x^(1/101) as 2^(log2(x) / 101) with log2 from the series of
// atanh (m in [sqrt(1/2), sqrt(2)), to t^11) and 2^f from its Taylor
// polynomial (f in [0, 1), to f^9).  Relative error below 1e-7.
__host__ __device__ __forceinline__ float phong_pow100(float x) {
  const float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8, x32 = x16 * x16, x64 = x32 * x32;
  return (x64 * x32) * x4;
}
__host__ __device__ __forceinline__ float phong_root101(float x) {
  if (!(x > 0.0f)) return 0.0f;
  if (x >= 1.0f) return 1.0f;
  uint32_t bits;
  __builtin_memcpy(&bits, &x, 4);
  if (bits < 0x00800000u) return 0.4214777f;  // subnormal inputs: x^(1/101) of 2^-126
  int32_t e = (int32_t)(bits >> 23) - 127;
  uint32_t mb = (bits & 0x7fffffu) | 0x3f800000u;
  if (mb >= 0x3fb504f3u) {  // m >= sqrt(2): halve it
    mb -= 0x00800000u;
    e += 1;
  }
  float m;
  __builtin_memcpy(&m, &mb, 4);
  // log2(m) = 2/ln2 * atanh(t), t = (m - 1) / (m + 1)
  const float t = (m - 1.0f) / (m + 1.0f), t2 = t * t;
  float a = 0.09090909090909091f;
  a = a * t2 + 0.1111111111111111f;
  a = a * t2 + 0.14285714285714285f;
  a = a * t2 + 0.2f;
  a = a * t2 + 0.3333333333333333f;
  a = a * t2 + 1.0f;
  const float l2 = (float)e + (a * t) * 2.8853900817779268f;
  const float y = l2 / 101.0f;  // in (-1.26, 0)
  const float k = floorf(y), f = y - k;
  float p = 1.0178086009239699e-07f;
  p = p * f + 1.3215486790144307e-06f;
  p = p * f + 1.5252733804059841e-05f;
  p = p * f + 0.00015403530393381609f;
  p = p * f + 0.0013333558146428443f;
  p = p * f + 0.009618129107628477f;
  p = p * f + 0.05550410866482158f;
  p = p * f + 0.2402265069591007f;
  p = p * f + 0.6931471805599453f;
  p = p * f + 1.0f;
  const int32_t ki = (int32_t)k;  // -2 or -1
  const uint32_t sb = (uint32_t)(ki + 127) << 23;
  float scale;
  __builtin_memcpy(&scale, &sb, 4);
  return p * scale;
}

// --------------------------------------------------------------- texture --
// RGBA8 texture with its box-filtered mip chain: level 0 (w x h), then
// level l of ((W+1)/2) x ((H+1)/2) texels, `levels` levels after level 0
// (floor(log2(min(w, h))), texture_data, src/material.h:27-35).
struct DevTex {
  const uchar4 *texels;
  int32_t w, h, levels, pad;
};

// Mip level of a footprint: ceil(log2(diff * max(w, h))) clamped to
// [0, levels], taken as the number of i < levels with 2^i < diff*max(w,h)
// (the loop `for (i = 0; i < lod; ++i)` of sample_bilin_lod,
// src/material.h:88-99; compared exactly, without log2f's last-ulp rounding
// near powers of two); a NaN footprint takes the coarsest level, as
// fminf(NaN, max_mm) does there.
__host__ __device__ __forceinline__ int tex_level(float diff, int32_t w, int32_t h, int32_t levels) {
  const float x = diff * (float)(w > h ? w : h);
  if (x != x) return levels;
  int l = 0;
  float p = 1.0f;
  while (l < levels && p < x) {
    ++l;
    p = p * 2.0f;
  }
  return l;
}

// sample_bilin_lod (src/material.h:88-137): wrap (s, t) into [0, 1), pick
// the level, x = s W, y = (1 - t) H; blend the texel holding (x, y) with its
// neighbours on the side of the texel centre the point lies, weights
// |neighbour centre - point|.  Declared difference: addressing repeats
// (the reference leaves a neighbour index of exactly W or H, and y = H for
// t = 0, unwrapped, which reads the next row or the next level); texture
// coordinates that are NaN, infinite or at least 2^23 in magnitude wrap to 0.
// The four texels a lookup blends (indices into the chain) and their
// weights: d value / d texel.
struct TexFetch {
  uint32_t idx[4];
  float w[4];
};

// c / 255 for an 8-bit channel value c, correctly rounded: c * RN(1/255)
// corrected by one fma step (the residual c - 255 q0 is exact).  Equal to
// the IEEE division for all 256 values (checked exhaustively in exact
// rational arithmetic by tests/test_textures.py), at three instructions
// instead of a division's ten.
__device__ __forceinline__ float div255(float c) {
  const float r = 1.0f / 255.0f;
  const float q0 = c * r;
  return __builtin_fmaf(__builtin_fmaf(-255.0f, q0, c), r, q0);
}
// ((i % n) + n) % n for i in [-1, 2n) and n >= 1, by selects instead of
// integer divisions.
__device__ __forceinline__ int32_t wrap_index(int32_t i, int32_t n) {
  i = i < 0 ? i + n : i;
  i = i >= n ? i - n : i;
  return i >= n ? i - n : i;
}
__device__ __forceinline__ f3 tex_sample_lod(const DevTex &T, float s, float t, float diff, TexFetch &tf) {
  int32_t W = T.w, H = T.h;
  const int lv = tex_level(diff, T.w, T.h, T.levels);
  uint32_t off = 0;
  for (int i = 0; i < lv; ++i) {
    off += (uint32_t)W * (uint32_t)H;
    W = (W + 1) / 2;
    H = (H + 1) / 2;
  }
  if (!(fabsf(s) < 0x1p23f)) s = 0.0f;  // NaN, infinite or integral: wraps to 0
  if (!(fabsf(t) < 0x1p23f)) t = 0.0f;
  if (s < 0.0f) s = s + (-truncf(s) + 1.0f);
  if (s >= 1.0f) s = s - truncf(s);
  if (t < 0.0f) t = t + (-truncf(t) + 1.0f);
  if (t >= 1.0f) t = t - truncf(t);
  const float x = s * (float)W;
  const float y = (1.0f - t) * (float)H;
  int32_t nx = (int32_t)x, ny = (int32_t)y;
  int32_t ox = nx + 1, oy = ny + 1;
  if (x - floorf(x) < 0.5f) ox = nx - 1;
  if (y - floorf(y) < 0.5f) oy = ny - 1;
  const float wx = fabsf(((float)ox + 0.5f) - x);
  const float wy = fabsf(((float)oy + 0.5f) - y);
  // s, 1 - t in [0, 1] make x in [0, W], y in [0, H]: nx, ny in [0, W], [0, H]
  // and ox, oy in [-1, W + 1], [-1, H + 1], which wrap_index maps as the
  // modulo arithmetic of the reference's lookup does
  nx = wrap_index(nx, W);
  ny = wrap_index(ny, H);
  ox = wrap_index(ox, W);
  oy = wrap_index(oy, H);
  const uchar4 *L = T.texels + off;
  const uchar4 a00 = L[ny * W + nx], a01 = L[ny * W + ox], a10 = L[oy * W + nx], a11 = L[oy * W + ox];
  tf.idx[0] = off + (uint32_t)(ny * W + nx);
  tf.idx[1] = off + (uint32_t)(ny * W + ox);
  tf.idx[2] = off + (uint32_t)(oy * W + nx);
  tf.idx[3] = off + (uint32_t)(oy * W + ox);
  tf.w[0] = wy * wx;
  tf.w[1] = wy * (1.0f - wx);
  tf.w[2] = (1.0f - wy) * wx;
  tf.w[3] = (1.0f - wy) * (1.0f - wx);
  const f3 c00 = f3{div255((float)a00.x), div255((float)a00.y), div255((float)a00.z)};
  const f3 c01 = f3{div255((float)a01.x), div255((float)a01.y), div255((float)a01.z)};
  const f3 c10 = f3{div255((float)a10.x), div255((float)a10.y), div255((float)a10.z)};
  const f3 c11 = f3{div255((float)a11.x), div255((float)a11.y), div255((float)a11.z)};
  const float vx = 1.0f - wx, vy = 1.0f - wy;
  const f3 r0 = f3{wx * c00.x + vx * c01.x, wx * c00.y + vx * c01.y, wx * c00.z + vx * c01.z};
  const f3 r1 = f3{wx * c10.x + vx * c11.x, wx * c10.y + vx * c11.y, wx * c10.z + vx * c11.z};
  return f3{wy * r0.x + vy * r1.x, wy * r0.y + vy * r1.y, wy * r0.z + vy * r1.z};
}

// The material with its diffuse texture's value tx applied
// (material_t::diffuseColor, src/material.h:155-168): Lambert scales the
// diffuse colour; principled scales the base colour, and the quantities
// derived from it are recomputed as on the host (mon2lin by pow22).
__device__ __forceinline__ void apply_texel(DevMat &m, f3 tx) {
  if (!m.principled) {
    m.diffuse[0] = m.diffuse[0] * tx.x;
    m.diffuse[1] = m.diffuse[1] * tx.y;
    m.diffuse[2] = m.diffuse[2] * tx.z;
    return;
  }
  m.color[0] = m.color[0] * tx.x;
  m.color[1] = m.color[1] * tx.y;
  m.color[2] = m.color[2] * tx.z;
  float cd[3];
  for (int k = 0; k < 3; ++k) cd[k] = pow22(m.color[k]);
  const float lum = 0.3f * cd[0] + 0.6f * cd[1] + 0.1f * cd[2];
  const float sp = m.specular * 0.08f;
  for (int k = 0; k < 3; ++k) {
    const float tint = lum > 0.f ? cd[k] / lum : 1.f;
    const float mt = (1.0f - m.spec_tint) * 1.0f + m.spec_tint * tint;
    const float x = sp * mt;
    m.cspec0[k] = (1.0f - m.metallic) * x + m.metallic * cd[k];
    m.csheen[k] = (1.0f - m.sheen_tint) * 1.0f + m.sheen_tint * tint;
    m.cdlin[k] = cd[k];
  }
}

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
