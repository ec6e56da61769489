// gfx950 kernels of the wavefront path tracer and the C-ABI (include/gi_hip.h).
//
// Data layout in HBM (all arrays 16-byte aligned, structure of arrays):
//   WideNode   nodes[]      128 B per 4-wide node (one cache line)
//   HalfNode   hnodes[]     the traversal's copy: f16 child planes, 80 of 128 B
//   IsectTri   itris[]      48 B per triangle, BVH leaf order
//   ShadeTri   stris[]      128 B per triangle, original order (shading)
//   ray queue  ray_o[], ray_d[], tp[]  float4 each: (o, halton index),
//              (d, path slot), (throughput, 0); compacted per bounce
//   hits[]     float4 (t, beta, gamma, triangle id) per queue entry
//   shadow     sh_o[] (o, max_t), sh_d[] (d, slot), sh_c[] (contribution)
//   path_l[]   float4 per path slot: the path's accumulated colour
//   images     float4 per pixel per view: running mean over samples
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "bvh_gpu.h"
#include "gi_device.h"
#include "gi_internal.h"

using namespace gi;
using namespace gi::dev;

namespace {

// ------------------------------------------------------------- records ----
struct alignas(16) ShadeTri {
  float4 a, b, c;     // a.w = material index (bits)
  float4 na, nb, nc;
  float4 tab;         // ta.xy, tb.xy
  float4 tcm;         // tc.xy, pad
};
static_assert(sizeof(ShadeTri) == 128, "shade tri");

struct alignas(16) DevView {
  float U[3], V[3], W[3], pos[3], vd[3];
  float width, height, fw, fh, aperture, focus;
  int32_t w, h, spp;
  uint32_t index_base;
  float4 *image;
  uint32_t pix_base;  // offset of the view's pixels in views-concatenated images
  int32_t centre;     // 1: primary rays through the pixel centres (GI_FLAG_PIXEL_CENTRE)
};

constexpr int kMaxChunks = 64;
// A pass holds chunks of path slots: samples s0 .. s0+k-1 of every pixel of
// one view, pixel-major (slot_base + p * k + j is sample s0 + j of pixel p).
struct PassChunk {
  int32_t view, s0, k, npix;
  uint32_t slot_base, pix_base;
  int32_t pad[2];
};
struct PassDesc {
  int32_t nchunks;
  uint32_t total_slots;
  uint32_t total_pix;
  int32_t pad;
  PassChunk chunk[kMaxChunks];
};

struct alignas(16) DevLight {
  float center[3], power;
  float dir[3], w;
  float right[3], h;
  float up[3];
  int32_t type;  // GI_LIGHT_RECT / GI_LIGHT_SKY (scale in col[0])
  float col[3], pad2;
};

// A placed mesh: world-to-object and object-to-world rows (3x4), the root
// of the mesh's BVH, the virtual id of its triangle 0, the ShadeTri index of
// its triangle 0 and the material override (-1 = the triangles' own).
struct alignas(16) DevInstance {
  float4 inv[3];
  float4 fwd[3];
  uint32_t root, vbase, stri;
  int32_t material;
  uint32_t pad[4];
};
static_assert(sizeof(DevInstance) == 128, "instance record");

// The instance tree: root of the tree over the instances' world boxes and
// the pieces of the world tree's top (leaves: instance refs or references
// into the world tree), the instances, and the number of world triangles
// (ids below it are world triangles).
struct InstScene {
  const DevInstance *inst;
  const uint32_t *vbase;  // inst[i].vbase, packed for the search by id
  uint32_t tlas_root, n_world;
  int32_t n_inst;
  // vtab[b]: the instance holding virtual id n_world + (b << vshift); a
  // block of 2^vshift ids is no longer than the smallest instanced mesh, so
  // at most the next instance also starts inside it.  Null: search vbase.
  const uint32_t *vtab;
  uint32_t vshift;
};

// Sky map of the scene's sky light and the tables of its Distribution2D.
struct SkyDev {
  const float4 *rgb;                  // w*h texels (r, g, b, 0)
  const float *ycdf, *ypdf;           // h+1, h
  const float *xcdf, *xpdf;           // h*(w+1), h*w
  int32_t w, h, light;                // light: index of the sky light, -1 = none
  float scale;
};

struct ShadeParams {
  const ShadeTri *tris;
  const DevMat *mats;
  const DevLight *lights;
  int32_t nlights;
  float power_total;
  HaltonDev H;
  float off_b, off_s;
  float bg[3];
  int32_t depth, rr_start;
  SkyDev sky;
  InstScene is;
  const DevTex *tex;       // textures (diffuse, by DevMat::texture)
  float4 *rdiff;           // ray differentials per path slot, or null (no texture in use):
  uint32_t rdiff_stride;   // planes [0..3] = upper origin, upper dir, right origin, right dir
  const float4 *ks;        // per material: specular colour (GI_FLAG_PHONG renders)
  const float *lsum;       // running sums of the light powers: lsum[k] = power of lights 0..k-1
  const PassDesc *pass;    // the pass being shaded and the views: the first vertex's ray
  const DevView *views;    // differentials are the camera's, computed from the slot's pixel
};

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }

// Append of two predicates per thread across the whole block: one atomic
// per counter per block (device-wide atomics on one address serialise, so
// they are kept few).  Called uniformly by every thread of the block;
// blockDim <= 256.
__device__ __forceinline__ void block_append2(bool p0, uint32_t *c0, bool p1, uint32_t *c1, uint32_t &q0,
                                              uint32_t &q1) {
  __shared__ uint32_t s_n[2][4];
  __shared__ uint32_t s_base[2];
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint64_t m0 = __ballot(p0), m1 = __ballot(p1);
  const uint64_t below = (1ull << lane) - 1ull;
  if (lane == 0) {
    s_n[0][wave] = (uint32_t)__popcll(m0);
    s_n[1][wave] = (uint32_t)__popcll(m1);
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const uint32_t k = threadIdx.x, nw = blockDim.x >> 6;
    uint32_t tot = 0;
    for (uint32_t w = 0; w < nw; ++w) tot += s_n[k][w];
    s_base[k] = tot ? atomicAdd(k ? c1 : c0, tot) : 0u;
  }
  __syncthreads();
  uint32_t o0 = s_base[0], o1 = s_base[1];
  for (uint32_t w = 0; w < wave; ++w) {
    o0 += s_n[0][w];
    o1 += s_n[1][w];
  }
  q0 = o0 + (uint32_t)__popcll(m0 & below);
  q1 = o1 + (uint32_t)__popcll(m1 & below);
  __syncthreads();  // s_n / s_base are rewritten by the next call
}

// -------------------------------------------------------------- traversal --
// Ordered 4-wide BVH traversal, one ray per lane, persistent.  The queue is
// cut into spans of S consecutive rays; wave w of W owns spans w, w + W,
// w + 2W, ... (no atomics) and takes their rays in order.  A wave keeps its
// lanes busy: whenever at least C of them have finished their rays (C, the
// refill threshold, is a tuning parameter: 8..64), the idle lanes take the
// wave's next rays, one each.  S is the largest power of two from C to 256
// that still gives every wave 16 spans: successive refills of a wave take
// neighbouring rays, whose traversals are more alike than those of chunks
// W spans apart, and the last spans still balance the waves' work.  Interior nodes are processed in an inner
// loop in which a lane that reaches a leaf parks it and goes on with its
// stack (speculative traversal, Aila & Laine 2009), until every lane with
// interior work left holds a parked leaf; then the parked leaves are
// tested, and the wave refills.
//
// Stack: kLdsDepth entries per lane in LDS, interleaved by lane (entry k of
// lane l at (k << shift) + l, with 1 << shift >= blockDim: conflict free),
// continued in a per-thread global column.  The capacity covers 3 entries
// per level of the tree; an overflow (impossible by construction) is
// counted and turns into GI_ERR_LIMIT on the host.
//
// Box test on the half-precision nodes (HalfNode): per node and axis
// b = fma(org, 1/d, -o/d) and s' = scale/d (exact: a power of two), then
// each child plane's distance is fma(q, s', b) with the f16 offset q
// (v_fma_mix_f32: one rounding); entry <= exit * (1 + 2^-21).  The relative
// slack covers the rounding of the reciprocal, of o/d and of the fmas
// relative to the distance.  The parts proportional to coordinates are
// covered by padding: the f32 boxes are padded by 2^-20 of their largest
// coordinate magnitude, and the f16 planes are rounded outward past a
// further 2^-18 of the node's coordinate magnitude, which covers the
// rounding of b (2^-24 of |org - o| / |d|).  Box culling is therefore
// conservative and the result is the closest hit = minimum (t, id) over
// all triangles with 0 < t <= max_t; any-hit stops at the first such
// triangle.
// Occupancy target of the traversal kernels (waves per SIMD): 6 caps them
// at 80 VGPRs, which they fit without spilling (the default launch, 6 blocks
// of 4 waves per CU, runs 6 per SIMD).  At 7 (72 VGPRs) the two-level
// kernels spill.
#ifndef GI_TRACE_WAVES_PER_EU
#define GI_TRACE_WAVES_PER_EU 6
#endif

constexpr uint32_t kNoRay = 0xffffffffu;
// The traversal addresses a node's planes by a 32-bit byte offset (index
// << 7): at most 2^25 wide nodes.
constexpr uint64_t kMaxWideNodes = (uint64_t)1 << 25;
constexpr float kSlabSlack = 1.0f + 0x1p-21f;
// The interior-node loop hands over to the leaf tests once at most this
// many lanes of the wave are still looking for their first leaf (and some
// lane holds one): the last searchers would otherwise keep the whole wave
// in the loop at a few active lanes while the others' parked leaves wait.
#ifndef GI_SPEC_LANES
#define GI_SPEC_LANES 12
#endif
constexpr uint32_t kSpecLanes = GI_SPEC_LANES;

typedef __attribute__((address_space(3))) uint32_t lds_u32;

// Largest scene (wide nodes + intersection triangles) that the traversal
// kernels copy into each block's LDS: nodes and triangles are then read
// with ds_read_b128 instead of from L2.  10 KB keeps 6 blocks of 256 lanes
// (16 KB of stack each) within a CU's 160 KB.
constexpr size_t kLdsSceneMaxBytes = 10240;

template <int kLdsDepth>
struct Stack {
  lds_u32 *lds;    // this lane's LDS column: entry k at lds[k << shift]
  uint32_t *glb;   // this thread's global column: entry k at glb[k * glb_stride]
  uint32_t shift, glb_stride;
  int cap, sp;
  __device__ __forceinline__ void put(int pos, uint32_t v) {
    if (pos < kLdsDepth) lds[(uint32_t)pos << shift] = v;
    else glb[(uint32_t)(pos - kLdsDepth) * glb_stride] = v;
  }
  __device__ __forceinline__ uint32_t take() {
    --sp;
    return sp < kLdsDepth ? lds[(uint32_t)sp << shift] : glb[(uint32_t)(sp - kLdsDepth) * glb_stride];
  }
};

// Sign bits of the inverse direction (bit 0 x, 1 y, 2 z): on an axis with
// a positive inverse the near plane is lo and the far plane hi, with a
// negative one the other way round (fma rounding is monotone in the plane,
// so these are exactly the min and max of the two plane distances).
// rcp_clamped never returns zero, and -0 directions give negative inverses.
__device__ __forceinline__ uint32_t dir_signs(f3 idir) {
  return (__float_as_uint(idir.x) >> 31) | ((__float_as_uint(idir.y) >> 31) << 1) |
         ((__float_as_uint(idir.z) >> 31) << 2);
}

// v_fma_mix_f32: the f16 in the low (mix_lo) or high (mix_hi) half of h2,
// converted exactly, times s plus b, rounded once in f32.
__device__ __forceinline__ float mix_lo(uint32_t h2, float s, float b) {
  float r;
  asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h2), "v"(s), "v"(b));
  return r;
}
__device__ __forceinline__ float mix_hi(uint32_t h2, float s, float b) {
  float r;
  asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h2), "v"(s), "v"(b));
  return r;
}

// v_max3_f32 / v_min3_f32 on plane distances (never NaN: an empty slot's
// +inf offset times a finite nonzero s' gives an infinity), without the
// quieting the compiler adds to values it did not produce itself.
__device__ __forceinline__ float max3_raw(float a, float b, float c) {
  float r;
  asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
__device__ __forceinline__ float min3_raw(float a, float b, float c) {
  float r;
  asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// LDS column stride of a block: the next power of two >= blockDim.
__host__ __device__ __forceinline__ uint32_t lds_shift_for(uint32_t block) {
  uint32_t s = 0;
  while ((1u << s) < block) ++s;
  return s;
}

template <int kLdsDepth>
__device__ __forceinline__ Stack<kLdsDepth> make_stack(uint32_t *smem, uint32_t *gstack, uint32_t cap) {
  const uint32_t shift = blockDim.x <= 1 ? 0u : 32u - (uint32_t)__clz((int)(blockDim.x - 1));
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  return Stack<kLdsDepth>{(lds_u32 *)(smem + threadIdx.x), gstack + gtid, shift, gridDim.x * blockDim.x,
                          (int)cap, 0};
}

// The persistent traversal loop over queue entries [0, count).
//   start(i, o, d, tmax) loads ray i;
//   finish(i, t, u, v, id) stores its result (id == GI_INVALID_REF: no hit).
// kInst: two-level traversal from the instance tree's root.  An instance
// reference is parked like a leaf and the lane goes on with its stack;
// taking it up pushes the lane's current reference and kReturnRef, moves
// the ray into the mesh's object space and continues at the mesh's root.
// The return marker is parked too, but the lane holds still (the ray's
// space changes when it is taken up): that reloads the world ray and pops.
// Triangle ids found in a mesh are offset by the instance's virtual base.
__device__ __forceinline__ bool special_ref(uint32_t ref) {
  return (ref & kLeafBit) && ((ref >> 27) & 15u) == 0u && ref != kDoneRef && ref != kEmptyRef;
}

// kCount: the lane also counts its rays, interior-node visits, triangle
// tests and instance entries, and the wave adds them to stats[0..3] at the
// end (traversal statistics; a separate instantiation).
// cs: log2 of the refill threshold (8..64 idle lanes).
// kLdsScene: the block's LDS holds a copy of the f32 wide nodes (8 float4 each) at
// lnodes and of the triangles at ltris; they are read from there.
template <bool kAny, bool kInst, int kLdsDepth, bool kCount, bool kLdsScene, class Start, class Finish>
__device__ __forceinline__ void trace_persistent(uint32_t count, uint32_t cs, const HalfNode *__restrict__ nodes,
                                                 const float4 *__restrict__ itris, const float4 *lnodes,
                                                 const float4 *ltris, Stack<kLdsDepth> &st, uint32_t *overflow,
                                                 InstScene IS, unsigned long long *stats, Start start,
                                                 Finish finish) {
  uint32_t c_rays = 0, c_nodes = 0, c_tris = 0, c_inst = 0;
  const uint32_t lane = __lane_id();
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const uint32_t chunk_n = 1u << cs;  // idle lanes a refill needs
  // span length 2^ss (wave-uniform): 16 spans per wave at least, 256 rays at most
  uint32_t ss = cs;
  while (ss < 8u && (((uint64_t)waves * 16u) << (ss + 1)) <= count) ++ss;
  // the start and the end of span k, both clamped to count
  const auto span_range = [&](uint32_t k, uint32_t &b, uint32_t &e) {
    const uint64_t b64 = (uint64_t)k << ss;
    b = b64 >= count ? count : (uint32_t)b64;
    e = min(b + (1u << ss), count);
  };
  uint32_t span = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);  // wave-uniform: current span
  uint32_t cur, end;  // the wave's next ray, the end of its current span
  span_range(span, cur, end);
  uint32_t ray = kNoRay, ref = kDoneRef, parked = kDoneRef;
  f3 o{0.f, 0.f, 0.f}, d{0.f, 0.f, 0.f}, idir{0.f, 0.f, 0.f}, oid{0.f, 0.f, 0.f};
  uint32_t neg = 0;  // dir_signs of the ray
  float tmax = 0.f, bt = FLT_MAX, bu = 0.f, bv = 0.f;
  uint32_t bid = GI_INVALID_REF;
  uint32_t vbase = 0;  // virtual id base of the current space (0: world)
  for (;;) {
    // ---- refill: the idle lanes, in lane order, take the wave's next rays:
    // the rest of its current span, then the start of its next one
    const uint64_t idle = __ballot(ray == kNoRay);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (n_idle >= chunk_n && cur < count) {
      const uint32_t rem = end - cur;  // >= 1
      uint32_t nb, ne;
      span_range(span + waves, nb, ne);
      const uint32_t n_take = min(n_idle, rem + (ne - nb));
      if (ray == kNoRay) {
        const uint32_t r = (uint32_t)__popcll(idle & below);
        if (r < n_take) {
          ray = r < rem ? cur + r : nb + (r - rem);
          if constexpr (kCount) ++c_rays;
          start(ray, o, d, tmax);
          idir = f3{rcp_clamped(d.x), rcp_clamped(d.y), rcp_clamped(d.z)};
          oid = f3{o.x * idir.x, o.y * idir.y, o.z * idir.z};
          neg = dir_signs(idir);
          bt = FLT_MAX;
          bu = bv = 0.0f;
          bid = GI_INVALID_REF;
          ref = kInst ? IS.tlas_root : 0u;
          vbase = 0;
          st.sp = 0;
        }
      }
      if (n_take < rem) {
        cur += n_take;
      } else {
        span += waves;
        cur = nb + (n_take - rem);
        end = ne;
        if (cur >= end) {
          span += waves;
          span_range(span, cur, end);
        }
      }
    }
    if (__ballot(ray != kNoRay) == 0) {
      if (cur >= count) break;
      continue;
    }
    if (ray == kNoRay) continue;
    // ---- interior nodes.  A lane that reaches a leaf parks it (one slot)
    // and goes on with its stack, so it keeps doing interior work while the
    // other lanes look for their leaves; the loop ends once every lane with
    // interior work left has a parked leaf (or none has interior work), or
    // once at most kSpecLanes lanes still look for one.
    for (;;) {
      if (!(ref & kLeafBit)) {
        if constexpr (kCount) ++c_nodes;
        const float lim = kAny ? tmax : smin(bt, tmax);
        float dist[4];
        uint32_t cref[4];
        const uint32_t base = ref << 7;
        const bool gx = neg & 1u, gy = neg & 2u, gz = neg & 4u;
        if constexpr (kLdsScene) {
          // the small-scene copy in LDS holds the f32 wide nodes: LDS reads
          // are cheap, so only the near and the far plane of each axis are
          // read (at offsets by the direction signs), with no conversion
          const char *np = reinterpret_cast<const char *>(lnodes);  // in LDS: derived from the shared base
          const uint32_t onx = base | (gx ? 16u : 0u), ony = base | (gy ? 48u : 32u), onz = base | (gz ? 80u : 64u);
          const float4 nxv = *reinterpret_cast<const float4 *>(np + onx);
          const float4 fxv = *reinterpret_cast<const float4 *>(np + (onx ^ 16u));
          const float4 nyv = *reinterpret_cast<const float4 *>(np + ony);
          const float4 fyv = *reinterpret_cast<const float4 *>(np + (ony ^ 16u));
          const float4 nzv = *reinterpret_cast<const float4 *>(np + onz);
          const float4 fzv = *reinterpret_cast<const float4 *>(np + (onz ^ 16u));
          const float4 c = *reinterpret_cast<const float4 *>(np + (base | 96u));
          cref[0] = __float_as_uint(c.x);
          cref[1] = __float_as_uint(c.y);
          cref[2] = __float_as_uint(c.z);
          cref[3] = __float_as_uint(c.w);
          const float nx[4] = {nxv.x, nxv.y, nxv.z, nxv.w}, fx[4] = {fxv.x, fxv.y, fxv.z, fxv.w};
          const float ny[4] = {nyv.x, nyv.y, nyv.z, nyv.w}, fy[4] = {fyv.x, fyv.y, fyv.z, fyv.w};
          const float nz[4] = {nzv.x, nzv.y, nzv.z, nzv.w}, fz[4] = {fzv.x, fzv.y, fzv.z, fzv.w};
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const float tn = fmaxf(fmaxf(fmaf(nx[k], idir.x, -oid.x), fmaf(ny[k], idir.y, -oid.y)),
                                   fmaxf(fmaf(nz[k], idir.z, -oid.z), 0.0f));
            const float tf = fminf(fminf(fmaf(fx[k], idir.x, -oid.x), fmaf(fy[k], idir.y, -oid.y)),
                                   fminf(fmaf(fz[k], idir.z, -oid.z), lim));
            dist[k] = tn <= tf * kSlabSlack ? tn : INFINITY;  // empty slots: +inf boxes, never hit
          }
        } else {
          // origin and scale, the f16 planes of each axis, the references
          const char *np = reinterpret_cast<const char *>(nodes);
          const float4 org = *reinterpret_cast<const float4 *>(np + base);
          const uint4 qx = *reinterpret_cast<const uint4 *>(np + (base | 16u));
          const uint4 qy = *reinterpret_cast<const uint4 *>(np + (base | 32u));
          const uint4 qz = *reinterpret_cast<const uint4 *>(np + (base | 48u));
          const uint4 ch = *reinterpret_cast<const uint4 *>(np + (base | 64u));
          cref[0] = ch.x;
          cref[1] = ch.y;
          cref[2] = ch.z;
          cref[3] = ch.w;
          const float sx = idir.x * org.w, sy = idir.y * org.w, sz = idir.z * org.w;
          const float bx = fmaf(org.x, idir.x, -oid.x), by = fmaf(org.y, idir.y, -oid.y),
                      bz = fmaf(org.z, idir.z, -oid.z);
          // near / far plane pairs (children 0|1 and 2|3) of each axis
          const uint32_t nx0 = gx ? qx.z : qx.x, nx1 = gx ? qx.w : qx.y, fx0 = gx ? qx.x : qx.z,
                         fx1 = gx ? qx.y : qx.w;
          const uint32_t ny0 = gy ? qy.z : qy.x, ny1 = gy ? qy.w : qy.y, fy0 = gy ? qy.x : qy.z,
                         fy1 = gy ? qy.y : qy.w;
          const uint32_t nz0 = gz ? qz.z : qz.x, nz1 = gz ? qz.w : qz.y, fz0 = gz ? qz.x : qz.z,
                         fz1 = gz ? qz.y : qz.w;
#define GI_SLAB(k, NX, NY, NZ, FX, FY, FZ, MIX)                                                          \
  {                                                                                                  \
    const float tn = max3_raw(MIX(NX, sx, bx), MIX(NY, sy, by), max3_raw(MIX(NZ, sz, bz), 0.0f, 0.0f)); \
    const float tf = min3_raw(MIX(FX, sx, bx), MIX(FY, sy, by), min3_raw(MIX(FZ, sz, bz), lim, lim));  \
    dist[k] = tn <= tf * kSlabSlack ? tn : INFINITY; /* empty slots: +inf boxes, never hit */       \
  }
          GI_SLAB(0, nx0, ny0, nz0, fx0, fy0, fz0, mix_lo)
          GI_SLAB(1, nx0, ny0, nz0, fx0, fy0, fz0, mix_hi)
          GI_SLAB(2, nx1, ny1, nz1, fx1, fy1, fz1, mix_lo)
          GI_SLAB(3, nx1, ny1, nz1, fx1, fy1, fz1, mix_hi)
#undef GI_SLAB
        }
        // the child references are needed only after the sort; consuming
        // them here (an empty register-only asm) keeps their load issued
        // with the planes' instead of sunk behind the sort, where it would
        // add one memory round trip to every node visit
        asm volatile("" : "+v"(cref[0]), "+v"(cref[1]), "+v"(cref[2]), "+v"(cref[3]));
        // sort (dist, ref) ascending: network (0,1)(2,3)(0,2)(1,3)(1,2)
#define GI_CSWAP(i, j)                                \
  {                                                 \
    const bool sw = dist[j] < dist[i];              \
    const float di = dist[i], dj = dist[j];         \
    const uint32_t ri = cref[i], rj = cref[j];      \
    dist[i] = sw ? dj : di;                         \
    dist[j] = sw ? di : dj;                         \
    cref[i] = sw ? rj : ri;                         \
    cref[j] = sw ? ri : rj;                         \
  }
        GI_CSWAP(0, 1) GI_CSWAP(2, 3) GI_CSWAP(0, 2) GI_CSWAP(1, 3) GI_CSWAP(1, 2)
#undef GI_CSWAP
        if (dist[0] != INFINITY) {
          // the hit children beyond the nearest, cref[1..nh], go on the stack
          // farthest first, so the nearer ones are popped first
          const int nh = (dist[1] != INFINITY ? 1 : 0) + (dist[2] != INFINITY ? 1 : 0) +
                         (dist[3] != INFINITY ? 1 : 0);
          const int sp = st.sp;
          if (sp + nh > st.cap) {
            atomicAdd(overflow, 1u);
          } else {
            if (nh >= 1) st.put(sp + nh - 1, cref[1]);
            if (nh >= 2) st.put(sp + nh - 2, cref[2]);
            if (nh >= 3) st.put(sp, cref[3]);
            st.sp = sp + nh;
          }
          ref = cref[0];
        } else {
          ref = st.sp > 0 ? st.take() : kDoneRef;
        }
      }
      if (parked == kDoneRef && ref != kDoneRef && (ref & kLeafBit)) {
        parked = ref;
        // a leaf or an instance lets the lane go on with its stack; the
        // return marker changes the ray's space, so the lane holds still
        ref = (kInst && ref == kReturnRef) ? kDoneRef : (st.sp > 0 ? st.take() : kDoneRef);
      }
      // lanes still looking for a leaf: the loop runs while some do, until
      // few are left and some lane has a parked leaf to test (the searchers
      // go on from their stacks after the leaf tests)
      const uint64_t searching = __ballot(!(ref & kLeafBit) && parked == kDoneRef);
      if (searching == 0 ||
          ((uint32_t)__popcll(searching) <= kSpecLanes && __ballot(parked != kDoneRef) != 0))
        break;
    }
    // ---- the parked leaf, instance or return marker
    if (kInst && special_ref(parked)) {
      if (parked == kReturnRef) {
        start(ray, o, d, tmax);  // back to the world ray
        vbase = 0;
        ref = st.sp > 0 ? st.take() : kDoneRef;
      } else {
        // the world reference the lane went on to waits under the marker
        const DevInstance &I = IS.inst[(parked & kMaxLeafFirst) - 1u];
        const int sp = st.sp, keep = ref != kDoneRef ? 1 : 0;
        if (sp + keep + 1 > st.cap) {
          atomicAdd(overflow, 1u);
          if (!keep) ref = st.sp > 0 ? st.take() : kDoneRef;
        } else {
          if constexpr (kCount) ++c_inst;
          if (keep) st.put(sp, ref);
          st.put(sp + keep, kReturnRef);
          st.sp = sp + keep + 1;
          const float4 r0 = I.inv[0], r1 = I.inv[1], r2 = I.inv[2];
          const f3 oo = f3{((r0.x * o.x + r0.y * o.y) + r0.z * o.z) + r0.w,
                           ((r1.x * o.x + r1.y * o.y) + r1.z * o.z) + r1.w,
                           ((r2.x * o.x + r2.y * o.y) + r2.z * o.z) + r2.w};
          const f3 dd = f3{(r0.x * d.x + r0.y * d.y) + r0.z * d.z, (r1.x * d.x + r1.y * d.y) + r1.z * d.z,
                           (r2.x * d.x + r2.y * d.y) + r2.z * d.z};
          o = oo;
          d = dd;
          vbase = I.vbase;
          ref = I.root;
        }
      }
      idir = f3{rcp_clamped(d.x), rcp_clamped(d.y), rcp_clamped(d.z)};
      oid = f3{o.x * idir.x, o.y * idir.y, o.z * idir.z};
      neg = dir_signs(idir);
      parked = kDoneRef;
    } else if (parked != kDoneRef) {
      // two triangle buffers, refilled alternately: triangle j+2 is loaded
      // into j's registers once j is tested, so no register copies
      const uint32_t cnt = (parked >> 27) & 15u;
      const auto tp = [&]() {
        if constexpr (kLdsScene) return ltris + 3 * (parked & kMaxLeafFirst);
        else return itris + 3 * (parked & kMaxLeafFirst);
      }();
      float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), b0 = a0, c0 = a0, a1 = a0, b1 = a0, c1 = a0;
      if (cnt > 0) {
        a0 = tp[0];
        b0 = tp[1];
        c0 = tp[2];
      }
      if (cnt > 1) {
        a1 = tp[3];
        b1 = tp[4];
        c1 = tp[5];
      }
      // tests one triangle; true: an any-hit ray is done
      const auto test = [&](const float4 &ta, const float4 &t1, const float4 &t2) -> bool {
        float t, u, v;
        if constexpr (kCount) ++c_tris;
        if (intersect_tri(xyz(ta), xyz(t1), xyz(t2), o, d, t, u, v) && t <= tmax) {
          const uint32_t id = fbits(ta.w) + vbase;
          if (kAny) {
            bt = t; bu = u; bv = v; bid = id;
            ref = kDoneRef;  // any hit ends the ray
            return true;
          }
          if (t < bt || (t == bt && id < bid)) {
            bt = t; bu = u; bv = v; bid = id;
          }
        }
        return false;
      };
      for (uint32_t j = 0; j < cnt; j += 2) {
        if (test(a0, b0, c0)) break;
        if (j + 2 < cnt) {
          a0 = tp[3 * j + 6];
          b0 = tp[3 * j + 7];
          c0 = tp[3 * j + 8];
        }
        if (j + 1 >= cnt) break;
        if (test(a1, b1, c1)) break;
        if (j + 3 < cnt) {
          a1 = tp[3 * j + 9];
          b1 = tp[3 * j + 10];
          c1 = tp[3 * j + 11];
        }
      }
      parked = kDoneRef;
    }
    if (ref == kDoneRef) {
      finish(ray, bt, bu, bv, bid);
      ray = kNoRay;
    }
  }
  if constexpr (kCount) {
    // every lane of the wave leaves the loop together
    uint32_t c[4] = {c_rays, c_nodes, c_tris, c_inst};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      for (int off = 32; off > 0; off >>= 1) c[k] += (uint32_t)__shfl_xor((int)c[k], off);
      if (lane == 0 && c[k]) atomicAdd(stats + k, (unsigned long long)c[k]);
    }
  }
}

constexpr int kMinLdsDepth = 8;  // smallest instantiated LDS stack; sizes the global part

// Closest-hit kernel over a ray queue.  kHasTmax: max_t in ray_o.w (the
// primitive entry point); otherwise ray_o.w carries the Halton index and
// max_t is FLT_MAX (camera and bounce rays, src/hybrid-pt.cpp:558).
// kLdsScene: copy the scene (n_nodes wide nodes, n_tris triangles) into
// LDS after the stack (stack_words 32-bit words), then traverse it there.
template <bool kLdsScene>
__device__ __forceinline__ void stage_scene(uint32_t *smem, uint32_t stack_words, const WideNode *nodes,
                                            uint32_t n_nodes, const float4 *itris, uint32_t n_tris,
                                            const float4 *&lnodes, const float4 *&ltris) {
  if constexpr (kLdsScene) {
    float4 *dst = reinterpret_cast<float4 *>(smem + stack_words);
    const float4 *srcn = reinterpret_cast<const float4 *>(nodes);
    for (uint32_t i = threadIdx.x; i < 8u * n_nodes; i += blockDim.x) dst[i] = srcn[i];
    for (uint32_t i = threadIdx.x; i < 3u * n_tris; i += blockDim.x) dst[8u * n_nodes + i] = itris[i];
    __syncthreads();
    lnodes = dst;
    ltris = dst + 8u * n_nodes;
  }
}

template <bool kHasTmax, bool kInst, int kLdsDepth, bool kCount, bool kLdsScene>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(GI_TRACE_WAVES_PER_EU))) void k_trace_closest(
    const uint32_t *__restrict__ count_ptr, uint32_t count_host, const float4 *__restrict__ ray_o,
    const float4 *__restrict__ ray_d, float4 *__restrict__ hits, const HalfNode *__restrict__ nodes,
    const WideNode *__restrict__ wnodes, const float4 *__restrict__ itris, uint32_t *__restrict__ gstack, uint32_t stack_cap, uint32_t *overflow,
    InstScene IS, unsigned long long *stats, uint32_t chunk_shift, uint32_t stack_words, uint32_t n_nodes,
    uint32_t n_tris) {
  extern __shared__ uint32_t smem[];
  const uint32_t count = count_ptr ? *count_ptr : count_host;
  Stack<kLdsDepth> st = make_stack<kLdsDepth>(smem, gstack, stack_cap);
  const float4 *lnodes = nullptr, *ltris = nullptr;
  stage_scene<kLdsScene>(smem, stack_words, wnodes, n_nodes, itris, n_tris, lnodes, ltris);
  trace_persistent<false, kInst, kLdsDepth, kCount, kLdsScene>(
      count, chunk_shift, nodes, itris, lnodes, ltris, st, overflow, IS, stats,
      [&](uint32_t i, f3 &o, f3 &d, float &tmax) {
        const float4 ro = ray_o[i], rd = ray_d[i];
        o = xyz(ro);
        d = xyz(rd);
        tmax = kHasTmax ? ro.w : FLT_MAX;
      },
      [&](uint32_t i, float t, float u, float v, uint32_t id) { hits[i] = make_float4(t, u, v, bitsf(id)); });
}

// Any-hit kernel.  kDeposit: shadow rays of the renderer -- an unoccluded
// ray adds its precomputed NEE contribution to its path's colour (the
// !is.valid() branch of src/hybrid-pt.cpp:355-365).  Otherwise writes the
// occlusion byte (primitive entry point).
template <bool kDeposit, bool kInst, int kLdsDepth, bool kCount, bool kLdsScene>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(GI_TRACE_WAVES_PER_EU))) void k_trace_any(
    const uint32_t *__restrict__ count_ptr, uint32_t count_host, const float4 *__restrict__ ray_o,
    const float4 *__restrict__ ray_d, const float4 *__restrict__ contrib, float4 *__restrict__ path_l,
    uint8_t *__restrict__ occluded, const HalfNode *__restrict__ nodes, const WideNode *__restrict__ wnodes,
    const float4 *__restrict__ itris,
    uint32_t *__restrict__ gstack, uint32_t stack_cap, uint32_t *overflow, float4 *__restrict__ rec,
    uint32_t rec_stride, int32_t b, InstScene IS, unsigned long long *stats, uint32_t chunk_shift,
    uint32_t stack_words, uint32_t n_nodes, uint32_t n_tris) {
  extern __shared__ uint32_t smem[];
  const uint32_t count = count_ptr ? *count_ptr : count_host;
  Stack<kLdsDepth> st = make_stack<kLdsDepth>(smem, gstack, stack_cap);
  const float4 *lnodes = nullptr, *ltris = nullptr;
  stage_scene<kLdsScene>(smem, stack_words, wnodes, n_nodes, itris, n_tris, lnodes, ltris);
  trace_persistent<true, kInst, kLdsDepth, kCount, kLdsScene>(
      count, chunk_shift, nodes, itris, lnodes, ltris, st, overflow, IS, stats,
      [&](uint32_t i, f3 &o, f3 &d, float &tmax) {
        const float4 ro = ray_o[i], rd = ray_d[i];
        o = xyz(ro);
        d = xyz(rd);
        tmax = ro.w;
      },
      [&](uint32_t i, float, float, float, uint32_t id) {
        const bool occ = id != GI_INVALID_REF;
        if (kDeposit) {
          if (!occ) {
            const uint32_t slot = fbits(ray_d[i].w);
            const float4 c = contrib[i];
            float4 l = path_l[slot];
            l.x = l.x + c.x;
            l.y = l.y + c.y;
            l.z = l.z + c.z;
            path_l[slot] = l;
            if (rec) {
              float4 &r = rec[(size_t)b * rec_stride + slot];
              r = make_float4(c.x, c.y, c.z, r.w);
            }
          }
        } else {
          occluded[i] = occ ? 1 : 0;
        }
      });
}

__global__ void k_pack_rays(const float *__restrict__ orig, const float *__restrict__ dir,
                            const float *__restrict__ maxt, uint32_t n, float4 *__restrict__ ray_o,
                            float4 *__restrict__ ray_d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ray_o[i] = make_float4(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2], maxt[i]);
  ray_d[i] = make_float4(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2], bitsf(i));
}

// ------------------------------------------------------------- camera ----
// Jittered pinhole ray (src/raygen.cu:17-58) or thin-lens ray
// (src/raygen.cu:99-117; lens jitter uniform on [-1/2,1/2]^2).
__device__ __forceinline__ void camera_ray(const DevView &v, int32_t x, int32_t y, float jx, float jy, float lx,
                                           float ly, f3 &o, f3 &d) {
  float u_s = ((float(x) + jx) / v.fw) * 2.0f - 1.0f;
  float v_s = ((float(y) + jy) / v.fh) * 2.0f - 1.0f;
  u_s = v.width * u_s;
  v_s = v.height * v_s;
  const f3 U = ld3(v.U), V = ld3(v.V), W = ld3(v.W);
  d = normalize(add(add(scl(U, u_s), scl(V, v_s)), W));
  o = ld3(v.pos);
  if (v.aperture > 0.0f) {
    const f3 vd = ld3(v.vd);
    const f3 focal = add(o, scl(d, (1.0f / dot(d, vd)) * v.focus));
    const float ox = lx - 0.5f, oy = ly - 0.5f;
    const f3 jp = add(add(o, scl(U, ox * v.aperture)), scl(V, oy * v.aperture));
    d = normalize(sub(focal, jp));
    o = jp;
  }
}

__device__ __forceinline__ int find_chunk(const PassDesc *pass, uint32_t slot) {
  int c = 0;
  while (c + 1 < pass->nchunks && slot >= pass->chunk[c + 1].slot_base) ++c;
  return c;
}

// One level of the box-filtered mip chain (k::mipmap, src/material.cu:235-256):
// texel (x, y) of the dw x dh level is the truncating integer mean, per
// channel, of the (up to) 2x2 texels of the sw x sh level below it that
// exist.  (The reference zeroes the wrong weight at an odd right edge and
// reads an unset value; the missing texel's weight is the one dropped here.)
__global__ __launch_bounds__(256) void k_mipmap(const uchar4 *__restrict__ src, int32_t sw, int32_t sh,
                                                uchar4 *__restrict__ dst, int32_t dw, int32_t dh) {
  const uint32_t total = (uint32_t)dw * (uint32_t)dh;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int32_t x = (int32_t)(i % (uint32_t)dw), y = (int32_t)(i / (uint32_t)dw);
    const int32_t X = 2 * x, Y = 2 * y;
    const bool rx = X + 1 < sw, ry = Y + 1 < sh;
    const uchar4 a = src[Y * sw + X];
    uint32_t r = a.x, g = a.y, b = a.z, w = a.w, cnt = 1;
    if (rx) {
      const uchar4 c = src[Y * sw + X + 1];
      r += c.x; g += c.y; b += c.z; w += c.w; ++cnt;
    }
    if (ry) {
      const uchar4 c = src[(Y + 1) * sw + X];
      r += c.x; g += c.y; b += c.z; w += c.w; ++cnt;
    }
    if (rx && ry) {
      const uchar4 c = src[(Y + 1) * sw + X + 1];
      r += c.x; g += c.y; b += c.z; w += c.w; ++cnt;
    }
    dst[i] = make_uchar4((unsigned char)(r / cnt), (unsigned char)(g / cnt), (unsigned char)(b / cnt),
                         (unsigned char)(w / cnt));
  }
}

// The coarse levels' replicas summed into the chain's gradient: coarse
// texel c0 + j of the replicas is chain texel dst + j.
__global__ __launch_bounds__(256) void k_tex_rep_sum(float *__restrict__ g, uint32_t dst, const float *__restrict__ rep,
                                                     uint32_t c0, uint32_t n, uint32_t n_coarse, uint32_t n_rep) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    float r = 0.f, gg = 0.f, bb = 0.f;
    for (uint32_t k = 0; k < n_rep; ++k) {
      const float *q = rep + 3 * ((size_t)k * n_coarse + c0 + j);
      r += q[0];
      gg += q[1];
      bb += q[2];
    }
    float *o = g + 3 * ((size_t)dst + j);
    o[0] = o[0] + r;
    o[1] = o[1] + gg;
    o[2] = o[2] + bb;
  }
}

// Texel gradients of mip level l+1 passed down to level l: the chain is the
// box filter's mean of the (up to) 2x2 texels below (the truncation of the
// integer mean is not differentiated), so each child gets its parent's
// gradient over the parent's child count.  Run from the coarsest level down.
__global__ __launch_bounds__(256) void k_tex_grad_down(float *__restrict__ g, uint32_t child_off, int32_t cw,
                                                       int32_t ch, uint32_t parent_off, int32_t pw) {
  const uint32_t total = (uint32_t)cw * (uint32_t)ch;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int32_t x = (int32_t)(i % (uint32_t)cw), y = (int32_t)(i / (uint32_t)cw);
    const int32_t px = x >> 1, py = y >> 1;
    const float cnt = (float)((2 * px + 1 < cw ? 2 : 1) * (2 * py + 1 < ch ? 2 : 1));
    const float *gp = g + 3 * ((size_t)parent_off + (size_t)py * pw + px);
    float *gc = g + 3 * ((size_t)child_off + i);
    gc[0] = gc[0] + gp[0] / cnt;
    gc[1] = gc[1] + gp[1] / cnt;
    gc[2] = gc[2] + gp[2] / cnt;
  }
}

// Primary rays of every path slot of a pass and the colour reset (new_path:
// generate_rays and the throughput / colour reset, src/gpu-pt.h:110-118).
// The first vertex's throughput (1) and ray differentials (the camera's) are
// not stored: the shading of vertex 0 takes them as known, which spares 80
// of the 128 bytes per slot this kernel would write.
__global__ __launch_bounds__(256) void k_raygen(const PassDesc *__restrict__ pass, const DevView *__restrict__ views,
                                                HaltonDev H, float4 *__restrict__ ray_o, float4 *__restrict__ ray_d,
                                                float4 *__restrict__ path_l, float4 *__restrict__ rec,
                                                uint32_t rec_stride, int32_t depth) {
  __shared__ uint16_t s_halton[kHaltonLdsEntries];
  H = halton_in_lds(H, kDimJitterX, kDimVertex0, s_halton);
  const uint32_t total = pass->total_slots;
  for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < total; slot += gridDim.x * blockDim.x) {
    const PassChunk &ck = pass->chunk[find_chunk(pass, slot)];
    const DevView &v = views[ck.view];
    const uint32_t local = slot - ck.slot_base;
    // pixel-major: the chunk's k samples of a pixel are consecutive slots,
    // so a wave's rays start in a few pixels (coherent first bounces, and
    // the adjoint's texel gradients of one pixel meet in one wave)
    const uint32_t s = (uint32_t)ck.s0 + local % (uint32_t)ck.k;
    const uint32_t p = local / (uint32_t)ck.k;
    const uint32_t n = v.index_base + p * (uint32_t)v.spp + s;
    f3 o, d;
    // jittered (setup_jittered_shirley / _dof_shirley) or through the pixel
    // centre (setup_shirley_lens_rays, src/dofrays.cu:21-22)
    const float jx = v.centre ? 0.5f : halton(n, kDimJitterX, H), jy = v.centre ? 0.5f : halton(n, kDimJitterY, H);
    camera_ray(v, (int32_t)(p % (uint32_t)v.w), (int32_t)(p / (uint32_t)v.w), jx, jy, halton(n, kDimLensX, H),
               halton(n, kDimLensY, H), o, d);
    ray_o[slot] = make_float4(o.x, o.y, o.z, bitsf(n));
    ray_d[slot] = make_float4(d.x, d.y, d.z, bitsf(slot));
    path_l[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (rec)
      for (int32_t b = 0; b < depth; ++b) rec[(size_t)b * rec_stride + slot] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// The ray differentials of a path's primary ray: pinhole rays from the
// camera position through the centres of pixels (x, y+1) ("upper") and
// (x+1, y) ("right"); rta's camera_ray_generator_shirley, which makes them
// in the reference, is external (src/raygen.h:122-130).
__device__ __forceinline__ void primary_differentials(const PassDesc *pass, const DevView *views, uint32_t slot,
                                                      f3 &uo, f3 &ud, f3 &ro, f3 &rd) {
  const PassChunk &ck = pass->chunk[find_chunk(pass, slot)];
  DevView pv = views[ck.view];
  const uint32_t p = (slot - ck.slot_base) / (uint32_t)ck.k;  // pixel-major slots
  pv.aperture = 0.0f;
  const int32_t x = (int32_t)(p % (uint32_t)pv.w), y = (int32_t)(p / (uint32_t)pv.w);
  camera_ray(pv, x, y + 1, 0.5f, 0.5f, 0.5f, 0.5f, uo, ud);
  camera_ray(pv, x + 1, y, 0.5f, 0.5f, 0.5f, 0.5f, ro, rd);
}

// One path vertex: NEE (pixel_generate_arealight_sample_cpu,
// src/arealight-sampler-cpu.cpp:25-119) and the BRDF sample of the bounce
// (compute_path_contribution_and_bounce, src/hybrid-pt.cpp:450-577).  The
// NEE contribution is evaluated with the lobe the bounce sample selects (as
// evaluate_material does after sample_material).  Shared by the forward
// shading kernel and the adjoint replay, so both see identical values.
// Sky light sample for NEE (pixel_generate_arealight_sample_cpu with
// USE_SKY_SAMPLING, src/arealight-sampler-cpu.cpp:87-96; sky_light::sample,
// src/lights.h:121-139): (u, v) from the map's Distribution2D; direction
// (sin t cos p, sin t sin p, cos t) with t = u pi / h, p = v 2 pi / w, turned
// into the frame around the normal (make_tangential); pdf = pdf_uv w h /
// (2 pi^2 sin t); contribution = scale * L / pdf * |dir . N| with the
// unnormalised NEE normal, as for rect lights.
__device__ __forceinline__ void sky_sample(const SkyDev &S, float x0, float x1, f3 Ni, f3 &dir, f3 &contrib) {
  float yp, xp;
  const float u = cdf_sample(S.ycdf, S.ypdf, S.h, x0, yp);
  const int row = min(max((int)u, 0), S.h - 1);
  const float v = cdf_sample(S.xcdf + (size_t)row * (S.w + 1), S.xpdf + (size_t)row * S.w, S.w, x1, xp);
  const float pdf = yp * xp;
  float sth, cth, sph, cph;
  sincos2pi(u / (2.0f * (float)S.h), sth, cth);
  sincos2pi(v / (float)S.w, sph, cph);
  const f3 dm = f3{sth * cph, sth * sph, cth};
  const float outpdf = (float)((double)(pdf * (float)(S.w * S.h)) / (2.0 * M_PI * M_PI * (double)sth));
  const int ix = min(max((int)v, 0), S.w - 1);
  const float4 L = S.rgb[(size_t)row * S.w + ix];
  dir = tangential(dm, normalize(Ni));
  const float a = 1.0f / outpdf;
  const float cosn = fabsf(dot(dir, Ni));
  contrib = f3{((S.scale * L.x) * a) * cosn, ((S.scale * L.y) * a) * cosn, ((S.scale * L.z) * a) * cosn};
}

// Sky radiance seen by a ray that leaves the scene (evaluateSkyLight,
// src/hybrid-pt.cpp:175-189): theta = acos(d.y), phi = atan2(d.z, d.x) in
// [0, 2 pi), texel (int(t h), int(s w)); row h (d.y = -1) is black as the
// reference's range check makes it; the column is clamped to w-1.  Unscaled.
__device__ __forceinline__ f3 sky_eval(const SkyDev &S, f3 d) {
  const float yc = smin(smax(d.y, -1.0f), 1.0f);
  const float st = sqrtf((1.0f - yc) * (1.0f + yc));
  const float t = 2.0f * atan2_turns(st, yc);
  float sx = atan2_turns(d.z, d.x);
  if (sx < 0.0f) sx = sx + 1.0f;
  const int iy = (int)(t * (float)S.h), ix = min((int)(sx * (float)S.w), S.w - 1);
  if (iy < 0 || iy >= S.h || ix < 0) return f3{0.f, 0.f, 0.f};
  const float4 L = S.rgb[(size_t)iy * S.w + ix];
  return f3{L.x, L.y, L.z};
}

struct VertexEval {
  f3 P, N, org_dir, wi, wo_ts, wi_ts, xp;
  f3 sd, so, psc, fnee;
  f3 tpmul;       // Phong mode: the throughput factor of the bounce
  bool absorbed;  // Phong mode: no lobe chosen, or the lobe's direction is below the surface
  float len, gk, pdf;
  int type, light;
  uint32_t mat;
  bool glass;
  DevMat m;  // the material at this vertex (its texture applied)
  bool textured;  // the material has a texture: tx was applied, tf says where from
  f3 tx;
  TexFetch tf;
};

// reflect(d, N) = d - 2 (d.N) N
__device__ __forceinline__ f3 reflect_dir(f3 d, f3 N) {
  const float k = 2.0f * dot(d, N);
  return f3{d.x - N.x * k, d.y - N.y * k, d.z - N.z * k};
}

// kTex: some material has a texture (P.rdiff is set): the ray differentials
// are carried along and ve.m holds the material with its texel applied;
// otherwise the material is read from P.mats where it is used.
template <bool kTex>
__device__ __forceinline__ const DevMat &vertex_mat(const ShadeParams &P, const VertexEval &ve) {
  return kTex ? ve.m : P.mats[ve.mat];
}

// gpu_pt's material model (compute_path_contribution_and_bounce,
// src/gpu-pt-kernels.cu:89-199): Lambert plus a Phong lobe of shininess 100
// around the mirror direction R of the incoming ray about the (unflipped)
// shading normal.  kd = diffuse colour (times the diffuse texture's texel),
// ks = specular colour; both are divided by their sum of all six channels
// when it exceeds 1.  NEE weight f = kd / pi + 101 ks / (2 pi) max(R.l, 0)^100
// (the NEE contribution already holds the cosine).  The bounce picks the
// diffuse lobe when xi_z <= pd = sum(kd), the specular one when xi_z <=
// pd + ps (ps = sum(ks); with the reference's renormalisation ps /= pd + ps
// after pd /= pd + ps when pd + ps > 1), otherwise the path is absorbed; the
// direction is cos^n distributed around N (n = 1) or R (n = 100) --
// z = xi_x^(1/(n+1)) -- and a specular direction below the surface is
// absorbed too.  Throughput: TP * |dir.N| f_lobe / P_lobe * 2 pi / ((n + 1)
// z^n), with no clamp.
template <bool kTex>
__device__ __forceinline__ void phong_vertex(const ShadeParams &P, HaltonDev H, uint32_t n, int dim, f3 N,
                                             f3 org_dir, VertexEval &ve) {
  f3 kd = ld3(P.mats[ve.mat].diffuse);
  if (kTex && ve.textured) kd = mul(kd, ve.tx);
  const float4 ks4 = P.ks[ve.mat];
  f3 ks = f3{ks4.x, ks4.y, ks4.z};
  const float sum = ((((kd.x + kd.y) + kd.z) + ks.x) + ks.y) + ks.z;
  if (sum > 1.0f) {
    kd = divs(kd, sum);
    ks = divs(ks, sum);
  }
  const f3 R = reflect_dir(org_dir, N);
  const f3 l = normalize(ve.sd);
  const float cl = phong_pow100(smax(dot(R, l), 0.0f));
  ve.fnee = add(scl(kd, kInvPi), scl(ks, kPhongNorm * cl));
  ve.glass = false;
  ve.N = N;
  ve.xp = f3{halton(n, dim + 3, H), halton(n, dim + 4, H), halton(n, dim + 5, H)};
  float pd = (kd.x + kd.y) + kd.z;
  float ps = (ks.x + ks.y) + ks.z;
  if (pd + ps > 1.0f) {
    pd = pd / (pd + ps);
    ps = ps / (pd + ps);
  }
  ve.absorbed = true;
  ve.tpmul = f3{0.f, 0.f, 0.f};
  ve.type = kTypeDiffuse;
  const bool dif = ve.xp.z <= pd;
  if (!dif && !(ve.xp.z <= pd + ps)) return;
  const f3 axis = dif ? N : R;
  f3 Tt, Bt;
  tangent_frame(axis, Tt, Bt);
  const float z = dif ? sqrtf(ve.xp.x) : phong_root101(ve.xp.x);
  const float r = sqrtf(1.0f - z * z);
  float sn, cs;
  sincos2pi(ve.xp.y, sn, cs);
  f3 dir = add(add(scl(Tt, r * cs), scl(Bt, r * sn)), scl(axis, z));
  const float dn = dot(dir, N);
  f3 f;
  float den;
  if (dif) {
    f = scl(kd, kInvPi);
    den = 2.0f * z;
  } else {
    if (dn < 0.0f) return;
    ve.type = kTypeSpecRefl;
    f = scl(ks, kPhongNorm * phong_pow100(smax(dot(R, dir), 0.0f)));
    den = 101.0f * phong_pow100(z);
  }
  if (!(den > 0.0f)) return;  // xi_x = 0 (or z^100 below the float range): absorbed
  const float w = (1.0f / (dif ? pd : ps)) * (kTwoPi / den);
  ve.tpmul = scl(scl(f, fabsf(dn)), w);
  ve.wi = normalize(dir);
  ve.absorbed = false;
}

template <bool kTex, bool kPhong = false>
// H: the sampler (its tables in global memory or an LDS copy)
__device__ __forceinline__ void eval_vertex(const ShadeParams &P, HaltonDev H, uint32_t n, int dim, uint32_t slot,
                                            f3 org_dir, float4 h, VertexEval &ve) {
  uint32_t tid = fbits(h.w);
  const DevInstance *I = nullptr;
  if (tid >= P.is.n_world) {
    // instance triangle: the last instance whose virtual base is <= tid
    int lo;
    if (P.is.vtab) {
      lo = (int)P.is.vtab[(tid - P.is.n_world) >> P.is.vshift];
      if (lo + 1 < P.is.n_inst && P.is.vbase[lo + 1] <= tid) ++lo;
    } else {  // no table (more than 2^24 blocks): binary search
      lo = 0;
      int hi = P.is.n_inst - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (P.is.vbase[mid] <= tid) lo = mid;
        else hi = mid - 1;
      }
    }
    I = &P.is.inst[lo];
    tid = I->stri + (tid - I->vbase);
  }
  const ShadeTri &T = P.tris[tid];
  const float beta = h.y, gamma = h.z;
  const f3 bc = f3{(1.0f - beta) - gamma, beta, gamma};
  f3 Pp = add(add(scl(xyz(T.a), bc.x), scl(xyz(T.b), bc.y)), scl(xyz(T.c), bc.z));
  f3 Ni = add(add(scl(xyz(T.na), bc.x), scl(xyz(T.nb), bc.y)), scl(xyz(T.nc), bc.z));
  ve.mat = fbits(T.a.w);
  if (I) {
    // object to world: positions by the forward rows, normals by the
    // inverse transpose (columns of the inverse rows)
    const float4 f0 = I->fwd[0], f1 = I->fwd[1], f2 = I->fwd[2];
    const float4 r0 = I->inv[0], r1 = I->inv[1], r2 = I->inv[2];
    Pp = f3{((f0.x * Pp.x + f0.y * Pp.y) + f0.z * Pp.z) + f0.w, ((f1.x * Pp.x + f1.y * Pp.y) + f1.z * Pp.z) + f1.w,
            ((f2.x * Pp.x + f2.y * Pp.y) + f2.z * Pp.z) + f2.w};
    Ni = f3{(r0.x * Ni.x + r1.x * Ni.y) + r2.x * Ni.z, (r0.y * Ni.x + r1.y * Ni.y) + r2.y * Ni.z,
            (r0.z * Ni.x + r1.z * Ni.y) + r2.z * Ni.z};
    if (I->material >= 0) ve.mat = (uint32_t)I->material;
  }
  if (kTex) ve.m = P.mats[ve.mat];
  const DevMat &m = vertex_mat<kTex>(P, ve);
  const f3 N0 = normalize(Ni);
  if (kTex) {
    // computeRayDiffs (src/hybrid-pt.cpp:300-328): the differentials cross
    // the hit triangle's plane; their texture coordinates give the lookup's
    // footprint (material_t::diffuseColor, src/material.h:160-166); the
    // origins move to the crossings and the upper direction is reflected
    // about the shading normal (the right direction is kept, as there)
    const uint32_t st = P.rdiff_stride;
    f3 A = xyz(T.a), B = xyz(T.b), C = xyz(T.c);
    if (I) {
      const float4 f0 = I->fwd[0], f1 = I->fwd[1], f2 = I->fwd[2];
      A = f3{((f0.x * A.x + f0.y * A.y) + f0.z * A.z) + f0.w, ((f1.x * A.x + f1.y * A.y) + f1.z * A.z) + f1.w,
             ((f2.x * A.x + f2.y * A.y) + f2.z * A.z) + f2.w};
      B = f3{((f0.x * B.x + f0.y * B.y) + f0.z * B.z) + f0.w, ((f1.x * B.x + f1.y * B.y) + f1.z * B.z) + f1.w,
             ((f2.x * B.x + f2.y * B.y) + f2.z * B.z) + f2.w};
      C = f3{((f0.x * C.x + f0.y * C.y) + f0.z * C.z) + f0.w, ((f1.x * C.x + f1.y * C.y) + f1.z * C.z) + f1.w,
             ((f2.x * C.x + f2.y * C.y) + f2.z * C.z) + f2.w};
    }
    const f3 e1 = sub(B, A), e2 = sub(C, A);
    // the first vertex's differentials are the camera's (not stored); later
    // ones come from the previous vertex
    const bool first = dim == kDimVertex0;
    f3 duo, dud, dro, drd;
    if (first) {
      primary_differentials(P.pass, P.views, slot, duo, dud, dro, drd);
    } else {
      duo = xyz(P.rdiff[slot]);
      dud = xyz(P.rdiff[st + slot]);
      dro = xyz(P.rdiff[2 * st + slot]);
      drd = xyz(P.rdiff[3 * st + slot]);
    }
    float ub, ug, rb, rg;
    plane_bary(A, e1, e2, duo, dud, ub, ug);
    plane_bary(A, e1, e2, dro, drd, rb, rg);
    const float u0 = (1.0f - ub) - ug, r0 = (1.0f - rb) - rg;
    if (m.texture >= 0) {
      const float4 tab = T.tab, tcm = T.tcm;
      const float tcx = (tab.x * bc.x + tab.z * bc.y) + tcm.x * bc.z;
      const float tcy = (tab.y * bc.x + tab.w * bc.y) + tcm.y * bc.z;
      const float ux = (tab.x * u0 + tab.z * ub) + tcm.x * ug, uy = (tab.y * u0 + tab.w * ub) + tcm.y * ug;
      const float rx = (tab.x * r0 + tab.z * rb) + tcm.x * rg, ry = (tab.y * r0 + tab.w * rb) + tcm.y * rg;
      float dx = fabsf(tcx - ux), dy = fabsf(tcy - uy);
      dx = fmaxf(fabsf(tcx - rx), dx);
      dy = fmaxf(fabsf(tcy - ry), dy);
      ve.tx = tex_sample_lod(P.tex[m.texture], tcx, tcy, fmaxf(dx, dy), ve.tf);
      apply_texel(ve.m, ve.tx);
    }
    ve.textured = m.texture >= 0;
    const f3 uo = add(add(scl(A, u0), scl(B, ub)), scl(C, ug));
    const f3 ro = add(add(scl(A, r0), scl(B, rb)), scl(C, rg));
    const f3 ud = reflect_dir(dud, N0);
    P.rdiff[slot] = make_float4(uo.x, uo.y, uo.z, 0.0f);
    P.rdiff[st + slot] = make_float4(ud.x, ud.y, ud.z, 0.0f);
    P.rdiff[2 * st + slot] = make_float4(ro.x, ro.y, ro.z, 0.0f);
    if (first) P.rdiff[3 * st + slot] = make_float4(drd.x, drd.y, drd.z, 0.0f);  // kept from here on
  }
  ve.org_dir = org_dir;
  ve.P = Pp;

  // NEE: power-proportional light choice, uniform point on the rect
  const f3 xl = f3{halton(n, dim + 0, H), halton(n, dim + 1, H), halton(n, dim + 2, H)};
  // the first light k with choice <= lsum[k + 1], the last one if none:
  // what the linear scan over running sums picks, by bisection (the sums
  // only grow)
  const float choice = xl.z * P.power_total;
  int k = 0, hi = P.nlights - 1;
  while (k < hi) {
    const int mid = (k + hi) >> 1;
    if (choice > P.lsum[mid + 1]) k = mid + 1;
    else hi = mid;
  }
  ve.light = k;
  const DevLight &lt = P.lights[k];
  f3 sd, contrib;
  float len, factor = 0.0f;
  if (lt.type == GI_LIGHT_SKY) {
    sky_sample(P.sky, xl.x, xl.y, Ni, sd, contrib);
    len = FLT_MAX;
    sd = normalize(sd);
  } else {
    const f3 ldir = ld3(lt.dir);
    const float offx = (xl.x - 0.5f) * lt.w, offy = (xl.y - 0.5f) * lt.h;
    const f3 ls = add(add(ld3(lt.center), scl(ld3(lt.right), offx)), scl(ld3(lt.up), offy));
    sd = sub(ls, Pp);
    len = length(sd);
    sd = divs(sd, len);
    const float ndotl = smax(dot(Ni, sd), 0.0f);
    const float lcos = smax(dot(ldir, neg(sd)), 0.0f);
    factor = (((lt.w * lt.h) * ndotl) * lcos) / (len * len);
    contrib = scl(ld3(lt.col), factor);
    sd = normalize(sd);
  }
  ve.sd = sd;
  ve.len = len;
  ve.so = add(Pp, scl(sd, P.off_s));
  const float sel = P.power_total / lt.power;
  ve.psc = scl(contrib, sel);
  ve.gk = factor * sel;

  if constexpr (kPhong) {
    phong_vertex<kTex>(P, H, n, dim, N0, org_dir, ve);
    return;
  }
  // bounce: tangent frame, normal flip, BRDF sample
  f3 N = N0;
  f3 Tt, Bt;
  tangent_frame(N, Tt, Bt);
  if (dot(org_dir, N) > 0.0f) N = neg(N);
  ve.N = N;
  const f3 inv = neg(org_dir);
  ve.wo_ts = f3{dot(inv, Tt), dot(inv, Bt), dot(inv, N)};
  ve.xp = f3{halton(n, dim + 3, H), halton(n, dim + 4, H), halton(n, dim + 5, H)};
  ve.pdf = 1.0f;
  ve.type = brdf_sample(m, ve.wo_ts, ve.xp, ve.wi_ts, ve.pdf);
  ve.wi = normalize(add(add(scl(Tt, ve.wi_ts.x), scl(Bt, ve.wi_ts.y)), scl(N, ve.wi_ts.z)));
  ve.glass = m.principled && m.opacity < 1.0f;
  ve.fnee = ve.glass ? f3{0.f, 0.f, 0.f} : brdf_eval(m, ve.type, inv, normalize(sd), N);
}

// Next throughput: min(1, TP * f / pdf) per channel (src/hybrid-pt.cpp:559-562);
// clamp bit c set where channel c was clamped.
__device__ __forceinline__ f3 next_throughput(f3 TP, f3 f, float pdf, uint32_t &clamp) {
  f3 t = f3{(TP.x * f.x) / pdf, (TP.y * f.y) / pdf, (TP.z * f.z) / pdf};
  clamp = 0;
  if (t.x > 1.0f) { t.x = 1.0f; clamp |= 1u; }
  if (t.y > 1.0f) { t.y = 1.0f; clamp |= 2u; }
  if (t.z > 1.0f) { t.z = 1.0f; clamp |= 4u; }
  return t;
}

// Per-vertex record of the adjoint's forward pass: the radiance the vertex
// adds to its path (NEE if unoccluded, background on a white miss) and flag
// bits (0-2: throughput clamp per channel, 3: background term).
constexpr uint32_t kRecBackground = 8u;

// Shading of path vertex b.  The NEE contribution travels with the shadow
// ray; the shadow kernel adds it when unoccluded.  rec != nullptr: also
// record the vertex for the adjoint (see kRecBackground).
template <bool kTex, bool kPhong = false>
__device__ __forceinline__ void shade_vertices(
    int32_t b, const uint32_t *__restrict__ count_ptr, const ShadeParams &P, const float4 *__restrict__ ray_o,
    const float4 *__restrict__ ray_d, const float4 *__restrict__ tp_in, const float4 *__restrict__ hits,
    float4 *__restrict__ path_l, float4 *__restrict__ nro, float4 *__restrict__ nrd, float4 *__restrict__ ntp,
    uint32_t *__restrict__ next_count, float4 *__restrict__ sho, float4 *__restrict__ shd, float4 *__restrict__ shc,
    uint32_t *__restrict__ shadow_count, float4 *__restrict__ rec, uint32_t rec_stride) {
  const int dim = kDimVertex0 + kDimsPerVertex * b;
  // the vertex's sample dimensions, read from an LDS copy of their tables;
  // the textured kernel keeps reading them from global memory: with the LDS
  // reads its schedule spills 11 registers instead of 2 and runs 5 % slower
  __shared__ uint16_t s_halton[kHaltonLdsEntries];
  const HaltonDev H = kTex ? P.H : halton_in_lds(P.H, dim, kDimsPerVertex, s_halton);
  const uint32_t count = *count_ptr;
  const bool last = b + 1 >= P.depth;
  for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    bool emit_next = false, emit_shadow = false;
    float4 o_next, d_next, tp_next, s_o, s_d, s_c;
    if (i < count) {
      const float4 ro = ray_o[i], rd = ray_d[i], h = hits[i];
      const float4 tpv = b == 0 ? make_float4(1.f, 1.f, 1.f, 0.f) : tp_in[i];  // a new path's throughput
      const uint32_t n = fbits(ro.w), slot = fbits(rd.w), id = fbits(h.w);
      const f3 TP = xyz(tpv);
      const f3 org_dir = normalize(xyz(rd));
      // the adjoint's record also keeps the vertex's hit (its second half),
      // so that the replay need not trace the same ray again
      if (rec) rec[(size_t)(P.depth + b) * rec_stride + slot] = h;
      if (id == GI_INVALID_REF) {
        // handle_no_hit with NO_BACKGROUND (src/hybrid-pt.cpp:196-206,567-574)
        if (TP.x == 1.0f && TP.y == 1.0f && TP.z == 1.0f) {
          float4 l = path_l[slot];
          const float bx = TP.x * P.bg[0], by = TP.y * P.bg[1], bz = TP.z * P.bg[2];
          l.x = l.x + bx;
          l.y = l.y + by;
          l.z = l.z + bz;
          path_l[slot] = l;
          if (rec) rec[(size_t)b * rec_stride + slot] = make_float4(bx, by, bz, bitsf(kRecBackground));
        } else if (P.sky.light >= 0) {
          const f3 Ls = sky_eval(P.sky, org_dir);
          float4 l = path_l[slot];
          l.x = l.x + TP.x * Ls.x;
          l.y = l.y + TP.y * Ls.y;
          l.z = l.z + TP.z * Ls.z;
          path_l[slot] = l;
        }
      } else {
        VertexEval ve;
        eval_vertex<kTex, kPhong>(P, H, n, dim, slot, org_dir, h, ve);
        const DevMat &m = vertex_mat<kTex>(P, ve);
        if (!ve.glass) {
          const f3 C = mul(ve.fnee, mul(TP, ve.psc));
          if (C.x > 0.0f || C.y > 0.0f || C.z > 0.0f) {
            emit_shadow = true;
            s_o = make_float4(ve.so.x, ve.so.y, ve.so.z, ve.len);
            s_d = make_float4(ve.sd.x, ve.sd.y, ve.sd.z, bitsf(slot));
            s_c = make_float4(C.x, C.y, C.z, 0.0f);
          }
        }
        uint32_t clamp = 0;
        if (!last && !(kPhong && ve.absorbed)) {
          f3 np, t;
          if constexpr (kPhong) {
            // gpu_pt: offset along the new direction, no throughput clamp
            np = add(ve.P, scl(ve.wi, P.off_b));
            t = mul(TP, ve.tpmul);
          } else {
            const f3 f = brdf_eval(m, ve.type, neg(org_dir), ve.wi, ve.N);
            np = ve.glass ? add(ve.P, scl(org_dir, P.off_b)) : add(ve.P, scl(ve.N, P.off_b));
            t = next_throughput(TP, f, ve.pdf, clamp);
          }
          bool alive = true;
          if (P.rr_start > 0 && b + 1 >= P.rr_start) {
            const float q = smax(t.x, smax(t.y, t.z));
            const float xr = halton(n, dim + 6, H);
            if (!(xr < q)) alive = false;
            else t = divs(t, q);
          }
          if (alive) {
            emit_next = true;
            o_next = make_float4(np.x, np.y, np.z, bitsf(n));
            d_next = make_float4(ve.wi.x, ve.wi.y, ve.wi.z, bitsf(slot));
            tp_next = make_float4(t.x, t.y, t.z, 0.0f);
          }
        }
        if (rec) rec[(size_t)b * rec_stride + slot] = make_float4(0.f, 0.f, 0.f, bitsf(clamp));
      }
    }
    uint32_t qn, qs;
    block_append2(emit_next, next_count, emit_shadow, shadow_count, qn, qs);
    if (emit_next) {
      nro[qn] = o_next;
      nrd[qn] = d_next;
      ntp[qn] = tp_next;
    }
    if (emit_shadow) {
      sho[qs] = s_o;
      shd[qs] = s_d;
      shc[qs] = s_c;
    }
  }
}

// Both shading kernels are held to 4 waves per SIMD (128 VGPRs): the
// untextured one fits without spilling, the textured one spills two
// registers; at the compiler's own choice they take 137 and 177
// VGPRs (3 waves), which measured slower.
#define GI_SHADE_ARGS                                                                                              \
  int32_t b, const uint32_t *__restrict__ count_ptr, ShadeParams P, const float4 *__restrict__ ray_o,             \
      const float4 *__restrict__ ray_d, const float4 *__restrict__ tp_in, const float4 *__restrict__ hits,        \
      float4 *__restrict__ path_l, float4 *__restrict__ nro, float4 *__restrict__ nrd, float4 *__restrict__ ntp,  \
      uint32_t *__restrict__ next_count, float4 *__restrict__ sho, float4 *__restrict__ shd,                      \
      float4 *__restrict__ shc, uint32_t *__restrict__ shadow_count, float4 *__restrict__ rec, uint32_t rec_stride
#define GI_SHADE_PASS                                                                                               \
  b, count_ptr, P, ray_o, ray_d, tp_in, hits, path_l, nro, nrd, ntp, next_count, sho, shd, shc, shadow_count, rec, \
      rec_stride
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_shade(GI_SHADE_ARGS) {
  shade_vertices<false>(GI_SHADE_PASS);
}
// gpu_pt's Lambert + Phong model (GI_FLAG_PHONG), untextured and textured
__global__ __launch_bounds__(256) void k_shade_phong(GI_SHADE_ARGS) { shade_vertices<false, true>(GI_SHADE_PASS); }
__global__ __launch_bounds__(256) void k_shade_phong_tex(GI_SHADE_ARGS) { shade_vertices<true, true>(GI_SHADE_PASS); }
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_shade_tex(GI_SHADE_ARGS) {
  shade_vertices<true>(GI_SHADE_PASS);
}
#undef GI_SHADE_ARGS
#undef GI_SHADE_PASS

// ------------------------------------------------------------- adjoint ----
// Derivatives d f / d (color r, g, b, roughness, metallic) of the BRDF value
// for the lobe `type`, for the frozen directions wo, wi.  Lambert: d/d rho.
struct alignas(16) DevMatGrad {
  float dcd[3];   // d Cdlin / d color = 2.2 color^1.2
  float lum;      // .3 Cd.r + .6 Cd.g + .1 Cd.b
  float tint[3];  // Cdlin / lum
  float spec_tint;
  float sheen_tint, aspect, pad0, pad1;
};

__device__ __forceinline__ float f3c(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
__device__ __forceinline__ float dtint(const DevMatGrad &g, int c, int i) {
  if (!(g.lum > 0.0f)) return 0.0f;
  const float w = i == 0 ? 0.3f : (i == 1 ? 0.6f : 0.1f);
  return ((c == i ? 1.0f : 0.0f) - g.tint[c] * w) / g.lum;
}

__device__ __forceinline__ void brdf_grad(const DevMat &m, const DevMatGrad &g, int type, f3 wo, f3 wi, f3 N,
                                          float df[5][3]) {
#pragma unroll
  for (int k = 0; k < 5; ++k) df[k][0] = df[k][1] = df[k][2] = 0.0f;
  if (!m.principled) {
    const float a = fabsf(dot(wi, N)) * kInvPi;
    df[0][0] = a;
    df[1][1] = a;
    df[2][2] = a;
    return;
  }
  const float ndl = dot(N, wi), ndv = dot(N, wo);
  if (ndl < 0.0f || ndv < 0.0f) return;
  const f3 H = normalize(add(wi, wo));
  const float ldh = dot(wi, H);
  if (type == kTypeDisneyRefl) {
    const float ndh = dot(N, H);
    const float fh = schlick(ldh), mm = m.metallic, sp08 = m.specular * 0.08f, st = g.spec_tint;
    float cg = 0.0f;
    if (m.clearcoat > 0.0f) {
      const float dr = gtr1(ndh, m.cc_a, m.cc_log_a2);
      const float fr = mixf(0.04f, 1.0f, fh);
      const float gr = smith_ggx(ndl, 0.25f) * smith_ggx(ndv, 0.25f);
      cg = 0.25f * m.clearcoat * gr * fr * dr;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float mt = (1.0f - st) + st * g.tint[c];
      const float cs0 = (1.0f - mm) * sp08 * mt + mm * m.cdlin[c];
      const float fs = (1.0f - fh) * cs0 + fh;
      df[4][c] = (fs + cg) * ndl + mm * ndl * (1.0f - fh) * (m.cdlin[c] - sp08 * mt);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const float dcs0 = mm * (c == i ? 1.0f : 0.0f) + (1.0f - mm) * sp08 * st * dtint(g, c, i);
        df[i][c] = mm * ndl * (1.0f - fh) * dcs0 * g.dcd[i];
      }
    }
    return;
  }
  if (type != kTypeDiffuse) return;
  const float fl = schlick(ndl), fv = schlick(ndv), r = m.roughness, mm = m.metallic, sub = m.subsurface;
  const float fd90 = 0.5f + 2.0f * ldh * ldh * r;
  const float fd = mixf(1.0f, fd90, fl) * mixf(1.0f, fd90, fv);
  const float fh = schlick(ldh);
  const float fss90 = ldh * ldh * r;
  const float fss = mixf(1.0f, fss90, fl) * mixf(1.0f, fss90, fv);
  const float q = 1.0f / (ndl + ndv) - 0.5f;
  const float ss = 1.25f * (fss * q + 0.5f);
  const float k = kInvPi * mixf(fd, ss, sub);
  const float dfd = 2.0f * ldh * ldh * (fl * mixf(1.0f, fd90, fv) + mixf(1.0f, fd90, fl) * fv);
  const float dss = 1.25f * q * ldh * ldh * (fl * mixf(1.0f, fss90, fv) + mixf(1.0f, fss90, fl) * fv);
  const float dk = kInvPi * ((1.0f - sub) * dfd + sub * dss);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float csh = (1.0f - g.sheen_tint) + g.sheen_tint * g.tint[c];
    df[3][c] = dk * m.cdlin[c] * (1.0f - mm) * ndl;
    df[4][c] = -(k * m.cdlin[c] + fh * m.sheen * csh) * ndl;
#pragma unroll
    for (int i = 0; i < 3; ++i)
      df[i][c] = (1.0f - mm) * ndl * (k * (c == i ? 1.0f : 0.0f) + fh * m.sheen * g.sheen_tint * dtint(g, c, i)) *
                 g.dcd[i];
  }
}

// GGX roughness derivatives of a sampled specular lobe: d log(lweight)/dr
// (the reference's specular "pdf", src/principledMaterial.h:158-175) and
// the score d log D(h)/dr of the half-vector density D(h) cos(theta_h).
__device__ __forceinline__ void gtr2_grads(const DevMat &m, const DevMatGrad &g, f3 V, f3 L, f3 h, float &dlogw,
                                           float &score) {
  const float r = m.roughness, r2 = r * r, aspect = g.aspect;
  const float ax = m.ax, ay = m.ay;
  const float dax = (r2 / aspect > 0.001f) ? 2.0f * r / aspect : 0.0f;
  const float day = (r2 * aspect > 0.001f) ? 2.0f * r * aspect : 0.0f;
  const float qq = ax * ay * 0.5f + 0.5f;
  const float drg2 = 4.0f * qq * qq * qq * 0.5f * (dax * ay + ax * day);
  const float dlm = dot(L, h), dvm = dot(V, h);
  const float tl = sqrtf(1.0f - dlm * dlm) / dlm, tv = sqrtf(1.0f - dvm * dvm) / dvm;
  const float xl = (dot(L, h) / L.z) > 0.0f ? 1.0f : 0.0f, xv = (dot(V, h) / V.z) > 0.0f ? 1.0f : 0.0f;
  const float sl = sqrtf(1.0f + m.roughg2 * tl * tl), sv = sqrtf(1.0f + m.roughg2 * tv * tv);
  const float g1l = xl * 2.0f / (1.0f + sl), g1v = xv * 2.0f / (1.0f + sv);
  const float lw = (fabsf(dot(L, h)) * g1l * g1v) / (fabsf(L.z) * fabsf(h.z));
  dlogw = 0.0f;
  if (lw >= 0.01f && g1l > 0.0f && g1v > 0.0f) {
    const float dg1l = -(1.0f / (1.0f + sl)) * (tl * tl / (2.0f * sl));
    const float dg1v = -(1.0f / (1.0f + sv)) * (tv * tv / (2.0f * sv));
    dlogw = (dg1l + dg1v) * drg2;
  }
  const float Q = h.x * h.x / (ax * ax) + h.y * h.y / (ay * ay) + h.z * h.z;
  score = (-1.0f / ax + 4.0f * h.x * h.x / (ax * ax * ax * Q)) * dax +
          (-1.0f / ay + 4.0f * h.y * h.y / (ay * ay * ay * Q)) * day;
}

struct AdjParams {
  const float *dL;          // adjoint image, 3 floats per pixel, views concatenated
  const float4 *rec;        // per-vertex records of the forward pass
  uint32_t rec_stride;
  const DevMatGrad *mg;
  float *g_mats;            // 5 per material
  float *g_em;              // 3 per light
  int32_t nmats, nlights;
  const PassDesc *pass;
  const DevView *views;
  float *g_tex;               // texel gradients, RGB per texel of every chain
  const uint32_t *tex_goff;   // first texel of texture i in g_tex (in texels)
  // replicas of the coarse levels' gradients (levels of <= kCoarseTexels
  // texels, where many lanes across the chip add to few texels): replica
  // r of coarse texel c at g_rep[3 (r n_coarse + c)], r from the block
  float *g_rep;
  const uint32_t *tex_coarse;  // chain index of texture i's first replicated texel
  const uint32_t *tex_gc;      // texture i's first coarse texel in the replicas' indexing
  uint32_t n_coarse;
};

// Coarse-level replication of the texel gradients: levels of at most
// kCoarseTexels texels get kGradReplicas copies.
constexpr uint32_t kCoarseTexels = 16384;
constexpr uint32_t kGradReplicas = 64;

// Adjoint replay of path vertex b: the same hit, NEE point and BRDF sample
// as the forward pass (same Halton indices), the visibility read off the
// recorded contribution, and the gradient terms of the frozen-sample
// estimator (NEE, throughput chain up to the next clamp, score of the lobe
// choice and of the GGX half-vector, emission).  Gradients gather in LDS per
// block and go to global memory with one atomic per parameter per block.
// Sum over the 64 lanes of a wave (butterfly); every lane gets the total.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Texel gradients of one lookup per lane (has: the lane has one): g (RGB)
// times each of the four bilinear weights, added at texel gi[k] (index into
// g_tex in texels).  Called by the whole wave.  Lanes whose texel equals the
// first pending lane's are summed across the wave first and added by one
// lane, for up to 8 such groups of at least 2 lanes: the samples of a pixel
// share a wave (pixel-major slots) and mostly its texels, coarse mip levels
// are shared by many lanes, and global atomics are what this pass spends
// its time on.  ta[k] is where texel gi[k]'s RGB gradient is added (a
// coarse texel's replica).
__device__ __forceinline__ void scatter_texel_grads(bool has, float *const ta[4], const uint32_t gi[4],
                                                    const float w[4], const float g[3]) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    bool pend = has;
    for (int round = 0; round < 8; ++round) {
      const uint64_t act = __ballot(pend);
      if (act == 0) break;
      const int leader = __ffsll((unsigned long long)act) - 1;
      const uint32_t key = (uint32_t)__shfl((int)gi[k], leader, 64);
      const bool mine = pend && gi[k] == key;
      if (__popcll(__ballot(mine)) < 2) break;
      const float r = wave_sum(mine ? g[0] * w[k] : 0.0f);
      const float gg = wave_sum(mine ? g[1] * w[k] : 0.0f);
      const float bb = wave_sum(mine ? g[2] * w[k] : 0.0f);
      if (lane == (uint32_t)leader) {
        atomicAdd(ta[k] + 0, r);
        atomicAdd(ta[k] + 1, gg);
        atomicAdd(ta[k] + 2, bb);
      }
      pend = pend && !mine;
    }
    if (pend) {
      atomicAdd(ta[k] + 0, g[0] * w[k]);
      atomicAdd(ta[k] + 1, g[1] * w[k]);
      atomicAdd(ta[k] + 2, g[2] * w[k]);
    }
  }
}

// kTex: the ray differentials are carried (as in the forward pass) and a
// textured vertex's base colour b = colour * texel: the colour gradient is
// dL/db * texel, and dL/db * colour goes to the four blended texels with
// their bilinear weights (global atomics into the chain's gradient).
template <bool kTex>
__global__ __launch_bounds__(256) void k_shade_adjoint(int32_t b, const uint32_t *__restrict__ count_ptr,
                                                       ShadeParams P, AdjParams A, const float4 *__restrict__ ray_o,
                                                       const float4 *__restrict__ ray_d,
                                                       const float4 *__restrict__ tp_in, float4 *__restrict__ nro,
                                                       float4 *__restrict__ nrd, float4 *__restrict__ ntp,
                                                       uint32_t *__restrict__ next_count) {
  extern __shared__ float sg[];  // 5*nmats + 3*nlights
  const int ng = 5 * A.nmats + 3 * A.nlights;
  for (int k = threadIdx.x; k < ng; k += blockDim.x) sg[k] = 0.0f;
  __syncthreads();
  const uint32_t count = *count_ptr;
  const int dim = kDimVertex0 + kDimsPerVertex * b;
  const bool last = b + 1 >= P.depth;
  for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    bool emit_next = false;
    float4 o_next, d_next, tp_next;
    bool tex_has = false;  // texel gradient of this lane's lookup, scattered below by the whole wave
    uint32_t tex_gi[4] = {0u, 0u, 0u, 0u};
    float *tex_a[4] = {nullptr, nullptr, nullptr, nullptr};
    float tex_w[4] = {0.f, 0.f, 0.f, 0.f}, tex_g[3] = {0.f, 0.f, 0.f};
    if (i < count) {
      const float4 ro = ray_o[i], rd = ray_d[i];
      const float4 tpv = b == 0 ? make_float4(1.f, 1.f, 1.f, 0.f) : tp_in[i];  // a new path's throughput
      const uint32_t n = fbits(ro.w), slot = fbits(rd.w);
      // the hit the forward pass recorded for this path's vertex b (the
      // replay's rays are the forward pass's, bit for bit)
      const float4 h = A.rec[(size_t)(P.depth + b) * A.rec_stride + slot];
      const uint32_t id = fbits(h.w);
      const f3 TP = xyz(tpv);
      const f3 org_dir = normalize(xyz(rd));
      if (id != GI_INVALID_REF) {
        VertexEval ve;
        eval_vertex<kTex>(P, P.H, n, dim, slot, org_dir, h, ve);
        const DevMat &m = vertex_mat<kTex>(P, ve);
        DevMatGrad mg = A.mg[ve.mat];
        if (kTex && ve.textured && m.principled) {
          // the derivative quantities of the textured base colour b:
          // d Cdlin / d b = 2.2 b^1.2 = 2.2 Cdlin / b
          const float lum = 0.3f * m.cdlin[0] + 0.6f * m.cdlin[1] + 0.1f * m.cdlin[2];
          mg.lum = lum;
          for (int c = 0; c < 3; ++c) {
            mg.tint[c] = lum > 0.f ? m.cdlin[c] / lum : 1.f;
            mg.dcd[c] = m.color[c] > 0.f ? (2.2f * m.cdlin[c]) / m.color[c] : 0.f;
          }
        }
        const f3 inv = neg(org_dir);
        // adjoint weight of the path's pixel: dL/dI / spp
        const PassChunk &ck = A.pass->chunk[find_chunk(A.pass, slot)];
        const DevView &v = A.views[ck.view];
        const uint32_t p = (slot - ck.slot_base) / (uint32_t)ck.k;  // pixel-major slots
        const float *w = A.dL + 3 * ((size_t)v.pix_base + p);
        const float inv_spp = 1.0f / (float)v.spp;
        const f3 delta = f3{w[0] * inv_spp, w[1] * inv_spp, w[2] * inv_spp};
        // recorded contributions from this vertex on: S (all) and R (after
        // this vertex, up to the next throughput clamp per channel)
        const float4 r0 = A.rec[(size_t)b * A.rec_stride + slot];
        f3 S = xyz(r0);
        f3 R = f3{0.f, 0.f, 0.f};
        uint32_t open = 7u & ~fbits(r0.w);  // channels whose chain continues past b
        for (int bb = b + 1; bb < P.depth; ++bb) {
          const float4 r = A.rec[(size_t)bb * A.rec_stride + slot];
          S = add(S, xyz(r));
          const uint32_t fl = fbits(r.w);
          if (!(fl & kRecBackground)) {
            if (open & 1u) R.x += r.x;
            if (open & 2u) R.y += r.y;
            if (open & 4u) R.z += r.z;
          }
          open &= ~fl;
        }
        const bool vis = r0.x != 0.0f || r0.y != 0.0f || r0.z != 0.0f;
        float dfn[5][3], dfb[5][3];
        float score[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, dlogw = 0.0f;
        if (m.principled && !ve.glass) {
          score[4] = ve.xp.z < m.metallic ? 1.0f / m.metallic : -1.0f / (1.0f - m.metallic);
          if (ve.type == kTypeDisneyRefl) {
            float sn, cs;
            sincos2pi(ve.xp.x, sn, cs);
            const float fac = sqrtf(ve.xp.y / (1.0f - ve.xp.y));
            const f3 hv = normalize(f3{fac * (m.ax * cs), fac * (m.ay * sn), 1.0f});
            gtr2_grads(m, mg, ve.wo_ts, ve.wi_ts, hv, dlogw, score[3]);
          }
        }
        if (!ve.glass) brdf_grad(m, mg, ve.type, inv, normalize(ve.sd), ve.N, dfn);
        f3 f = f3{0.f, 0.f, 0.f};
        uint32_t clamp = 7u;
        if (!last) {
          f = brdf_eval(m, ve.type, inv, ve.wi, ve.N);
          brdf_grad(m, mg, ve.type, inv, ve.wi, ve.N, dfb);
          const f3 t = next_throughput(TP, f, ve.pdf, clamp);
          const f3 np = ve.glass ? add(ve.P, scl(org_dir, P.off_b)) : add(ve.P, scl(ve.N, P.off_b));
          emit_next = true;
          o_next = make_float4(np.x, np.y, np.z, bitsf(n));
          d_next = make_float4(ve.wi.x, ve.wi.y, ve.wi.z, bitsf(slot));
          tp_next = make_float4(t.x, t.y, t.z, 0.0f);
        }
        const float sdot = dot(delta, S);
        const float dl[3] = {delta.x, delta.y, delta.z};
        const float Tc[3] = {TP.x, TP.y, TP.z}, Pc[3] = {ve.psc.x, ve.psc.y, ve.psc.z};
        const float Rc[3] = {R.x, R.y, R.z}, Fc[3] = {f.x, f.y, f.z};
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          float gq = sdot * score[q];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            if (vis && !ve.glass) gq += dl[c] * dfn[q][c] * Tc[c] * Pc[c];
            if (!last && Fc[c] != 0.0f && Rc[c] != 0.0f)
              gq += dl[c] * Rc[c] * (dfb[q][c] / Fc[c] - (q == 3 ? dlogw : 0.0f));
          }
          if (kTex && ve.textured && q < 3) {
            // b_q = colour_q * tx_q
            const DevMat &m0 = P.mats[ve.mat];
            tex_g[q] = gq * (m.principled ? m0.color[q] : m0.diffuse[q]);
            gq = gq * f3c(ve.tx, q);
          }
          if (gq != 0.0f) atomicAdd(&sg[5 * ve.mat + q], gq);
        }
        if (kTex && ve.textured) {
          tex_has = tex_g[0] != 0.0f || tex_g[1] != 0.0f || tex_g[2] != 0.0f;
          const uint32_t base = A.tex_goff[m.texture], coarse = A.tex_coarse[m.texture];
          const uint32_t rep = (blockIdx.x % kGradReplicas) * A.n_coarse + A.tex_gc[m.texture];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const uint32_t ix = ve.tf.idx[k];
            tex_gi[k] = base + ix;
            tex_w[k] = ve.tf.w[k];
            tex_a[k] = ix >= coarse ? A.g_rep + 3 * ((size_t)rep + (ix - coarse)) : A.g_tex + 3 * (size_t)(base + ix);
          }
        }
        if (vis && !ve.glass) {
          atomicAdd(&sg[5 * A.nmats + 3 * ve.light + 0], dl[0] * ve.fnee.x * Tc[0] * ve.gk);
          atomicAdd(&sg[5 * A.nmats + 3 * ve.light + 1], dl[1] * ve.fnee.y * Tc[1] * ve.gk);
          atomicAdd(&sg[5 * A.nmats + 3 * ve.light + 2], dl[2] * ve.fnee.z * Tc[2] * ve.gk);
        }
      }
    }
    if (kTex) scatter_texel_grads(tex_has, tex_a, tex_gi, tex_w, tex_g);
    uint32_t qn, unused;
    block_append2(emit_next, next_count, false, next_count, qn, unused);
    if (emit_next) {
      nro[qn] = o_next;
      nrd[qn] = d_next;
      ntp[qn] = tp_next;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < ng; k += blockDim.x) {
    const float v = sg[k];
    if (v != 0.0f) {
      if (k < 5 * A.nmats) atomicAdd(&A.g_mats[k], v);
      else atomicAdd(&A.g_em[k - 5 * A.nmats], v);
    }
  }
}

// Running mean over the pass's samples in sample order
// (combine_color_samples, src/util-kernels.cu:44-51).  A block takes 256
// pixels, one per thread.  The slots are pixel-major, so the block's samples
// are read kAccJ at a time into LDS with line-sized loads (consecutive lanes
// read a pixel's consecutive samples) and each thread then folds its pixel's
// kAccJ samples in order from LDS (rows padded by one float4 against bank
// conflicts).  Blocks must be 256 threads.
constexpr int kAccJ = 8;
__global__ __launch_bounds__(256) void k_accumulate(const PassDesc *__restrict__ pass,
                                                    const DevView *__restrict__ views,
                                                    const float4 *__restrict__ path_l) {
  __shared__ float4 tile[256 * (kAccJ + 1)];
  __shared__ uint32_t s_slot[256];
  __shared__ int32_t s_k[256];
  __shared__ int32_t s_kmax;
  const uint32_t total = pass->total_pix, t = threadIdx.x;
  for (uint32_t g0 = blockIdx.x * 256u; g0 < total; g0 += gridDim.x * 256u) {
    const uint32_t g = g0 + t;
    const bool live = g < total;
    int c = 0;
    if (live)
      while (c + 1 < pass->nchunks && g >= pass->chunk[c + 1].pix_base) ++c;
    const PassChunk &ck = pass->chunk[c];
    const uint32_t p = live ? g - ck.pix_base : 0u;
    if (t == 0) s_kmax = 0;
    s_slot[t] = live ? ck.slot_base + p * (uint32_t)ck.k : 0u;  // pixel-major slots
    s_k[t] = live ? ck.k : 0;
    __syncthreads();
    if (live) atomicMax(&s_kmax, ck.k);
    float4 acc = live ? views[ck.view].image[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    const int kmax = s_kmax;
    for (int j0 = 0; j0 < kmax; j0 += kAccJ) {
#pragma unroll
      for (int i = 0; i < kAccJ; ++i) {
        const uint32_t e = (uint32_t)i * 256u + t, q = e / kAccJ, j = e % kAccJ;
        if (j0 + (int)j < s_k[q]) tile[q * (kAccJ + 1) + j] = path_l[s_slot[q] + (uint32_t)j0 + j];
      }
      __syncthreads();
      if (live) {
        for (int j = j0; j < min(j0 + kAccJ, ck.k); ++j) {
          const float4 l = tile[t * (kAccJ + 1) + (uint32_t)(j - j0)];
          const float sj = float(ck.s0 + j), s1 = float(ck.s0 + j + 1);
          acc.x = (sj * acc.x + l.x) / s1;
          acc.y = (sj * acc.y + l.y) / s1;
          acc.z = (sj * acc.z + l.z) / s1;
        }
      }
      __syncthreads();
    }
    if (live) views[ck.view].image[p] = acc;
  }
}

// Per-pass queue counters: count[0] = all slots, everything else zero.
// A kernel rather than a memset so the sequence can sit in a hipGraph.
__global__ void k_pass_setup(const PassDesc *__restrict__ pass, uint32_t *__restrict__ counters, int32_t n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) counters[i] = (i == 0) ? pass->total_slots : 0u;
}

__global__ void k_fill4(float4 *__restrict__ p, uint32_t n, float4 v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_unpack_image(const float4 *__restrict__ img, float *__restrict__ rgb, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 v = img[i];
  rgb[3 * i] = v.x;
  rgb[3 * i + 1] = v.y;
  rgb[3 * i + 2] = v.z;
}

// ------------------------------------------------------------- buffers ----
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      bytes = 0;
      if (e != hipSuccess) return e;
    }
    hipError_t e = hipMalloc(&p, std::max<size_t>(n, 256));
    if (e == hipSuccess) bytes = std::max<size_t>(n, 256);
    return e;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

}  // namespace

// ======================================================== no-alloc scratch ==
// Scratch memory of gi_trace_closest_device_noalloc, for callers that cannot
// allocate device memory: a file-scope device array that every GPU the
// library runs on holds until the process ends.  Layout (bytes):
//   [0, 256)                        the traversal-stack overflow counter
//   [256, 256 + 32 * kNoAllocRays)  the packed rays: ray_o then ray_d, float4
//   then kNoAllocStackWords words   the global part of the traversal stacks
// Host code never names the array in a launch or copy (in host code the
// name is a stand-in, not a device address): gi_create resolves its address
// on each device once, with hipGetSymbolAddress, and launches use that.
// Calls that overlap on one GPU share the array and race.
constexpr uint32_t kNoAllocRays = 1u << 19;                 // 524288 rays per call
constexpr size_t kNoAllocStackWords = (size_t)4 << 20;      // 16 MiB of stack columns
constexpr size_t kNoAllocRayBytes = (size_t)kNoAllocRays * 2 * sizeof(float4);
constexpr size_t kScratchBytes = 256 + kNoAllocRayBytes + kNoAllocStackWords * sizeof(uint32_t);
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The scratch array's address on `device` (current), resolved once per
// device.
hipError_t scratch_address(int device, void **out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}
}  // namespace

// ================================================================ context ==
struct gi_ctx {
  int device = 0;
  int num_cus = 256;
  size_t lds_per_cu = 160 * 1024;
  hipStream_t stream = nullptr;
  std::string err;
  // scene
  std::vector<gi_tri> tris;
  std::vector<gi_material> mats;
  std::vector<gi_light> lights;
  std::unique_ptr<gi_bvh> bvh;
  int64_t ntri = 0, nnodes = 0;
  int32_t wide_depth = 0;
  bool have_bvh = false;
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, diag = 0.f, power_total = 0.f;
  float clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};  // bounds of the triangle centroids
  DBuf d_stris, d_itris, d_nodes, d_hnodes, d_mats, d_lights;  // d_hnodes: HalfNode copy of d_nodes
  DBuf d_ks;  // specular colour per material (float4), for GI_FLAG_PHONG
  DBuf d_lsum;  // running sums of the light powers (n_lights + 1 floats)
  // Halton tables
  uint32_t hal_seed = 0;
  int32_t hal_dims = 0;
  DBuf d_hdims, d_hchunks;
  // traversal stack + overflow flag
  DBuf d_gstack, d_overflow;
  // traversal statistics (gi_set_traversal_stats): 8 counters
  bool count_traversal = false;
  DBuf d_tstats;
  DBuf d_sky_rgb, d_sky_x, d_sky_y;  // sky map, per-row pdf|cdf, row pdf|cdf
  // instancing: meshes (object space), instances, the instance tree
  std::vector<std::vector<gi_tri>> meshes;
  DBuf d_inst, d_inst_vbase, d_inst_vtab;
  uint32_t inst_vshift = 0;
  bool use_vtab = false;
  // textures: one buffer per texture (level 0 + mip chain), their records
  struct TexHost {
    int32_t w, h, levels;
    size_t texels;
  };
  std::vector<DBuf> tex_bufs;
  std::vector<TexHost> texs;
  DBuf d_tex;
  bool textured = false;  // some material names a texture
  DBuf d_gtex, d_gtex_off, d_grep;            // adjoint: texel gradients of the chains, coarse replicas
  std::vector<std::vector<float>> tex_grad;   // level-0 texel gradients of the last adjoint
  int64_t world_nnodes = 0;
  int32_t world_depth = 0, n_inst = 0;
  // range of the material indices shading can reach: world triangles, and
  // with instances the instanced meshes' triangles and the overrides
  int32_t mat_lo = 0, mat_hi = 0, inst_mat_lo = 0, inst_mat_hi = 0;
  uint32_t tlas_root = 0;
  int32_t sky_light = -1, sky_w = 0, sky_h = 0;
  uint32_t stack_cap = 0;
  int32_t trace_grid = 0;
  // render state
  std::vector<gi_view> views;
  gi_render_opts opts{};
  std::vector<DBuf> images;
  DBuf d_views, d_passes, d_counters;
  DBuf d_cam_view, d_cam_pass;  // gi_generate_camera_rays' own view and pass
  // the executable graph of the last GI_FLAG_GRAPH render and its launch key
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uint8_t> graph_key;
  int graph_launches = 0;
  DBuf d_matgrad, d_rec, d_dl, d_gmats, d_gem;
  DBuf q_o[2], q_d[2], q_tp[2], q_hit, s_o, s_d, s_c, path_l;
  DBuf rdiff;  // 4 planes of float4 per slot: ray differentials (textures in use)
  size_t slot_cap = 0;
  // primitives
  DBuf p_orig, p_dir, p_maxt, p_ro, p_rd, p_hit, p_occ;
  // progressive
  int32_t prog_sample = 0, prog_max = 0;
  bool prog_active = false;
  gi_counters counters{};
  gi_tuning tuning{256, 6, 256, 0, 16, 0, 4, 32};
  std::vector<hipEvent_t> events;
  uint8_t *scratch = nullptr;  // this device's scratch array (gi_create)
};

namespace {

#define GI_HIP(ctx, expr)                                                          \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);              \
      return e_ == hipErrorOutOfMemory ? GI_ERR_OUT_OF_MEMORY : GI_ERR_HIP;        \
    }                                                                              \
  } while (0)


// Every copy goes through the context stream and is synchronised: the
// stream is non-blocking (not ordered after the null stream), and a pageable
// host-to-device hipMemcpy may return before its DMA has landed.
hipError_t copy_sync(gi_ctx *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  if (bytes == 0) return hipSuccess;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(ctx->stream);
}

int fail(gi_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

float rect_power(const gi_light &l) {
  // add-rectlight, src/lights.cpp:92
  return l.wh[0] * l.wh[1] * (l.col[0] + l.col[1] + l.col[2]) * .333f;
}

// host copies of the device tangent frame (same formula, same rounding)
void host_frame(const float *n, float *t, float *b) {
  float sign = std::copysign(1.0f, n[2]);
  float a = -1.0f / (sign + n[2]);
  float bb = n[0] * n[1] * a;
  float T[3] = {1.0f + sign * n[0] * n[0] * a, sign * bb, -sign * n[0]};
  float B[3] = {bb, sign + n[1] * n[1] * a, -n[1]};
  float it = 1.0f / std::sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]);
  float ib = 1.0f / std::sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
  for (int k = 0; k < 3; ++k) {
    t[k] = T[k] * it;
    b[k] = B[k] * ib;
  }
}
void host_tangential(const float *v, const float *n, float *out) {
  float t[3], b[3];
  host_frame(n, t, b);
  for (int k = 0; k < 3; ++k) out[k] = (t[k] * v[0] + b[k] * v[1]) + n[k] * v[2];
}

DevMat make_devmat(const gi_material &m) {
  DevMat d;
  std::memset(&d, 0, sizeof(d));
  for (int k = 0; k < 3; ++k) d.diffuse[k] = m.diffuse_color[k];
  d.principled = m.principled;
  float cd[3];
  for (int k = 0; k < 3; ++k) cd[k] = std::pow(m.color[k], 2.2f);  // mon2lin
  float lum = 0.3f * cd[0] + 0.6f * cd[1] + 0.1f * cd[2];
  float tint[3];
  for (int k = 0; k < 3; ++k) tint[k] = lum > 0.f ? cd[k] / lum : 1.f;
  const float sp = m.specular * 0.08f;
  for (int k = 0; k < 3; ++k) {
    float mt = (1.0f - m.specular_tint) * 1.0f + m.specular_tint * tint[k];
    float x = sp * mt;
    d.cspec0[k] = (1.0f - m.metallic) * x + m.metallic * cd[k];
    d.csheen[k] = (1.0f - m.sheen_tint) * 1.0f + m.sheen_tint * tint[k];
    d.cdlin[k] = cd[k];
    d.color[k] = m.color[k];
  }
  d.metallic = m.metallic;
  d.roughness = m.roughness;
  d.sheen = m.sheen;
  d.subsurface = m.subsurface;
  d.clearcoat = m.clearcoat;
  d.cc_a = (1.0f - m.clearcoat_gloss) * 0.1f + m.clearcoat_gloss * 0.001f;
  d.cc_log_a2 = std::log(d.cc_a * d.cc_a);
  d.opacity = m.opacity;
  d.specular = m.specular;
  d.texture = m.texture;
  d.spec_tint = m.specular_tint;
  d.sheen_tint = m.sheen_tint;
  float aspect = std::sqrt(1.0f - m.anisotropic * 0.9f);
  float r2 = m.roughness * m.roughness;
  d.ax = std::max(0.001f, r2 / aspect);
  d.ay = std::max(0.001f, r2 * aspect);
  float rg = d.ax * d.ay * 0.5f + 0.5f;
  rg = rg * rg;
  d.roughg2 = rg * rg;
  return d;
}

// Host-side quantities of the adjoint's material derivatives.
DevMatGrad make_devmatgrad(const gi_material &m) {
  DevMatGrad g;
  std::memset(&g, 0, sizeof(g));
  float cd[3];
  for (int k = 0; k < 3; ++k) {
    cd[k] = std::pow(m.color[k], 2.2f);
    g.dcd[k] = m.color[k] > 0.f ? 2.2f * std::pow(m.color[k], 1.2f) : 0.f;
  }
  if (!m.principled)
    for (int k = 0; k < 3; ++k) g.dcd[k] = 1.0f;  // Lambert: the parameter is rho itself
  g.lum = 0.3f * cd[0] + 0.6f * cd[1] + 0.1f * cd[2];
  for (int k = 0; k < 3; ++k) g.tint[k] = g.lum > 0.f ? cd[k] / g.lum : 1.f;
  g.spec_tint = m.specular_tint;
  g.sheen_tint = m.sheen_tint;
  g.aspect = std::sqrt(1.0f - m.anisotropic * 0.9f);
  return g;
}

int ensure_halton(gi_ctx *ctx, uint32_t seed, int32_t ndims) {
  if (ctx->hal_dims >= ndims && ctx->hal_seed == seed && ctx->d_hdims.p) return GI_OK;
  std::vector<HaltonDim> dims;
  std::vector<uint16_t> perms;  // chunk tables
  halton_tables(seed, ndims, dims, perms);
  // ray generation and each shading launch copy the tables of their
  // dimensions into LDS: kDimVertex0, then kDimsPerVertex at a time
  for (int32_t d0 = 0; d0 < ndims;) {
    const int32_t nd = std::min(d0 == 0 ? kDimVertex0 : kDimsPerVertex, ndims - d0);
    if (dims[d0 + nd - 1].chunk_off + dims[d0 + nd - 1].P - dims[d0].chunk_off > kHaltonLdsEntries)
      return fail(ctx, GI_ERR_LIMIT,
                  "sampler tables of dimensions " + std::to_string(d0) + ".." + std::to_string(d0 + nd - 1) +
                      " exceed the LDS copy");
    d0 += nd;
  }
  GI_HIP(ctx, ctx->d_hdims.ensure(dims.size() * sizeof(HaltonDim)));
  GI_HIP(ctx, ctx->d_hchunks.ensure(perms.size() * sizeof(uint16_t)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_hdims.p, dims.data(), dims.size() * sizeof(HaltonDim), hipMemcpyHostToDevice));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_hchunks.p, perms.data(), perms.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  ctx->hal_seed = seed;
  ctx->hal_dims = ndims;
  return GI_OK;
}

int ensure_gstack(gi_ctx *ctx);

ShadeTri make_shadetri(const gi_tri &t) {
  ShadeTri s;
  int32_t mi = t.material_index;
  float mf;
  std::memcpy(&mf, &mi, 4);
  s.a = make_float4(t.a[0], t.a[1], t.a[2], mf);
  s.b = make_float4(t.b[0], t.b[1], t.b[2], 0.f);
  s.c = make_float4(t.c[0], t.c[1], t.c[2], 0.f);
  s.na = make_float4(t.na[0], t.na[1], t.na[2], 0.f);
  s.nb = make_float4(t.nb[0], t.nb[1], t.nb[2], 0.f);
  s.nc = make_float4(t.nc[0], t.nc[1], t.nc[2], 0.f);
  s.tab = make_float4(t.ta[0], t.ta[1], t.tb[0], t.tb[1]);
  s.tcm = make_float4(t.tc[0], t.tc[1], 0.f, 0.f);
  return s;
}

InstScene inst_scene(const gi_ctx *ctx) {
  return InstScene{ctx->d_inst.as<DevInstance>(), ctx->d_inst_vbase.as<uint32_t>(), ctx->tlas_root,
                   (uint32_t)ctx->ntri, ctx->n_inst, ctx->use_vtab ? ctx->d_inst_vtab.as<uint32_t>() : nullptr,
                   ctx->inst_vshift};
}
bool instanced(const gi_ctx *ctx) { return ctx->n_inst > 0; }

// ---- half-precision traversal nodes (HalfNode) from the wide nodes.
// Per node: the origin is the min corner of its children's boxes, the
// margin on each axis 2^-18 of that axis's largest coordinate magnitude,
// and the scale s the power of two that brings the node's extent plus
// margins below 2^14 (f16's largest finite value is 65504; being floating
// point, f16 keeps small offsets precise at any s).  Nodes keep their
// index, so child references and the 128-byte addressing carry over.  Each plane offset is
// computed in double ((bound -/+ margin - origin) / s, exact up to 2^-53 of
// the coordinates, far inside the margin) and rounded outward to f16.
__device__ __forceinline__ uint16_t f16_step(uint16_t b, bool up) {
  if ((b & 0x7fffu) == 0u) return up ? (uint16_t)0x0001u : (uint16_t)0x8001u;  // from +-0
  if ((b & 0x7fffu) == 0x7c00u) return b;                                       // infinities stay
  return (up != (bool)(b & 0x8000u)) ? (uint16_t)(b + 1u) : (uint16_t)(b - 1u);
}
__device__ __forceinline__ uint16_t f16_outward(double x, bool up) {
  const _Float16 h = (_Float16)(float)x;
  const uint16_t b = __builtin_bit_cast(uint16_t, h);
  const double v = (double)(float)h;
  return (up ? v < x : v > x) ? f16_step(b, up) : b;
}
__device__ __forceinline__ double f16_value(uint32_t h2, int half) {
  return (double)(float)__builtin_bit_cast(_Float16, (uint16_t)(half ? h2 >> 16 : h2 & 0xffffu));
}

struct HalfFrame {
  double org[3], margin[3], scale;
};
__device__ __forceinline__ HalfFrame half_frame(const WideNode &w) {
  const float *lo[3] = {w.lox, w.loy, w.loz}, *hi[3] = {w.hix, w.hiy, w.hiz};
  HalfFrame f;
  double ext = 0.0;
  for (int a = 0; a < 3; ++a) {
    double mn = INFINITY, mx = -INFINITY;
    for (int k = 0; k < 4; ++k)
      if (lo[a][k] != INFINITY) {
        mn = fmin(mn, (double)lo[a][k]);
        mx = fmax(mx, (double)hi[a][k]);
      }
    if (!(mn <= mx)) mn = mx = 0.0;  // no child (a lone root without children)
    f.org[a] = mn;                   // a float: the origin is stored exactly
    f.margin[a] = ldexp(fmax(fabs(mn), fabs(mx)), -18);
    ext = fmax(ext, mx - mn + 2.0 * f.margin[a]);
  }
  int e = 0;
  if (ext > 0.0) frexp(ext, &e);  // ext < 2^e
  // 2^-60..2^40: scale/d then neither overflows nor underflows for the
  // inverse directions rcp_clamped gives (|1/d| <= 2^80) and any |d| below
  // 2^89; a node wider than 2^54 gets +inf hi planes (still conservative)
  f.scale = ldexp(1.0, min(max(e - 14, -60), 40));
  return f;
}

__global__ void k_half_nodes(const WideNode *__restrict__ in, HalfNode *__restrict__ out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const WideNode w = in[i];
  const HalfFrame f = half_frame(w);
  const float *lo[3] = {w.lox, w.loy, w.loz}, *hi[3] = {w.hix, w.hiy, w.hiz};
  uint32_t q[3][4];
  for (int a = 0; a < 3; ++a) {
    uint32_t ql[4], qh[4];
    for (int k = 0; k < 4; ++k) {
      if (lo[a][k] == INFINITY) {  // empty slot
        ql[k] = qh[k] = 0x7c00u;
      } else {
        ql[k] = f16_outward(((double)lo[a][k] - f.margin[a] - f.org[a]) / f.scale, false);
        qh[k] = f16_outward(((double)hi[a][k] + f.margin[a] - f.org[a]) / f.scale, true);
      }
    }
    q[a][0] = ql[0] | (ql[1] << 16);
    q[a][1] = ql[2] | (ql[3] << 16);
    q[a][2] = qh[0] | (qh[1] << 16);
    q[a][3] = qh[2] | (qh[3] << 16);
  }
  float4 *o = reinterpret_cast<float4 *>(out + i);
  o[0] = make_float4((float)f.org[0], (float)f.org[1], (float)f.org[2], (float)f.scale);
  for (int a = 0; a < 3; ++a)
    o[1 + a] = make_float4(__uint_as_float(q[a][0]), __uint_as_float(q[a][1]), __uint_as_float(q[a][2]),
                           __uint_as_float(q[a][3]));
  o[4] = make_float4(__uint_as_float(w.child[0]), __uint_as_float(w.child[1]), __uint_as_float(w.child[2]),
                     __uint_as_float(w.child[3]));
  o[5] = o[6] = o[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Checks every half node against its wide node: same references; an empty
// slot is +inf on every plane; otherwise the decoded planes org + q * s
// (exact in double) enclose the f32 box widened by the node's margins.
// Counts the failing child slots.
__global__ void k_check_half_nodes(const WideNode *__restrict__ in, const HalfNode *__restrict__ hn, uint32_t n,
                                   unsigned long long *bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const WideNode w = in[i];
  const HalfNode h = hn[i];
  const HalfFrame f = half_frame(w);
  const float *lo[3] = {w.lox, w.loy, w.loz}, *hi[3] = {w.hix, w.hiy, w.hiz};
  uint32_t nbad = 0;
  for (int k = 0; k < 4; ++k) {
    bool ok = h.child[k] == w.child[k];
    for (int a = 0; a < 3; ++a) {
      const double ql = f16_value(h.q[a][k >> 1], k & 1), qh = f16_value(h.q[a][2 + (k >> 1)], k & 1);
      if (lo[a][k] == INFINITY) {
        ok = ok && ql == INFINITY && qh == INFINITY;
      } else {
        const double dl = (double)h.org[a] + ql * (double)h.scale, dh = (double)h.org[a] + qh * (double)h.scale;
        ok = ok && (double)h.scale == f.scale && dl <= (double)lo[a][k] - f.margin[a] &&
             dh >= (double)hi[a][k] + f.margin[a];
      }
    }
    nbad += ok ? 0u : 1u;
  }
  if (nbad) atomicAdd(bad, (unsigned long long)nbad);
}

// Builds the traversal's half-node copy of the current wide tree.
int make_half_nodes(gi_ctx *ctx) {
  GI_HIP(ctx, ctx->d_hnodes.ensure((size_t)std::max<int64_t>(1, ctx->nnodes) * sizeof(HalfNode)));
  if (ctx->nnodes > 0) {
    hipLaunchKernelGGL(k_half_nodes, dim3((unsigned)((ctx->nnodes + 255) / 256)), dim3(256), 0, ctx->stream,
                       ctx->d_nodes.as<WideNode>(), ctx->d_hnodes.as<HalfNode>(), (uint32_t)ctx->nnodes);
    GI_HIP(ctx, hipGetLastError());
  }
  GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GI_OK;
}

// State that follows a new wide tree on the device.
int finish_bvh(gi_ctx *ctx, int64_t nnodes, int32_t wide_depth) {
  if ((uint64_t)nnodes > kMaxWideNodes) {
    ctx->have_bvh = false;
    return fail(ctx, GI_ERR_LIMIT, "more than 2^25 BVH nodes");
  }
  ctx->nnodes = ctx->world_nnodes = nnodes;
  ctx->wide_depth = ctx->world_depth = wide_depth;
  ctx->n_inst = 0;  // a new world tree drops the instances (gi_upload_instances again)
  // at most three pushes per wide level on the current path
  ctx->stack_cap = (uint32_t)(3 * (wide_depth + 1) + 4);
  ctx->trace_grid = ctx->num_cus * std::max(1, ctx->tuning.trace_blocks_per_cu);
  int rc = ensure_gstack(ctx);
  if (rc) return rc;
  rc = make_half_nodes(ctx);
  if (rc) return rc;
  ctx->have_bvh = true;
  return GI_OK;
}

int upload_wide(gi_ctx *ctx) {
  std::vector<WideNode> wn;
  int32_t wd = 0;
  collapse_wide_nodes(
      *ctx->bvh, [](const Bvh2Node &n) { return kLeafBit | ((uint32_t)n.count << 27) | (uint32_t)n.first; }, 0u, wn,
      wd, kCollapseMergeLimit, ctx->tuning.node_width);
  GI_HIP(ctx, ctx->d_nodes.ensure(wn.size() * sizeof(WideNode)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_nodes.p, wn.data(), wn.size() * sizeof(WideNode), hipMemcpyHostToDevice));
  // the leaf-ordered intersection triangles are gathered on the device from
  // the triangle table through the permutation (16 MB up for 4 M triangles,
  // instead of 192 MB of gathered records)
  const std::vector<uint32_t> &perm = ctx->bvh->perm;
  GI_HIP(ctx, ctx->d_itris.ensure(std::max<size_t>(1, perm.size()) * sizeof(IsectTri)));
  if (!perm.empty()) {
    DBuf d_perm;
    GI_HIP(ctx, d_perm.ensure(perm.size() * sizeof(uint32_t)));
    GI_HIP(ctx, copy_sync(ctx, d_perm.p, perm.data(), perm.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    GI_HIP(ctx, launch_leaf_tris(ctx->stream, ctx->d_stris.as<float4>(), (uint32_t)(sizeof(ShadeTri) / sizeof(float4)),
                                 (uint32_t)perm.size(), d_perm.as<uint32_t>(), ctx->d_itris.as<float4>()));
    GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return finish_bvh(ctx, (int64_t)wn.size(), wd);
}

// GI_OK, or GI_ERR_LIMIT when some traversal stack overflowed since the
// counter was last cleared, or the copy's own error.
int check_overflow(gi_ctx *ctx) {
  uint32_t v = 0;
  GI_HIP(ctx, copy_sync(ctx, &v, ctx->d_overflow.p, 4, hipMemcpyDeviceToHost));
  return v != 0 ? fail(ctx, GI_ERR_LIMIT, "traversal stack overflow") : GI_OK;
}

// Global part of the traversal stacks: one column per thread of the
// largest (persistent) traversal grid.  Allocated before any launch or
// graph capture; launches never exceed that grid.
int ensure_gstack(gi_ctx *ctx) {
  const int block = std::max(64, (std::min(256, ctx->tuning.trace_block) / 64) * 64);
  const size_t threads = (size_t)ctx->num_cus * (size_t)std::max(1, ctx->tuning.trace_blocks_per_cu) * block;
  const uint32_t gdepth = ctx->stack_cap > (uint32_t)kMinLdsDepth ? ctx->stack_cap - kMinLdsDepth : 1;
  GI_HIP(ctx, ctx->d_gstack.ensure(threads * gdepth * sizeof(uint32_t)));
  return GI_OK;
}

int stack_lds_of(const gi_ctx *ctx);
bool lds_scene_of(const gi_ctx *ctx);
size_t trace_lds_bytes(int lds_depth, int block);

// Traversal blocks per CU that are resident at once: the tuned count,
// capped by the waves the register budget allows (GI_TRACE_WAVES_PER_EU per
// SIMD, 4 SIMDs) and by the LDS a block takes (its stack, plus the scene
// copy on the small-scene path), so that the persistent grid never exceeds
// what the chip holds at once (a block that had to wait for another to
// finish would leave its chunks of rays to the end of the launch).
int resident_trace_blocks(const gi_ctx *ctx, int block) {
  const int by_waves = GI_TRACE_WAVES_PER_EU * 4 / std::max(1, block / 64);
  const bool lds_scene = !ctx->count_traversal && stack_lds_of(ctx) == 16 && lds_scene_of(ctx);
  const size_t lds = trace_lds_bytes(ctx->count_traversal ? 16 : stack_lds_of(ctx), block) +
                     (lds_scene ? (size_t)ctx->nnodes * sizeof(WideNode) + (size_t)ctx->ntri * sizeof(IsectTri) : 0);
  const int by_lds = (int)std::min<size_t>(64, ctx->lds_per_cu / std::max<size_t>(1, lds));
  return std::max(1, std::min({ctx->tuning.trace_blocks_per_cu, by_waves, by_lds}));
}

int trace_grid_for(gi_ctx *ctx, uint64_t n, int &grid, int &block) {
  block = std::max(64, (std::min(256, ctx->tuning.trace_block) / 64) * 64);
  const int max_grid = ctx->num_cus * resident_trace_blocks(ctx, block);
  uint64_t need = (n + (uint64_t)block - 1) / (uint64_t)block;
  grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(need, (uint64_t)max_grid));
  return GI_OK;
}

int check_scene(gi_ctx *ctx, bool need_shading) {
  if (!ctx) return GI_ERR_INVALID_ARG;
  if (!ctx->have_bvh || ctx->ntri <= 0) return fail(ctx, GI_ERR_STATE, "no triangles / BVH uploaded");
  if (need_shading) {
    if (ctx->mats.empty()) return fail(ctx, GI_ERR_STATE, "no materials uploaded");
    if (ctx->lights.empty()) return fail(ctx, GI_ERR_STATE, "no lights uploaded");
    if (ctx->sky_light >= 0 && ctx->sky_w <= 0) return fail(ctx, GI_ERR_STATE, "sky light without a map (gi_upload_sky)");
    int32_t lo = ctx->mat_lo, hi = ctx->mat_hi;
    if (instanced(ctx)) {
      lo = std::min(lo, ctx->inst_mat_lo);
      hi = std::max(hi, ctx->inst_mat_hi);
    }
    if (lo < 0 || hi >= (int32_t)ctx->mats.size())
      return fail(ctx, GI_ERR_STATE, "a triangle or instance material index is outside the uploaded materials");
    for (const gi_material &m : ctx->mats)
      if (m.texture < -1 || m.texture >= (int32_t)ctx->texs.size())
        return fail(ctx, GI_ERR_STATE, "a material names a texture that was not uploaded (gi_upload_texture)");
  }
  return GI_OK;
}

// LDS bytes of a traversal block: kLds entries per lane, lane stride the
// next power of two >= block.
size_t trace_lds_bytes(int lds_depth, int block) {
  return (size_t)lds_depth * ((size_t)1 << lds_shift_for((uint32_t)block)) * sizeof(uint32_t);
}

InstScene inst_scene(const gi_ctx *ctx);

// log2 of the traversal kernels' refill threshold (gi_tuning.refill_lanes)
uint32_t chunk_shift_of(const gi_ctx *ctx) {
  const int r = ctx->tuning.refill_lanes;
  return r == 8 ? 3u : r == 16 ? 4u : r == 64 ? 6u : 5u;
}

// The world tree and its triangles fit the LDS copy (non-instanced scenes).
bool lds_scene_of(const gi_ctx *ctx) {
  return !instanced(ctx) &&
         (size_t)ctx->nnodes * sizeof(WideNode) + (size_t)ctx->ntri * sizeof(IsectTri) <= kLdsSceneMaxBytes;
}

// The global part of the traversal stacks and the overflow counter a launch
// uses: the context's own, or (gi_trace_closest_device_noalloc) the
// library's scratch array with the grid capped to the stack columns it holds.
struct TraceScratch {
  uint32_t *gstack = nullptr;
  uint32_t *overflow = nullptr;
  int max_grid = 0;  // 0: no cap beyond the persistent grid
};

template <bool kHasTmax, bool kInst, int kLds, bool kCount = false, bool kLdsScene = false>
void launch_closest_t(gi_ctx *ctx, int grid, int block, const uint32_t *count_ptr, uint32_t count_host,
                      const float4 *ro, const float4 *rd, float4 *hits, const TraceScratch &ts) {
  const size_t stack_bytes = trace_lds_bytes(kLds, block);
  const size_t scene_bytes =
      kLdsScene ? (size_t)ctx->nnodes * sizeof(WideNode) + (size_t)ctx->ntri * sizeof(IsectTri) : 0;
  hipLaunchKernelGGL((k_trace_closest<kHasTmax, kInst, kLds, kCount, kLdsScene>), dim3(grid), dim3(block),
                     stack_bytes + scene_bytes, ctx->stream, count_ptr, count_host, ro, rd, hits,
                     ctx->d_hnodes.as<HalfNode>(), ctx->d_nodes.as<WideNode>(), ctx->d_itris.as<float4>(),
                     ts.gstack ? ts.gstack : ctx->d_gstack.as<uint32_t>(), ctx->stack_cap,
                     ts.overflow ? ts.overflow : ctx->d_overflow.as<uint32_t>(), inst_scene(ctx),
                     kCount ? ctx->d_tstats.as<unsigned long long>() : nullptr, chunk_shift_of(ctx),
                     (uint32_t)(stack_bytes / 4), (uint32_t)ctx->nnodes, (uint32_t)ctx->ntri);
}

template <bool kDeposit, bool kInst, int kLds, bool kCount = false, bool kLdsScene = false>
void launch_any_t(gi_ctx *ctx, int grid, int block, const uint32_t *count_ptr, uint32_t count_host,
                  const float4 *ro, const float4 *rd, const float4 *contrib, float4 *pl, uint8_t *occ, float4 *rec,
                  uint32_t rec_stride, int32_t b) {
  const size_t stack_bytes = trace_lds_bytes(kLds, block);
  const size_t scene_bytes =
      kLdsScene ? (size_t)ctx->nnodes * sizeof(WideNode) + (size_t)ctx->ntri * sizeof(IsectTri) : 0;
  hipLaunchKernelGGL((k_trace_any<kDeposit, kInst, kLds, kCount, kLdsScene>), dim3(grid), dim3(block),
                     stack_bytes + scene_bytes, ctx->stream, count_ptr, count_host, ro, rd, contrib, pl, occ,
                     ctx->d_hnodes.as<HalfNode>(), ctx->d_nodes.as<WideNode>(), ctx->d_itris.as<float4>(),
                     ctx->d_gstack.as<uint32_t>(),
                     ctx->stack_cap, ctx->d_overflow.as<uint32_t>(), rec, rec_stride, b, inst_scene(ctx),
                     kCount ? ctx->d_tstats.as<unsigned long long>() + 4 : nullptr, chunk_shift_of(ctx),
                     (uint32_t)(stack_bytes / 4), (uint32_t)ctx->nnodes, (uint32_t)ctx->ntri);
}

int stack_lds_of(const gi_ctx *ctx) {
  const int s = ctx->tuning.stack_lds;
  return s == 8 || s == 32 ? s : 16;
}

bool instanced(const gi_ctx *ctx);

int launch_closest(gi_ctx *ctx, bool has_tmax, const uint32_t *count_ptr, uint32_t count_host, uint64_t n_bound,
                   const float4 *ro, const float4 *rd, float4 *hits, const TraceScratch &ts = TraceScratch{}) {
  int grid, block;
  int rc = trace_grid_for(ctx, n_bound, grid, block);
  if (rc) return rc;
  if (ts.max_grid > 0) grid = std::min(grid, ts.max_grid);
  const int L = stack_lds_of(ctx);
  const bool I = instanced(ctx);
#define GI_CL(T, II, LL) launch_closest_t<T, II, LL>(ctx, grid, block, count_ptr, count_host, ro, rd, hits, ts)
#define GI_CL_L(T, II) \
  if (ctx->count_traversal) \
    launch_closest_t<T, II, 16, true>(ctx, grid, block, count_ptr, count_host, ro, rd, hits, ts); \
  else if (!(II) && L == 16 && lds_scene_of(ctx)) \
    launch_closest_t<T, false, 16, false, true>(ctx, grid, block, count_ptr, count_host, ro, rd, hits, ts); \
  else if (L == 8) GI_CL(T, II, 8); \
  else if (L == 32) GI_CL(T, II, 32); \
  else GI_CL(T, II, 16)
  if (has_tmax) {
    if (I) { GI_CL_L(true, true); } else { GI_CL_L(true, false); }
  } else {
    if (I) { GI_CL_L(false, true); } else { GI_CL_L(false, false); }
  }
#undef GI_CL_L
#undef GI_CL
  GI_HIP(ctx, hipGetLastError());
  return GI_OK;
}

int launch_any(gi_ctx *ctx, bool deposit, const uint32_t *count_ptr, uint32_t count_host, uint64_t n_bound,
               const float4 *ro, const float4 *rd, const float4 *contrib, float4 *pl, uint8_t *occ,
               float4 *rec = nullptr, uint32_t rec_stride = 0, int32_t b = 0) {
  int grid, block;
  int rc = trace_grid_for(ctx, n_bound, grid, block);
  if (rc) return rc;
  const int L = stack_lds_of(ctx);
  const bool I = instanced(ctx);
#define GI_ANY(D, II, LL) \
  launch_any_t<D, II, LL>(ctx, grid, block, count_ptr, count_host, ro, rd, contrib, pl, occ, rec, rec_stride, b)
#define GI_ANY_L(D, II) \
  if (ctx->count_traversal) \
    launch_any_t<D, II, 16, true>(ctx, grid, block, count_ptr, count_host, ro, rd, contrib, pl, occ, rec, rec_stride, \
                                  b); \
  else if (!(II) && L == 16 && lds_scene_of(ctx)) \
    launch_any_t<D, false, 16, false, true>(ctx, grid, block, count_ptr, count_host, ro, rd, contrib, pl, occ, rec, \
                                            rec_stride, b); \
  else if (L == 8) GI_ANY(D, II, 8); \
  else if (L == 32) GI_ANY(D, II, 32); \
  else GI_ANY(D, II, 16)
  if (deposit) {
    if (I) { GI_ANY_L(true, true); } else { GI_ANY_L(true, false); }
  } else {
    if (I) { GI_ANY_L(false, true); } else { GI_ANY_L(false, false); }
  }
#undef GI_ANY_L
#undef GI_ANY
  GI_HIP(ctx, hipGetLastError());
  return GI_OK;
}

// ---------------------------------------------------------------- render --
struct PassPlan {
  std::vector<PassDesc> passes;
  size_t max_slots = 0;
};

// Pack (view, sample range) chunks into passes of at most `target` slots.
// A pass holds at most one chunk per view (the accumulate kernel updates
// each pixel once per pass, in sample order).  batched: views advance
// round-robin and chunks of different views share passes; otherwise every
// pass holds one chunk and views are rendered one after the other (the
// per-view wavefront).
void plan_passes(const std::vector<gi_view> &views, const std::vector<int32_t> &s_begin,
                 const std::vector<int32_t> &s_end, bool batched, size_t target, int32_t spp_per_pass,
                 PassPlan &plan) {
  plan.passes.clear();
  plan.max_slots = 0;
  PassDesc cur;
  std::memset(&cur, 0, sizeof(cur));
  auto flush = [&]() {
    if (cur.nchunks == 0) return;
    plan.max_slots = std::max<size_t>(plan.max_slots, cur.total_slots);
    plan.passes.push_back(cur);
    std::memset(&cur, 0, sizeof(cur));
  };
  auto holds = [&](int32_t v) {
    for (int c = 0; c < cur.nchunks; ++c)
      if (cur.chunk[c].view == v) return true;
    return false;
  };
  std::vector<int32_t> s(s_begin);
  auto add = [&](size_t v) {
    const size_t npix = (size_t)views[v].w * (size_t)views[v].h;
    int32_t k = spp_per_pass > 0 ? spp_per_pass : (int32_t)std::max<size_t>(1, target / std::max<size_t>(1, npix));
    k = std::min(k, s_end[v] - s[v]);
    const size_t slots = npix * (size_t)k;
    if (cur.nchunks > 0 && (!batched || holds((int32_t)v) || cur.nchunks == kMaxChunks ||
                            cur.total_slots + slots > std::max(target, npix)))
      flush();
    PassChunk &c = cur.chunk[cur.nchunks++];
    c.view = (int32_t)v;
    c.s0 = s[v];
    c.k = k;
    c.npix = (int32_t)npix;
    c.slot_base = cur.total_slots;
    c.pix_base = cur.total_pix;
    cur.total_slots += (uint32_t)slots;
    cur.total_pix += (uint32_t)npix;
    s[v] += k;
  };
  if (batched) {
    for (bool more = true; more;) {
      more = false;
      for (size_t v = 0; v < views.size(); ++v)
        if (s[v] < s_end[v]) {
          add(v);
          more = true;
        }
    }
  } else {
    for (size_t v = 0; v < views.size(); ++v)
      while (s[v] < s_end[v]) add(v);
  }
  flush();
}

int ensure_workspace(gi_ctx *ctx, size_t slots) {
  if (slots <= ctx->slot_cap && ctx->path_l.p) {
    // ray differentials only while a material has a texture
    if (ctx->textured) GI_HIP(ctx, ctx->rdiff.ensure(4 * ctx->slot_cap * sizeof(float4)));
    return GI_OK;
  }
  const size_t b4 = slots * sizeof(float4);
  if (ctx->textured) GI_HIP(ctx, ctx->rdiff.ensure(4 * b4));
  for (int i = 0; i < 2; ++i) {
    GI_HIP(ctx, ctx->q_o[i].ensure(b4));
    GI_HIP(ctx, ctx->q_d[i].ensure(b4));
    GI_HIP(ctx, ctx->q_tp[i].ensure(b4));
  }
  GI_HIP(ctx, ctx->q_hit.ensure(b4));
  GI_HIP(ctx, ctx->s_o.ensure(b4));
  GI_HIP(ctx, ctx->s_d.ensure(b4));
  GI_HIP(ctx, ctx->s_c.ensure(b4));
  GI_HIP(ctx, ctx->path_l.ensure(b4));
  ctx->slot_cap = slots;
  return GI_OK;
}

ShadeParams shade_params(gi_ctx *ctx, int32_t depth) {
  ShadeParams P;
  P.tris = ctx->d_stris.as<ShadeTri>();
  P.mats = ctx->d_mats.as<DevMat>();
  P.lights = ctx->d_lights.as<DevLight>();
  P.nlights = (int32_t)ctx->lights.size();
  P.power_total = ctx->power_total;
  P.H = HaltonDev{ctx->d_hdims.as<HaltonDim>(), ctx->d_hchunks.as<uint16_t>(), 0u};
  P.off_b = ctx->opts.offset_bounce > 0.f ? ctx->opts.offset_bounce : 1e-4f * ctx->diag;
  P.off_s = ctx->opts.offset_shadow > 0.f ? ctx->opts.offset_shadow : 1e-4f * ctx->diag;
  for (int k = 0; k < 3; ++k) P.bg[k] = ctx->opts.background[k];
  P.depth = depth;
  P.rr_start = ctx->opts.rr_start;
  P.sky = SkyDev{ctx->d_sky_rgb.as<float4>(), ctx->d_sky_y.as<float>() + ctx->sky_h, ctx->d_sky_y.as<float>(),
                 ctx->d_sky_x.as<float>() + (size_t)ctx->sky_w * ctx->sky_h, ctx->d_sky_x.as<float>(),
                 ctx->sky_w, ctx->sky_h, ctx->sky_light,
                 ctx->sky_light >= 0 ? ctx->lights[(size_t)ctx->sky_light].col[0] : 0.0f};
  P.is = inst_scene(ctx);
  P.tex = ctx->d_tex.as<DevTex>();
  P.rdiff = ctx->textured ? ctx->rdiff.as<float4>() : nullptr;
  P.rdiff_stride = (uint32_t)ctx->slot_cap;
  P.ks = ctx->d_ks.as<float4>();
  P.lsum = ctx->d_lsum.as<float>();
  P.pass = nullptr;  // set per pass
  P.views = ctx->d_views.as<DevView>();
  return P;
}

hipEvent_t ev(gi_ctx *ctx, size_t i) {
  while (ctx->events.size() <= i) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ctx->events.push_back(e);
  }
  return ctx->events[i];
}

// Everything the launches of enqueue_passes take as arguments or launch
// shapes: buffer pointers, scalars, tuning, the plan.  A captured graph is
// replayed only while this key is unchanged (buffer contents -- materials,
// lights, views, pass descriptors -- are read by the kernels from device
// memory, so they may change between replays).
template <class T> void key_put(std::vector<uint8_t> &k, const T &v) {
  const uint8_t *p = reinterpret_cast<const uint8_t *>(&v);
  k.insert(k.end(), p, p + sizeof(T));
}
std::vector<uint8_t> launch_key(gi_ctx *ctx, const PassPlan &plan, int32_t depth, const ShadeParams &P) {
  std::vector<uint8_t> k;
  key_put(k, depth);
  key_put(k, ctx->tuning);
  key_put(k, ctx->count_traversal);
  key_put(k, instanced(ctx));
  key_put(k, lds_scene_of(ctx));
  key_put(k, ctx->stack_cap);
  key_put(k, ctx->nnodes);
  key_put(k, ctx->ntri);
  key_put(k, ctx->num_cus);
  key_put(k, ctx->slot_cap);
  for (const void *q : {ctx->d_passes.p, ctx->d_views.p, ctx->d_counters.p, ctx->q_o[0].p, ctx->q_o[1].p,
                        ctx->q_d[0].p, ctx->q_d[1].p, ctx->q_tp[0].p, ctx->q_tp[1].p, ctx->q_hit.p, ctx->s_o.p,
                        ctx->s_d.p, ctx->s_c.p, ctx->path_l.p, ctx->d_hnodes.p, ctx->d_nodes.p, ctx->d_itris.p, ctx->d_gstack.p,
                        ctx->d_overflow.p, ctx->d_tstats.p})
    key_put(k, q);
  // shading parameters, field by field (the struct has padding)
  for (const void *q : {(const void *)P.tris, (const void *)P.mats, (const void *)P.lights, (const void *)P.H.dims,
                        (const void *)P.H.chunks, (const void *)P.sky.rgb, (const void *)P.sky.ycdf,
                        (const void *)P.sky.ypdf, (const void *)P.sky.xcdf, (const void *)P.sky.xpdf,
                        (const void *)P.is.inst, (const void *)P.is.vbase, (const void *)P.is.vtab,
                        (const void *)P.tex, (const void *)P.rdiff, (const void *)P.ks, (const void *)P.lsum,
                        (const void *)P.views})
    key_put(k, q);
  key_put(k, ctx->opts.flags & GI_FLAG_PHONG);  // picks the shading kernel
  key_put(k, P.nlights);
  key_put(k, P.power_total);
  key_put(k, P.off_b);
  key_put(k, P.off_s);
  for (int i = 0; i < 3; ++i) key_put(k, P.bg[i]);
  key_put(k, P.depth);
  key_put(k, P.rr_start);
  key_put(k, P.sky.w);
  key_put(k, P.sky.h);
  key_put(k, P.sky.light);
  key_put(k, P.sky.scale);
  key_put(k, P.is.tlas_root);
  key_put(k, P.is.n_world);
  key_put(k, P.is.n_inst);
  key_put(k, P.is.vshift);
  key_put(k, P.rdiff_stride);
  // launch shapes follow the passes' sizes (PassDesc is zero-initialised)
  const uint8_t *pp = reinterpret_cast<const uint8_t *>(plan.passes.data());
  k.insert(k.end(), pp, pp + plan.passes.size() * sizeof(PassDesc));
  return k;
}

void drop_graph(gi_ctx *ctx) {
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  ctx->graph_exec = nullptr;
  ctx->graph_key.clear();
}

// The launches of the passes on the context stream.  sync_counts: read
// every bounce's queue length back to the host and size the launches
// exactly (the reference's synchronous, per-view flow); otherwise every
// launch uses the persistent grid and reads its length from device memory,
// and nothing synchronises.  mark(cat) records a timing event after a
// launch (cat -1 opens an interval), or is null.
template <class Mark>
int enqueue_passes(gi_ctx *ctx, const PassPlan &plan, int32_t depth, bool sync_counts, const ShadeParams &SP,
                   int &launches, Mark mark) {
  const int32_t ncnt = 2 * depth + 1;  // ray queue counts [0..depth], shadow counts
  const size_t npass = plan.passes.size();
  const PassDesc *dpass = ctx->d_passes.as<PassDesc>();
  const DevView *dviews = ctx->d_views.as<DevView>();
  const int sblock = std::min(256, std::max(64, (ctx->tuning.shade_block / 64) * 64));
  const int persist = ctx->num_cus * 8;
  int rc = GI_OK;
  for (size_t pi = 0; pi < npass; ++pi) {
    const PassDesc &pd = plan.passes[pi];
    const PassDesc *dp = dpass + pi;
    ShadeParams SPp = SP;
    SPp.pass = dp;
    uint32_t *cnt = ctx->d_counters.as<uint32_t>() + pi * ncnt;
    hipLaunchKernelGGL(k_pass_setup, dim3(1), dim3(256), 0, ctx->stream, dp, cnt, ncnt);
    const int rg = (int)std::min<uint64_t>((pd.total_slots + 255) / 256, (uint64_t)persist);
    hipLaunchKernelGGL(k_raygen, dim3(std::max(1, rg)), dim3(256), 0, ctx->stream, dp, dviews, SP.H,
                       ctx->q_o[0].as<float4>(), ctx->q_d[0].as<float4>(), ctx->path_l.as<float4>(),
                       (float4 *)nullptr, 0u, depth);
    uint32_t host_count = pd.total_slots;
    for (int32_t b = 0; b < depth; ++b) {
      const int cur = b & 1, nxt = cur ^ 1;
      if (sync_counts && b > 0) {
        GI_HIP(ctx, hipMemcpyAsync(&host_count, cnt + b, 4, hipMemcpyDeviceToHost, ctx->stream));
        GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (host_count == 0) break;
      }
      const uint64_t bound = sync_counts ? host_count : pd.total_slots;
      if ((rc = mark(-1))) return rc;
      rc = launch_closest(ctx, false, cnt + b, 0, bound, ctx->q_o[cur].as<float4>(),
                          ctx->q_d[cur].as<float4>(), ctx->q_hit.as<float4>());
      if (rc) return rc;
      ++launches;
      if ((rc = mark(0))) return rc;
      const int sg = (int)std::max<uint64_t>(1, std::min<uint64_t>((bound + sblock - 1) / sblock, (uint64_t)persist));
      const bool phong = (ctx->opts.flags & GI_FLAG_PHONG) != 0;
      hipLaunchKernelGGL(phong ? (SP.rdiff ? k_shade_phong_tex : k_shade_phong) : (SP.rdiff ? k_shade_tex : k_shade),
                         dim3(sg), dim3(sblock), 0, ctx->stream, b,
                         cnt + b, SPp, ctx->q_o[cur].as<float4>(), ctx->q_d[cur].as<float4>(),
                         ctx->q_tp[cur].as<float4>(), ctx->q_hit.as<float4>(), ctx->path_l.as<float4>(),
                         ctx->q_o[nxt].as<float4>(), ctx->q_d[nxt].as<float4>(), ctx->q_tp[nxt].as<float4>(),
                         cnt + b + 1, ctx->s_o.as<float4>(), ctx->s_d.as<float4>(), ctx->s_c.as<float4>(),
                         cnt + depth + 1 + b, (float4 *)nullptr, 0u);
      GI_HIP(ctx, hipGetLastError());
      if ((rc = mark(2))) return rc;
      uint32_t shadow_host = 0;
      uint64_t sbound = bound;
      if (sync_counts) {
        GI_HIP(ctx, hipMemcpyAsync(&shadow_host, cnt + depth + 1 + b, 4, hipMemcpyDeviceToHost, ctx->stream));
        GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
        sbound = shadow_host;
      }
      if (sbound > 0) {
        rc = launch_any(ctx, true, cnt + depth + 1 + b, 0, sbound, ctx->s_o.as<float4>(), ctx->s_d.as<float4>(),
                        ctx->s_c.as<float4>(), ctx->path_l.as<float4>(), nullptr);
        if (rc) return rc;
      }
      if ((rc = mark(1))) return rc;
    }
    const int ag = (int)std::max<uint64_t>(1, std::min<uint64_t>((pd.total_pix + 255) / 256, (uint64_t)persist));
    hipLaunchKernelGGL(k_accumulate, dim3(ag), dim3(256), 0, ctx->stream, dp, dviews, ctx->path_l.as<float4>());
    GI_HIP(ctx, hipGetLastError());
  }
  return GI_OK;
}

// Run the passes: per-view with host-read queue lengths (sync_counts), or
// every launch on device-side counts; use_graph: the launches as one
// hipGraph.  The executable graph is kept and relaunched by later renders
// whose launches take the same arguments (launch_key); otherwise it is
// captured and instantiated again (counters.ms_graph_build).
int run_passes(gi_ctx *ctx, const PassPlan &plan, int32_t depth, bool sync_counts, bool use_graph, bool timing) {
  const int32_t ncnt = 2 * depth + 1;
  const size_t npass = plan.passes.size();
  GI_HIP(ctx, ctx->d_passes.ensure(npass * sizeof(PassDesc)));
  GI_HIP(ctx, hipMemcpyAsync(ctx->d_passes.p, plan.passes.data(), npass * sizeof(PassDesc), hipMemcpyHostToDevice,
                             ctx->stream));
  GI_HIP(ctx, ctx->d_counters.ensure(npass * ncnt * sizeof(uint32_t)));
  int rc = ensure_workspace(ctx, plan.max_slots);
  if (rc) return rc;
  if ((rc = ensure_gstack(ctx))) return rc;
  // the stack-overflow counter starts at zero for every render (a fill
  // kernel rather than a memset node, ahead of any capture)
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->d_overflow.as<uint32_t>(), 1u, 0u);
  GI_HIP(ctx, hipGetLastError());
  ShadeParams SP = shade_params(ctx, depth);
  double ms_c = 0, ms_a = 0, ms_s = 0;
  int launches = 0;
  if (use_graph) {
    std::vector<uint8_t> key = launch_key(ctx, plan, depth, SP);
    if (!ctx->graph_exec || key != ctx->graph_key) {
      drop_graph(ctx);
      const auto g0 = std::chrono::steady_clock::now();
      GI_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      rc = enqueue_passes(ctx, plan, depth, false, SP, launches, [](int) { return GI_OK; });
      hipGraph_t graph = nullptr;
      // the capture always ends, so the stream is usable after an error
      const hipError_t ee = hipStreamEndCapture(ctx->stream, &graph);
      if (rc == GI_OK && ee != hipSuccess)
        rc = fail(ctx, GI_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
      if (rc == GI_OK) {
        const hipError_t ei = hipGraphInstantiate(&ctx->graph_exec, graph, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
          ctx->graph_exec = nullptr;
          rc = fail(ctx, GI_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
      }
      if (graph) (void)hipGraphDestroy(graph);
      if (rc) return rc;
      ctx->graph_key = std::move(key);
      ctx->counters.ms_graph_build +=
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - g0).count();
    } else {
      launches = ctx->graph_launches;
    }
    ctx->graph_launches = launches;
    GI_HIP(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
  } else {
    size_t evi = 0;
    std::vector<std::pair<size_t, int>> marks;  // (event index, category)
    auto mark = [&](int cat) -> int {
      if (!timing) return GI_OK;
      hipEvent_t e = ev(ctx, evi);
      if (!e) return fail(ctx, GI_ERR_HIP, "hipEventCreate");
      GI_HIP(ctx, hipEventRecord(e, ctx->stream));
      marks.push_back({evi++, cat});
      return GI_OK;
    };
    if ((rc = enqueue_passes(ctx, plan, depth, sync_counts, SP, launches, mark))) return rc;
    GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 1; i < marks.size(); ++i) {
      if (marks[i].second < 0) continue;
      float ms = 0.f;
      GI_HIP(ctx, hipEventElapsedTime(&ms, ctx->events[marks[i - 1].first], ctx->events[marks[i].first]));
      if (marks[i].second == 0) ms_c += ms;
      else if (marks[i].second == 1) ms_a += ms;
      else ms_s += ms;
    }
  }
  GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // telemetry
  std::vector<uint32_t> hc(npass * ncnt);
  GI_HIP(ctx, copy_sync(ctx, hc.data(), ctx->d_counters.p, hc.size() * 4, hipMemcpyDeviceToHost));
  for (size_t pi = 0; pi < npass; ++pi) {
    for (int32_t b = 0; b < depth; ++b) {
      ctx->counters.rays_closest += hc[pi * ncnt + b];
      ctx->counters.rays_shadow += hc[pi * ncnt + depth + 1 + b];
    }
  }
  ctx->counters.ms_trace_closest += ms_c;
  ctx->counters.ms_trace_any += ms_a;
  ctx->counters.ms_shade += ms_s;
  ctx->counters.launches_closest += launches;
  ctx->counters.passes += (int32_t)npass;
  return check_overflow(ctx);
}

// Adjoint: per pass, a forward pass in record mode (the per-vertex
// contributions and clamp bits of every path) and the replay of the same
// paths -- same Halton indices, closest-hit rays re-traced, no shadow rays --
// that accumulates the gradients.  Device-side queue lengths, persistent
// grids, no host synchronisation until the end.
int run_adjoint_passes(gi_ctx *ctx, const PassPlan &plan, int32_t depth, AdjParams A) {
  const int32_t ncnt = 2 * depth + 1;
  const size_t npass = plan.passes.size();
  GI_HIP(ctx, ctx->d_passes.ensure(npass * sizeof(PassDesc)));
  GI_HIP(ctx, hipMemcpyAsync(ctx->d_passes.p, plan.passes.data(), npass * sizeof(PassDesc), hipMemcpyHostToDevice,
                             ctx->stream));
  GI_HIP(ctx, ctx->d_counters.ensure(2 * npass * ncnt * sizeof(uint32_t)));
  int rc = ensure_workspace(ctx, plan.max_slots);
  if (rc) return rc;
  if ((rc = ensure_gstack(ctx))) return rc;
  // per path and vertex: the radiance record, then the hit
  GI_HIP(ctx, ctx->d_rec.ensure((size_t)2 * depth * plan.max_slots * sizeof(float4)));
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->d_overflow.as<uint32_t>(), 1u, 0u);
  GI_HIP(ctx, hipGetLastError());
  ShadeParams SP = shade_params(ctx, depth);
  const PassDesc *dpass = ctx->d_passes.as<PassDesc>();
  const DevView *dviews = ctx->d_views.as<DevView>();
  float4 *rec = ctx->d_rec.as<float4>();
  const uint32_t stride = (uint32_t)plan.max_slots;
  A.rec = rec;
  A.rec_stride = stride;
  A.views = dviews;
  const int sblock = std::min(256, std::max(64, (ctx->tuning.shade_block / 64) * 64));
  const int persist = ctx->num_cus * 8;
  const size_t lds_adj = (size_t)(5 * A.nmats + 3 * A.nlights) * sizeof(float);
  for (size_t pi = 0; pi < npass; ++pi) {
    const PassDesc &pd = plan.passes[pi];
    const PassDesc *dp = dpass + pi;
    A.pass = dp;
    ShadeParams SPp = SP;
    SPp.pass = dp;
    const int rg = (int)std::min<uint64_t>((pd.total_slots + 255) / 256, (uint64_t)persist);
    const int sg = (int)std::max<uint64_t>(1, std::min<uint64_t>((pd.total_slots + sblock - 1) / sblock,
                                                                 (uint64_t)persist));
    for (int phase = 0; phase < 2; ++phase) {
      uint32_t *cnt = ctx->d_counters.as<uint32_t>() + (2 * pi + phase) * ncnt;
      hipLaunchKernelGGL(k_pass_setup, dim3(1), dim3(256), 0, ctx->stream, dp, cnt, ncnt);
      hipLaunchKernelGGL(k_raygen, dim3(std::max(1, rg)), dim3(256), 0, ctx->stream, dp, dviews, SP.H,
                         ctx->q_o[0].as<float4>(), ctx->q_d[0].as<float4>(), ctx->path_l.as<float4>(),
                         phase == 0 ? rec : (float4 *)nullptr, stride, depth);
      for (int32_t b = 0; b < depth; ++b) {
        const int cur = b & 1, nxt = cur ^ 1;
        if (phase == 0) {
          // the replay (phase 1) reads these hits from the record
          rc = launch_closest(ctx, false, cnt + b, 0, pd.total_slots, ctx->q_o[cur].as<float4>(),
                              ctx->q_d[cur].as<float4>(), ctx->q_hit.as<float4>());
          if (rc) return rc;
          hipLaunchKernelGGL(SP.rdiff ? k_shade_tex : k_shade, dim3(sg), dim3(sblock), 0, ctx->stream, b,
                             cnt + b, SPp,
                             ctx->q_o[cur].as<float4>(), ctx->q_d[cur].as<float4>(), ctx->q_tp[cur].as<float4>(),
                             ctx->q_hit.as<float4>(), ctx->path_l.as<float4>(), ctx->q_o[nxt].as<float4>(),
                             ctx->q_d[nxt].as<float4>(), ctx->q_tp[nxt].as<float4>(), cnt + b + 1,
                             ctx->s_o.as<float4>(), ctx->s_d.as<float4>(), ctx->s_c.as<float4>(),
                             cnt + depth + 1 + b, rec, stride);
          GI_HIP(ctx, hipGetLastError());
          rc = launch_any(ctx, true, cnt + depth + 1 + b, 0, pd.total_slots, ctx->s_o.as<float4>(),
                          ctx->s_d.as<float4>(), ctx->s_c.as<float4>(), ctx->path_l.as<float4>(), nullptr, rec,
                          stride, b);
          if (rc) return rc;
        } else {
          hipLaunchKernelGGL(SP.rdiff ? k_shade_adjoint<true> : k_shade_adjoint<false>, dim3(sg), dim3(sblock),
                             lds_adj, ctx->stream, b, cnt + b, SPp, A,
                             ctx->q_o[cur].as<float4>(), ctx->q_d[cur].as<float4>(), ctx->q_tp[cur].as<float4>(),
                             ctx->q_o[nxt].as<float4>(), ctx->q_d[nxt].as<float4>(),
                             ctx->q_tp[nxt].as<float4>(), cnt + b + 1);
          GI_HIP(ctx, hipGetLastError());
        }
      }
    }
  }
  GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<uint32_t> hc(2 * npass * ncnt);
  GI_HIP(ctx, copy_sync(ctx, hc.data(), ctx->d_counters.p, hc.size() * 4, hipMemcpyDeviceToHost));
  for (size_t k = 0; k < 2 * npass; k += 2)  // the forward (recording) phases trace; the replays do not
    for (int32_t b = 0; b < depth; ++b) {
      ctx->counters.rays_closest += hc[k * ncnt + b];
      ctx->counters.rays_shadow += hc[k * ncnt + depth + 1 + b];
    }
  ctx->counters.passes += (int32_t)npass;
  return check_overflow(ctx);
}

// Camera basis and image-plane extent of one view (src/raygen.cu:17-39);
// false (with the reason in err) for a view that cannot be rendered.  The
// image pointer, index base and pixel base are left to the caller.
bool make_devview(const gi_view &w, DevView &d, std::string &err) {
  std::memset(&d, 0, sizeof(d));
  if (w.w <= 0 || w.h <= 0 || w.spp <= 0) {
    err = "view size / spp must be > 0";
    return false;
  }
  if (!(w.fovy > 0.f && w.fovy < 180.f)) {
    err = "fovy must be in (0, 180)";
    return false;
  }
  const float dl = w.dir[0] * w.dir[0] + w.dir[1] * w.dir[1] + w.dir[2] * w.dir[2];
  if (!(dl > 0.f) || !std::isfinite(dl)) {
    err = "view dir";
    return false;
  }
  float half = w.fovy / 2.0f;
  d.height = (float)std::tan(3.14159265358979323846 * (double)half / 180.0);
  d.width = ((float)w.w / (float)w.h) * d.height;
  float vl = std::sqrt(w.dir[0] * w.dir[0] + w.dir[1] * w.dir[1] + w.dir[2] * w.dir[2]);
  float W[3] = {w.dir[0] / vl, w.dir[1] / vl, w.dir[2] / vl};
  float txw[3] = {w.up[1] * W[2] - w.up[2] * W[1], w.up[2] * W[0] - w.up[0] * W[2], w.up[0] * W[1] - w.up[1] * W[0]};
  float tl = std::sqrt(txw[0] * txw[0] + txw[1] * txw[1] + txw[2] * txw[2]);
  if (!(tl > 0.f)) {
    err = "view up parallel to dir";
    return false;
  }
  float U[3] = {txw[0] / tl, txw[1] / tl, txw[2] / tl};
  float V[3] = {W[1] * U[2] - W[2] * U[1], W[2] * U[0] - W[0] * U[2], W[0] * U[1] - W[1] * U[0]};
  for (int k = 0; k < 3; ++k) {
    d.U[k] = U[k]; d.V[k] = V[k]; d.W[k] = W[k];
    d.pos[k] = w.pos[k]; d.vd[k] = w.dir[k];
  }
  d.fw = (float)w.w;
  d.fh = (float)w.h;
  d.aperture = w.aperture;
  d.focus = w.focus_distance;
  d.w = w.w; d.h = w.h; d.spp = w.spp;
  return true;
}

// The render state of a new render or progressive sequence: every view is
// validated before any state changes; the views, options and images are
// committed only when all of them (and their buffers) are in place.  On
// failure the context holds no views (gi_adjoint, gi_progress_step and
// gi_get_image then report the missing render) and no progressive state.
int setup_views(gi_ctx *ctx, const gi_view *views, int32_t n_views, const gi_render_opts *opts) {
  ctx->views.clear();
  ctx->prog_active = false;
  if (!views || n_views <= 0 || !opts) return fail(ctx, GI_ERR_INVALID_ARG, "views/opts");
  if (opts->max_depth < 1 || opts->max_depth > kMaxDepth)
    return fail(ctx, GI_ERR_INVALID_ARG, "max_depth must be in [1, 32]");
  std::vector<DevView> dv(n_views);
  uint64_t total = 0;
  uint32_t base = 0, pix_base = 0;
  std::string err;
  for (int32_t v = 0; v < n_views; ++v) {
    const gi_view &w = views[v];
    if (!make_devview(w, dv[v], err)) return fail(ctx, GI_ERR_INVALID_ARG, err);
    dv[v].centre = (opts->flags & GI_FLAG_PIXEL_CENTRE) ? 1 : 0;
    total += (uint64_t)w.w * (uint64_t)w.h * (uint64_t)w.spp;
    if (total >= (1ull << 32)) return fail(ctx, GI_ERR_LIMIT, "more than 2^32 samples in one render");
    dv[v].index_base = base;
    base += (uint32_t)((uint64_t)w.w * w.h * w.spp);
    dv[v].pix_base = pix_base;
    pix_base += (uint32_t)((uint64_t)w.w * w.h);
  }
  int rc = ensure_halton(ctx, opts->seed, dims_for_depth(opts->max_depth));
  if (rc) return rc;
  if (ctx->images.size() < (size_t)n_views) ctx->images.resize(n_views);
  for (int32_t v = 0; v < n_views; ++v) {
    const size_t npix = (size_t)views[v].w * views[v].h;
    GI_HIP(ctx, ctx->images[v].ensure(npix * sizeof(float4)));
    dv[v].image = ctx->images[v].as<float4>();
    const uint32_t nb = (uint32_t)((npix + 255) / 256);
    hipLaunchKernelGGL(k_fill4, dim3(nb), dim3(256), 0, ctx->stream, dv[v].image, (uint32_t)npix,
                       make_float4(0.f, 0.f, 0.f, 0.f));
    GI_HIP(ctx, hipGetLastError());
  }
  GI_HIP(ctx, ctx->d_views.ensure(dv.size() * sizeof(DevView)));
  GI_HIP(ctx, hipMemcpyAsync(ctx->d_views.p, dv.data(), dv.size() * sizeof(DevView), hipMemcpyHostToDevice,
                             ctx->stream));
  GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->views.assign(views, views + n_views);
  ctx->opts = *opts;
  return GI_OK;
}

size_t target_slots(gi_ctx *ctx) {
  const int t = ctx->tuning.target_slots;
  // default: 128 Mi paths in flight (22 GB of path state, 30 GB with ray
  // differentials), so a 1280x720x64 render or a 3840x2160x16 view is one
  // pass and every launch is as wide as it gets
  return (size_t)(t > 0 ? std::min(t, 256) : 128) << 20;
}

}  // namespace

// ================================================================= C ABI ==
extern "C" {

int gi_abi_version(void) { return GI_ABI_VERSION; }

const char *gi_status_string(int s) {
  switch (s) {
    case GI_OK: return "ok";
    case GI_ERR_INVALID_ARG: return "invalid argument";
    case GI_ERR_NO_DEVICE: return "no HIP device";
    case GI_ERR_HIP: return "HIP runtime error";
    case GI_ERR_STATE: return "invalid call order";
    case GI_ERR_OUT_OF_MEMORY: return "out of device memory";
    case GI_ERR_LIMIT: return "size beyond a documented limit";
    case GI_ERR_UNKNOWN_NAME: return "unknown name";
    default: return "unknown status";
  }
}

int gi_create(gi_ctx **out, int device) {
  if (!out) return GI_ERR_INVALID_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GI_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return GI_ERR_INVALID_ARG;
  auto ctx = std::make_unique<gi_ctx>();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) return GI_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return GI_ERR_NO_DEVICE;
  ctx->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  if (prop.maxSharedMemoryPerMultiProcessor > 0) ctx->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return GI_ERR_NO_DEVICE;
  void *sp = nullptr;
  if (scratch_address(device, &sp) != hipSuccess) return GI_ERR_HIP;
  ctx->scratch = static_cast<uint8_t *>(sp);
  if (ctx->d_overflow.ensure(256) != hipSuccess) return GI_ERR_OUT_OF_MEMORY;
  if (hipMemsetAsync(ctx->d_overflow.p, 0, 4, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return GI_ERR_HIP;
  *out = ctx.release();
  return GI_OK;
}

void gi_destroy(gi_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  drop_graph(ctx);
  for (auto e : ctx->events) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *gi_last_error(const gi_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int gi_bvh_build(const gi_tri *tris, int64_t n, int32_t max_leaf, gi_bvh **out) {
  if (!out || n < 0 || (n > 0 && !tris)) return GI_ERR_INVALID_ARG;
  if (n > (int64_t)kMaxLeafFirst) return GI_ERR_LIMIT;
  auto b = std::make_unique<gi_bvh>();
  build_bvh2(tris, n, max_leaf > 0 ? max_leaf : 4, *b);
  *out = b.release();
  return GI_OK;
}

void gi_bvh_free(gi_bvh *bvh) { delete bvh; }

int gi_bvh_sizes(const gi_bvh *bvh, int64_t *n_nodes, int64_t *n_indices) {
  if (!bvh || !n_nodes || !n_indices) return GI_ERR_INVALID_ARG;
  *n_nodes = (int64_t)bvh->nodes.size();
  *n_indices = (int64_t)bvh->perm.size();
  return GI_OK;
}

int gi_bvh_export2(const gi_bvh *bvh, float *nodes, uint32_t *indices) {
  if (!bvh || !nodes || (!indices && !bvh->perm.empty())) return GI_ERR_INVALID_ARG;
  for (size_t i = 0; i < bvh->nodes.size(); ++i) {
    const Bvh2Node &n = bvh->nodes[i];
    float *p = nodes + 8 * i;
    int32_t x, y;
    if (n.left >= 0) {
      x = (n.left << 1) | 1;
      y = n.right;
    } else {
      x = n.count << 1;
      y = n.first;
    }
    std::memcpy(&p[0], &x, 4);
    std::memcpy(&p[1], &y, 4);
    p[2] = n.lo[0]; p[3] = n.lo[1]; p[4] = n.lo[2];
    p[5] = n.hi[0]; p[6] = n.hi[1]; p[7] = n.hi[2];
  }
  if (!bvh->perm.empty()) std::memcpy(indices, bvh->perm.data(), bvh->perm.size() * 4);
  return GI_OK;
}

int gi_upload_triangles(gi_ctx *ctx, const gi_tri *tris, int64_t n) {
  if (!ctx || !tris || n <= 0) return fail(ctx, GI_ERR_INVALID_ARG, "triangles");
  if (n > (int64_t)kMaxLeafFirst) return fail(ctx, GI_ERR_LIMIT, "more than 2^27 triangles");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  ctx->tris.assign(tris, tris + n);
  ctx->ntri = n;
  ctx->have_bvh = false;
  ctx->bvh.reset();
  ctx->meshes.clear();  // a new scene: meshes and instances are uploaded again
  ctx->n_inst = 0;
  ctx->mat_lo = INT32_MAX;
  ctx->mat_hi = INT32_MIN;
  for (int64_t i = 0; i < n; ++i) {
    ctx->mat_lo = std::min(ctx->mat_lo, tris[i].material_index);
    ctx->mat_hi = std::max(ctx->mat_hi, tris[i].material_index);
  }
  for (int k = 0; k < 3; ++k) {
    ctx->lo[k] = ctx->clo[k] = FLT_MAX;
    ctx->hi[k] = ctx->chi[k] = -FLT_MAX;
  }
  std::vector<ShadeTri> st(n);
  for (int64_t i = 0; i < n; ++i) {
    const gi_tri &t = tris[i];
    for (int k = 0; k < 3; ++k) {
      ctx->lo[k] = std::min(ctx->lo[k], std::min(t.a[k], std::min(t.b[k], t.c[k])));
      ctx->hi[k] = std::max(ctx->hi[k], std::max(t.a[k], std::max(t.b[k], t.c[k])));
      const float cen = (t.a[k] + t.b[k] + t.c[k]) * (1.0f / 3.0f);
      ctx->clo[k] = std::min(ctx->clo[k], cen);
      ctx->chi[k] = std::max(ctx->chi[k], cen);
    }
    st[i] = make_shadetri(t);
  }
  float dx = ctx->hi[0] - ctx->lo[0], dy = ctx->hi[1] - ctx->lo[1], dz = ctx->hi[2] - ctx->lo[2];
  ctx->diag = std::sqrt(dx * dx + dy * dy + dz * dz);
  GI_HIP(ctx, ctx->d_stris.ensure(n * sizeof(ShadeTri)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_stris.p, st.data(), n * sizeof(ShadeTri), hipMemcpyHostToDevice));
  return GI_OK;
}

int gi_upload_materials(gi_ctx *ctx, const gi_material *mats, int32_t n) {
  if (!ctx || !mats || n <= 0) return fail(ctx, GI_ERR_INVALID_ARG, "materials");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  ctx->mats.assign(mats, mats + n);
  ctx->textured = false;
  for (int32_t i = 0; i < n; ++i) ctx->textured = ctx->textured || mats[i].texture >= 0;
  std::vector<DevMat> dm(n);
  std::vector<DevMatGrad> dg(n);
  for (int32_t i = 0; i < n; ++i) {
    dm[i] = make_devmat(mats[i]);
    dg[i] = make_devmatgrad(mats[i]);
  }
  GI_HIP(ctx, ctx->d_mats.ensure(n * sizeof(DevMat)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_mats.p, dm.data(), n * sizeof(DevMat), hipMemcpyHostToDevice));
  GI_HIP(ctx, ctx->d_matgrad.ensure(n * sizeof(DevMatGrad)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_matgrad.p, dg.data(), n * sizeof(DevMatGrad), hipMemcpyHostToDevice));
  std::vector<float4> ks(n);
  for (int32_t i = 0; i < n; ++i)
    ks[i] = make_float4(mats[i].specular_color[0], mats[i].specular_color[1], mats[i].specular_color[2], 0.0f);
  GI_HIP(ctx, ctx->d_ks.ensure(n * sizeof(float4)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_ks.p, ks.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  return GI_OK;
}

int gi_upload_lights(gi_ctx *ctx, const gi_light *lights, int32_t n) {
  if (!ctx || !lights || n <= 0) return fail(ctx, GI_ERR_INVALID_ARG, "lights");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  std::vector<DevLight> dl(n);
  float total = 0.f;
  int32_t sky = -1;
  for (int32_t i = 0; i < n; ++i) {
    const gi_light &l = lights[i];
    if (l.type != GI_LIGHT_RECT && l.type != GI_LIGHT_SKY)
      return fail(ctx, GI_ERR_INVALID_ARG, "light type must be GI_LIGHT_RECT or GI_LIGHT_SKY");
    if (l.type == GI_LIGHT_SKY) {
      if (sky >= 0) return fail(ctx, GI_ERR_INVALID_ARG, "at most one sky light");
      sky = i;
    }
    if (!(l.power > 0.f)) return fail(ctx, GI_ERR_INVALID_ARG, "light power must be > 0");
    DevLight &d = dl[i];
    std::memset(&d, 0, sizeof(d));
    d.type = l.type;
    for (int k = 0; k < 3; ++k) {
      d.center[k] = l.center[k];
      d.dir[k] = l.dir[k];
      d.col[k] = l.col[k];
    }
    d.power = l.power;
    d.w = l.wh[0];
    d.h = l.wh[1];
    const float ex[3] = {1, 0, 0}, ey[3] = {0, 1, 0};
    host_tangential(ex, l.dir, d.right);
    host_tangential(ey, l.dir, d.up);
    total += l.power;
  }
  // running sums of the powers in light order, each the float sum the
  // linear selection (pixel_generate_arealight_sample_cpu,
  // src/arealight-sampler-cpu.cpp:33-39) forms: light k is picked when
  // choice <= sums[k + 1] for the first time
  std::vector<float> sums(n + 1);
  sums[0] = 0.0f;
  for (int32_t i = 0; i < n; ++i) sums[i + 1] = sums[i] + lights[i].power;
  ctx->lights.assign(lights, lights + n);
  ctx->power_total = total;
  if (sky != ctx->sky_light) ctx->sky_w = ctx->sky_h = 0;  // a new sky light needs its map
  ctx->sky_light = sky;
  GI_HIP(ctx, ctx->d_lights.ensure(n * sizeof(DevLight)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_lights.p, dl.data(), n * sizeof(DevLight), hipMemcpyHostToDevice));
  GI_HIP(ctx, ctx->d_lsum.ensure((size_t)(n + 1) * sizeof(float)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_lsum.p, sums.data(), (size_t)(n + 1) * sizeof(float), hipMemcpyHostToDevice));
  return GI_OK;
}

int gi_upload_sky(gi_ctx *ctx, const float *rgb, int32_t w, int32_t h) {
  if (!ctx || !rgb || w <= 0 || h <= 0) return fail(ctx, GI_ERR_INVALID_ARG, "sky map");
  if (ctx->sky_light < 0) return fail(ctx, GI_ERR_STATE, "upload a GI_LIGHT_SKY light first");
  if ((int64_t)w * h > (1 << 26)) return fail(ctx, GI_ERR_LIMIT, "sky map larger than 2^26 texels");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  // initSkylight (src/lights.h:107-116) and Distribution1D::init
  // (src/lights.h:27-44), in float, in order
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
  const size_t n = (size_t)w * h;
  std::vector<float> imp(n), rows(h), xpdf(n), xcdf((size_t)(w + 1) * h), ypdf(h), ycdf(h + 1);
  std::vector<float4> tex(n);
  for (int32_t y = 0; y < h; ++y)
    for (int32_t x = 0; x < w; ++x) {
      const float *c = rgb + 3 * ((size_t)y * w + x);
      imp[(size_t)y * w + x] = std::sin((float)(M_PI * (y + 0.5f) / h)) * ((c[0] + c[1]) + c[2]);
      tex[(size_t)y * w + x] = make_float4(c[0], c[1], c[2], 0.f);
    }
  for (int32_t y = 0; y < h; ++y) rows[y] = init(&imp[(size_t)y * w], w, &xpdf[(size_t)y * w], &xcdf[(size_t)y * (w + 1)]);
  init(rows.data(), h, ypdf.data(), ycdf.data());
  GI_HIP(ctx, ctx->d_sky_rgb.ensure(n * sizeof(float4)));
  GI_HIP(ctx, ctx->d_sky_x.ensure((n + (size_t)(w + 1) * h) * sizeof(float)));
  GI_HIP(ctx, ctx->d_sky_y.ensure((size_t)(2 * h + 1) * sizeof(float)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_sky_rgb.p, tex.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_sky_x.p, xpdf.data(), n * sizeof(float), hipMemcpyHostToDevice));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_sky_x.as<float>() + n, xcdf.data(), xcdf.size() * sizeof(float),
                        hipMemcpyHostToDevice));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_sky_y.p, ypdf.data(), h * sizeof(float), hipMemcpyHostToDevice));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_sky_y.as<float>() + h, ycdf.data(), (h + 1) * sizeof(float),
                        hipMemcpyHostToDevice));
  ctx->sky_w = w;
  ctx->sky_h = h;
  return GI_OK;
}

int gi_build_bvh(gi_ctx *ctx, int32_t max_leaf) {
  if (!ctx || ctx->tris.empty()) return fail(ctx, GI_ERR_STATE, "upload triangles first");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  ctx->bvh = std::make_unique<gi_bvh>();
  build_bvh2(ctx->tris.data(), (int64_t)ctx->tris.size(), max_leaf > 0 ? max_leaf : 4, *ctx->bvh);
  return upload_wide(ctx);
}

int gi_build_bvh_device(gi_ctx *ctx, int32_t max_leaf) {
  return gi_build_bvh_device_ex(ctx, max_leaf, GI_BVH_LBVH, 0);
}

int gi_build_bvh_device_ex(gi_ctx *ctx, int32_t max_leaf, int32_t method, int32_t radius) {
  if (!ctx || ctx->tris.empty()) return fail(ctx, GI_ERR_STATE, "upload triangles first");
  const int ml = max_leaf > 0 ? max_leaf : 4;
  if (ml > kMaxLeafTris) return fail(ctx, GI_ERR_INVALID_ARG, "max_leaf must be at most 15");
  if (method != GI_BVH_LBVH && method != GI_BVH_PLOC && method != GI_BVH_SAH)
    return fail(ctx, GI_ERR_INVALID_ARG, "unknown BVH method");
  const int rad = radius == 0 ? 16 : radius;
  if (method == GI_BVH_PLOC && (rad < 1 || rad > 64)) return fail(ctx, GI_ERR_INVALID_ARG, "PLOC radius must be in 1..64");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  const uint32_t n = (uint32_t)ctx->ntri;
  GI_HIP(ctx, ctx->d_nodes.ensure(std::max<size_t>(1, n) * sizeof(WideNode)));
  GI_HIP(ctx, ctx->d_itris.ensure(std::max<size_t>(1, n) * sizeof(IsectTri)));
  ctx->bvh.reset();
  ctx->have_bvh = false;
  LbvhResult r;
  std::string err;
  if (!device_bvh_build(ctx->stream, ctx->d_stris.as<float4>(), (uint32_t)(sizeof(ShadeTri) / sizeof(float4)), n,
                        ctx->clo, ctx->chi, ml,
                        method == GI_BVH_PLOC  ? DeviceBvhMethod::kPloc
                        : method == GI_BVH_SAH ? DeviceBvhMethod::kSah
                                               : DeviceBvhMethod::kLbvh,
                        rad, ctx->tuning.node_width, ctx->d_nodes.as<WideNode>(), ctx->d_itris.as<float4>(), r, err))
    return fail(ctx, GI_ERR_HIP, err);
  return finish_bvh(ctx, r.nnodes, r.depth);
}

int gi_upload_texture(gi_ctx *ctx, const uint8_t *rgba, int32_t w, int32_t h, int32_t *tex_id) {
  if (!ctx || !rgba || w <= 0 || h <= 0 || !tex_id) return fail(ctx, GI_ERR_INVALID_ARG, "texture");
  if ((int64_t)w * h > ((int64_t)1 << 28)) return fail(ctx, GI_ERR_LIMIT, "texture of more than 2^28 texels");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  int32_t levels = 0;
  while ((2 << levels) <= std::min(w, h)) ++levels;  // floor(log2(min(w, h)))
  std::vector<size_t> off(levels + 1);
  size_t total = 0;
  int32_t W = w, H = h;
  for (int32_t l = 0; l <= levels; ++l) {
    off[l] = total;
    total += (size_t)W * H;
    W = (W + 1) / 2;
    H = (H + 1) / 2;
  }
  DBuf buf;
  GI_HIP(ctx, buf.ensure(total * sizeof(uchar4)));
  GI_HIP(ctx, copy_sync(ctx, buf.p, rgba, (size_t)w * h * sizeof(uchar4), hipMemcpyHostToDevice));
  W = w;
  H = h;
  for (int32_t l = 0; l < levels; ++l) {
    const int32_t dw = (W + 1) / 2, dh = (H + 1) / 2;
    const uint32_t n = (uint32_t)dw * (uint32_t)dh;
    hipLaunchKernelGGL(k_mipmap, dim3(std::min<uint32_t>((n + 255) / 256, 4096)), dim3(256), 0, ctx->stream,
                       buf.as<uchar4>() + off[l], W, H, buf.as<uchar4>() + off[l + 1], dw, dh);
    GI_HIP(ctx, hipGetLastError());
    W = dw;
    H = dh;
  }
  GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->tex_bufs.emplace_back();
  std::swap(ctx->tex_bufs.back().p, buf.p);
  std::swap(ctx->tex_bufs.back().bytes, buf.bytes);
  ctx->texs.push_back({w, h, levels, total});
  std::vector<DevTex> dt(ctx->texs.size());
  for (size_t i = 0; i < dt.size(); ++i)
    dt[i] = DevTex{ctx->tex_bufs[i].as<uchar4>(), ctx->texs[i].w, ctx->texs[i].h, ctx->texs[i].levels, 0};
  GI_HIP(ctx, ctx->d_tex.ensure(dt.size() * sizeof(DevTex)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_tex.p, dt.data(), dt.size() * sizeof(DevTex), hipMemcpyHostToDevice));
  *tex_id = (int32_t)ctx->texs.size() - 1;
  return GI_OK;
}

int gi_texture_info(gi_ctx *ctx, int32_t tex_id, int32_t *levels, int64_t *texels) {
  if (!ctx || tex_id < 0 || tex_id >= (int32_t)ctx->texs.size()) return fail(ctx, GI_ERR_INVALID_ARG, "texture id");
  if (levels) *levels = ctx->texs[(size_t)tex_id].levels;
  if (texels) *texels = (int64_t)ctx->texs[(size_t)tex_id].texels;
  return GI_OK;
}

int gi_texture_grad(gi_ctx *ctx, int32_t tex_id, float *d_rgb) {
  if (!ctx || !d_rgb || tex_id < 0 || tex_id >= (int32_t)ctx->texs.size())
    return fail(ctx, GI_ERR_INVALID_ARG, "texture id");
  if ((size_t)tex_id >= ctx->tex_grad.size() || ctx->tex_grad[(size_t)tex_id].empty())
    return fail(ctx, GI_ERR_STATE, "no texture gradient: gi_adjoint of a textured scene first");
  std::memcpy(d_rgb, ctx->tex_grad[(size_t)tex_id].data(), ctx->tex_grad[(size_t)tex_id].size() * sizeof(float));
  return GI_OK;
}

int gi_download_texture(gi_ctx *ctx, int32_t tex_id, uint8_t *rgba) {
  if (!ctx || !rgba || tex_id < 0 || tex_id >= (int32_t)ctx->texs.size())
    return fail(ctx, GI_ERR_INVALID_ARG, "texture id");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  GI_HIP(ctx, copy_sync(ctx, rgba, ctx->tex_bufs[(size_t)tex_id].p, ctx->texs[(size_t)tex_id].texels * sizeof(uchar4),
                        hipMemcpyDeviceToHost));
  return GI_OK;
}

int gi_upload_mesh(gi_ctx *ctx, const gi_tri *tris, int64_t n, int32_t *mesh_id) {
  if (!ctx || !tris || n <= 0 || !mesh_id) return fail(ctx, GI_ERR_INVALID_ARG, "mesh");
  if (n > (int64_t)kMaxLeafFirst) return fail(ctx, GI_ERR_LIMIT, "mesh of more than 2^27 triangles");
  ctx->meshes.emplace_back(tris, tris + n);
  *mesh_id = (int32_t)ctx->meshes.size() - 1;
  return GI_OK;
}

// Inverse of the affine map x -> A x + t (A = rows of m, t = column 3), in
// double, as 3 rows of (A^-1 | -A^-1 t) rounded to float.  false if singular.
bool invert_affine(const float *m, float *inv) {
  const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
  const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
  const double r[9] = {(e * i - f * h) / det, (c * h - b * i) / det, (b * f - c * e) / det,
                       (f * g - d * i) / det, (a * i - c * g) / det, (c * d - a * f) / det,
                       (d * h - e * g) / det, (b * g - a * h) / det, (a * e - b * d) / det};
  const double t[3] = {m[3], m[7], m[11]};
  for (int row = 0; row < 3; ++row) {
    for (int col = 0; col < 3; ++col) inv[4 * row + col] = (float)r[3 * row + col];
    inv[4 * row + 3] = (float)-(r[3 * row] * t[0] + r[3 * row + 1] * t[1] + r[3 * row + 2] * t[2]);
  }
  return true;
}

int gi_upload_instances(gi_ctx *ctx, const gi_instance *inst, int32_t n) {
  if (!ctx || !inst || n <= 0) return fail(ctx, GI_ERR_INVALID_ARG, "instances");
  if (!ctx->have_bvh) return fail(ctx, GI_ERR_STATE, "build the world BVH before gi_upload_instances");
  if ((uint32_t)n > kMaxInstances) return fail(ctx, GI_ERR_LIMIT, "too many instances");
  const int32_t nm = (int32_t)ctx->meshes.size();
  uint64_t vcount = (uint64_t)ctx->ntri;
  int32_t mlo = INT32_MAX, mhi = INT32_MIN;
  for (int32_t i = 0; i < n; ++i) {
    if (inst[i].mesh < 0 || inst[i].mesh >= nm) return fail(ctx, GI_ERR_INVALID_ARG, "instance mesh id");
    if (inst[i].material < -1) return fail(ctx, GI_ERR_INVALID_ARG, "instance material");
    const std::vector<gi_tri> &mt = ctx->meshes[(size_t)inst[i].mesh];
    vcount += mt.size();
    if (inst[i].material >= 0) {
      mlo = std::min(mlo, inst[i].material);
      mhi = std::max(mhi, inst[i].material);
    } else {
      for (const gi_tri &t : mt) {
        mlo = std::min(mlo, t.material_index);
        mhi = std::max(mhi, t.material_index);
      }
    }
  }
  if (vcount >= 0xffffffffull) return fail(ctx, GI_ERR_LIMIT, "more than 2^32 - 1 triangles with instances");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  const uint32_t n_world = (uint32_t)ctx->ntri;
  // per mesh: its BVH (object space), nodes and triangles appended after the world's
  std::vector<WideNode> xnodes;
  std::vector<IsectTri> xit;
  std::vector<ShadeTri> xst;
  std::vector<uint32_t> root(nm), stri(nm);
  std::vector<float> mbox(6 * (size_t)nm);
  int32_t mesh_depth = 0;
  uint32_t node_base = (uint32_t)ctx->world_nnodes, tri_base = n_world;
  for (int32_t m = 0; m < nm; ++m) {
    const std::vector<gi_tri> &mt = ctx->meshes[(size_t)m];
    gi_bvh b;
    build_bvh2(mt.data(), (int64_t)mt.size(), 4, b);
    std::vector<IsectTri> it;
    leaf_order_tris(b, mt.data(), it);
    std::vector<WideNode> wn;
    int32_t depth = 0;
    const uint32_t tb = tri_base;
    collapse_wide_nodes(
        b, [tb](const Bvh2Node &nd) { return kLeafBit | ((uint32_t)nd.count << 27) | ((uint32_t)nd.first + tb); },
        node_base, wn, depth, kCollapseMergeLimit);
    root[m] = node_base;
    stri[m] = tri_base;
    std::memcpy(&mbox[6 * m], b.nodes[0].lo, 12);
    std::memcpy(&mbox[6 * m + 3], b.nodes[0].hi, 12);
    mesh_depth = std::max(mesh_depth, depth);
    xnodes.insert(xnodes.end(), wn.begin(), wn.end());
    xit.insert(xit.end(), it.begin(), it.end());
    for (const gi_tri &t : mt) xst.push_back(make_shadetri(t));
    node_base += (uint32_t)wn.size();
    tri_base += (uint32_t)mt.size();
  }
  if (tri_base > kMaxLeafFirst) return fail(ctx, GI_ERR_LIMIT, "world + mesh triangles beyond 2^27");
  // instances: transforms, virtual bases, world boxes (mesh box corners
  // mapped in double, rounded outwards, then padded as the builder pads)
  std::vector<DevInstance> di(n);
  std::vector<float> boxes(6 * (size_t)n);
  uint32_t vbase = n_world;
  for (int32_t i = 0; i < n; ++i) {
    const gi_instance &g = inst[i];
    DevInstance &D = di[i];
    std::memset(&D, 0, sizeof(D));
    float inv[12];
    if (!invert_affine(g.xform, inv)) return fail(ctx, GI_ERR_INVALID_ARG, "singular instance transform");
    for (int r = 0; r < 3; ++r) {
      D.fwd[r] = make_float4(g.xform[4 * r], g.xform[4 * r + 1], g.xform[4 * r + 2], g.xform[4 * r + 3]);
      D.inv[r] = make_float4(inv[4 * r], inv[4 * r + 1], inv[4 * r + 2], inv[4 * r + 3]);
    }
    D.root = root[(size_t)g.mesh];
    D.vbase = vbase;
    D.stri = stri[(size_t)g.mesh];
    D.material = g.material;
    vbase += (uint32_t)ctx->meshes[(size_t)g.mesh].size();
    const float *mb = &mbox[6 * (size_t)g.mesh];
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int c = 0; c < 8; ++c) {
      const double p[3] = {mb[(c & 1) ? 3 : 0], mb[(c & 2) ? 4 : 1], mb[(c & 4) ? 5 : 2]};
      for (int r = 0; r < 3; ++r) {
        const double w = (double)g.xform[4 * r] * p[0] + (double)g.xform[4 * r + 1] * p[1] +
                         (double)g.xform[4 * r + 2] * p[2] + (double)g.xform[4 * r + 3];
        lo[r] = std::min(lo[r], w);
        hi[r] = std::max(hi[r], w);
      }
    }
    float *bx = &boxes[6 * (size_t)i];
    for (int r = 0; r < 3; ++r) {
      bx[r] = std::nextafter((float)lo[r], -FLT_MAX);
      bx[3 + r] = std::nextafter((float)hi[r], FLT_MAX);
    }
    pad_box(bx, bx + 3);
  }
  // the world tree's top: the child slots of its upper wide nodes, opened
  // largest surface area first until there are about as many pieces as
  // instances (16..4096); each piece (box, reference) is a primitive of the
  // instance tree, so world subtrees and instances are visited in one
  // distance order (one world leaf under the instance tree measured 1.45x
  // slower on 2000 instances over a 4.2M-triangle world)
  std::vector<float> pbox;
  std::vector<uint32_t> pref;
  auto add_children = [&](const WideNode &nd) {
    for (int k = 0; k < 4; ++k) {
      if (nd.child[k] == kEmptyRef) continue;
      pbox.insert(pbox.end(), {nd.lox[k], nd.loy[k], nd.loz[k], nd.hix[k], nd.hiy[k], nd.hiz[k]});
      pref.push_back(nd.child[k]);
    }
  };
  {
    WideNode nd;
    GI_HIP(ctx, copy_sync(ctx, &nd, ctx->d_nodes.p, sizeof(WideNode), hipMemcpyDeviceToHost));
    add_children(nd);
    const size_t target = std::min<size_t>(4096, std::max<size_t>(16, (size_t)n));
    while (pref.size() < target) {
      int64_t best = -1;
      float ba = -1.0f;
      for (size_t i = 0; i < pref.size(); ++i) {
        if (pref[i] & kLeafBit) continue;
        const float *b = &pbox[6 * i];
        const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
        const float a = dx * dy + dy * dz + dz * dx;
        if (a > ba) {
          ba = a;
          best = (int64_t)i;
        }
      }
      if (best < 0) break;
      GI_HIP(ctx, copy_sync(ctx, &nd, ctx->d_nodes.as<WideNode>() + pref[(size_t)best], sizeof(WideNode),
                            hipMemcpyDeviceToHost));
      pbox.erase(pbox.begin() + 6 * best, pbox.begin() + 6 * best + 6);
      pref.erase(pref.begin() + best);
      add_children(nd);
    }
  }
  boxes.insert(boxes.end(), pbox.begin(), pbox.end());
  gi_bvh tb;
  build_bvh2_boxes(boxes.data(), (int64_t)n + (int64_t)pref.size(), 1, tb);
  std::vector<WideNode> tn;
  int32_t tlas_depth = 0;
  const uint32_t tlas_base = node_base;
  collapse_wide_nodes(
      tb,
      [&](const Bvh2Node &nd) {
        const uint32_t prim = tb.perm[(size_t)nd.first];
        return prim >= (uint32_t)n ? pref[prim - (uint32_t)n] : (kLeafBit | (prim + 1u));
      },
      tlas_base, tn, tlas_depth, 1);  // instance-tree leaves hold one instance or piece each
  // combined device arrays: world part copied on the device, the rest uploaded
  const size_t nw = (size_t)ctx->world_nnodes, nx = xnodes.size(), nt = tn.size();
  DBuf nodes, itris, stris;
  GI_HIP(ctx, nodes.ensure((nw + nx + nt) * sizeof(WideNode)));
  GI_HIP(ctx, itris.ensure((size_t)tri_base * sizeof(IsectTri)));
  GI_HIP(ctx, stris.ensure((size_t)tri_base * sizeof(ShadeTri)));
  GI_HIP(ctx, hipMemcpyAsync(nodes.p, ctx->d_nodes.p, nw * sizeof(WideNode), hipMemcpyDeviceToDevice, ctx->stream));
  GI_HIP(ctx, hipMemcpyAsync(itris.p, ctx->d_itris.p, (size_t)n_world * sizeof(IsectTri), hipMemcpyDeviceToDevice,
                             ctx->stream));
  GI_HIP(ctx, hipMemcpyAsync(stris.p, ctx->d_stris.p, (size_t)n_world * sizeof(ShadeTri), hipMemcpyDeviceToDevice,
                             ctx->stream));
  if (nx) GI_HIP(ctx, copy_sync(ctx, nodes.as<WideNode>() + nw, xnodes.data(), nx * sizeof(WideNode), hipMemcpyHostToDevice));
  GI_HIP(ctx, copy_sync(ctx, nodes.as<WideNode>() + nw + nx, tn.data(), nt * sizeof(WideNode), hipMemcpyHostToDevice));
  if (!xit.empty()) {
    GI_HIP(ctx, copy_sync(ctx, itris.as<IsectTri>() + n_world, xit.data(), xit.size() * sizeof(IsectTri),
                          hipMemcpyHostToDevice));
    GI_HIP(ctx, copy_sync(ctx, stris.as<ShadeTri>() + n_world, xst.data(), xst.size() * sizeof(ShadeTri),
                          hipMemcpyHostToDevice));
  }
  GI_HIP(ctx, ctx->d_inst.ensure((size_t)n * sizeof(DevInstance)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_inst.p, di.data(), (size_t)n * sizeof(DevInstance), hipMemcpyHostToDevice));
  std::vector<uint32_t> vb(n);
  for (int32_t i = 0; i < n; ++i) vb[i] = di[i].vbase;
  GI_HIP(ctx, ctx->d_inst_vbase.ensure((size_t)n * sizeof(uint32_t)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_inst_vbase.p, vb.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice));
  // block table of the virtual ids: blocks of 2^shift ids, 2^shift no more
  // than the smallest instanced mesh (each mesh has >= 1 triangle), so a
  // block holds the start of at most one more instance than its entry
  {
    const uint64_t vtotal = (uint64_t)vcount - (uint64_t)n_world;
    size_t min_mesh = SIZE_MAX;
    for (int32_t i = 0; i < n; ++i) min_mesh = std::min(min_mesh, ctx->meshes[(size_t)inst[i].mesh].size());
    uint32_t shift = 0;
    while (shift < 31 && ((size_t)2 << shift) <= min_mesh) ++shift;
    const size_t nblk = (size_t)((vtotal + ((uint64_t)1 << shift) - 1) >> shift);
    ctx->use_vtab = nblk <= ((size_t)1 << 24);  // else the shading searches
    if (ctx->use_vtab) {
      std::vector<uint32_t> tab(std::max<size_t>(1, nblk));
      int32_t cur = 0;
      for (size_t blk = 0; blk < nblk; ++blk) {
        const uint64_t id = (uint64_t)n_world + ((uint64_t)blk << shift);
        while (cur + 1 < n && (uint64_t)vb[cur + 1] <= id) ++cur;
        tab[blk] = (uint32_t)cur;
      }
      GI_HIP(ctx, ctx->d_inst_vtab.ensure(tab.size() * sizeof(uint32_t)));
      GI_HIP(ctx,
             copy_sync(ctx, ctx->d_inst_vtab.p, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    ctx->inst_vshift = shift;
  }
  if ((uint64_t)(nw + nx + nt) > kMaxWideNodes)
    return fail(ctx, GI_ERR_LIMIT, "more than 2^25 BVH nodes (world, meshes and instance tree)");
  std::swap(ctx->d_nodes.p, nodes.p);
  std::swap(ctx->d_nodes.bytes, nodes.bytes);
  std::swap(ctx->d_itris.p, itris.p);
  std::swap(ctx->d_itris.bytes, itris.bytes);
  std::swap(ctx->d_stris.p, stris.p);
  std::swap(ctx->d_stris.bytes, stris.bytes);
  ctx->nnodes = (int64_t)(nw + nx + nt);
  ctx->tlas_root = tlas_base;
  ctx->n_inst = n;
  ctx->inst_mat_lo = mlo;
  ctx->inst_mat_hi = mhi;
  // on the stack at once: instance-tree levels, world levels below a piece,
  // a kept reference and the return marker, mesh levels (three per level)
  ctx->wide_depth = tlas_depth + 1 + ctx->world_depth + 1 + mesh_depth;
  ctx->stack_cap = (uint32_t)(3 * (ctx->wide_depth + 1) + 4 + 2);
  const int rc = ensure_gstack(ctx);
  if (rc) return rc;
  return make_half_nodes(ctx);
}

int gi_upload_bvh2(gi_ctx *ctx, const float *nodes, int64_t n_nodes, const uint32_t *indices, int64_t n_indices) {
  if (!ctx || !nodes || n_nodes <= 0 || (!indices && n_indices > 0) || n_indices < 0)
    return fail(ctx, GI_ERR_INVALID_ARG, "bvh2 arrays");
  if (ctx->tris.empty()) return fail(ctx, GI_ERR_STATE, "upload triangles first");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  auto b = std::make_unique<gi_bvh>();
  std::string err;
  if (!import_bvh2(nodes, n_nodes, indices, n_indices, (int64_t)ctx->tris.size(), *b, err))
    return fail(ctx, GI_ERR_INVALID_ARG, err);
  ctx->bvh = std::move(b);
  return upload_wide(ctx);
}

int gi_trace_closest_device(gi_ctx *ctx, const float *d_orig, const float *d_dir, const float *d_max_t, int64_t n,
                            gi_hit *d_out) {
  int rc = check_scene(ctx, false);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!d_orig || !d_dir || !d_max_t || !d_out))) return fail(ctx, GI_ERR_INVALID_ARG, "rays");
  if (n == 0) return GI_OK;
  if (n >= (int64_t)1 << 32) return fail(ctx, GI_ERR_LIMIT, "more than 2^32 rays");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  if ((rc = ensure_gstack(ctx))) return rc;
  GI_HIP(ctx, ctx->p_ro.ensure(n * sizeof(float4)));
  GI_HIP(ctx, ctx->p_rd.ensure(n * sizeof(float4)));
  // the stack-overflow counter counts this call only
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->d_overflow.as<uint32_t>(), 1u, 0u);
  hipLaunchKernelGGL(k_pack_rays, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, ctx->stream, d_orig, d_dir,
                     d_max_t, (uint32_t)n, ctx->p_ro.as<float4>(), ctx->p_rd.as<float4>());
  GI_HIP(ctx, hipGetLastError());
  rc = launch_closest(ctx, true, nullptr, (uint32_t)n, (uint64_t)n, ctx->p_ro.as<float4>(), ctx->p_rd.as<float4>(),
                      reinterpret_cast<float4 *>(d_out));
  if (rc) return rc;
  GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return check_overflow(ctx);
}

int gi_trace_any_device(gi_ctx *ctx, const float *d_orig, const float *d_dir, const float *d_max_t, int64_t n,
                        uint8_t *d_occluded) {
  int rc = check_scene(ctx, false);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!d_orig || !d_dir || !d_max_t || !d_occluded)))
    return fail(ctx, GI_ERR_INVALID_ARG, "rays");
  if (n == 0) return GI_OK;
  if (n >= (int64_t)1 << 32) return fail(ctx, GI_ERR_LIMIT, "more than 2^32 rays");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  if ((rc = ensure_gstack(ctx))) return rc;
  GI_HIP(ctx, ctx->p_ro.ensure(n * sizeof(float4)));
  GI_HIP(ctx, ctx->p_rd.ensure(n * sizeof(float4)));
  // the stack-overflow counter counts this call only
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->d_overflow.as<uint32_t>(), 1u, 0u);
  hipLaunchKernelGGL(k_pack_rays, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, ctx->stream, d_orig, d_dir,
                     d_max_t, (uint32_t)n, ctx->p_ro.as<float4>(), ctx->p_rd.as<float4>());
  GI_HIP(ctx, hipGetLastError());
  rc = launch_any(ctx, false, nullptr, (uint32_t)n, (uint64_t)n, ctx->p_ro.as<float4>(), ctx->p_rd.as<float4>(),
                  nullptr, nullptr, d_occluded);
  if (rc) return rc;
  GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return check_overflow(ctx);
}

int gi_trace_closest_device_noalloc(gi_ctx *ctx, const float *d_orig, const float *d_dir, const float *d_max_t,
                                    int64_t n, gi_hit *d_out) {
  int rc = check_scene(ctx, false);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!d_orig || !d_dir || !d_max_t || !d_out))) return fail(ctx, GI_ERR_INVALID_ARG, "rays");
  if (n == 0) return GI_OK;
  if (n > (int64_t)kNoAllocRays)
    return fail(ctx, GI_ERR_LIMIT, "gi_trace_closest_device_noalloc: at most 524288 rays per call");
  if (!ctx->scratch) return fail(ctx, GI_ERR_STATE, "no scratch array on this device");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  uint32_t *overflow = reinterpret_cast<uint32_t *>(ctx->scratch);
  float4 *ro = reinterpret_cast<float4 *>(ctx->scratch + 256);
  float4 *rd = ro + kNoAllocRays;
  uint32_t *gstack = reinterpret_cast<uint32_t *>(ctx->scratch + 256 + kNoAllocRayBytes);
  // one global stack column per thread, as deep as the tree needs beyond
  // the smallest LDS part; the grid is capped to the columns the array holds
  int grid, block;
  if ((rc = trace_grid_for(ctx, (uint64_t)n, grid, block))) return rc;
  const size_t gdepth = ctx->stack_cap > (uint32_t)kMinLdsDepth ? ctx->stack_cap - kMinLdsDepth : 1;
  const size_t max_grid = kNoAllocStackWords / (gdepth * (size_t)block);
  if (max_grid < 1) return fail(ctx, GI_ERR_LIMIT, "gi_trace_closest_device_noalloc: BVH too deep for the scratch stack");
  TraceScratch ts;
  ts.gstack = gstack;
  ts.overflow = overflow;
  ts.max_grid = (int)std::min<size_t>(max_grid, (size_t)INT32_MAX);
  // the counter keeps the last call's value: this call starts it at zero
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, overflow, 1u, 0u);
  hipLaunchKernelGGL(k_pack_rays, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, ctx->stream, d_orig, d_dir,
                     d_max_t, (uint32_t)n, ro, rd);
  GI_HIP(ctx, hipGetLastError());
  rc = launch_closest(ctx, true, nullptr, (uint32_t)n, (uint64_t)n, ro, rd, reinterpret_cast<float4 *>(d_out), ts);
  if (rc) return rc;
  uint32_t ov = 0;
  GI_HIP(ctx, hipMemcpyAsync(&ov, overflow, 4, hipMemcpyDeviceToHost, ctx->stream));
  GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ov != 0 ? fail(ctx, GI_ERR_LIMIT, "traversal stack overflow") : GI_OK;
}

int gi_trace_closest(gi_ctx *ctx, const float *orig, const float *dir, const float *max_t, int64_t n, gi_hit *out) {
  int rc = check_scene(ctx, false);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!orig || !dir || !max_t || !out))) return fail(ctx, GI_ERR_INVALID_ARG, "rays");
  if (n == 0) return GI_OK;
  GI_HIP(ctx, hipSetDevice(ctx->device));
  GI_HIP(ctx, ctx->p_orig.ensure(n * 12));
  GI_HIP(ctx, ctx->p_dir.ensure(n * 12));
  GI_HIP(ctx, ctx->p_maxt.ensure(n * 4));
  GI_HIP(ctx, ctx->p_hit.ensure(n * sizeof(gi_hit)));
  GI_HIP(ctx, copy_sync(ctx, ctx->p_orig.p, orig, n * 12, hipMemcpyHostToDevice));
  GI_HIP(ctx, copy_sync(ctx, ctx->p_dir.p, dir, n * 12, hipMemcpyHostToDevice));
  GI_HIP(ctx, copy_sync(ctx, ctx->p_maxt.p, max_t, n * 4, hipMemcpyHostToDevice));
  rc = gi_trace_closest_device(ctx, ctx->p_orig.as<float>(), ctx->p_dir.as<float>(), ctx->p_maxt.as<float>(), n,
                               ctx->p_hit.as<gi_hit>());
  if (rc) return rc;
  GI_HIP(ctx, copy_sync(ctx, out, ctx->p_hit.p, n * sizeof(gi_hit), hipMemcpyDeviceToHost));
  return GI_OK;
}

int gi_trace_any(gi_ctx *ctx, const float *orig, const float *dir, const float *max_t, int64_t n, uint8_t *occluded) {
  int rc = check_scene(ctx, false);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!orig || !dir || !max_t || !occluded))) return fail(ctx, GI_ERR_INVALID_ARG, "rays");
  if (n == 0) return GI_OK;
  GI_HIP(ctx, hipSetDevice(ctx->device));
  GI_HIP(ctx, ctx->p_orig.ensure(n * 12));
  GI_HIP(ctx, ctx->p_dir.ensure(n * 12));
  GI_HIP(ctx, ctx->p_maxt.ensure(n * 4));
  GI_HIP(ctx, ctx->p_occ.ensure(n));
  GI_HIP(ctx, copy_sync(ctx, ctx->p_orig.p, orig, n * 12, hipMemcpyHostToDevice));
  GI_HIP(ctx, copy_sync(ctx, ctx->p_dir.p, dir, n * 12, hipMemcpyHostToDevice));
  GI_HIP(ctx, copy_sync(ctx, ctx->p_maxt.p, max_t, n * 4, hipMemcpyHostToDevice));
  rc = gi_trace_any_device(ctx, ctx->p_orig.as<float>(), ctx->p_dir.as<float>(), ctx->p_maxt.as<float>(), n,
                           ctx->p_occ.as<uint8_t>());
  if (rc) return rc;
  GI_HIP(ctx, copy_sync(ctx, occluded, ctx->p_occ.p, n, hipMemcpyDeviceToHost));
  return GI_OK;
}

int gi_generate_camera_rays(gi_ctx *ctx, const gi_view *v, int32_t sample, uint32_t seed, float *orig, float *dir,
                            float *max_t) {
  if (!ctx || !v || !orig || !dir || !max_t) return fail(ctx, GI_ERR_INVALID_ARG, "camera rays");
  if (sample < 0 || sample >= v->spp) return fail(ctx, GI_ERR_INVALID_ARG, "sample index");
  if ((uint64_t)v->w * (uint64_t)v->h * (uint64_t)v->spp >= (1ull << 32))
    return fail(ctx, GI_ERR_LIMIT, "more than 2^32 samples in one view");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  // the view, pass and Halton state of this call are its own: the last
  // render's views, options and images stay as they are
  DevView dv;
  std::string err;
  if (!make_devview(*v, dv, err)) return fail(ctx, GI_ERR_INVALID_ARG, err);
  int rc = ensure_halton(ctx, seed, dims_for_depth(1));
  if (rc) return rc;
  const uint32_t npix = (uint32_t)((uint64_t)v->w * v->h);
  PassDesc pd;
  std::memset(&pd, 0, sizeof(pd));
  pd.nchunks = 1;
  pd.total_slots = npix;
  pd.total_pix = npix;
  pd.chunk[0] = PassChunk{0, sample, 1, (int32_t)npix, 0, 0, {0, 0}};
  GI_HIP(ctx, ctx->d_cam_pass.ensure(sizeof(PassDesc)));
  GI_HIP(ctx, ctx->d_cam_view.ensure(sizeof(DevView)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_cam_pass.p, &pd, sizeof(pd), hipMemcpyHostToDevice));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_cam_view.p, &dv, sizeof(dv), hipMemcpyHostToDevice));
  rc = ensure_workspace(ctx, npix);
  if (rc) return rc;
  HaltonDev H{ctx->d_hdims.as<HaltonDim>(), ctx->d_hchunks.as<uint16_t>(), 0u};
  hipLaunchKernelGGL(k_raygen, dim3((npix + 255) / 256), dim3(256), 0, ctx->stream, ctx->d_cam_pass.as<PassDesc>(),
                     ctx->d_cam_view.as<DevView>(), H, ctx->q_o[0].as<float4>(), ctx->q_d[0].as<float4>(),
                     ctx->path_l.as<float4>(), (float4 *)nullptr, 0u, 1);
  GI_HIP(ctx, hipGetLastError());
  std::vector<float4> ro(npix), rd(npix);
  GI_HIP(ctx, hipMemcpyAsync(ro.data(), ctx->q_o[0].p, npix * 16, hipMemcpyDeviceToHost, ctx->stream));
  GI_HIP(ctx, hipMemcpyAsync(rd.data(), ctx->q_d[0].p, npix * 16, hipMemcpyDeviceToHost, ctx->stream));
  GI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (uint32_t i = 0; i < npix; ++i) {
    orig[3 * i] = ro[i].x; orig[3 * i + 1] = ro[i].y; orig[3 * i + 2] = ro[i].z;
    dir[3 * i] = rd[i].x; dir[3 * i + 1] = rd[i].y; dir[3 * i + 2] = rd[i].z;
    max_t[i] = FLT_MAX;
  }
  return GI_OK;
}

int gi_render(gi_ctx *ctx, const gi_view *views, int32_t n_views, const gi_render_opts *opts) {
  int rc = check_scene(ctx, true);
  if (rc) return rc;
  GI_HIP(ctx, hipSetDevice(ctx->device));
  auto t0 = std::chrono::steady_clock::now();
  std::memset(&ctx->counters, 0, sizeof(ctx->counters));
  rc = setup_views(ctx, views, n_views, opts);
  if (rc) return rc;
  ctx->prog_active = false;
  std::vector<int32_t> sb(n_views, 0), se(n_views);
  for (int32_t v = 0; v < n_views; ++v) {
    se[v] = views[v].spp;
    ctx->counters.samples += (uint64_t)views[v].w * views[v].h * views[v].spp;
  }
  const bool batched = (opts->flags & GI_FLAG_BATCHED) != 0;
  const bool graph = (opts->flags & GI_FLAG_GRAPH) != 0;
  const int32_t spp_pp = opts->samples_per_pass > 0 ? opts->samples_per_pass : ctx->tuning.samples_per_pass;
  PassPlan plan;
  plan_passes(ctx->views, sb, se, batched, target_slots(ctx), spp_pp, plan);
  rc = run_passes(ctx, plan, opts->max_depth, !batched, graph, !graph);
  if (rc) return rc;
  ctx->counters.ms_total =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return GI_OK;
}

int gi_progress_begin(gi_ctx *ctx, const gi_view *views, int32_t n_views, const gi_render_opts *opts) {
  int rc = check_scene(ctx, true);
  if (rc) return rc;
  GI_HIP(ctx, hipSetDevice(ctx->device));
  std::memset(&ctx->counters, 0, sizeof(ctx->counters));
  rc = setup_views(ctx, views, n_views, opts);
  if (rc) return rc;
  ctx->prog_sample = 0;
  ctx->prog_max = 0;
  for (int32_t v = 0; v < n_views; ++v) ctx->prog_max = std::max(ctx->prog_max, views[v].spp);
  ctx->prog_active = true;
  return GI_OK;
}

int gi_progress_step(gi_ctx *ctx, int32_t *done) {
  if (!ctx || !done) return fail(ctx, GI_ERR_INVALID_ARG, "progress_step");
  if (!ctx->prog_active) return fail(ctx, GI_ERR_STATE, "gi_progress_begin first");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  // gi_generate_camera_rays may have re-seeded the tables since
  int hrc = ensure_halton(ctx, ctx->opts.seed, dims_for_depth(ctx->opts.max_depth));
  if (hrc) return hrc;
  const int32_t nv = (int32_t)ctx->views.size();
  std::vector<int32_t> sb(nv), se(nv);
  for (int32_t v = 0; v < nv; ++v) {
    sb[v] = std::min(ctx->prog_sample, ctx->views[v].spp);
    se[v] = std::min(ctx->prog_sample + 1, ctx->views[v].spp);
    ctx->counters.samples += (uint64_t)ctx->views[v].w * ctx->views[v].h * (uint64_t)(se[v] - sb[v]);
  }
  PassPlan plan;
  plan_passes(ctx->views, sb, se, true, target_slots(ctx), 1, plan);
  if (!plan.passes.empty()) {
    int rc = run_passes(ctx, plan, ctx->opts.max_depth, false, false, true);
    if (rc) return rc;
  }
  ++ctx->prog_sample;
  *done = ctx->prog_sample >= ctx->prog_max ? 1 : 0;
  if (*done) ctx->prog_active = false;
  return GI_OK;
}

int gi_get_image(gi_ctx *ctx, int32_t view, float *rgb) {
  if (!ctx || !rgb) return fail(ctx, GI_ERR_INVALID_ARG, "get_image");
  if (view < 0 || view >= (int32_t)ctx->views.size()) return fail(ctx, GI_ERR_INVALID_ARG, "view index");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  const size_t npix = (size_t)ctx->views[view].w * ctx->views[view].h;
  std::vector<float4> tmp(npix);
  GI_HIP(ctx, copy_sync(ctx, tmp.data(), ctx->images[view].p, npix * sizeof(float4), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < npix; ++i) {
    rgb[3 * i] = tmp[i].x;
    rgb[3 * i + 1] = tmp[i].y;
    rgb[3 * i + 2] = tmp[i].z;
  }
  return GI_OK;
}

int gi_get_image_device(gi_ctx *ctx, int32_t view, float **d_rgb) {
  if (!ctx || !d_rgb) return fail(ctx, GI_ERR_INVALID_ARG, "get_image_device");
  if (view < 0 || view >= (int32_t)ctx->views.size()) return fail(ctx, GI_ERR_INVALID_ARG, "view index");
  *d_rgb = ctx->images[view].as<float>();
  return GI_OK;
}

int gi_get_counters(const gi_ctx *ctx, gi_counters *out) {
  if (!ctx || !out) return GI_ERR_INVALID_ARG;
  *out = ctx->counters;
  return GI_OK;
}

int gi_set_traversal_stats(gi_ctx *ctx, int32_t enable) {
  if (!ctx) return GI_ERR_INVALID_ARG;
  if (enable) {
    GI_HIP(ctx, hipSetDevice(ctx->device));
    GI_HIP(ctx, ctx->d_tstats.ensure(sizeof(gi_traversal_stats)));
    const gi_traversal_stats zero{};
    GI_HIP(ctx, copy_sync(ctx, ctx->d_tstats.p, &zero, sizeof(zero), hipMemcpyHostToDevice));
  }
  ctx->count_traversal = enable != 0;
  return GI_OK;
}

int gi_check_half_nodes(gi_ctx *ctx, int64_t *bad_slots) {
  if (!ctx || !bad_slots) return fail(ctx, GI_ERR_INVALID_ARG, "check_half_nodes");
  if (!ctx->have_bvh) return fail(ctx, GI_ERR_STATE, "no BVH");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  DBuf cnt;
  GI_HIP(ctx, cnt.ensure(sizeof(unsigned long long)));
  const unsigned long long zero = 0;
  GI_HIP(ctx, copy_sync(ctx, cnt.p, &zero, sizeof(zero), hipMemcpyHostToDevice));
  if (ctx->nnodes > 0) {
    hipLaunchKernelGGL(k_check_half_nodes, dim3((unsigned)((ctx->nnodes + 255) / 256)), dim3(256), 0, ctx->stream,
                       ctx->d_nodes.as<WideNode>(), ctx->d_hnodes.as<HalfNode>(), (uint32_t)ctx->nnodes,
                       cnt.as<unsigned long long>());
    GI_HIP(ctx, hipGetLastError());
  }
  unsigned long long n = 0;
  GI_HIP(ctx, copy_sync(ctx, &n, cnt.p, sizeof(n), hipMemcpyDeviceToHost));
  *bad_slots = (int64_t)n;
  return GI_OK;
}

int gi_get_traversal_stats(gi_ctx *ctx, gi_traversal_stats *out) {
  if (!ctx || !out) return fail(ctx, GI_ERR_INVALID_ARG, "traversal stats");
  if (!ctx->d_tstats.p) return fail(ctx, GI_ERR_STATE, "traversal statistics were never enabled");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  GI_HIP(ctx, copy_sync(ctx, out, ctx->d_tstats.p, sizeof(*out), hipMemcpyDeviceToHost));
  return GI_OK;
}

int gi_set_tuning(gi_ctx *ctx, const gi_tuning *t) {
  if (!ctx || !t) return fail(ctx, GI_ERR_INVALID_ARG, "tuning");
  gi_tuning n = *t;
  if (n.trace_block <= 0) n.trace_block = 256;
  if (n.trace_blocks_per_cu <= 0) n.trace_blocks_per_cu = 6;
  if (n.shade_block <= 0) n.shade_block = 256;
  if (n.stack_lds <= 0) n.stack_lds = 16;
  // every kernel is compiled with __launch_bounds__(256)
  if (n.trace_block % 64 || n.trace_block > 256 || n.shade_block % 64 || n.shade_block > 256)
    return fail(ctx, GI_ERR_INVALID_ARG, "block sizes must be multiples of 64, at most 256");
  if (n.stack_lds != 8 && n.stack_lds != 16 && n.stack_lds != 32)
    return fail(ctx, GI_ERR_INVALID_ARG, "stack_lds must be 8, 16 or 32");
  if (n.target_slots < 0 || n.target_slots > 256 || n.trace_blocks_per_cu > 64 || n.samples_per_pass < 0)
    return fail(ctx, GI_ERR_INVALID_ARG, "target_slots in [0,256], trace_blocks_per_cu <= 64");
  if (n.node_width == 0) n.node_width = 4;
  if (n.refill_lanes == 0) n.refill_lanes = 32;
  if (n.node_width < 2 || n.node_width > 4) return fail(ctx, GI_ERR_INVALID_ARG, "node_width must be 2, 3 or 4");
  if (n.refill_lanes != 8 && n.refill_lanes != 16 && n.refill_lanes != 32 && n.refill_lanes != 64)
    return fail(ctx, GI_ERR_INVALID_ARG, "refill_lanes must be 8, 16, 32 or 64");
  const bool recollapse = n.node_width != ctx->tuning.node_width && ctx->bvh && ctx->have_bvh;
  if (recollapse && ctx->n_inst > 0)
    return fail(ctx, GI_ERR_STATE, "node_width changes the world tree: set it before gi_upload_instances");
  ctx->tuning = n;
  if (recollapse) {
    GI_HIP(ctx, hipSetDevice(ctx->device));
    return upload_wide(ctx);
  }
  return GI_OK;
}

int gi_get_tuning(const gi_ctx *ctx, gi_tuning *t) {
  if (!ctx || !t) return GI_ERR_INVALID_ARG;
  *t = ctx->tuning;
  return GI_OK;
}

int gi_adjoint(gi_ctx *ctx, const float *dL_dimage, gi_material_grad *d_mats, float *d_emission) {
  int rc = check_scene(ctx, true);
  if (rc) return rc;
  if (!dL_dimage || !d_mats || !d_emission) return fail(ctx, GI_ERR_INVALID_ARG, "adjoint buffers");
  if (ctx->views.empty()) return fail(ctx, GI_ERR_STATE, "gi_render first");
  if (ctx->opts.rr_start > 0) return fail(ctx, GI_ERR_INVALID_ARG, "the adjoint does not support Russian roulette");
  if (ctx->opts.flags & GI_FLAG_PHONG)
    return fail(ctx, GI_ERR_INVALID_ARG, "the adjoint does not support the Phong material model");
  if (ctx->sky_light >= 0) return fail(ctx, GI_ERR_INVALID_ARG, "the adjoint does not support sky lights");
  const int32_t nm = (int32_t)ctx->mats.size(), nl = (int32_t)ctx->lights.size();
  if ((size_t)(5 * nm + 3 * nl) * sizeof(float) > 48 * 1024)
    return fail(ctx, GI_ERR_LIMIT, "adjoint: 5*materials + 3*lights must fit 48 KiB of LDS");
  GI_HIP(ctx, hipSetDevice(ctx->device));
  // the replay needs the render's own Halton points (gi_generate_camera_rays
  // may have re-seeded the tables since)
  rc = ensure_halton(ctx, ctx->opts.seed, dims_for_depth(ctx->opts.max_depth));
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  std::memset(&ctx->counters, 0, sizeof(ctx->counters));
  const int32_t nv = (int32_t)ctx->views.size();
  size_t total_pix = 0;
  std::vector<int32_t> sb(nv, 0), se(nv);
  for (int32_t v = 0; v < nv; ++v) {
    total_pix += (size_t)ctx->views[v].w * ctx->views[v].h;
    se[v] = ctx->views[v].spp;
    ctx->counters.samples += (uint64_t)ctx->views[v].w * ctx->views[v].h * ctx->views[v].spp;
  }
  GI_HIP(ctx, ctx->d_dl.ensure(total_pix * 3 * sizeof(float)));
  GI_HIP(ctx, copy_sync(ctx, ctx->d_dl.p, dL_dimage, total_pix * 3 * sizeof(float), hipMemcpyHostToDevice));
  GI_HIP(ctx, ctx->d_gmats.ensure((size_t)5 * nm * sizeof(float)));
  GI_HIP(ctx, ctx->d_gem.ensure((size_t)3 * nl * sizeof(float)));
  GI_HIP(ctx, hipMemsetAsync(ctx->d_gmats.p, 0, (size_t)5 * nm * sizeof(float), ctx->stream));
  GI_HIP(ctx, hipMemsetAsync(ctx->d_gem.p, 0, (size_t)3 * nl * sizeof(float), ctx->stream));
  const int32_t spp_pp = ctx->opts.samples_per_pass > 0 ? ctx->opts.samples_per_pass : ctx->tuning.samples_per_pass;
  PassPlan plan;
  // the record holds 2 float4 per path and vertex: passes of at most
  // 32 GiB of record (128 Mi paths up to depth 8)
  const size_t rec_cap = ((size_t)32 << 30) / ((size_t)2 * ctx->opts.max_depth * sizeof(float4));
  plan_passes(ctx->views, sb, se, true, std::min(target_slots(ctx), std::max<size_t>(rec_cap, (size_t)1 << 20)),
              spp_pp, plan);
  AdjParams A;
  std::memset(&A, 0, sizeof(A));
  A.dL = ctx->d_dl.as<float>();
  A.mg = ctx->d_matgrad.as<DevMatGrad>();
  A.g_mats = ctx->d_gmats.as<float>();
  A.g_em = ctx->d_gem.as<float>();
  A.nmats = nm;
  A.nlights = nl;
  // texel gradients: RGB per texel of every texture's chain
  std::vector<uint32_t> goff(ctx->texs.size() + 1, 0);
  for (size_t t = 0; t < ctx->texs.size(); ++t) goff[t + 1] = goff[t] + (uint32_t)ctx->texs[t].texels;
  const size_t gtex = ctx->textured ? (size_t)goff.back() : 0;
  // coarse levels (<= kCoarseTexels texels) and their replicas
  std::vector<uint32_t> coarse(ctx->texs.size() + 1, 0), gc(ctx->texs.size() + 1, 0);
  for (size_t t = 0; t < ctx->texs.size(); ++t) {
    const gi_ctx::TexHost &T = ctx->texs[t];
    uint32_t o = 0;
    int32_t W = T.w, H = T.h;
    for (int32_t l = 0; l <= T.levels && (uint64_t)W * H > kCoarseTexels; ++l) {
      o += (uint32_t)W * (uint32_t)H;  // o <= texels: at most every level is skipped
      W = (W + 1) / 2;
      H = (H + 1) / 2;
    }
    coarse[t] = o;
    gc[t + 1] = gc[t] + ((uint32_t)T.texels - o);
  }
  const size_t nrep = (size_t)kGradReplicas * gc[ctx->texs.size()] * 3;
  if (gtex) {
    GI_HIP(ctx, ctx->d_gtex.ensure(gtex * 3 * sizeof(float)));
    GI_HIP(ctx, hipMemsetAsync(ctx->d_gtex.p, 0, gtex * 3 * sizeof(float), ctx->stream));
    GI_HIP(ctx, ctx->d_grep.ensure(nrep * sizeof(float)));
    GI_HIP(ctx, hipMemsetAsync(ctx->d_grep.p, 0, nrep * sizeof(float), ctx->stream));
    std::vector<uint32_t> tab(goff);
    tab.insert(tab.end(), coarse.begin(), coarse.end());
    tab.insert(tab.end(), gc.begin(), gc.end());
    GI_HIP(ctx, ctx->d_gtex_off.ensure(tab.size() * sizeof(uint32_t)));
    GI_HIP(ctx, copy_sync(ctx, ctx->d_gtex_off.p, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    A.g_tex = ctx->d_gtex.as<float>();
    A.tex_goff = ctx->d_gtex_off.as<uint32_t>();
    A.tex_coarse = A.tex_goff + goff.size();
    A.tex_gc = A.tex_coarse + coarse.size();
    A.g_rep = ctx->d_grep.as<float>();
    A.n_coarse = gc[ctx->texs.size()];
  }
  rc = run_adjoint_passes(ctx, plan, ctx->opts.max_depth, A);
  if (rc) return rc;
  ctx->tex_grad.assign(ctx->texs.size(), std::vector<float>());
  for (size_t t = 0; gtex && t < ctx->texs.size(); ++t) {
    const gi_ctx::TexHost &T = ctx->texs[t];
    std::vector<uint32_t> loff(T.levels + 1);
    std::vector<int32_t> lw(T.levels + 1), lh(T.levels + 1);
    uint32_t o = 0;
    int32_t W = T.w, H = T.h;
    for (int32_t l = 0; l <= T.levels; ++l) {
      loff[l] = o;
      lw[l] = W;
      lh[l] = H;
      o += (uint32_t)W * (uint32_t)H;
      W = (W + 1) / 2;
      H = (H + 1) / 2;
    }
    float *g = ctx->d_gtex.as<float>() + 3 * (size_t)goff[t];
    {
      const uint32_t n = gc[t + 1] - gc[t];
      hipLaunchKernelGGL(k_tex_rep_sum, dim3(std::min<uint32_t>((n + 255) / 256, 4096)), dim3(256), 0, ctx->stream, g,
                         coarse[t], ctx->d_grep.as<float>(), gc[t], n, gc[ctx->texs.size()], kGradReplicas);
      GI_HIP(ctx, hipGetLastError());
    }
    for (int32_t l = T.levels - 1; l >= 0; --l) {
      const uint32_t n = (uint32_t)lw[l] * (uint32_t)lh[l];
      hipLaunchKernelGGL(k_tex_grad_down, dim3(std::min<uint32_t>((n + 255) / 256, 4096)), dim3(256), 0,
                         ctx->stream, g, loff[l], lw[l], lh[l], loff[l + 1], lw[l + 1]);
      GI_HIP(ctx, hipGetLastError());
    }
    ctx->tex_grad[t].resize((size_t)T.w * T.h * 3);
    GI_HIP(ctx, copy_sync(ctx, ctx->tex_grad[t].data(), g, ctx->tex_grad[t].size() * sizeof(float),
                          hipMemcpyDeviceToHost));
  }
  std::vector<float> gm((size_t)5 * nm), ge((size_t)3 * nl);
  GI_HIP(ctx, copy_sync(ctx, gm.data(), ctx->d_gmats.p, gm.size() * sizeof(float), hipMemcpyDeviceToHost));
  GI_HIP(ctx, copy_sync(ctx, ge.data(), ctx->d_gem.p, ge.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int32_t i = 0; i < nm; ++i) {
    for (int k = 0; k < 3; ++k) d_mats[i].d_color[k] = gm[5 * i + k];
    d_mats[i].d_roughness = gm[5 * i + 3];
    d_mats[i].d_metallic = gm[5 * i + 4];
  }
  std::memcpy(d_emission, ge.data(), ge.size() * sizeof(float));
  ctx->counters.ms_total =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return GI_OK;
}

float gi_halton_sample(uint64_t n, int32_t dim, uint32_t seed) {
  if (dim < 0 || dim >= 1024) return -1.0f;
  return halton_host(n, dim, seed);
}

float gi_radical_inverse(uint64_t index, int32_t base) {
  // gi::halton, src/util.cpp:22-32
  if (base < 2) return -1.0f;
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

int32_t gi_halton_prime(int32_t dim) {
  if (dim < 0 || dim >= 1024) return -1;
  return prime_of(dim);
}

int gi_halton_permutation(int32_t dim, uint32_t seed, uint16_t *perm) {
  if (dim < 0 || dim >= 1024 || !perm) return GI_ERR_INVALID_ARG;
  halton_permutation(dim, seed, perm);
  return GI_OK;
}

float gi_sky_light_power(float scale, float diameter) {
  // l.power = s*4*M_PI*(d*.5)*(d*.5) (src/lights.cpp:104), evaluated in double
  return (float)((double)(scale * 4.0f) * M_PI * ((double)diameter * 0.5) * ((double)diameter * 0.5));
}

float gi_rect_light_power(const gi_light *l) { return l ? rect_power(*l) : 0.0f; }

}  // extern "C"
