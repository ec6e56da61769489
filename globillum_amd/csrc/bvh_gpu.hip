// Device BVH build.  The binary tree is either a linear BVH (Karras 2012,
// "Maximizing parallelism in the construction of BVHs, octrees and k-d
// trees") over 30-bit Morton codes of the triangle centroids with a box
// refit level by level from the deepest, or a PLOC clustering over the
// 63-bit Morton order (below).  Either is then collapsed into the traversal
// kernels' 4-wide nodes by the host builder's surface-area-cost dynamic
// program (which subtrees become leaves of at most max_leaf triangles, and
// which binary nodes fill each wide node's slots): costs bottom up, one
// launch per radix-tree level or clustering round, then the nodes top down
// level by level -- the same node format, leaf order conventions and
// collapse as the host builder's collapse_wide.  The device counterpart of
// the reference's GPU LBVH construction (src/main.cpp:218-225).
//
// Boxes: a leaf box is the triangle's box padded by 2^-20 of its largest
// coordinate magnitude / extent, as the host builder pads; an interior box
// is the union of its children's.  Every box therefore contains its
// triangles with the margin the traversal's conservative slab test needs,
// and closest-hit results do not depend on which builder made the tree.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bvh_gpu.h"

namespace gi {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ uint32_t quant10(float x) {
  const float q = x * 1024.0f;
  return q <= 0.0f ? 0u : (q >= 1023.0f ? 1023u : (uint32_t)q);
}

// Morton key of each triangle's centroid in the centroid bounds.
__global__ void k_morton(const float4 *__restrict__ tri, uint32_t stride, uint32_t n, float3 lo, float3 scale,
                         uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 a = tri[(size_t)i * stride], b = tri[(size_t)i * stride + 1], c = tri[(size_t)i * stride + 2];
  const float cx = (a.x + b.x + c.x) * (1.0f / 3.0f), cy = (a.y + b.y + c.y) * (1.0f / 3.0f),
              cz = (a.z + b.z + c.z) * (1.0f / 3.0f);
  const uint32_t x = quant10((cx - lo.x) * scale.x), y = quant10((cy - lo.y) * scale.y),
                 z = quant10((cz - lo.z) * scale.z);
  keys[i] = (spread10(x) << 2) | (spread10(y) << 1) | spread10(z);
  vals[i] = i;
}

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__device__ __forceinline__ uint32_t quant21(float x) {
  const float q = x * 2097152.0f;
  return q <= 0.0f ? 0u : (q >= 2097151.0f ? 2097151u : (uint32_t)q);
}

// 63-bit Morton key (21 bits per axis), for the clustering builder.
__global__ void k_morton64(const float4 *__restrict__ tri, uint32_t stride, uint32_t n, float3 lo, float3 scale,
                           uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 a = tri[(size_t)i * stride], b = tri[(size_t)i * stride + 1], c = tri[(size_t)i * stride + 2];
  const float cx = (a.x + b.x + c.x) * (1.0f / 3.0f), cy = (a.y + b.y + c.y) * (1.0f / 3.0f),
              cz = (a.z + b.z + c.z) * (1.0f / 3.0f);
  const uint32_t x = quant21((cx - lo.x) * scale.x), y = quant21((cy - lo.y) * scale.y),
                 z = quant21((cz - lo.z) * scale.z);
  keys[i] = (spread21(x) << 2) | (spread21(y) << 1) | spread21(z);
  vals[i] = i;
}

// Common-prefix length of sorted keys i and j (index bits break ties).
__device__ __forceinline__ int delta(const uint32_t *__restrict__ k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = k[i], b = k[j];
  return a == b ? 32 + __clz(i ^ j) : __clz((int)(a ^ b));
}

// Internal node i of the radix tree over n sorted keys.  Children are coded
// c >= 0: internal node c; c < 0: leaf ~c (a sorted position).
__global__ void k_karras(const uint32_t *__restrict__ keys, int n, int32_t *__restrict__ left,
                         int32_t *__restrict__ right, int32_t *__restrict__ parent, int32_t *__restrict__ leaf_parent,
                         uint32_t *__restrict__ first, uint32_t *__restrict__ last) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n - 1) return;
  const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(keys, n, i, j);
  int s = 0;
  for (int div = 2;; div *= 2) {
    const int t = (l + div - 1) / div;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  const int g = i + s * d + (d < 0 ? -1 : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const int32_t lc = lo == g ? ~g : g, rc = hi == g + 1 ? ~(g + 1) : g + 1;
  left[i] = lc;
  right[i] = rc;
  first[i] = (uint32_t)lo;
  last[i] = (uint32_t)hi;
  if (lc < 0) leaf_parent[~lc] = i;
  else parent[lc] = i;
  if (rc < 0) leaf_parent[~rc] = i;
  else parent[rc] = i;
  if (i == 0) parent[0] = -1;
}

struct Box {
  float lo[3], hi[3];
};

// The padded box of triangle t (the host builder's padding).
__device__ __forceinline__ Box leaf_box(const float4 *__restrict__ tri, uint32_t stride, size_t t) {
  const float4 a = tri[t * stride], b = tri[t * stride + 1], c = tri[t * stride + 2];
  Box bx;
  bx.lo[0] = fminf(a.x, fminf(b.x, c.x));
  bx.lo[1] = fminf(a.y, fminf(b.y, c.y));
  bx.lo[2] = fminf(a.z, fminf(b.z, c.z));
  bx.hi[0] = fmaxf(a.x, fmaxf(b.x, c.x));
  bx.hi[1] = fmaxf(a.y, fmaxf(b.y, c.y));
  bx.hi[2] = fmaxf(a.z, fmaxf(b.z, c.z));
  float m = 0.0f;
  for (int q = 0; q < 3; ++q) m = fmaxf(m, fmaxf(fabsf(bx.lo[q]), fmaxf(fabsf(bx.hi[q]), bx.hi[q] - bx.lo[q])));
  const float pad = m * 0x1p-20f + 1e-30f;
  for (int q = 0; q < 3; ++q) {
    bx.lo[q] -= pad;
    bx.hi[q] += pad;
  }
  return bx;
}

__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t *counter) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return 0;
  const uint32_t lane = __lane_id();
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = (uint32_t)__shfl((int)base, leader);
  return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

// Reserves `want` consecutive entries of *counter for every lane of the
// wave, in lane order; returns the lane's first entry.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t want, uint32_t *counter) {
  const uint32_t lane = __lane_id();
  uint32_t incl = want;  // inclusive prefix sum over the lanes
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t v = (uint32_t)__shfl_up((int)incl, d);
    if (lane >= (uint32_t)d) incl += v;
  }
  const uint32_t total = (uint32_t)__shfl((int)incl, 63);
  uint32_t base = 0;
  if (total != 0 && lane == 0) base = atomicAdd(counter, total);
  base = (uint32_t)__shfl((int)base, 0);
  return base + incl - want;
}

// Leaf boxes in sorted order.
__global__ void k_leaf_boxes(const float4 *__restrict__ tri, uint32_t stride, uint32_t n,
                             const uint32_t *__restrict__ order, float4 *__restrict__ leafbox) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const Box b = leaf_box(tri, stride, order[k]);
  leafbox[2 * (size_t)k] = make_float4(b.lo[0], b.lo[1], b.lo[2], 0.0f);
  leafbox[2 * (size_t)k + 1] = make_float4(b.hi[0], b.hi[1], b.hi[2], 0.0f);
}

// One breadth-first level of the radix tree: the interior children of the
// level's nodes, appended to the next level.
__global__ void k_tree_level(const int32_t *__restrict__ left, const int32_t *__restrict__ right,
                             const int32_t *__restrict__ level, uint32_t count, int32_t *__restrict__ next,
                             uint32_t *next_count) {
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < count; i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;  // uniform per block: wave_append needs every lane
    const int32_t p = i < count ? level[i] : 0;
    const int32_t l = i < count ? left[p] : -1, r = i < count ? right[p] : -1;
    const uint32_t ql = wave_append(l >= 0, next_count);
    if (l >= 0) next[ql] = l;
    const uint32_t qr = wave_append(r >= 0, next_count);
    if (r >= 0) next[qr] = r;
  }
}

// Boxes of one level's nodes from their children's (the deeper levels are
// done by earlier launches).
__global__ void k_refit_level(const int32_t *__restrict__ left, const int32_t *__restrict__ right,
                              const int32_t *__restrict__ level, uint32_t count, const float4 *__restrict__ leafbox,
                              float4 *__restrict__ nodebox) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int32_t p = level[i], l = left[p], r = right[p];
  const float4 *bl = l < 0 ? leafbox + 2 * (size_t)~l : nodebox + 2 * (size_t)l;
  const float4 *br = r < 0 ? leafbox + 2 * (size_t)~r : nodebox + 2 * (size_t)r;
  const float4 al = bl[0], ah = bl[1], cl = br[0], ch = br[1];
  nodebox[2 * (size_t)p] = make_float4(fminf(al.x, cl.x), fminf(al.y, cl.y), fminf(al.z, cl.z), 0.0f);
  nodebox[2 * (size_t)p + 1] = make_float4(fmaxf(ah.x, ch.x), fmaxf(ah.y, ch.y), fmaxf(ah.z, ch.z), 0.0f);
}

// Binary -> wide collapse by surface-area cost, the dynamic program of
// Ylitie, Karras and Laine 2017 ("Efficient incoherent ray traversal on GPUs
// through compressed wide BVHs") that the host builder runs, on the device.
// cost[p] holds C[p][i], i = 1..4: the least cost of binary subtree p when
// it may fill up to i child slots of its parent -- one slot holds it as a
// wide node (area * kCollapseNodeCost + the best split of `width` slots
// between its children) or, at most merge_limit triangles, as one leaf
// (area * count); more slots open it and split them between its children.
// A binary leaf (one triangle) costs its area in every slot count.  meta[p]
// packs the decisions the emission replays: bit 0 as-leaf, bits 1..3 "C[p][i]
// opens p" for i = 2..4, bits 4+2(i-2) the left share of the best split of
// i slots.  Device-tree interior nodes always hold contiguous leaf ranges
// (first..last), so every subtree within merge_limit may become a leaf.
// Children c < 0 are one-triangle leaves at sorted position ~c (radix tree,
// clustering); nodes whose left child is kLeafNode are leaves of the range
// first..last (binned SAH).
struct CollapseArgs {
  const int32_t *left, *right;
  const uint32_t *first, *last;
  const float *leafbox, *nodebox;
  float4 *cost;
  uint32_t *meta;
  int merge_limit;
  int width;  // child slots filled per wide node (2..4)
};

__device__ __forceinline__ float box_area(const float *p) {
  const float dx = p[4] - p[0], dy = p[5] - p[1], dz = p[6] - p[2];
  const float a = dx * dy + dy * dz + dz * dx;
  return a > 0.0f && a < FLT_MAX ? a : 0.0f;  // NaN and degenerate boxes cost nothing
}

constexpr int32_t kLeafNode = INT32_MIN;

__device__ __forceinline__ bool leaf_node(const CollapseArgs &A, int32_t c) {
  return c < 0 || A.left[c] == kLeafNode;
}

__device__ __forceinline__ float4 child_cost(const CollapseArgs &A, int32_t c) {
  if (!leaf_node(A, c)) return A.cost[c];
  const float a = c < 0 ? box_area(A.leafbox + 8 * (size_t)~c)
                        : box_area(A.nodebox + 8 * (size_t)c) * (float)(A.last[c] - A.first[c] + 1u);
  return make_float4(a, a, a, a);
}

__device__ __forceinline__ float at(const float4 &v, int i) {  // C[.][i], i = 1..4
  return i == 1 ? v.x : (i == 2 ? v.y : (i == 3 ? v.z : v.w));
}

// Costs of the binary interior nodes list[i] (or base + i when list is
// null), whose interior children have theirs already: one launch per
// radix-tree level from the deepest, or per clustering round in order.
__global__ void k_collapse_cost(CollapseArgs A, const int32_t *__restrict__ list, uint32_t base, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int32_t p = list ? list[i] : (int32_t)(base + i);
  if (A.left[p] == kLeafNode) return;  // leaves' costs come from their boxes
  const float4 cl = child_cost(A, A.left[p]), cr = child_cost(A, A.right[p]);
  const int W = A.width;
  float D[5] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX};
  uint32_t meta = 0;
  for (int j = 2; j <= W; ++j) {
    int best = 1;
    for (int k = 1; k < j; ++k) {
      const float c = at(cl, k) + at(cr, j - k);
      if (c < D[j]) {
        D[j] = c;
        best = k;
      }
    }
    meta |= (uint32_t)best << (4 + 2 * (j - 2));
  }
  const float area = box_area(A.nodebox + 8 * (size_t)p);
  const uint32_t cnt = A.last[p] - A.first[p] + 1u;
  const float node_cost = area * (float)kCollapseNodeCost + D[W];
  const float leaf_cost = cnt <= (uint32_t)A.merge_limit ? area * (float)cnt : FLT_MAX;
  const bool leaf = leaf_cost <= node_cost;
  float c[5];
  c[1] = leaf ? leaf_cost : node_cost;
  meta |= leaf ? 1u : 0u;
  for (int k = 2; k <= 4; ++k) {
    const bool open = k <= W && D[k] < c[k - 1];
    c[k] = open ? D[k] : c[k - 1];
    meta |= (open ? 1u : 0u) << (k - 1);
  }
  A.cost[p] = make_float4(c[1], c[2], c[3], c[4]);
  A.meta[p] = meta;
}

__device__ __forceinline__ uint32_t count_of(const CollapseArgs &A, int32_t c) {
  return c < 0 ? 1u : A.last[c] - A.first[c] + 1u;
}

// Whether child c of a wide node is a wide node itself (else a leaf).
__device__ __forceinline__ bool stays_interior(const CollapseArgs &A, int32_t c) {
  return !leaf_node(A, c) && !(A.meta[c] & 1u);
}

// The binary nodes that fill the i slots given to subtree n, depth first
// (left share first), appended to kids.
__device__ __forceinline__ void take(const CollapseArgs &A, int32_t n, int i, int32_t (&kids)[4], int &nk) {
  int32_t sn[4];
  int si[4], sp = 0;
  sn[0] = n;
  si[0] = i;
  sp = 1;
  while (sp > 0) {
    --sp;
    const int32_t m = sn[sp];
    int k = si[sp];
    const bool leaf = leaf_node(A, m);
    const uint32_t mm = leaf ? 0u : A.meta[m];
    while (k > 1 && !leaf && !((mm >> (k - 1)) & 1u)) --k;
    if (k == 1 || leaf) {
      if (nk < 4) kids[nk++] = m;
      continue;
    }
    const int s = (int)((mm >> (4 + 2 * (k - 2))) & 3u);
    if (sp + 2 > 4) continue;  // unreachable: at most 3 entries are pending
    sn[sp] = A.right[m];
    si[sp++] = k - s;
    sn[sp] = A.left[m];
    si[sp++] = s;
  }
}

// One level of the collapse: frontier[i] (a binary interior node) becomes
// wide node base + i; its children that stay interior are appended to next
// (wide nodes next_base + position).
__global__ void k_collapse(CollapseArgs A, const int32_t *__restrict__ frontier, uint32_t count, uint32_t base,
                           WideNode *__restrict__ nodes, int32_t *__restrict__ next, uint32_t *next_count,
                           uint32_t next_base) {
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < count; i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;  // the loop is uniform per block: wave_reserve needs every lane
    const bool active = i < count;
    int32_t kids[4] = {0, 0, 0, 0};
    int nk = 0;
    if (active) {
      const int32_t b = frontier[i];
      const int s = (int)((A.meta[b] >> (4 + 2 * (A.width - 2))) & 3u);
      take(A, A.left[b], s, kids, nk);
      take(A, A.right[b], A.width - s, kids, nk);
    }
    // the children that stay interior take consecutive entries of the next
    // level (siblings side by side, in the frontier's order)
    uint32_t grow = 0;
    for (int k = 0; k < nk; ++k) grow += stays_interior(A, kids[k]) ? 1u : 0u;
    uint32_t q = wave_reserve(grow, next_count);
    if (!active) continue;
    WideNode nd;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k >= nk) {
        set_empty_slot(nd, k);
        nd.pad[k] = 0;
        continue;
      }
      const int32_t c = kids[k];
      const float *p = c < 0 ? A.leafbox + 8 * (size_t)~c : A.nodebox + 8 * (size_t)c;
      nd.lox[k] = p[0];
      nd.loy[k] = p[1];
      nd.loz[k] = p[2];
      nd.hix[k] = p[4];
      nd.hiy[k] = p[5];
      nd.hiz[k] = p[6];
      if (stays_interior(A, c)) {
        next[q] = c;
        nd.child[k] = next_base + q++;
      } else nd.child[k] = kLeafBit | (count_of(A, c) << 27) | (c < 0 ? (uint32_t)~c : A.first[c]);
      nd.pad[k] = 0;
    }
    nodes[base + i] = nd;
  }
}

// ---------------------------------------------------------------------------
// PLOC (Meister & Bittner 2018, "Parallel locally-ordered clustering for
// bounding volume hierarchy construction").  Clusters start as the leaves in
// Morton order.  Each round, every cluster finds the neighbour within
// `radius` positions whose union box has the smallest surface area; mutual
// nearest neighbours merge into a new interior node at the lower position,
// and a scan compacts the surviving clusters in order.  Pairs are ranked by
// the 64-bit key (area bits, hash of the two positions), then by position:
// a strict total order that both ends of a pair compute alike, so the
// globally smallest pair is always mutual and every round merges.  The hash
// spreads ties (equal boxes) so that about a third of them merge per round
// instead of one pair.
//
// Interior nodes are numbered in merge order (the root is n - 2).  A leaf's
// slot in the final leaf order is the number of leaves left of it in a
// depth-first walk; the collapse then sees contiguous leaf ranges, as with
// the radix tree.

constexpr int kPlocMaxRadius = 64;

__device__ __forceinline__ uint32_t pair_hash(uint32_t a, uint32_t b) {
  uint32_t x = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
  x ^= x >> 15;
  x *= 0x2C1B3C6Du;
  x ^= x >> 12;
  x *= 0x297A2D39u;
  x ^= x >> 15;
  return x;
}

// Leaf boxes in sorted order, and the first clusters (one leaf each).
__global__ void k_ploc_init(const float4 *__restrict__ tri, uint32_t stride, uint32_t n,
                            const uint32_t *__restrict__ order, float4 *__restrict__ leafbox,
                            float4 *__restrict__ cbox, int32_t *__restrict__ cl) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const Box b = leaf_box(tri, stride, order[k]);
  const float4 lo = make_float4(b.lo[0], b.lo[1], b.lo[2], 0.0f), hi = make_float4(b.hi[0], b.hi[1], b.hi[2], 0.0f);
  leafbox[2 * (size_t)k] = lo;
  leafbox[2 * (size_t)k + 1] = hi;
  cbox[2 * (size_t)k] = lo;
  cbox[2 * (size_t)k + 1] = hi;
  cl[k] = ~(int32_t)k;
}

__device__ __forceinline__ float union_area(float4 alo, float4 ahi, float4 blo, float4 bhi) {
  const float dx = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
  const float dy = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
  const float dz = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
  return fabsf(dx * dy + dy * dz + dz * dx);  // |.|: +0 and -0 rank alike
}

// Nearest neighbour of every cluster within `radius` positions.  The block's
// clusters and `radius` on either side are staged in LDS.
__global__ void __launch_bounds__(kBlock) k_ploc_nn(const float4 *__restrict__ cbox, const uint32_t *__restrict__ ctl,
                                                    int radius, int32_t *__restrict__ nn) {
  __shared__ float4 slo[kBlock + 2 * kPlocMaxRadius], shi[kBlock + 2 * kPlocMaxRadius];
  const uint32_t cnt = ctl[0];  // clusters this round (the grid covers an upper bound)
  const int start = (int)(blockIdx.x * kBlock);
  if (start >= (int)cnt) return;  // whole block past the clusters
  const int lo = max(0, start - radius), hi = min((int)cnt, start + kBlock + radius);
  for (int k = lo + (int)threadIdx.x; k < hi; k += kBlock) {
    slo[k - lo] = cbox[2 * (size_t)k];
    shi[k - lo] = cbox[2 * (size_t)k + 1];
  }
  __syncthreads();
  const int i = start + (int)threadIdx.x;
  if (i >= (int)cnt) return;
  const float4 alo = slo[i - lo], ahi = shi[i - lo];
  uint64_t best = ~0ull;
  int bj = -1;
  const int jlo = max(0, i - radius), jhi = min((int)cnt - 1, i + radius);
  for (int j = jlo; j <= jhi; ++j) {
    if (j == i) continue;
    const float a = union_area(alo, ahi, slo[j - lo], shi[j - lo]);
    const uint64_t key = ((uint64_t)__float_as_uint(a) << 32) | pair_hash((uint32_t)min(i, j), (uint32_t)max(i, j));
    if (key < best) {  // ascending j: a full tie keeps the lower position
      best = key;
      bj = j;
    }
  }
  nn[i] = bj;
}

// flags[i]: low word 1 if cluster i survives the round (it does not merge
// into a lower position), high word 1 if it starts a merge.
__global__ void k_ploc_flags(const int32_t *__restrict__ nn, const uint32_t *__restrict__ ctl, uint32_t bound,
                             uint64_t *__restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bound) return;
  if (i >= ctl[0]) {  // past this round's clusters: nothing for the scan
    flags[i] = 0;
    return;
  }
  const int32_t m = nn[i];
  const bool mutual = m >= 0 && nn[m] == (int32_t)i;
  const uint64_t lead = mutual && (int32_t)i < m, keep = !(mutual && (int32_t)i > m);
  flags[i] = keep | (lead << 32);
}

__device__ __forceinline__ uint32_t leaves_of(const uint32_t *__restrict__ ncount, int32_t c) {
  return c < 0 ? 1u : ncount[c];
}

// Merges the mutual pairs into interior nodes ctl[1] + (merge rank) and
// writes the surviving clusters in order.  ctl = {clusters, nodes so far}
// of this round; totals = the same for the next round, whose node count is
// also logged to *round_nodes (the round's nodes are the range between two
// log entries, and their children come from earlier rounds).
__global__ void k_ploc_merge(const int32_t *__restrict__ cl, const float4 *__restrict__ cbox,
                             const int32_t *__restrict__ nn, const uint64_t *__restrict__ flags,
                             const uint64_t *__restrict__ pre, const uint32_t *__restrict__ ctl,
                             int32_t *__restrict__ left, int32_t *__restrict__ right, int32_t *__restrict__ parent,
                             int32_t *__restrict__ leaf_parent, uint32_t *__restrict__ ncount,
                             float4 *__restrict__ nodebox, int32_t *__restrict__ cl_out,
                             float4 *__restrict__ cbox_out, uint32_t *__restrict__ totals,
                             uint32_t *__restrict__ round_nodes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t cnt = ctl[0], node_base = ctl[1];  // clusters and nodes so far, this round
  if (i >= cnt) return;
  const uint64_t f = flags[i], p = pre[i];
  const uint32_t keep = (uint32_t)f, lead = (uint32_t)(f >> 32);
  const uint32_t slot = (uint32_t)p, rank = (uint32_t)(p >> 32);
  if (i == cnt - 1) {
    totals[0] = slot + keep;
    totals[1] = node_base + rank + lead;
    *round_nodes = node_base + rank + lead;
  }
  if (!keep) return;
  float4 lo = cbox[2 * (size_t)i], hi = cbox[2 * (size_t)i + 1];
  int32_t c = cl[i];
  if (lead) {
    const int32_t m = nn[i];
    const int32_t d = cl[m];
    const float4 blo = cbox[2 * (size_t)m], bhi = cbox[2 * (size_t)m + 1];
    lo = make_float4(fminf(lo.x, blo.x), fminf(lo.y, blo.y), fminf(lo.z, blo.z), 0.0f);
    hi = make_float4(fmaxf(hi.x, bhi.x), fmaxf(hi.y, bhi.y), fmaxf(hi.z, bhi.z), 0.0f);
    const int32_t node = (int32_t)(node_base + rank);
    left[node] = c;
    right[node] = d;
    ncount[node] = leaves_of(ncount, c) + leaves_of(ncount, d);
    nodebox[2 * (size_t)node] = lo;
    nodebox[2 * (size_t)node + 1] = hi;
    if (c < 0) leaf_parent[~c] = node;
    else parent[c] = node;
    if (d < 0) leaf_parent[~d] = node;
    else parent[d] = node;
    c = node;
  }
  cl_out[slot] = c;
  cbox_out[2 * (size_t)slot] = lo;
  cbox_out[2 * (size_t)slot + 1] = hi;
}

// Depth-first slot of every leaf: the leaves of the left siblings on its
// way to the root.
__global__ void k_ploc_slot(const int32_t *__restrict__ left, const int32_t *__restrict__ right,
                            const int32_t *__restrict__ parent, const int32_t *__restrict__ leaf_parent,
                            const uint32_t *__restrict__ ncount, uint32_t n, uint32_t *__restrict__ slot) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int32_t c = ~(int32_t)k, p = leaf_parent[k];
  uint32_t s = 0;
  while (p >= 0) {
    if (right[p] == c) s += leaves_of(ncount, left[p]);
    c = p;
    p = parent[p];
  }
  slot[k] = s;
}

// Leaf ranges of the interior nodes (each node's range starts at its
// leftmost leaf), the leaves in slot order, and leaf children renamed to
// their slots.
__global__ void k_ploc_ranges(const int32_t *__restrict__ left, const int32_t *__restrict__ parent,
                              const int32_t *__restrict__ leaf_parent, const uint32_t *__restrict__ ncount,
                              const uint32_t *__restrict__ slot, const uint32_t *__restrict__ order,
                              const float4 *__restrict__ leafbox, uint32_t n, uint32_t *__restrict__ first,
                              uint32_t *__restrict__ last, uint32_t *__restrict__ order_out,
                              float4 *__restrict__ leafbox_out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = slot[k];
  order_out[s] = order[k];
  leafbox_out[2 * (size_t)s] = leafbox[2 * (size_t)k];
  leafbox_out[2 * (size_t)s + 1] = leafbox[2 * (size_t)k + 1];
  int32_t c = ~(int32_t)k, p = leaf_parent[k];
  while (p >= 0 && left[p] == c) {
    first[p] = s;
    last[p] = s + ncount[p] - 1u;
    c = p;
    p = parent[p];
  }
}

__global__ void k_ploc_rename(int32_t *__restrict__ left, int32_t *__restrict__ right, const uint32_t *__restrict__ slot,
                              uint32_t ni) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ni) return;
  const int32_t l = left[p], r = right[p];
  if (l < 0) left[p] = ~(int32_t)slot[~l];
  if (r < 0) right[p] = ~(int32_t)slot[~r];
}

// Intersection triangles in leaf (sorted) order: a with the original id in
// w, e1 = b - a, e2 = c - a (the host builder's rounding).
__global__ void k_itris(const float4 *__restrict__ tri, uint32_t stride, uint32_t n, const uint32_t *__restrict__ order,
                        float4 *__restrict__ itris) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t t = order[k];
  const float4 a = tri[(size_t)t * stride], b = tri[(size_t)t * stride + 1], c = tri[(size_t)t * stride + 2];
  itris[3 * (size_t)k] = make_float4(a.x, a.y, a.z, __uint_as_float(t));
  itris[3 * (size_t)k + 1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
  itris[3 * (size_t)k + 2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
}

__global__ void k_zero(uint32_t *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

inline uint32_t blocks(size_t n) { return (uint32_t)std::max<size_t>(1, (n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------------------
// Binned SAH, top down, one tree level per step: the host builder's
// algorithm (Builder in bvh_build.cpp) with every split decision taken on
// the device.  A node of at most max_leaf triangles is a leaf.  Below depth
// kSahMaxDepth, its triangles' centroids are binned into kSahBins bins per
// axis over the node's centroid box; the split of least
// area(left) * count(left) + area(right) * count(right) wins (the first in
// axis, then bin order), unless 1 + cost / area(node) >= count with at most
// 15 triangles, which makes a leaf.  No valid split (equal centroids) or
// depth kSahMaxDepth and beyond: the object median by (centroid on the
// longest centroid axis, triangle index).  Bins hold boxes as order-
// preserving keys (min / max commute), so each decision -- and the tree's
// shape -- is the host builder's; the triangles within a node may be in
// another order (a stable partition here, std::partition there).
//
// Per level: nodes bin their triangles in chunks of kSahChunk (one
// workgroup each, bins in LDS; nodes of several chunks merge them in global
// memory), decide, get their children numbered by a scan, and a scan over
// all triangles moves each node's triangles to its children's ranges
// (median nodes: a segmented sort).  Children's boxes come from the bins.
constexpr uint32_t kSahChunk = 4096;   // triangles binned per workgroup
constexpr int kSahFields = 13;         // per bin: box lo, hi, centroid lo, hi (keys), count
constexpr int kSahBinWords = 3 * kSahBins * kSahFields;
enum : int32_t { kSahPending = 0, kSahLeaf = 1, kSahSplit = 2, kSahMedian = 3 };

// Order-preserving keys of floats (unsigned compare = float compare).
__device__ __forceinline__ uint32_t okey(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ofloat(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// A node's decision, per position in its level's list.
struct SahDec {
  int32_t kind, axis, bin;
  uint32_t nleft;   // triangles of the left child
  float lo, scale;  // the split axis' bin mapping: bin = (centroid - lo) * scale
  uint32_t child;   // position of the left child in the next level's list
  int32_t node;
  float4 box[8];    // left box lo, hi; right box; left centroid box; right centroid box
};

struct SahArgs {
  float4 *box[2];    // triangle boxes (lo, hi) in the current order, double-buffered
  uint32_t *perm[2]; // triangle index
  int32_t *pos[2];   // position of the triangle's node in this level's list; -1 in a leaf
  int32_t *left, *right;
  uint32_t *nb, *ne;  // node ranges [nb, ne)
  float4 *bb, *cb;    // node box and centroid box (lo, hi), unpadded
  float *nodebox;     // padded node box, 8 floats per node
  SahDec *dec;
  uint32_t *leaf_perm;  // triangle index at each position, written when its node becomes a leaf
  int max_leaf;
};

__device__ __forceinline__ float sah_area(const float *lo, const float *hi) {  // Box::area
  if (lo[0] > hi[0]) return 0.0f;
  const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// The padded copy of a node box (the host builder's pad_box).
__device__ __forceinline__ void write_padded(float *out, const float4 lo, const float4 hi) {
  const float l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
  float m = 0.0f;
  for (int k = 0; k < 3; ++k) m = fmaxf(m, fmaxf(fabsf(l[k]), fabsf(h[k])));
  for (int k = 0; k < 3; ++k) m = fmaxf(m, h[k] - l[k]);
  const float p = m * 0x1p-20f + 1e-30f;
  out[0] = l[0] - p;
  out[1] = l[1] - p;
  out[2] = l[2] - p;
  out[3] = 0.0f;
  out[4] = h[0] + p;
  out[5] = h[1] + p;
  out[6] = h[2] + p;
  out[7] = 0.0f;
}

__device__ __forceinline__ float comp(const float4 &v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ __forceinline__ int sah_bin(float c, float lo, float scale) {
  const int b = (int)((c - lo) * scale);
  return b < 0 ? 0 : (b > kSahBins - 1 ? kSahBins - 1 : b);
}

// Triangle boxes (min / max of the three vertices) in input order.
__global__ void k_sah_init(const float4 *__restrict__ tri, uint32_t stride, uint32_t n, SahArgs S) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 a = tri[(size_t)i * stride], b = tri[(size_t)i * stride + 1], c = tri[(size_t)i * stride + 2];
  S.box[0][2 * (size_t)i] = make_float4(fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y),
                                        fminf(fminf(a.z, b.z), c.z), 0.0f);
  S.box[0][2 * (size_t)i + 1] = make_float4(fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y),
                                            fmaxf(fmaxf(a.z, b.z), c.z), 0.0f);
  S.perm[0][i] = i;
  S.pos[0][i] = 0;
}

__device__ __forceinline__ void bin_identity(uint32_t *sb, int w) {
  const int f = w % kSahFields;
  sb[w] = (f < 3 || (f >= 6 && f < 9)) ? 0xFFFFFFFFu : 0u;
}

// Box and centroid box of the triangles i0, i0 + step, ... < i1, merged
// into r (12 order keys: box lo, hi, centroid lo, hi; in LDS or global
// memory) with one atomic per key per thread.
__device__ __forceinline__ void reduce_boxes(const float4 *__restrict__ box, uint32_t i0, uint32_t i1, uint32_t step,
                                             uint32_t *r) {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (uint32_t i = i0; i < i1; i += step) {
    const float4 a = box[2 * (size_t)i], b = box[2 * (size_t)i + 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float l = comp(a, k), h = comp(b, k), c = 0.5f * (l + h);
      lo[k] = fminf(lo[k], l);
      hi[k] = fmaxf(hi[k], h);
      clo[k] = fminf(clo[k], c);
      chi[k] = fmaxf(chi[k], c);
    }
  }
  if (i0 >= i1) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    atomicMin(&r[k], okey(lo[k]));
    atomicMax(&r[3 + k], okey(hi[k]));
    atomicMin(&r[6 + k], okey(clo[k]));
    atomicMax(&r[9 + k], okey(chi[k]));
  }
}

__device__ __forceinline__ void write_node_boxes(const SahArgs &S, int32_t p, const uint32_t *r) {
  const float4 blo = make_float4(ofloat(r[0]), ofloat(r[1]), ofloat(r[2]), 0.0f);
  const float4 bhi = make_float4(ofloat(r[3]), ofloat(r[4]), ofloat(r[5]), 0.0f);
  S.bb[2 * (size_t)p] = blo;
  S.bb[2 * (size_t)p + 1] = bhi;
  S.cb[2 * (size_t)p] = make_float4(ofloat(r[6]), ofloat(r[7]), ofloat(r[8]), 0.0f);
  S.cb[2 * (size_t)p + 1] = make_float4(ofloat(r[9]), ofloat(r[10]), ofloat(r[11]), 0.0f);
  write_padded(S.nodebox + 8 * (size_t)p, blo, bhi);
}

__device__ __forceinline__ void reduce_identity(uint32_t *r, uint32_t w) { r[w] = (w % 6) < 3 ? 0xFFFFFFFFu : 0u; }

// Boxes of nodes ids[b] from their triangles (the children of median
// splits): one workgroup per node.
__global__ void __launch_bounds__(kBlock) k_sah_reduce(SahArgs S, int buf, const int32_t *__restrict__ ids) {
  __shared__ uint32_t r[12];
  const int32_t p = ids[blockIdx.x];
  if (threadIdx.x < 12) reduce_identity(r, threadIdx.x);
  __syncthreads();
  reduce_boxes(S.box[buf], S.nb[p] + threadIdx.x, S.ne[p], blockDim.x, r);
  __syncthreads();
  if (threadIdx.x == 0) write_node_boxes(S, p, r);
}

// The root's boxes over all n triangles: per workgroup in LDS, then into
// the 12 global keys g (identity first: k_sah_root_init).
__global__ void k_sah_root_init(uint32_t *g) {
  if (threadIdx.x < 12) reduce_identity(g, threadIdx.x);
}
__global__ void __launch_bounds__(kBlock) k_sah_root(const float4 *__restrict__ box, uint32_t n, uint32_t *g) {
  __shared__ uint32_t r[12];
  if (threadIdx.x < 12) reduce_identity(r, threadIdx.x);
  __syncthreads();
  reduce_boxes(box, blockIdx.x * blockDim.x + threadIdx.x, n, gridDim.x * blockDim.x, r);
  __syncthreads();
  if (threadIdx.x < 12) {
    const uint32_t v = r[threadIdx.x];
    if ((threadIdx.x % 6) < 3) atomicMin(&g[threadIdx.x], v);
    else atomicMax(&g[threadIdx.x], v);
  }
}
__global__ void k_sah_root_write(SahArgs S, const uint32_t *g) {
  if (threadIdx.x == 0) write_node_boxes(S, 0, g);
}

__device__ __forceinline__ int longest_axis(const float4 &clo, const float4 &chi) {
  int ax = 0;
  for (int k = 1; k < 3; ++k)
    if (comp(chi, k) - comp(clo, k) > comp(chi, ax) - comp(clo, ax)) ax = k;
  return ax;
}

// First step of a level: leaves by count, median nodes by depth; the rest
// are binned, by one wave (one chunk) or in ceil(count / kSahChunk) chunks,
// which they ask for: want = (1 << 32) | chunks when several.
__global__ void k_sah_plan(SahArgs S, const int32_t *__restrict__ list, uint32_t count, int depth,
                           uint64_t *__restrict__ want) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int32_t p = list[j];
  const uint32_t cnt = S.ne[p] - S.nb[p];
  SahDec &d = S.dec[j];
  d.node = p;
  d.axis = -1;
  d.bin = -1;
  uint64_t w = 0;
  if (cnt <= (uint32_t)S.max_leaf) {
    d.kind = kSahLeaf;
    S.left[p] = S.right[p] = kLeafNode;
  } else if (depth >= kSahMaxDepth) {
    d.kind = kSahMedian;
    d.axis = longest_axis(S.cb[2 * (size_t)p], S.cb[2 * (size_t)p + 1]);
  } else {
    d.kind = kSahPending;
    const uint64_t nch = (cnt + kSahChunk - 1) / kSahChunk;
    if (nch > 1) w = (1ull << 32) | nch;  // one chunk: k_sah_wave bins the node
  }
  want[j] = w;
}

// Bins of triangles i0..i1-1 into sb (LDS), by `threads` threads (this one
// is tid): each takes a run of consecutive triangles (neighbours in memory
// are mostly neighbours in space) and accumulates in registers while the
// bin stays the same -- one set of LDS atomics per run of a bin instead of
// one per triangle, whose same-address conflicts would serialise the wave.
__device__ __forceinline__ void bin_range(uint32_t *sb, const float4 *__restrict__ box, uint32_t i0, uint32_t i1,
                                          uint32_t tid, uint32_t threads, const float4 &clo, const bool (&valid)[3],
                                          const float (&scale)[3]) {
  const uint32_t run = (i1 - i0 + threads - 1) / threads;
  const uint32_t r0 = min(i1, i0 + tid * run), r1 = min(i1, r0 + run);
  int cur[3] = {-1, -1, -1};
  float acc[3][12];
  uint32_t num[3] = {0, 0, 0};
  auto flush = [&](int k) {
    if (!num[k]) return;
    uint32_t *q = sb + (k * kSahBins + cur[k]) * kSahFields;
    for (int f = 0; f < 3; ++f) {
      atomicMin(&q[f], okey(acc[k][f]));
      atomicMax(&q[3 + f], okey(acc[k][3 + f]));
      atomicMin(&q[6 + f], okey(acc[k][6 + f]));
      atomicMax(&q[9 + f], okey(acc[k][9 + f]));
    }
    atomicAdd(&q[12], num[k]);
    num[k] = 0;
  };
  for (uint32_t i = r0; i < r1; ++i) {
    const float4 l = box[2 * (size_t)i], h = box[2 * (size_t)i + 1];
    const float v[9] = {l.x, l.y, l.z, h.x, h.y, h.z, 0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (!valid[k]) continue;
      const int bin = sah_bin(v[6 + k], comp(clo, k), scale[k]);
      if (bin != cur[k]) {
        flush(k);
        cur[k] = bin;
      }
      if (num[k] == 0) {
        for (int f = 0; f < 9; ++f) acc[k][f] = v[f];
        for (int f = 0; f < 3; ++f) acc[k][9 + f] = v[6 + f];
      } else {
        for (int f = 0; f < 3; ++f) {
          acc[k][f] = fminf(acc[k][f], v[f]);
          acc[k][3 + f] = fmaxf(acc[k][3 + f], v[3 + f]);
          acc[k][6 + f] = fminf(acc[k][6 + f], v[6 + f]);
          acc[k][9 + f] = fmaxf(acc[k][9 + f], v[6 + f]);
        }
      }
      ++num[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) flush(k);
}

// Bin mapping of a node's centroid box: axes of positive extent, and
// kSahBins / extent.
__device__ __forceinline__ void bin_mapping(const float4 &clo, const float4 &chi, bool (&valid)[3], float (&scale)[3]) {
  for (int k = 0; k < 3; ++k) {
    const float ext = comp(chi, k) - comp(clo, k);
    valid[k] = ext > 0.0f;
    scale[k] = valid[k] ? (float)kSahBins / ext : 0.0f;
  }
}

// The split of node list[j] (node p) from its bins (sb, in LDS), by one
// wave (a 64-thread block).  Lane l holds bin l & 31 of an axis (axes 0 and
// 1 in the first round, 2 in the second); prefix and suffix unions over the
// bins (the host builder's left and right sweeps -- min, max and counts are
// exact in any order) give every candidate's cost
// area(bins 0..i) * count + area(bins i+1..31) * count, and the least cost
// wins, ties to the lower axis, then bin (the host's first strict minimum).
// Then leaf / split / median as the host builder decides.
__device__ void wave_decide(const SahArgs &S, const uint32_t *sb, uint32_t j, int32_t p) {
  const uint32_t lane = threadIdx.x;
  const float4 clo = S.cb[2 * (size_t)p], chi = S.cb[2 * (size_t)p + 1];
  const int i = (int)(lane & 31u);
  uint64_t best = ~0ull;
  float win[12];  // the winning lane's left (0..5) and right (6..11) boxes
  float wcen[12]; // and centroid boxes
  uint32_t wcount = 0;
  for (int round = 0; round < 2; ++round) {
    const int k = round == 0 ? (int)(lane >> 5) : 2;
    const bool live = round == 0 || lane < 32;
    const float ext = comp(chi, k) - comp(clo, k);
    const uint32_t *q = sb + (k * kSahBins + i) * kSahFields;
    const uint32_t c = live ? q[12] : 0u;
    float v[12];
    for (int f = 0; f < 12; ++f) {
      const bool mn = f < 3 || (f >= 6 && f < 9);
      v[f] = c ? ofloat(q[f]) : (mn ? FLT_MAX : -FLT_MAX);
    }
    float pv[12], sv[12];
    uint32_t pc = c, sc = c;
    for (int f = 0; f < 12; ++f) pv[f] = sv[f] = v[f];
    for (int d = 1; d < 32; d <<= 1) {
      const uint32_t up = (uint32_t)__shfl_up((int)pc, d, 32), dn = (uint32_t)__shfl_down((int)sc, d, 32);
      if (i >= d) pc += up;
      if (i + d < 32) sc += dn;
      for (int f = 0; f < 12; ++f) {
        const bool mn = f < 3 || (f >= 6 && f < 9);
        const float pu = __shfl_up(pv[f], d, 32), sd = __shfl_down(sv[f], d, 32);
        if (i >= d) pv[f] = mn ? fminf(pv[f], pu) : fmaxf(pv[f], pu);
        if (i + d < 32) sv[f] = mn ? fminf(sv[f], sd) : fmaxf(sv[f], sd);
      }
    }
    // right side of the split after bin i: the suffix from bin i + 1
    float rv[12];
    for (int f = 0; f < 12; ++f) rv[f] = __shfl_down(sv[f], 1, 32);
    const uint32_t rc = (uint32_t)__shfl_down((int)sc, 1, 32);
    uint64_t key = ~0ull;
    if (live && ext > 0.0f && i < kSahBins - 1 && pc != 0 && rc != 0) {
      const float cost = sah_area(pv, pv + 3) * (float)pc + sah_area(rv, rv + 3) * (float)rc;
      if (cost < FLT_MAX) key = ((uint64_t)__float_as_uint(cost) << 32) | (uint32_t)(k * kSahBins + i);
    }
    // the wave's least key
    uint64_t m = key;
    for (int d = 32; d >= 1; d >>= 1) {
      const uint64_t o = __shfl_xor(m, d);
      m = o < m ? o : m;
    }
    if (m < best) {
      best = m;
      if (key == m) {  // this lane wins (keys are unique per lane)
        for (int f = 0; f < 6; ++f) {
          win[f] = pv[f];
          win[6 + f] = rv[f];
          wcen[f] = pv[6 + f];
          wcen[6 + f] = rv[6 + f];
        }
        wcount = pc;
      }
    }
  }
  SahDec &d = S.dec[j];
  const uint32_t cnt = S.ne[p] - S.nb[p];
  if (best == ~0ull) {
    if (lane == 0) {
      d.kind = kSahMedian;
      d.axis = longest_axis(clo, chi);
    }
    return;
  }
  const int axis = (int)((uint32_t)best / kSahBins), split = (int)((uint32_t)best % kSahBins);
  const int owner = (axis == 2 ? 0 : 32 * axis) + split;  // the lane that holds the winner
  if ((int)lane != owner) return;
  const float cost = __uint_as_float((uint32_t)(best >> 32));
  const float4 plo = S.bb[2 * (size_t)p], phi = S.bb[2 * (size_t)p + 1];
  const float pl[3] = {plo.x, plo.y, plo.z}, ph[3] = {phi.x, phi.y, phi.z};
  const float sah = 1.0f + cost / fmaxf(sah_area(pl, ph), 1e-30f);
  if (sah >= (float)cnt && cnt <= (uint32_t)kMaxLeafTris) {
    d.kind = kSahLeaf;
    S.left[p] = S.right[p] = kLeafNode;
    return;
  }
  d.kind = kSahSplit;
  d.axis = axis;
  d.bin = split;
  d.nleft = wcount;
  d.lo = comp(clo, axis);
  d.scale = (float)kSahBins / (comp(chi, axis) - comp(clo, axis));
  d.box[0] = make_float4(win[0], win[1], win[2], 0.0f);
  d.box[1] = make_float4(win[3], win[4], win[5], 0.0f);
  d.box[2] = make_float4(win[6], win[7], win[8], 0.0f);
  d.box[3] = make_float4(win[9], win[10], win[11], 0.0f);
  d.box[4] = make_float4(wcen[0], wcen[1], wcen[2], 0.0f);
  d.box[5] = make_float4(wcen[3], wcen[4], wcen[5], 0.0f);
  d.box[6] = make_float4(wcen[6], wcen[7], wcen[8], 0.0f);
  d.box[7] = make_float4(wcen[9], wcen[10], wcen[11], 0.0f);
}

// Nodes of one chunk (at most kSahChunk triangles), one wave each: bins in
// LDS, then the decision.  Other positions of the list return at once.
__global__ void __launch_bounds__(64) k_sah_wave(SahArgs S, int buf, uint32_t count) {
  __shared__ uint32_t sb[kSahBinWords];
  const uint32_t j = blockIdx.x;
  if (j >= count || S.dec[j].kind != kSahPending) return;
  const int32_t p = S.dec[j].node;
  const uint32_t b = S.nb[p], e = S.ne[p];
  if (e - b > kSahChunk) return;  // several chunks: k_sah_bin_chunks and k_sah_big
  for (int w = threadIdx.x; w < kSahBinWords; w += blockDim.x) bin_identity(sb, w);
  __syncthreads();
  const float4 clo = S.cb[2 * (size_t)p], chi = S.cb[2 * (size_t)p + 1];
  bool valid[3];
  float scale[3];
  bin_mapping(clo, chi, valid, scale);
  bin_range(sb, S.box[buf], b, e, threadIdx.x, blockDim.x, clo, valid, scale);
  __syncthreads();
  wave_decide(S, sb, j, p);
}

// Chunks of the nodes of several chunks (workgroup = chunk): bins in LDS,
// merged into the node's global bins (slot = its rank among them).
__global__ void __launch_bounds__(kBlock) k_sah_bin_chunks(SahArgs S, int buf, uint32_t count,
                                                         const uint64_t *__restrict__ pre,
                                                         const uint32_t *__restrict__ tot,
                                                         uint32_t *__restrict__ gbins, uint32_t *__restrict__ big_pos) {
  __shared__ uint32_t sb[kSahBinWords];
  const uint32_t chunk = blockIdx.x;
  if (chunk >= tot[0]) return;  // the grid covers an upper bound
  // the last list position whose first chunk is at or before this one
  uint32_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((uint32_t)pre[mid] <= chunk) lo = mid;
    else hi = mid;
  }
  const uint32_t j = lo;
  const int32_t p = S.dec[j].node;
  const uint32_t b = S.nb[p], e = S.ne[p];
  const uint32_t c = chunk - (uint32_t)pre[j];
  for (int w = threadIdx.x; w < kSahBinWords; w += blockDim.x) bin_identity(sb, w);
  __syncthreads();
  const float4 clo = S.cb[2 * (size_t)p], chi = S.cb[2 * (size_t)p + 1];
  bool valid[3];
  float scale[3];
  bin_mapping(clo, chi, valid, scale);
  const uint32_t i0 = b + c * kSahChunk, i1 = min(e, i0 + kSahChunk);
  bin_range(sb, S.box[buf], i0, i1, threadIdx.x, blockDim.x, clo, valid, scale);
  __syncthreads();
  const uint32_t slot = (uint32_t)(pre[j] >> 32);
  uint32_t *g = gbins + (size_t)slot * kSahBinWords;
  for (int w = threadIdx.x; w < kSahBinWords; w += blockDim.x) {
    const uint32_t v = sb[w];
    const int f = w % kSahFields;
    if (f == 12) {
      if (v) atomicAdd(&g[w], v);
    } else if (f < 3 || (f >= 6 && f < 9)) {
      if (v != 0xFFFFFFFFu) atomicMin(&g[w], v);
    } else if (v != 0u) {
      atomicMax(&g[w], v);
    }
  }
  if (c == 0 && threadIdx.x == 0) big_pos[slot] = j;
}

// tot = {chunks, nodes of several chunks} of the level, from the scan of want.
__global__ void k_sah_totals(const uint64_t *__restrict__ pre, const uint64_t *__restrict__ want, uint32_t count,
                             uint32_t *__restrict__ tot) {
  if (threadIdx.x == 0) {
    const uint64_t t = pre[count - 1] + want[count - 1];
    tot[0] = (uint32_t)t;
    tot[1] = (uint32_t)(t >> 32);
  }
}

// Global bins of the level's several-chunk nodes (tot[1] of them) to identity.
__global__ void k_sah_bins_identity(uint32_t *__restrict__ gbins, const uint32_t *__restrict__ tot) {
  const size_t words = (size_t)tot[1] * kSahBinWords;
  for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < words; w += (size_t)gridDim.x * blockDim.x)
    gbins[w] = ((w % kSahFields) < 3 || ((w % kSahFields) >= 6 && (w % kSahFields) < 9)) ? 0xFFFFFFFFu : 0u;
}

// Decisions of the several-chunk nodes from their merged bins (one wave each).
__global__ void __launch_bounds__(64) k_sah_big(SahArgs S, const uint32_t *__restrict__ tot,
                                              const uint32_t *__restrict__ gbins,
                                              const uint32_t *__restrict__ big_pos) {
  __shared__ uint32_t sb[kSahBinWords];
  if (blockIdx.x >= tot[1]) return;  // the grid covers an upper bound
  const uint32_t j = big_pos[blockIdx.x];
  const uint32_t *g = gbins + (size_t)blockIdx.x * kSahBinWords;
  for (int w = threadIdx.x; w < kSahBinWords; w += blockDim.x) sb[w] = g[w];
  __syncthreads();
  wave_decide(S, sb, j, S.dec[j].node);
}

// want = (median << 32) | (split or median): children are numbered by its scan.
__global__ void k_sah_count(const SahArgs S, uint32_t count, uint64_t *__restrict__ want) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int32_t k = S.dec[j].kind;
  want[j] = ((uint64_t)(k == kSahMedian ? 1 : 0) << 32) | (k == kSahSplit || k == kSahMedian ? 1u : 0u);
}

// Children of the split and median nodes: nodes node_base + 2r and
// node_base + 2r + 1 (r = rank among them), positions 2r, 2r + 1 of the
// next level's list.  Median nodes are listed for the sort.
__global__ void k_sah_children(SahArgs S, uint32_t count, const uint64_t *__restrict__ pre, uint32_t node_base,
                               const uint32_t *__restrict__ flags, const uint32_t *__restrict__ scan,
                               int32_t *__restrict__ next_list, uint32_t *__restrict__ med_pos,
                               int32_t *__restrict__ med_kids, uint32_t *__restrict__ seg_begin,
                               uint32_t *__restrict__ seg_end) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  SahDec &d = S.dec[j];
  if (d.kind != kSahSplit && d.kind != kSahMedian) return;
  const uint32_t r = (uint32_t)pre[j];
  const int32_t p = d.node, id = (int32_t)(node_base + 2 * r);
  const uint32_t b = S.nb[p], e = S.ne[p];
  // a split's left count from the flags themselves (the bins' count is the
  // same), so that the partition stays inside the node's range
  const uint32_t mid = d.kind == kSahMedian ? b + (e - b) / 2 : b + (scan[e - 1] + flags[e - 1] - scan[b]);
  d.nleft = mid - b;
  d.child = 2 * r;
  S.left[p] = id;
  S.right[p] = id + 1;
  S.nb[id] = b;
  S.ne[id] = mid;
  S.nb[id + 1] = mid;
  S.ne[id + 1] = e;
  next_list[2 * r] = id;
  next_list[2 * r + 1] = id + 1;
  if (d.kind == kSahSplit) {
    for (int h = 0; h < 2; ++h) {
      const size_t c = (size_t)(id + h);
      S.bb[2 * c] = d.box[2 * h];
      S.bb[2 * c + 1] = d.box[2 * h + 1];
      S.cb[2 * c] = d.box[4 + 2 * h];
      S.cb[2 * c + 1] = d.box[4 + 2 * h + 1];
      write_padded(S.nodebox + 8 * c, d.box[2 * h], d.box[2 * h + 1]);
    }
  } else {
    const uint32_t m = (uint32_t)(pre[j] >> 32);
    med_pos[m] = j;
    med_kids[2 * m] = id;
    med_kids[2 * m + 1] = id + 1;
    seg_begin[m] = b;
    seg_end[m] = e;
  }
}

// flags[i] = 1 if triangle i goes to the left child of a split node.
__global__ void k_sah_flags(const SahArgs S, int buf, uint32_t n, uint32_t *__restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t j = S.pos[buf][i];
  uint32_t f = 0;
  if (j >= 0 && S.dec[j].kind == kSahSplit) {
    const SahDec &d = S.dec[j];
    const float l = comp(S.box[buf][2 * (size_t)i], d.axis), h = comp(S.box[buf][2 * (size_t)i + 1], d.axis);
    f = sah_bin(0.5f * (l + h), d.lo, d.scale) <= d.bin ? 1u : 0u;
  }
  flags[i] = f;
}

// Every triangle of the level into the other buffer: split nodes'
// triangles to their child's range (a stable partition by the scan of
// flags); a new leaf's triangles stay where they are for good (leaf_perm),
// and later levels only carry their "in a leaf" mark; median nodes'
// triangles get their sort keys (centroid on the axis, -0 as +0, then
// triangle index) for k_sah_median, which moves them.
__global__ void k_sah_scatter(SahArgs S, int buf, uint32_t n, const uint32_t *__restrict__ flags,
                              const uint32_t *__restrict__ scan, uint64_t *__restrict__ keys,
                              uint32_t *__restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int out = buf ^ 1;
  const int32_t j = S.pos[buf][i];
  if (j < 0) {
    S.pos[out][i] = -1;
    return;
  }
  const SahDec &d = S.dec[j];
  if (d.kind == kSahSplit) {
    const uint32_t b = S.nb[d.node], lb = scan[b], li = scan[i] - lb;
    uint32_t dst;
    int32_t pn;
    if (flags[i]) {
      dst = b + li;
      pn = (int32_t)d.child;
    } else {
      dst = b + d.nleft + (i - b) - li;
      pn = (int32_t)d.child + 1;
    }
    S.box[out][2 * (size_t)dst] = S.box[buf][2 * (size_t)i];
    S.box[out][2 * (size_t)dst + 1] = S.box[buf][2 * (size_t)i + 1];
    S.perm[out][dst] = S.perm[buf][i];
    S.pos[out][dst] = pn;
  } else if (d.kind == kSahMedian) {
    const float c = 0.5f * (comp(S.box[buf][2 * (size_t)i], d.axis) + comp(S.box[buf][2 * (size_t)i + 1], d.axis));
    keys[i] = ((uint64_t)okey(c + 0.0f) << 32) | S.perm[buf][i];
    vals[i] = i;
  } else {  // a leaf
    S.leaf_perm[i] = S.perm[buf][i];
    S.pos[out][i] = -1;
  }
}

// The triangles of the last level's nodes (all leaves) into leaf_perm.
__global__ void k_sah_final(SahArgs S, int buf, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && S.pos[buf][i] >= 0) S.leaf_perm[i] = S.perm[buf][i];
}

// Median nodes (one workgroup each): their triangles in sorted order, the
// first half to the left child.
__global__ void __launch_bounds__(kBlock) k_sah_median(SahArgs S, int buf, const uint32_t *__restrict__ med_pos,
                                                     const uint32_t *__restrict__ sorted) {
  const SahDec &d = S.dec[med_pos[blockIdx.x]];
  const uint32_t b = S.nb[d.node], e = S.ne[d.node];
  const int out = buf ^ 1;
  for (uint32_t q = b + threadIdx.x; q < e; q += blockDim.x) {
    const uint32_t src = sorted[q];
    S.box[out][2 * (size_t)q] = S.box[buf][2 * (size_t)src];
    S.box[out][2 * (size_t)q + 1] = S.box[buf][2 * (size_t)src + 1];
    S.perm[out][q] = S.perm[buf][src];
    S.pos[out][q] = (int32_t)d.child + (q < b + d.nleft ? 0 : 1);
  }
}

// Node ranges [nb, ne) -> first..last for the collapse.
__global__ void k_sah_last(uint32_t *__restrict__ ne, uint32_t nodes) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nodes) ne[p] -= 1u;
}


}  // namespace

#define GI_CK(x)                                                      \
  do {                                                                \
    hipError_t e_ = (x);                                              \
    if (e_ != hipSuccess) {                                           \
      err = std::string(#x) + ": " + hipGetErrorString(e_);           \
      for (void *p : mem) (void)hipFree(p);                           \
      return false;                                                   \
    }                                                                 \
  } while (0)

hipError_t launch_leaf_tris(hipStream_t s, const float4 *tri, uint32_t stride, uint32_t n, const uint32_t *order,
                            float4 *itris) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_itris, dim3(blocks(n)), dim3(kBlock), 0, s, tri, stride, n, order, itris);
  return hipGetLastError();
}

// One wide node whose only child is the leaf of triangles 0..count-1, with
// the padded box d_box (8 floats on the device).
static bool emit_single_leaf(hipStream_t s, const float *d_box, uint32_t count, WideNode *nodes, LbvhResult &res,
                      std::string &err) {
  float lb[8];
  hipError_t e = hipMemcpyAsync(lb, d_box, 32, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  WideNode nd;
  std::memset(&nd, 0, sizeof(nd));
  for (int k = 0; k < 4; ++k) {
    if (k > 0) {
      set_empty_slot(nd, k);
      continue;
    }
    nd.lox[k] = lb[0];
    nd.loy[k] = lb[1];
    nd.loz[k] = lb[2];
    nd.hix[k] = lb[4];
    nd.hiy[k] = lb[5];
    nd.hiz[k] = lb[6];
    nd.child[k] = kLeafBit | (count << 27);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(nodes, &nd, sizeof(nd), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    err = std::string("device bvh: ") + hipGetErrorString(e);
    return false;
  }
  res.nnodes = 1;
  res.depth = 0;
  return true;
}

// The wide nodes of the collapse of the binary tree under interior node
// root, level by level (front: two frontier buffers of the binary tree's
// interior node count; cnt: a device counter).
static bool emit_wide(hipStream_t s, const CollapseArgs &A, int32_t root, WideNode *nodes, int32_t *const front[2],
               uint32_t *cnt, LbvhResult &res, std::string &err) {
  hipError_t e = hipMemcpyAsync(front[0], &root, 4, hipMemcpyHostToDevice, s);
  uint32_t count = 1, base = 0, next_base = 1;
  int levels = 0;
  while (e == hipSuccess && count > 0) {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, cnt, (size_t)1);
    const uint32_t g = std::min<uint32_t>(blocks(count), 4096);
    hipLaunchKernelGGL(k_collapse, dim3(g), dim3(kBlock), 0, s, A, front[levels & 1], count, base, nodes,
                       front[(levels + 1) & 1], cnt, next_base);
    e = hipGetLastError();
    uint32_t nc = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&nc, cnt, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    ++levels;
    base = next_base;
    next_base += nc;
    count = nc;
  }
  if (e != hipSuccess) {
    err = std::string("device bvh collapse: ") + hipGetErrorString(e);
    return false;
  }
  res.nnodes = next_base;
  res.depth = levels - 1;
  return true;
}

// The binned-SAH tree, level by level (kernels above), then the collapse.
static bool device_sah_build(hipStream_t s, const float4 *tri, uint32_t stride, uint32_t n, int max_leaf, int W,
                             WideNode *nodes, float4 *itris, LbvhResult &res, std::string &err) {
  std::vector<void *> mem;
  auto alloc = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    mem.push_back(p);
    return p;
  };
  const size_t nn = 2 * (size_t)n;  // binary nodes: at most 2n - 1
  // a level lists the children of split nodes (more than max_leaf triangles each)
  const size_t lcap = std::min<size_t>(n, 2 * ((size_t)n / (max_leaf + 1)) + 2);
  SahArgs S;
  for (int b = 0; b < 2; ++b) {
    S.box[b] = (float4 *)alloc(32 * (size_t)n);
    S.perm[b] = (uint32_t *)alloc(4 * (size_t)n);
    S.pos[b] = (int32_t *)alloc(4 * (size_t)n);
  }
  S.left = (int32_t *)alloc(4 * nn);
  S.right = (int32_t *)alloc(4 * nn);
  S.nb = (uint32_t *)alloc(4 * nn);
  S.ne = (uint32_t *)alloc(4 * nn);
  S.bb = (float4 *)alloc(32 * nn);
  S.cb = (float4 *)alloc(32 * nn);
  S.nodebox = (float *)alloc(32 * nn);
  S.dec = (SahDec *)alloc(sizeof(SahDec) * lcap);
  S.leaf_perm = (uint32_t *)alloc(4 * (size_t)n);
  S.max_leaf = max_leaf;
  int32_t *lvl = (int32_t *)alloc(4 * nn);  // every level's list, one after another
  uint64_t *want = (uint64_t *)alloc(8 * lcap), *pre = (uint64_t *)alloc(8 * lcap);
  uint32_t *flags = (uint32_t *)alloc(4 * (size_t)n), *scan = (uint32_t *)alloc(4 * (size_t)n);
  const size_t max_big = (size_t)n / kSahChunk + 1;  // nodes of several chunks per level
  uint32_t *gbins = (uint32_t *)alloc(4 * (size_t)kSahBinWords * max_big);
  uint32_t *big_pos = (uint32_t *)alloc(4 * max_big);
  uint32_t *med_pos = (uint32_t *)alloc(4 * lcap), *seg_b = (uint32_t *)alloc(4 * lcap),
           *seg_e = (uint32_t *)alloc(4 * lcap);
  int32_t *med_kids = (int32_t *)alloc(8 * lcap);
  float4 *cost = (float4 *)alloc(16 * nn);
  uint32_t *meta = (uint32_t *)alloc(4 * nn);
  int32_t *front[2] = {(int32_t *)alloc(4 * (size_t)n), (int32_t *)alloc(4 * (size_t)n)};
  uint32_t *cnt = (uint32_t *)alloc(64);  // the collapse's counter; first the root's 12 box keys
  uint32_t *tot = (uint32_t *)alloc(16);  // a level's chunk and several-chunk node counts
  for (void *p : mem)
    if (!p) {
      err = "device bvh: out of device memory";
      for (void *q : mem) (void)hipFree(q);
      return false;
    }
  uint64_t *keys = nullptr, *keys2 = nullptr;  // median sort, allocated when a level needs it
  uint32_t *vals = nullptr, *vals2 = nullptr;
  void *sort_temp = nullptr;
  size_t sort_bytes = 0;
  size_t scan_bytes = 0, scan64_bytes = 0;
  GI_CK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flags, scan, (int)n, s));
  GI_CK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan64_bytes, want, pre, (int)lcap, s));
  void *scan_temp = alloc(std::max(scan_bytes, scan64_bytes));
  if (!scan_temp) {
    err = "device bvh: out of device memory (scan)";
    for (void *q : mem) (void)hipFree(q);
    return false;
  }
  hipLaunchKernelGGL(k_sah_init, dim3(blocks(n)), dim3(kBlock), 0, s, tri, stride, n, S);
  const int32_t root = 0;
  const uint32_t range[2] = {0u, n};
  GI_CK(hipMemcpyAsync(lvl, &root, 4, hipMemcpyHostToDevice, s));
  GI_CK(hipMemcpyAsync(S.nb, &range[0], 4, hipMemcpyHostToDevice, s));
  GI_CK(hipMemcpyAsync(S.ne, &range[1], 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_sah_root_init, dim3(1), dim3(64), 0, s, cnt);
  hipLaunchKernelGGL(k_sah_root, dim3(std::min<uint32_t>(blocks(n), 1024)), dim3(kBlock), 0, s, S.box[0], n, cnt);
  hipLaunchKernelGGL(k_sah_root_write, dim3(1), dim3(64), 0, s, S, cnt);
  GI_CK(hipGetLastError());
  std::vector<uint32_t> off{0, 1};  // level L's list: lvl[off[L] .. off[L + 1])
  uint32_t nodes_made = 1;
  int buf = 0;
  for (int depth = 0;; ++depth) {
    const uint32_t count = off.back() - off[off.size() - 2];
    if (count == 0) break;
    if (depth > kSahMaxDepth + 64 || count > lcap) {
      err = "device bvh: SAH build did not terminate";
      for (void *q : mem) (void)hipFree(q);
      return false;
    }
    const int32_t *list = lvl + off[off.size() - 2];
    hipLaunchKernelGGL(k_sah_plan, dim3(blocks(count)), dim3(kBlock), 0, s, S, list, count, depth, want);
    GI_CK(hipGetLastError());
    GI_CK(hipcub::DeviceScan::ExclusiveSum(scan_temp, scan64_bytes, want, pre, (int)count, s));
    // the chunk and several-chunk node counts stay on the device: the
    // launches cover their upper bounds (a several-chunk node has more than
    // kSahChunk triangles, and at most one partial chunk)
    hipLaunchKernelGGL(k_sah_totals, dim3(1), dim3(64), 0, s, pre, want, count, tot);
    const size_t big_bound = std::min<size_t>(max_big, count), chunk_bound = (size_t)n / kSahChunk + big_bound;
    if (n > kSahChunk) {
      hipLaunchKernelGGL(k_sah_bins_identity,
                         dim3((uint32_t)std::min<size_t>(blocks(big_bound * kSahBinWords), 4096)), dim3(kBlock), 0, s,
                         gbins, tot);
      hipLaunchKernelGGL(k_sah_bin_chunks, dim3((uint32_t)chunk_bound), dim3(kBlock), 0, s, S, buf, count, pre, tot,
                         gbins, big_pos);
      hipLaunchKernelGGL(k_sah_big, dim3((uint32_t)big_bound), dim3(64), 0, s, S, tot, gbins, big_pos);
    }
    hipLaunchKernelGGL(k_sah_wave, dim3(count), dim3(64), 0, s, S, buf, count);
    hipLaunchKernelGGL(k_sah_flags, dim3(blocks(n)), dim3(kBlock), 0, s, S, buf, n, flags);
    GI_CK(hipGetLastError());
    GI_CK(hipcub::DeviceScan::ExclusiveSum(scan_temp, scan_bytes, flags, scan, (int)n, s));
    hipLaunchKernelGGL(k_sah_count, dim3(blocks(count)), dim3(kBlock), 0, s, S, count, want);
    GI_CK(hipGetLastError());
    GI_CK(hipcub::DeviceScan::ExclusiveSum(scan_temp, scan64_bytes, want, pre, (int)count, s));
    uint64_t tail[2] = {0, 0};
    GI_CK(hipMemcpyAsync(&tail[0], pre + count - 1, 8, hipMemcpyDeviceToHost, s));
    GI_CK(hipMemcpyAsync(&tail[1], want + count - 1, 8, hipMemcpyDeviceToHost, s));
    GI_CK(hipStreamSynchronize(s));
    const uint64_t tot2 = tail[0] + tail[1];
    const uint32_t splits = (uint32_t)tot2, medians = (uint32_t)(tot2 >> 32);
    if (splits == 0) break;  // every node of the level is a leaf
    if ((size_t)nodes_made + 2 * (size_t)splits > nn - 1 || 2 * (size_t)splits > lcap) {
      err = "device bvh: SAH node count exceeded";
      for (void *q : mem) (void)hipFree(q);
      return false;
    }
    hipLaunchKernelGGL(k_sah_children, dim3(blocks(count)), dim3(kBlock), 0, s, S, count, pre, nodes_made, flags,
                       scan, lvl + off.back(), med_pos, med_kids, seg_b, seg_e);
    GI_CK(hipGetLastError());
    if (medians && !keys) {
      keys = (uint64_t *)alloc(8 * (size_t)n);
      keys2 = (uint64_t *)alloc(8 * (size_t)n);
      vals = (uint32_t *)alloc(4 * (size_t)n);
      vals2 = (uint32_t *)alloc(4 * (size_t)n);
      if (!keys || !keys2 || !vals || !vals2) {
        err = "device bvh: out of device memory (median sort)";
        for (void *q : mem) (void)hipFree(q);
        return false;
      }
    }
    hipLaunchKernelGGL(k_sah_scatter, dim3(blocks(n)), dim3(kBlock), 0, s, S, buf, n, flags, scan, keys, vals);
    GI_CK(hipGetLastError());
    if (medians) {
      size_t need = 0;
      GI_CK(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, need, keys, keys2, vals, vals2, (int)n,
                                                         (int)medians, seg_b, seg_e, 0, 64, s));
      if (need > sort_bytes) {
        sort_temp = alloc(need);
        sort_bytes = need;
        if (!sort_temp) {
          err = "device bvh: out of device memory (median sort)";
          for (void *q : mem) (void)hipFree(q);
          return false;
        }
      }
      GI_CK(hipcub::DeviceSegmentedRadixSort::SortPairs(sort_temp, need, keys, keys2, vals, vals2, (int)n,
                                                         (int)medians, seg_b, seg_e, 0, 64, s));
      hipLaunchKernelGGL(k_sah_median, dim3(medians), dim3(kBlock), 0, s, S, buf, med_pos, vals2);
      hipLaunchKernelGGL(k_sah_reduce, dim3(2 * medians), dim3(kBlock), 0, s, S, buf ^ 1, med_kids);
      GI_CK(hipGetLastError());
    }
    buf ^= 1;
    nodes_made += 2 * splits;
    off.push_back(off.back() + 2 * splits);
  }
  // leaf order
  hipLaunchKernelGGL(k_sah_final, dim3(blocks(n)), dim3(kBlock), 0, s, S, buf, n);
  hipLaunchKernelGGL(k_itris, dim3(blocks(n)), dim3(kBlock), 0, s, tri, stride, n, S.leaf_perm, itris);
  hipLaunchKernelGGL(k_sah_last, dim3(blocks(nodes_made)), dim3(kBlock), 0, s, S.ne, nodes_made);
  GI_CK(hipGetLastError());
  const CollapseArgs A{S.left, S.right, S.nb, S.ne, S.nodebox, S.nodebox, cost, meta, kCollapseMergeLimit, W};
  for (size_t L = off.size() - 1; L-- > 0;) {
    const uint32_t c = off[L + 1] - off[L];
    if (c) hipLaunchKernelGGL(k_collapse_cost, dim3(blocks(c)), dim3(kBlock), 0, s, A, lvl + off[L], 0u, c);
  }
  GI_CK(hipGetLastError());
  bool ok;
  if (nodes_made == 1) {
    ok = emit_single_leaf(s, S.nodebox, n, nodes, res, err);  // n <= max_leaf
  } else {
    ok = emit_wide(s, A, 0, nodes, front, cnt, res, err);
  }
  res.rounds = (int32_t)(off.size() - 2);  // binary levels below the root
  for (void *q : mem) (void)hipFree(q);
  return ok;
}

bool device_bvh_build(hipStream_t s, const float4 *tri, uint32_t stride, uint32_t n, const float clo[3],
                      const float chi[3], int max_leaf, DeviceBvhMethod method, int radius, int width, WideNode *nodes,
                      float4 *itris, LbvhResult &res, std::string &err) {
  std::vector<void *> mem;
  auto alloc = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    mem.push_back(p);
    return p;
  };
  if (n == 0 || max_leaf < 1 || max_leaf > kMaxLeafTris) {
    err = "device bvh: n > 0 and 1 <= max_leaf <= 15 required";
    return false;
  }
  if (method == DeviceBvhMethod::kPloc && (radius < 1 || radius > kPlocMaxRadius)) {
    err = "device bvh: PLOC radius must be in 1.." + std::to_string(kPlocMaxRadius);
    return false;
  }
  if (method == DeviceBvhMethod::kSah)
    return device_sah_build(s, tri, stride, n, max_leaf, std::max(2, std::min(4, width)), nodes, itris, res, err);
  const bool ploc = method == DeviceBvhMethod::kPloc;
  const size_t ni = n > 1 ? n - 1 : 1;
  // 30-bit keys for the radix tree, 63-bit keys for the clustering
  uint32_t *keys = (uint32_t *)alloc(8 * (size_t)n), *keys2 = (uint32_t *)alloc(8 * (size_t)n);
  uint32_t *vals = (uint32_t *)alloc(4 * (size_t)n), *vals2 = (uint32_t *)alloc(4 * (size_t)n);
  int32_t *left = (int32_t *)alloc(4 * ni), *right = (int32_t *)alloc(4 * ni), *parent = (int32_t *)alloc(4 * ni);
  int32_t *leaf_parent = (int32_t *)alloc(4 * (size_t)n);
  uint32_t *first = (uint32_t *)alloc(4 * ni), *last = (uint32_t *)alloc(4 * ni);
  int32_t *levels_buf = (int32_t *)alloc(4 * ni);  // radix tree nodes by breadth-first level
  float *leafbox = (float *)alloc(32 * (size_t)n), *nodebox = (float *)alloc(32 * ni);
  int32_t *front[2] = {(int32_t *)alloc(4 * ni), (int32_t *)alloc(4 * ni)};
  uint32_t *cnt = (uint32_t *)alloc(16);
  float4 *cost = (float4 *)alloc(16 * ni);  // the collapse's cost table and decisions
  uint32_t *meta = (uint32_t *)alloc(4 * ni);
  const int W = std::max(2, std::min(4, width));
  // PLOC: clusters (double-buffered), their boxes, neighbours, round flags
  // and their scan, node leaf counts, leaf slots, the reordered leaves.
  int32_t *cl[2] = {nullptr, nullptr}, *nn = nullptr;
  float4 *cbox[2] = {nullptr, nullptr}, *leafbox2 = nullptr;
  uint64_t *flags = nullptr, *pre = nullptr;
  uint32_t *ncount = nullptr, *slot = nullptr, *order2 = nullptr, *round_log = nullptr;
  constexpr int kRoundsPerSync = 4;
  if (ploc) {
    round_log = (uint32_t *)alloc(4 * ((size_t)n + kRoundsPerSync + 1));  // rounds <= n - 1 + kRoundsPerSync
    cl[0] = (int32_t *)alloc(4 * (size_t)n);
    cl[1] = (int32_t *)alloc(4 * (size_t)n);
    cbox[0] = (float4 *)alloc(32 * (size_t)n);
    cbox[1] = (float4 *)alloc(32 * (size_t)n);
    nn = (int32_t *)alloc(4 * (size_t)n);
    flags = (uint64_t *)alloc(8 * (size_t)n);
    pre = (uint64_t *)alloc(8 * (size_t)n);
    ncount = (uint32_t *)alloc(4 * ni);
    slot = (uint32_t *)alloc(4 * (size_t)n);
    order2 = (uint32_t *)alloc(4 * (size_t)n);
    leafbox2 = (float4 *)alloc(32 * (size_t)n);
  }
  for (void *p : mem)
    if (!p) {
      err = "device bvh: out of device memory";
      for (void *q : mem) (void)hipFree(q);
      return false;
    }
  float3 lo = make_float3(clo[0], clo[1], clo[2]);
  float3 sc;
  sc.x = chi[0] > clo[0] ? 1.0f / (chi[0] - clo[0]) : 0.0f;
  sc.y = chi[1] > clo[1] ? 1.0f / (chi[1] - clo[1]) : 0.0f;
  sc.z = chi[2] > clo[2] ? 1.0f / (chi[2] - clo[2]) : 0.0f;
  uint64_t *keys64 = (uint64_t *)keys, *keys64_2 = (uint64_t *)keys2;
  if (ploc) hipLaunchKernelGGL(k_morton64, dim3(blocks(n)), dim3(kBlock), 0, s, tri, stride, n, lo, sc, keys64, vals);
  else hipLaunchKernelGGL(k_morton, dim3(blocks(n)), dim3(kBlock), 0, s, tri, stride, n, lo, sc, keys, vals);
  GI_CK(hipGetLastError());
  size_t temp_bytes = 0, scan_bytes = 0;
  if (ploc)
    GI_CK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys64, keys64_2, vals, vals2, (int)n, 0, 63, s));
  else GI_CK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys, keys2, vals, vals2, (int)n, 0, 30, s));
  if (ploc) GI_CK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flags, pre, (int)n, s));
  void *temp = alloc(std::max(temp_bytes, scan_bytes));
  if (!temp) {
    err = "device bvh: out of device memory (sort)";
    for (void *q : mem) (void)hipFree(q);
    return false;
  }
  if (ploc) GI_CK(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys64, keys64_2, vals, vals2, (int)n, 0, 63, s));
  else GI_CK(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys2, vals, vals2, (int)n, 0, 30, s));
  const uint32_t *order = vals2;
  if (!ploc) {
    hipLaunchKernelGGL(k_leaf_boxes, dim3(blocks(n)), dim3(kBlock), 0, s, tri, stride, n, vals2, (float4 *)leafbox);
    GI_CK(hipGetLastError());
    if (n > 1) {
      hipLaunchKernelGGL(k_karras, dim3(blocks(ni)), dim3(kBlock), 0, s, keys2, (int)n, left, right, parent,
                         leaf_parent, first, last);
      GI_CK(hipGetLastError());
      // breadth-first levels (the host reads each level's size), then the
      // boxes level by level from the deepest: kernel boundaries order them
      std::vector<uint32_t> off{0, 1};
      const int32_t root0 = 0;
      GI_CK(hipMemcpyAsync(levels_buf, &root0, 4, hipMemcpyHostToDevice, s));
      while (off.back() > off[off.size() - 2] && off.back() < ni) {
        const uint32_t lo_i = off[off.size() - 2], cnt_l = off.back() - lo_i;
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, cnt, (size_t)1);
        hipLaunchKernelGGL(k_tree_level, dim3(std::min<uint32_t>(blocks(cnt_l), 4096)), dim3(kBlock), 0, s, left,
                           right, levels_buf + lo_i, cnt_l, levels_buf + off.back(), cnt);
        GI_CK(hipGetLastError());
        uint32_t nc = 0;
        GI_CK(hipMemcpyAsync(&nc, cnt, 4, hipMemcpyDeviceToHost, s));
        GI_CK(hipStreamSynchronize(s));
        if (nc > ni - off.back()) {
          err = "device bvh: radix tree levels exceed its node count";
          for (void *q : mem) (void)hipFree(q);
          return false;
        }
        off.push_back(off.back() + nc);
      }
      const CollapseArgs A{left, right, first, last, leafbox, nodebox, cost, meta, max_leaf, W};
      for (size_t L = off.size() - 1; L-- > 0;) {
        const uint32_t cnt_l = off[L + 1] - off[L];
        if (cnt_l == 0) continue;
        hipLaunchKernelGGL(k_refit_level, dim3(blocks(cnt_l)), dim3(kBlock), 0, s, left, right, levels_buf + off[L],
                           cnt_l, (const float4 *)leafbox, (float4 *)nodebox);
        hipLaunchKernelGGL(k_collapse_cost, dim3(blocks(cnt_l)), dim3(kBlock), 0, s, A, levels_buf + off[L], 0u,
                           cnt_l);
      }
      GI_CK(hipGetLastError());
    }
  } else {
    hipLaunchKernelGGL(k_ploc_init, dim3(blocks(n)), dim3(kBlock), 0, s, tri, stride, n, vals2, (float4 *)leafbox,
                       cbox[0], cl[0]);
    GI_CK(hipGetLastError());
    GI_CK(hipMemsetAsync(parent, 0xFF, 4 * ni, s));  // the root keeps parent -1
    // the cluster count and the nodes made so far live on the device
    // (ctl[0..1] for even rounds, ctl[2..3] for odd ones); the host reads
    // them back every kRoundsPerSync rounds and sizes the launches by the
    // last count it read, an upper bound
    uint32_t *ctl = cnt;
    const uint32_t ctl0[4] = {n, 0u, 0u, 0u};
    GI_CK(hipMemcpyAsync(ctl, ctl0, 16, hipMemcpyHostToDevice, s));
    GI_CK(hipMemcpyAsync(round_log, ctl0 + 1, 4, hipMemcpyHostToDevice, s));  // no nodes before round 0
    uint32_t bound = n;
    int rounds = 0, cur = 0;
    while (bound > 1) {
      if ((size_t)rounds > (size_t)n) {  // the round log holds n + kRoundsPerSync + 1 entries
        err = "device bvh: PLOC clustering stalled (too many rounds)";
        for (void *q : mem) (void)hipFree(q);
        return false;
      }
      for (int r = 0; r < kRoundsPerSync; ++r) {
        const uint32_t *c_in = ctl + 2 * (rounds & 1);
        uint32_t *c_out = ctl + 2 * ((rounds + 1) & 1);
        hipLaunchKernelGGL(k_ploc_nn, dim3(blocks(bound)), dim3(kBlock), 0, s, cbox[cur], c_in, radius, nn);
        hipLaunchKernelGGL(k_ploc_flags, dim3(blocks(bound)), dim3(kBlock), 0, s, nn, c_in, bound, flags);
        GI_CK(hipGetLastError());
        GI_CK(hipcub::DeviceScan::ExclusiveSum(temp, scan_bytes, flags, pre, (int)bound, s));
        hipLaunchKernelGGL(k_ploc_merge, dim3(blocks(bound)), dim3(kBlock), 0, s, cl[cur], cbox[cur], nn, flags, pre,
                           c_in, left, right, parent, leaf_parent, ncount, (float4 *)nodebox, cl[cur ^ 1],
                           cbox[cur ^ 1], c_out, round_log + rounds + 1);
        GI_CK(hipGetLastError());
        cur ^= 1;
        ++rounds;
      }
      uint32_t tot[2] = {0, 0};
      GI_CK(hipMemcpyAsync(tot, ctl + 2 * (rounds & 1), 8, hipMemcpyDeviceToHost, s));
      GI_CK(hipStreamSynchronize(s));
      // every round merges at least the globally closest pair
      if (tot[0] >= bound || tot[0] == 0 || tot[0] + tot[1] != n) {
        err = "device bvh: PLOC clustering stalled (non-finite vertex coordinates?)";
        for (void *q : mem) (void)hipFree(q);
        return false;
      }
      bound = tot[0];
    }
    res.rounds = rounds;
    if (n > 1) {
      hipLaunchKernelGGL(k_ploc_slot, dim3(blocks(n)), dim3(kBlock), 0, s, left, right, parent, leaf_parent, ncount, n,
                         slot);
      hipLaunchKernelGGL(k_ploc_ranges, dim3(blocks(n)), dim3(kBlock), 0, s, left, parent, leaf_parent, ncount, slot,
                         vals2, (const float4 *)leafbox, n, first, last, order2, leafbox2);
      hipLaunchKernelGGL(k_ploc_rename, dim3(blocks(n - 1)), dim3(kBlock), 0, s, left, right, slot, n - 1);
      GI_CK(hipGetLastError());
      order = order2;
      leafbox = (float *)leafbox2;
      // costs round by round: each round's nodes have their children from
      // earlier rounds
      std::vector<uint32_t> rn((size_t)rounds + 1);
      GI_CK(hipMemcpyAsync(rn.data(), round_log, 4 * rn.size(), hipMemcpyDeviceToHost, s));
      GI_CK(hipStreamSynchronize(s));
      const CollapseArgs A{left, right, first, last, leafbox, nodebox, cost, meta, max_leaf, W};
      for (int r = 0; r < rounds; ++r) {
        if (rn[r + 1] < rn[r] || rn[r + 1] > ni) {
          err = "device bvh: PLOC round log out of order";
          for (void *q : mem) (void)hipFree(q);
          return false;
        }
        const uint32_t c = rn[r + 1] - rn[r];
        if (c > 0) hipLaunchKernelGGL(k_collapse_cost, dim3(blocks(c)), dim3(kBlock), 0, s, A, nullptr, rn[r], c);
      }
      GI_CK(hipGetLastError());
    }
  }
  hipLaunchKernelGGL(k_itris, dim3(blocks(n)), dim3(kBlock), 0, s, tri, stride, n, order, itris);
  GI_CK(hipGetLastError());
  const CollapseArgs A{left, right, first, last, leafbox, nodebox, cost, meta, max_leaf, W};
  // one triangle: a root with a single leaf child; else the radix tree's
  // root is node 0 and PLOC's the last merge, n - 2
  const bool ok = n == 1 ? emit_single_leaf(s, leafbox, 1u, nodes, res, err)
                         : emit_wide(s, A, ploc ? (int32_t)(n - 2) : 0, nodes, front, cnt, res, err);
  for (void *q : mem) (void)hipFree(q);
  return ok;
}

}  // namespace gi
