// Device BVH build.  The binary tree is either a linear BVH (Karras 2012,
// "Maximizing parallelism in the construction of BVHs, octrees and k-d
// trees") over 30-bit Morton codes of the triangle centroids with a box
// refit level by level from the deepest, or a PLOC clustering over the
// 63-bit Morton order
// (below).  Either is then collapsed level by level into the traversal
// kernels' 4-wide nodes (children picked by opening the
// child of largest surface area, subtrees of at most max_leaf triangles
// becoming leaves) -- the same node format and leaf order conventions as
// the host builder's collapse_wide.  The device counterpart of the
// reference's GPU LBVH construction (src/main.cpp:218-225).
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

struct CollapseArgs {
  const int32_t *left, *right;
  const uint32_t *first, *last;
  const float *leafbox, *nodebox;
  int max_leaf;
  int width;  // child slots filled per wide node (2..4)
};

__device__ __forceinline__ uint32_t count_of(const CollapseArgs &A, int32_t c) {
  return c < 0 ? 1u : A.last[c] - A.first[c] + 1u;
}

// One level of the collapse: frontier[i] (a binary interior node) becomes
// wide node base + i; its interior children that stay interior are
// appended to next (wide nodes next_base + position).
__global__ void k_collapse(CollapseArgs A, const int32_t *__restrict__ frontier, uint32_t count, uint32_t base,
                           WideNode *__restrict__ nodes, int32_t *__restrict__ next, uint32_t *next_count,
                           uint32_t next_base) {
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < count; i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;  // the loop is uniform per block: wave_append needs every lane
    const bool active = i < count;
    int32_t kids[4] = {0, 0, 0, 0};
    int nk = 0;
    if (active) {
      const int32_t b = frontier[i];
      kids[0] = A.left[b];
      kids[1] = A.right[b];
      nk = 2;
      while (nk < A.width) {
        int best = -1;
        float ba = -1.0f;
        for (int k = 0; k < nk; ++k) {
          const int32_t c = kids[k];
          if (c < 0 || count_of(A, c) <= (uint32_t)A.max_leaf) continue;
          const float *p = A.nodebox + 8 * (size_t)c;
          const float dx = p[4] - p[0], dy = p[5] - p[1], dz = p[6] - p[2];
          const float area = dx * dy + dy * dz + dz * dx;
          if (area > ba) {
            ba = area;
            best = k;
          }
        }
        if (best < 0) break;
        const int32_t o = kids[best];
        kids[best] = A.left[o];
        kids[nk++] = A.right[o];
      }
    }
    WideNode nd;
    uint32_t qs[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int32_t c = kids[k];
      const bool grow = active && k < nk && c >= 0 && count_of(A, c) > (uint32_t)A.max_leaf;
      qs[k] = wave_append(grow, next_count);
      if (grow) next[qs[k]] = c;
    }
    if (!active) continue;
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
      const uint32_t cnt = count_of(A, c);
      if (c >= 0 && cnt > (uint32_t)A.max_leaf) nd.child[k] = next_base + qs[k];
      else nd.child[k] = kLeafBit | (cnt << 27) | (c < 0 ? (uint32_t)~c : A.first[c]);
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
// of this round; totals = the same for the next round.
__global__ void k_ploc_merge(const int32_t *__restrict__ cl, const float4 *__restrict__ cbox,
                             const int32_t *__restrict__ nn, const uint64_t *__restrict__ flags,
                             const uint64_t *__restrict__ pre, const uint32_t *__restrict__ ctl,
                             int32_t *__restrict__ left, int32_t *__restrict__ right, int32_t *__restrict__ parent,
                             int32_t *__restrict__ leaf_parent, uint32_t *__restrict__ ncount,
                             float4 *__restrict__ nodebox, int32_t *__restrict__ cl_out,
                             float4 *__restrict__ cbox_out, uint32_t *__restrict__ totals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t cnt = ctl[0], node_base = ctl[1];  // clusters and nodes so far, this round
  if (i >= cnt) return;
  const uint64_t f = flags[i], p = pre[i];
  const uint32_t keep = (uint32_t)f, lead = (uint32_t)(f >> 32);
  const uint32_t slot = (uint32_t)p, rank = (uint32_t)(p >> 32);
  if (i == cnt - 1) {
    totals[0] = slot + keep;
    totals[1] = node_base + rank + lead;
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
  // PLOC: clusters (double-buffered), their boxes, neighbours, round flags
  // and their scan, node leaf counts, leaf slots, the reordered leaves.
  int32_t *cl[2] = {nullptr, nullptr}, *nn = nullptr;
  float4 *cbox[2] = {nullptr, nullptr}, *leafbox2 = nullptr;
  uint64_t *flags = nullptr, *pre = nullptr;
  uint32_t *ncount = nullptr, *slot = nullptr, *order2 = nullptr;
  if (ploc) {
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
      for (size_t L = off.size() - 1; L-- > 0;) {
        const uint32_t cnt_l = off[L + 1] - off[L];
        if (cnt_l == 0) continue;
        hipLaunchKernelGGL(k_refit_level, dim3(blocks(cnt_l)), dim3(kBlock), 0, s, left, right, levels_buf + off[L],
                           cnt_l, (const float4 *)leafbox, (float4 *)nodebox);
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
    constexpr int kRoundsPerSync = 4;
    uint32_t *ctl = cnt;
    const uint32_t ctl0[4] = {n, 0u, 0u, 0u};
    GI_CK(hipMemcpyAsync(ctl, ctl0, 16, hipMemcpyHostToDevice, s));
    uint32_t bound = n;
    int rounds = 0, cur = 0;
    while (bound > 1) {
      for (int r = 0; r < kRoundsPerSync; ++r) {
        const uint32_t *c_in = ctl + 2 * (rounds & 1);
        uint32_t *c_out = ctl + 2 * ((rounds + 1) & 1);
        hipLaunchKernelGGL(k_ploc_nn, dim3(blocks(bound)), dim3(kBlock), 0, s, cbox[cur], c_in, radius, nn);
        hipLaunchKernelGGL(k_ploc_flags, dim3(blocks(bound)), dim3(kBlock), 0, s, nn, c_in, bound, flags);
        GI_CK(hipGetLastError());
        GI_CK(hipcub::DeviceScan::ExclusiveSum(temp, scan_bytes, flags, pre, (int)bound, s));
        hipLaunchKernelGGL(k_ploc_merge, dim3(blocks(bound)), dim3(kBlock), 0, s, cl[cur], cbox[cur], nn, flags, pre,
                           c_in, left, right, parent, leaf_parent, ncount, (float4 *)nodebox, cl[cur ^ 1],
                           cbox[cur ^ 1], c_out);
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
    }
  }
  hipLaunchKernelGGL(k_itris, dim3(blocks(n)), dim3(kBlock), 0, s, tri, stride, n, order, itris);
  GI_CK(hipGetLastError());
  if (n == 1) {
    // one triangle: a root with a single leaf child
    float lb[8];
    GI_CK(hipMemcpyAsync(lb, leafbox, 32, hipMemcpyDeviceToHost, s));
    GI_CK(hipStreamSynchronize(s));
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
      nd.child[k] = kLeafBit | (1u << 27);
    }
    GI_CK(hipMemcpyAsync(nodes, &nd, sizeof(nd), hipMemcpyHostToDevice, s));
    GI_CK(hipStreamSynchronize(s));
    res.nnodes = 1;
    res.depth = 0;
    for (void *q : mem) (void)hipFree(q);
    return true;
  }
  CollapseArgs A{left, right, first, last, leafbox, nodebox, max_leaf, std::max(2, std::min(4, width))};
  // the radix tree's root is node 0; PLOC's is the last merge, n - 2
  const int32_t root = ploc ? (int32_t)(n - 2) : 0;
  GI_CK(hipMemcpyAsync(front[0], &root, 4, hipMemcpyHostToDevice, s));
  uint32_t count = 1, base = 0, next_base = 1;
  int levels = 0;
  while (count > 0) {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, cnt, (size_t)1);
    const uint32_t g = std::min<uint32_t>(blocks(count), 4096);
    hipLaunchKernelGGL(k_collapse, dim3(g), dim3(kBlock), 0, s, A, front[levels & 1], count, base, nodes,
                       front[(levels + 1) & 1], cnt, next_base);
    GI_CK(hipGetLastError());
    uint32_t nc = 0;
    GI_CK(hipMemcpyAsync(&nc, cnt, 4, hipMemcpyDeviceToHost, s));
    GI_CK(hipStreamSynchronize(s));
    ++levels;
    base = next_base;
    next_base += nc;
    count = nc;
  }
  res.nnodes = next_base;
  res.depth = levels - 1;
  for (void *q : mem) (void)hipFree(q);
  return true;
}

}  // namespace gi
