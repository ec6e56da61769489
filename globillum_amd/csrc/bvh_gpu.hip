// Device BVH build: linear BVH (Karras 2012, "Maximizing parallelism in the
// construction of BVHs, octrees and k-d trees") over 30-bit Morton codes of
// the triangle centroids, bottom-up box refit, and a level-by-level collapse
// into the traversal kernels' 4-wide nodes (children picked by opening the
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

__device__ __forceinline__ float relaxed_load(const float *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void load_box(const float *p, Box &b) {
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = relaxed_load(p + k);
    b.hi[k] = relaxed_load(p + 4 + k);
  }
}
__device__ __forceinline__ void store_box(float *p, const Box &b) {
  for (int k = 0; k < 3; ++k) {
    __hip_atomic_store(p + k, b.lo[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 4 + k, b.hi[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Leaf boxes, then every interior box from the second of its two children
// to finish (release / acquire on the node's arrival counter).
__global__ void k_refit(const float4 *__restrict__ tri, uint32_t stride, int n, const uint32_t *__restrict__ order,
                        const int32_t *__restrict__ left, const int32_t *__restrict__ right,
                        const int32_t *__restrict__ parent, const int32_t *__restrict__ leaf_parent,
                        float *leafbox, float *nodebox, uint32_t *visit) {
  const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= n) return;
  const size_t t = order[k];
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
  store_box(leafbox + 8 * (size_t)k, bx);
  int p = n > 1 ? leaf_parent[k] : -1;
  while (p >= 0) {
    const uint32_t old = __hip_atomic_fetch_add(visit + p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0) return;  // the sibling subtree is not done yet: it continues
    Box l, r;
    const int32_t lc = left[p], rc = right[p];
    load_box(lc < 0 ? leafbox + 8 * (size_t)~lc : nodebox + 8 * (size_t)lc, l);
    load_box(rc < 0 ? leafbox + 8 * (size_t)~rc : nodebox + 8 * (size_t)rc, r);
    for (int q = 0; q < 3; ++q) {
      bx.lo[q] = fminf(l.lo[q], r.lo[q]);
      bx.hi[q] = fmaxf(l.hi[q], r.hi[q]);
    }
    store_box(nodebox + 8 * (size_t)p, bx);
    p = parent[p];
  }
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

struct CollapseArgs {
  const int32_t *left, *right;
  const uint32_t *first, *last;
  const float *leafbox, *nodebox;
  int max_leaf;
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
      while (nk < 4) {
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

bool lbvh_build(hipStream_t s, const float4 *tri, uint32_t stride, uint32_t n, const float clo[3], const float chi[3],
                int max_leaf, WideNode *nodes, float4 *itris, LbvhResult &res, std::string &err) {
  std::vector<void *> mem;
  auto alloc = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    mem.push_back(p);
    return p;
  };
  if (n == 0 || max_leaf < 1 || max_leaf > kMaxLeafTris) {
    err = "lbvh: n > 0 and 1 <= max_leaf <= 15 required";
    return false;
  }
  const size_t ni = n > 1 ? n - 1 : 1;
  uint32_t *keys = (uint32_t *)alloc(4 * (size_t)n), *keys2 = (uint32_t *)alloc(4 * (size_t)n);
  uint32_t *vals = (uint32_t *)alloc(4 * (size_t)n), *vals2 = (uint32_t *)alloc(4 * (size_t)n);
  int32_t *left = (int32_t *)alloc(4 * ni), *right = (int32_t *)alloc(4 * ni), *parent = (int32_t *)alloc(4 * ni);
  int32_t *leaf_parent = (int32_t *)alloc(4 * (size_t)n);
  uint32_t *first = (uint32_t *)alloc(4 * ni), *last = (uint32_t *)alloc(4 * ni), *visit = (uint32_t *)alloc(4 * ni);
  float *leafbox = (float *)alloc(32 * (size_t)n), *nodebox = (float *)alloc(32 * ni);
  int32_t *front[2] = {(int32_t *)alloc(4 * ni), (int32_t *)alloc(4 * ni)};
  uint32_t *cnt = (uint32_t *)alloc(16);
  for (void *p : mem)
    if (!p) {
      err = "lbvh: out of device memory";
      for (void *q : mem) (void)hipFree(q);
      return false;
    }
  float3 lo = make_float3(clo[0], clo[1], clo[2]);
  float3 sc;
  sc.x = chi[0] > clo[0] ? 1.0f / (chi[0] - clo[0]) : 0.0f;
  sc.y = chi[1] > clo[1] ? 1.0f / (chi[1] - clo[1]) : 0.0f;
  sc.z = chi[2] > clo[2] ? 1.0f / (chi[2] - clo[2]) : 0.0f;
  hipLaunchKernelGGL(k_morton, dim3(blocks(n)), dim3(kBlock), 0, s, tri, stride, n, lo, sc, keys, vals);
  GI_CK(hipGetLastError());
  size_t temp_bytes = 0;
  GI_CK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys, keys2, vals, vals2, (int)n, 0, 30, s));
  void *temp = alloc(temp_bytes);
  if (!temp) {
    err = "lbvh: out of device memory (sort)";
    for (void *q : mem) (void)hipFree(q);
    return false;
  }
  GI_CK(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys2, vals, vals2, (int)n, 0, 30, s));
  hipLaunchKernelGGL(k_zero, dim3(blocks(ni)), dim3(kBlock), 0, s, visit, ni);
  if (n > 1) {
    hipLaunchKernelGGL(k_karras, dim3(blocks(ni)), dim3(kBlock), 0, s, keys2, (int)n, left, right, parent,
                       leaf_parent, first, last);
    GI_CK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_refit, dim3(blocks(n)), dim3(kBlock), 0, s, tri, stride, (int)n, vals2, left, right, parent,
                     leaf_parent, leafbox, nodebox, visit);
  GI_CK(hipGetLastError());
  hipLaunchKernelGGL(k_itris, dim3(blocks(n)), dim3(kBlock), 0, s, tri, stride, n, vals2, itris);
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
  CollapseArgs A{left, right, first, last, leafbox, nodebox, max_leaf};
  const int32_t root = 0;
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
