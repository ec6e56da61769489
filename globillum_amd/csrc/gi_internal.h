// Internal declarations shared by the host sources of libgi_hip.so.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/gi_hip.h"

namespace gi {

// ---------------------------------------------------------------- BVH ----
// Binary BVH node of the host builder.  Leaves hold [first, first+count) of
// the builder's triangle permutation.
struct Bvh2Node {
  float lo[3], hi[3];
  int32_t left, right;   // children (inner), -1 for leaves
  int32_t first, count;  // leaf range, count == 0 for inner nodes
};

// 4-wide node streamed by the traversal kernels: the six child-box planes
// as float4 (one lane per child) plus the child references -- 112 bytes,
// padded to one 128-byte cache line.
//   ref >= 0                 : inner node index
//   ref & 0x80000000         : leaf, bits 27..30 = triangle count (1..15),
//                              bits 0..26 = first triangle (leaf order)
//   ref == kEmptyRef         : empty slot
struct alignas(128) WideNode {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  uint32_t child[4];
  uint32_t pad[4];
};
static_assert(sizeof(WideNode) == 128, "wide node must be one cache line");

// The traversal's copy of a WideNode, with the child planes in half
// precision: five 16-byte loads per visit instead of seven.
//   org, scale: the node box's min corner (f32) and a power of two s;
//   q[axis]: the four children's lo and hi planes on that axis as f16
//     offsets from org in units of s, packed lo0|lo1, lo2|lo3, hi0|hi1,
//     hi2|hi3 (low half first), rounded outward past a margin; empty slots
//     +inf (their box stays [+inf, +inf]);
//   child: the WideNode's child references.
// 80 bytes used of one 128-byte line, so nodes keep the WideNode indexing.
struct alignas(128) HalfNode {
  float org[3], scale;
  uint32_t q[3][4];
  uint32_t child[4];
  uint32_t pad[12];
};
static_assert(sizeof(HalfNode) == sizeof(WideNode), "half node keeps the wide node's stride");
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmptyRef = 0x80000000u;  // a leaf with zero triangles
constexpr int kMaxLeafTris = 15;
constexpr uint32_t kMaxLeafFirst = (1u << 27) - 1;

// An empty child slot: its box is [+inf, +inf] on every axis, so the slab
// test never reports it (each plane distance is +inf or -inf for any finite
// nonzero inverse direction) and the traversal needs no reference check.
#if defined(__HIP__)
__host__ __device__
#endif
inline void set_empty_slot(WideNode &nd, int k) {
  nd.lox[k] = nd.loy[k] = nd.loz[k] = __builtin_inff();
  nd.hix[k] = nd.hiy[k] = nd.hiz[k] = __builtin_inff();
  nd.child[k] = kEmptyRef;
}

// Triangle in leaf order for intersection: vertex a (w = original id as
// bits), edges b-a and c-a.  48 bytes.
struct alignas(16) IsectTri {
  float a[4];
  float e1[4];
  float e2[4];
};
static_assert(sizeof(IsectTri) == 48, "isect tri layout");

}  // namespace gi

struct gi_bvh {
  std::vector<gi::Bvh2Node> nodes;  // binary tree, root = 0
  std::vector<uint32_t> perm;       // leaf-order triangle ids
  int32_t max_depth = 0;
};

namespace gi {

// Binned-SAH build over the triangle list (host).  Boxes are padded by a
// relative 2^-20 so the float slab tests stay conservative.
// threads == 1 builds on the calling thread only; otherwise subtrees of at
// least 2^16 primitives near the root build on other threads.  Both give
// the same nodes in the same order.
void build_bvh2(const gi_tri *tris, int64_t n, int32_t max_leaf, gi_bvh &out, int32_t threads = 0);
// Import a binary BVH in the reference's 2 x float4 layout.
bool import_bvh2(const float *nodes, int64_t n_nodes, const uint32_t *indices, int64_t n_indices,
                 int64_t n_tris, gi_bvh &out, std::string &err);
// Collapse to the 4-wide layout (surface-area cost, dynamic program over
// slot counts) and emit the leaf-ordered intersection triangles.
void collapse_wide(const gi_bvh &b, const gi_tri *tris, std::vector<WideNode> &nodes,
                   std::vector<IsectTri> &leaf_tris, int32_t &wide_depth, int32_t width = 4);
// The pieces of collapse_wide: leaf-ordered triangles (a.w = index in tris),
// and the 4-wide nodes with interior references offset by node_base and
// leaf references from leaf_ref (called for leaves with count > 0).  The
// collapse may turn a binary subtree of at most merge_limit primitives (a
// contiguous range) into one leaf; 1 keeps the binary leaves as they are.
// width (2..4) is the number of child slots it fills per wide node.
#ifndef GI_COLLAPSE_MERGE
#define GI_COLLAPSE_MERGE 4
#endif
constexpr int32_t kCollapseMergeLimit = GI_COLLAPSE_MERGE;
// Cost of a wide-node visit (four box tests, one 128-byte fetch) in
// triangle tests, for the collapse's surface-area cost.
#ifndef GI_COLLAPSE_NODE_COST
#define GI_COLLAPSE_NODE_COST 1.0
#endif
constexpr double kCollapseNodeCost = GI_COLLAPSE_NODE_COST;
// Binned SAH (host builder and its device counterpart): centroid bins per
// axis, and the depth from which nodes split at the object median.
constexpr int kSahBins = 32;
constexpr int kSahMaxDepth = 48;
void leaf_order_tris(const gi_bvh &b, const gi_tri *tris, std::vector<IsectTri> &leaf_tris);
void collapse_wide_nodes(const gi_bvh &b, const std::function<uint32_t(const Bvh2Node &)> &leaf_ref,
                         uint32_t node_base, std::vector<WideNode> &nodes, int32_t &wide_depth,
                         int32_t merge_limit, int32_t width = 4);
// Relative box padding of the builders (2^-20 of the largest coordinate
// magnitude / extent): keeps the float slab tests conservative.
void pad_box(float *lo, float *hi);
// Binned-SAH build over boxes (lo xyz, hi xyz per primitive), split down to
// max_leaf primitives per leaf.
void build_bvh2_boxes(const float *boxes, int64_t n, int32_t max_leaf, gi_bvh &out, int32_t threads = 0);

// Special references (count 0): an empty slot, the end of a traversal, the
// return from an instance to world space, and instance i (kLeafBit | i + 1).
constexpr uint32_t kDoneRef = kLeafBit | kMaxLeafFirst;
constexpr uint32_t kReturnRef = kLeafBit | (kMaxLeafFirst - 1);
constexpr uint32_t kMaxInstances = kMaxLeafFirst - 2;

// ------------------------------------------------------------- Halton ----
// Per-dimension constants of the scrambled radical inverse, laid out for
// the device.  Digits are consumed m at a time (P = p^m <= 1024): chunk[c]
// for c < P holds the scrambled, reversed m-digit value
// sum_i perm[digit_i(c)] p^(m-1-i); magic = floor(2^64/P)+1 divides by P.
struct HaltonDim {
  uint32_t prime;       // p
  uint32_t chunk_off;   // offset of the chunk table in the uint16 pool
  uint32_t P;           // p^m
  uint32_t perm0;       // perm[0]
  uint64_t magic;       // floor(2^64 / P) + 1
  uint32_t kmax;        // chunks of any n < 2^32: the least K with P^K >= 2^32
  uint32_t pad;
};
static_assert(sizeof(HaltonDim) == 32, "halton dim layout");

int32_t prime_of(int32_t dim);
void halton_permutation(int32_t dim, uint32_t seed, uint16_t *perm);
void halton_tables(uint32_t seed, int32_t ndims, std::vector<HaltonDim> &dims,
                   std::vector<uint16_t> &chunks);
float halton_host(uint64_t n, int32_t dim, uint32_t seed);

// Sample dimension map: jitter x/y, lens x/y, then per path vertex b:
// NEE x, y, light choice; BSDF x, y, lobe; Russian roulette.
constexpr int kDimJitterX = 0, kDimJitterY = 1, kDimLensX = 2, kDimLensY = 3;
constexpr int kDimVertex0 = 4, kDimsPerVertex = 7;
constexpr int kMaxDepth = 32;
inline int dims_for_depth(int depth) { return kDimVertex0 + kDimsPerVertex * depth; }

}  // namespace gi
