// Host BVH construction: binned SAH (stands in for rta's "-b bsah",
// src/hl-main.cpp:377-379), import/export of the reference's 2 x float4
// node layout (src/tracers.cu:26-36), and the collapse to the 4-wide
// 128-byte nodes the gfx950 traversal kernels stream.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <functional>
#include <memory>
#include <future>
#include <numeric>

#include "gi_internal.h"

namespace gi {
namespace {

struct Box {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  void grow(const float *p) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], p[k]);
      hi[k] = std::max(hi[k], p[k]);
    }
  }
  void grow(const Box &b) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], b.lo[k]);
      hi[k] = std::max(hi[k], b.hi[k]);
    }
  }
  bool empty() const { return lo[0] > hi[0]; }
  float area() const {
    if (empty()) return 0.f;
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.f * (dx * dy + dy * dz + dz * dx);
  }
};

}  // namespace

// Relative padding that keeps float slab tests conservative for points on
// the triangles: 2^-20 of the largest coordinate magnitude / extent.
void pad_box(float *lo, float *hi) {
  float m = 0.f;
  for (int k = 0; k < 3; ++k) m = std::max(m, std::max(std::fabs(lo[k]), std::fabs(hi[k])));
  for (int k = 0; k < 3; ++k) m = std::max(m, hi[k] - lo[k]);
  float p = m * 0x1p-20f + 1e-30f;
  for (int k = 0; k < 3; ++k) {
    lo[k] -= p;
    hi[k] += p;
  }
}

namespace {

constexpr int kBins = kSahBins;
constexpr int kSahDepth = kSahMaxDepth;  // beyond this depth, split at the object median
// Subtrees of at least this many primitives, down to kParallelDepth, build
// their right half on another thread.
constexpr uint32_t kParallelMin = 1u << 16;
constexpr int kParallelDepth = 5;

struct Builder {
  const gi_tri *tris;
  std::vector<Box> tbox;
  std::vector<float> cen;  // 3 per triangle
  std::vector<uint32_t> &perm;
  std::vector<Bvh2Node> &nodes;
  int32_t max_leaf;
  int32_t max_depth = 0;
  bool sah_leaves = true;  // false: split down to max_leaf primitives per leaf
  bool parallel = false;   // build large right subtrees on other threads

  Builder(const gi_tri *t, int64_t n, int32_t ml, std::vector<uint32_t> &p, std::vector<Bvh2Node> &nd)
      : tris(t), tbox(n), cen(3 * n), perm(p), nodes(nd), max_leaf(ml) {
    for (int64_t i = 0; i < n; ++i) {
      tbox[i].grow(t[i].a);
      tbox[i].grow(t[i].b);
      tbox[i].grow(t[i].c);
      for (int k = 0; k < 3; ++k) cen[3 * i + k] = 0.5f * (tbox[i].lo[k] + tbox[i].hi[k]);
    }
  }
  // over arbitrary boxes (lo xyz, hi xyz per primitive)
  Builder(const float *boxes, int64_t n, int32_t ml, std::vector<uint32_t> &p, std::vector<Bvh2Node> &nd)
      : tris(nullptr), tbox(n), cen(3 * n), perm(p), nodes(nd), max_leaf(ml) {
    for (int64_t i = 0; i < n; ++i) {
      tbox[i].grow(boxes + 6 * i);
      tbox[i].grow(boxes + 6 * i + 3);
      for (int k = 0; k < 3; ++k) cen[3 * i + k] = 0.5f * (tbox[i].lo[k] + tbox[i].hi[k]);
    }
  }

  static int32_t make_leaf(std::vector<Bvh2Node> &nodes, int32_t node, uint32_t b, uint32_t e) {
    nodes[node].left = nodes[node].right = -1;
    nodes[node].first = (int32_t)b;
    nodes[node].count = (int32_t)(e - b);
    return node;
  }

  void build(int32_t node, uint32_t b, uint32_t e, int depth) { build(nodes, node, b, e, depth, max_depth); }

  // Subtree of node over perm[b, e) into `nodes`; children are appended in
  // the order of a depth-first build (left subtree complete before the
  // right), whichever thread builds them.
  void build(std::vector<Bvh2Node> &nodes, int32_t node, uint32_t b, uint32_t e, int depth, int32_t &maxd) {
    maxd = std::max(maxd, depth);
    Box bb, cb;
    for (uint32_t i = b; i < e; ++i) {
      bb.grow(tbox[perm[i]]);
      cb.grow(&cen[3 * perm[i]]);
    }
    std::memcpy(nodes[node].lo, bb.lo, 12);
    std::memcpy(nodes[node].hi, bb.hi, 12);
    pad_box(nodes[node].lo, nodes[node].hi);
    const uint32_t n = e - b;
    if (n <= (uint32_t)max_leaf) {
      make_leaf(nodes, node, b, e);
      return;
    }
    int axis = -1;
    int split_bin = -1;
    float best = FLT_MAX;
    const float leaf_cost = (float)n;
    uint32_t mid = b;
    if (depth < kSahDepth) {
      for (int k = 0; k < 3; ++k) {
        float ext = cb.hi[k] - cb.lo[k];
        if (!(ext > 0.f)) continue;
        Box bin[kBins];
        uint32_t cnt[kBins] = {0};
        float scale = (float)kBins / ext;
        for (uint32_t i = b; i < e; ++i) {
          int bi = std::min(kBins - 1, (int)((cen[3 * perm[i] + k] - cb.lo[k]) * scale));
          bin[bi].grow(tbox[perm[i]]);
          ++cnt[bi];
        }
        float ra[kBins];
        uint32_t rc[kBins];
        Box acc;
        uint32_t c = 0;
        for (int i = kBins - 1; i > 0; --i) {
          acc.grow(bin[i]);
          c += cnt[i];
          ra[i] = acc.area();
          rc[i] = c;
        }
        Box accl;
        uint32_t cl = 0;
        for (int i = 0; i < kBins - 1; ++i) {
          accl.grow(bin[i]);
          cl += cnt[i];
          if (cl == 0 || rc[i + 1] == 0) continue;
          float cost = accl.area() * (float)cl + ra[i + 1] * (float)rc[i + 1];
          if (cost < best) {
            best = cost;
            axis = k;
            split_bin = i;
          }
        }
      }
      float parent_area = bb.area();
      if (axis >= 0) {
        float sah = 1.0f + best / std::max(parent_area, 1e-30f);
        if (sah_leaves && sah >= leaf_cost && n <= (uint32_t)kMaxLeafTris) {
          make_leaf(nodes, node, b, e);
          return;
        }
        float ext = cb.hi[axis] - cb.lo[axis];
        float scale = (float)kBins / ext;
        auto it = std::partition(perm.begin() + b, perm.begin() + e, [&](uint32_t id) {
          int bi = std::min(kBins - 1, (int)((cen[3 * id + axis] - cb.lo[axis]) * scale));
          return bi <= split_bin;
        });
        mid = (uint32_t)(it - perm.begin());
      }
    }
    if (mid == b || mid == e) {
      // degenerate centroids or depth limit: object median on the longest axis
      int ax = 0;
      for (int k = 1; k < 3; ++k)
        if (cb.hi[k] - cb.lo[k] > cb.hi[ax] - cb.lo[ax]) ax = k;
      mid = (b + e) / 2;
      std::nth_element(perm.begin() + b, perm.begin() + mid, perm.begin() + e, [&](uint32_t x, uint32_t y) {
        float cx = cen[3 * x + ax], cy = cen[3 * y + ax];
        return cx < cy || (cx == cy && x < y);
      });
    }
    int32_t l = (int32_t)nodes.size();
    nodes.resize(nodes.size() + 2);
    nodes[node].left = l;
    nodes[node].right = l + 1;
    nodes[node].first = 0;
    nodes[node].count = 0;
    if (!(parallel && n >= kParallelMin && depth < kParallelDepth)) {
      build(nodes, l, b, mid, depth + 1, maxd);
      build(nodes, l + 1, mid, e, depth + 1, maxd);
      return;
    }
    // the right subtree on another thread, into its own array (its root at
    // 0), appended after the left subtree: the same nodes in the same order
    // as the sequential build
    std::vector<Bvh2Node> sub(1);
    sub.reserve(2 * (size_t)(e - mid));
    int32_t sub_depth = 0;
    auto right = std::async(std::launch::async, [&] { build(sub, 0, mid, e, depth + 1, sub_depth); });
    build(nodes, l, b, mid, depth + 1, maxd);
    right.get();
    maxd = std::max(maxd, sub_depth);
    const int32_t base = (int32_t)nodes.size();
    auto remap = [&](int32_t c) { return c < 0 ? c : (c == 0 ? l + 1 : base + c - 1); };
    nodes.resize((size_t)base + sub.size() - 1);
    for (size_t j = 0; j < sub.size(); ++j) {
      Bvh2Node nd = sub[j];
      if (nd.count == 0 && nd.left >= 0) {
        nd.left = remap(nd.left);
        nd.right = remap(nd.right);
      }
      nodes[j == 0 ? (size_t)l + 1 : (size_t)base + j - 1] = nd;
    }
  }
};

inline int32_t fbits(float f) {
  int32_t i;
  std::memcpy(&i, &f, 4);
  return i;
}
inline float bitsf(int32_t i) {
  float f;
  std::memcpy(&f, &i, 4);
  return f;
}

}  // namespace

void build_bvh2(const gi_tri *tris, int64_t n, int32_t max_leaf, gi_bvh &out, int32_t threads) {
  out.nodes.clear();
  out.perm.resize(n);
  std::iota(out.perm.begin(), out.perm.end(), 0u);
  max_leaf = std::max(1, std::min(max_leaf, kMaxLeafTris));
  out.nodes.reserve(2 * n + 1);
  out.nodes.resize(1);
  if (n == 0) {
    out.nodes[0] = Bvh2Node{{0, 0, 0}, {0, 0, 0}, -1, -1, 0, 0};
    out.max_depth = 0;
    return;
  }
  Builder bld(tris, n, max_leaf, out.perm, out.nodes);
  bld.parallel = threads != 1;
  bld.build(0, 0, (uint32_t)n, 0);
  out.max_depth = bld.max_depth;
}

void build_bvh2_boxes(const float *boxes, int64_t n, int32_t max_leaf, gi_bvh &out, int32_t threads) {
  out.nodes.clear();
  out.perm.resize(n);
  std::iota(out.perm.begin(), out.perm.end(), 0u);
  max_leaf = std::max(1, std::min(max_leaf, kMaxLeafTris));
  out.nodes.reserve(2 * n + 1);
  out.nodes.resize(1);
  if (n == 0) {
    out.nodes[0] = Bvh2Node{{0, 0, 0}, {0, 0, 0}, -1, -1, 0, 0};
    out.max_depth = 0;
    return;
  }
  Builder bld(boxes, n, max_leaf, out.perm, out.nodes);
  bld.sah_leaves = false;
  bld.parallel = threads != 1;
  bld.build(0, 0, (uint32_t)n, 0);
  out.max_depth = bld.max_depth;
}

bool import_bvh2(const float *nodes, int64_t n_nodes, const uint32_t *indices, int64_t n_indices,
                 int64_t n_tris, gi_bvh &out, std::string &err) {
  // Reference layout (src/tracers.cu:26-36): x bit0 = inner flag,
  // inner: left = x>>1, right = y; leaf: count = x>>1, offset = y;
  // box min = (n0.z, n0.w, n1.x), box max = (n1.y, n1.z, n1.w).
  // Leaves are re-expressed over a fresh permutation so each leaf owns a
  // contiguous range (the reference's index array may share entries).
  out.nodes.assign(n_nodes, Bvh2Node{});
  out.perm.clear();
  out.max_depth = 0;
  std::vector<std::pair<int64_t, int>> st{{0, 0}};
  std::vector<uint8_t> seen(n_nodes, 0);
  while (!st.empty()) {
    auto [i, depth] = st.back();
    st.pop_back();
    if (i < 0 || i >= n_nodes || seen[i]) {
      err = "bvh2: node index out of range or shared";
      return false;
    }
    seen[i] = 1;
    out.max_depth = std::max(out.max_depth, depth);
    const float *p = nodes + 8 * i;
    Bvh2Node &o = out.nodes[i];
    o.lo[0] = p[2]; o.lo[1] = p[3]; o.lo[2] = p[4];
    o.hi[0] = p[5]; o.hi[1] = p[6]; o.hi[2] = p[7];
    pad_box(o.lo, o.hi);  // keep the device slab test conservative
    int32_t x = fbits(p[0]), y = fbits(p[1]);
    if (x & 1) {
      o.left = x >> 1;
      o.right = y;
      o.first = o.count = 0;
      st.push_back({o.right, depth + 1});
      st.push_back({o.left, depth + 1});
    } else {
      int64_t cnt = x >> 1, off = y;
      if (cnt < 0 || off < 0 || off + cnt > n_indices) {
        err = "bvh2: leaf range outside the index array";
        return false;
      }
      if (cnt > kMaxLeafTris) {
        err = "bvh2: leaf with more than 15 triangles";
        return false;
      }
      o.left = o.right = -1;
      o.first = (int32_t)out.perm.size();
      o.count = (int32_t)cnt;
      for (int64_t k = 0; k < cnt; ++k) {
        if (indices[off + k] >= (uint64_t)n_tris) {
          err = "bvh2: triangle index out of range";
          return false;
        }
        out.perm.push_back(indices[off + k]);
      }
    }
  }
  return true;
}

void leaf_order_tris(const gi_bvh &b, const gi_tri *tris, std::vector<IsectTri> &leaf_tris) {
  // Leaf-ordered intersection triangles: e1 = b - a, e2 = c - a in float,
  // the same rounding the oracle's per-test subtraction produces; a.w holds
  // the triangle's index in tris.
  const size_t n = b.perm.size();
  leaf_tris.resize(n);
  auto fill = [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
      const gi_tri &t = tris[b.perm[i]];
      IsectTri &o = leaf_tris[i];
      for (int k = 0; k < 3; ++k) {
        o.a[k] = t.a[k];
        o.e1[k] = t.b[k] - t.a[k];
        o.e2[k] = t.c[k] - t.a[k];
      }
      o.a[3] = bitsf((int32_t)b.perm[i]);
      o.e1[3] = 0.f;
      o.e2[3] = 0.f;
    }
  };
  // a gather through the permutation: in chunks on other threads when large
  constexpr size_t kChunkTris = 1u << 18;
  std::vector<std::future<void>> parts;
  for (size_t lo = kChunkTris; lo < n; lo += kChunkTris)
    parts.push_back(std::async(std::launch::async, fill, lo, std::min(n, lo + kChunkTris)));
  fill(0, std::min(n, kChunkTris));
  for (auto &f : parts) f.get();
}

// Binary -> 4-wide collapse chosen by surface-area cost (the dynamic
// program of Ylitie, Karras and Laine 2017, "Efficient incoherent ray
// traversal on GPUs through compressed wide BVHs", for width 4).  C[n][i] is
// the least cost of binary subtree n when it may take up to i slots of its
// parent: one slot holds n as a wide node (area(n) * kCollapseNodeCost plus the best
// split of `width` (2..4) slots between its children) or, when its triangles are
// contiguous and at most merge_limit, as one leaf (area(n) * count); more
// slots open n and split them between its children.  Costs are in triangle
// tests; a wide-node visit (four box tests, one 128-byte fetch) counts as
// kCollapseNodeCost of them.

void collapse_wide_nodes(const gi_bvh &b, const std::function<uint32_t(const Bvh2Node &)> &leaf_ref,
                         uint32_t node_base, std::vector<WideNode> &wnodes, int32_t &wide_depth,
                         int32_t merge_limit, int32_t width) {
  wnodes.clear();
  wide_depth = 0;
  const int W = std::max(2, std::min(4, (int)width));  // slots filled per wide node
  const auto &N = b.nodes;
  const size_t nn = N.size();
  // post-order of the nodes reachable from the root: reverse index order
  // when every child follows its parent (the builders here allocate so),
  // else an explicit walk (imported trees)
  std::vector<int32_t> post;
  bool children_after = true;
  for (size_t i = 0; i < nn && children_after; ++i)
    if (N[i].left >= 0) children_after = N[i].left > (int32_t)i && N[i].right > (int32_t)i;
  if (children_after) {
    post.resize(nn);
    for (size_t i = 0; i < nn; ++i) post[i] = (int32_t)(nn - 1 - i);
  } else {
    post.reserve(nn);
    std::vector<std::pair<int32_t, bool>> st{{0, false}};
    while (!st.empty()) {
      const auto [i, open] = st.back();
      st.pop_back();
      if (open || N[i].left < 0) {
        post.push_back(i);
        continue;
      }
      st.push_back({i, true});
      st.push_back({N[i].right, false});
      st.push_back({N[i].left, false});
    }
  }
  // per node, written before it is read (children precede parents in post)
  constexpr double kInf = std::numeric_limits<double>::infinity();
  std::unique_ptr<double[]> C(new double[5 * nn]);      // C[5n + i], i = 1..4
  std::unique_ptr<int8_t[]> split(new int8_t[5 * nn]);  // left share of the best split of i slots
  std::unique_ptr<int8_t[]> opened(new int8_t[5 * nn]); // C[n][i] opens n (else it equals C[n][i-1])
  std::unique_ptr<uint8_t[]> as_leaf(new uint8_t[nn]), contiguous(new uint8_t[nn]);
  std::unique_ptr<int32_t[]> cnt(new int32_t[nn]), first(new int32_t[nn]);
  for (const int32_t n : post) {
    const Bvh2Node &nd = N[n];
    const double dx = (double)nd.hi[0] - nd.lo[0], dy = (double)nd.hi[1] - nd.lo[1], dz = (double)nd.hi[2] - nd.lo[2];
    double area = dx * dy + dy * dz + dz * dx;
    if (!(area > 0.0) || !std::isfinite(area)) area = 0.0;
    double *c = &C[5 * (size_t)n];
    if (nd.left < 0) {
      cnt[n] = nd.count;
      first[n] = nd.first;
      as_leaf[n] = 1;
      contiguous[n] = 1;
      for (int i = 1; i <= 4; ++i) {
        c[i] = area * nd.count;
        opened[5 * (size_t)n + i] = 0;
      }
      continue;
    }
    const int32_t l = nd.left, r = nd.right;
    cnt[n] = cnt[l] + cnt[r];
    first[n] = std::min(first[l], first[r]);
    contiguous[n] = contiguous[l] && contiguous[r] &&
                    (first[r] == first[l] + cnt[l] || first[l] == first[r] + cnt[r]);
    const double *cl = &C[5 * (size_t)l], *cr = &C[5 * (size_t)r];
    double D[5] = {kInf, kInf, kInf, kInf, kInf};
    for (int j = 0; j < 5; ++j) split[5 * (size_t)n + j] = 1;
    for (int j = 2; j <= W; ++j)
      for (int k = 1; k < j; ++k)
        if (cl[k] + cr[j - k] < D[j]) {
          D[j] = cl[k] + cr[j - k];
          split[5 * (size_t)n + j] = (int8_t)k;
        }
    const double node_cost = area * kCollapseNodeCost + D[W];
    const double leaf_cost =
        contiguous[n] && cnt[n] <= merge_limit && cnt[n] <= kMaxLeafTris ? area * cnt[n] : kInf;
    as_leaf[n] = leaf_cost <= node_cost;
    c[1] = std::min(leaf_cost, node_cost);
    for (int i = 2; i <= 4; ++i) {
      opened[5 * (size_t)n + i] = i <= W && D[i] < c[i - 1];
      c[i] = i <= W ? std::min(D[i], c[i - 1]) : c[i - 1];
    }
  }
  // the binary nodes that take the slots of a wide node opened at n
  // (at most 4: a fixed array, filled depth first)
  struct Kids {
    int32_t id[4];
    int n = 0;
  };
  auto take = [&](int32_t n0, int i0, Kids &out) {
    int32_t sn[4];
    int si[4], sp = 0;
    sn[sp] = n0;
    si[sp++] = i0;
    while (sp > 0) {
      const int32_t n = sn[--sp];
      int i = si[sp];
      while (i > 1 && N[n].left >= 0 && !opened[5 * (size_t)n + i]) --i;
      if (i == 1 || N[n].left < 0) {
        out.id[out.n++] = n;
        continue;
      }
      const int k = split[5 * (size_t)n + i];
      sn[sp] = N[n].right;  // right below left: the left share comes out first
      si[sp++] = i - k;
      sn[sp] = N[n].left;
      si[sp++] = k;
    }
  };
  std::vector<std::pair<int32_t, int32_t>> work;  // (binary node, wide slot), breadth first
  wnodes.reserve(nn / 2 + 1);
  work.reserve(nn / 2 + 1);
  std::vector<int32_t> wdepth;
  wdepth.reserve(nn / 2 + 1);
  wnodes.emplace_back();
  wdepth.push_back(0);
  work.push_back({0, 0});
  for (size_t w = 0; w < work.size(); ++w) {
    const int32_t bi = work[w].first, wi = work[w].second;
    Kids kids;
    if (N[bi].left < 0) {
      kids.id[kids.n++] = bi;  // root is a leaf: one child slot
    } else {
      const int k = split[5 * (size_t)bi + W];
      take(N[bi].left, k, kids);
      take(N[bi].right, W - k, kids);
    }
    WideNode nd;
    std::memset(&nd, 0, sizeof(nd));
    for (int k = 0; k < 4; ++k) {
      if (k >= kids.n) {
        set_empty_slot(nd, k);
        continue;
      }
      const int32_t c = kids.id[k];
      nd.lox[k] = N[c].lo[0]; nd.loy[k] = N[c].lo[1]; nd.loz[k] = N[c].lo[2];
      nd.hix[k] = N[c].hi[0]; nd.hiy[k] = N[c].hi[1]; nd.hiz[k] = N[c].hi[2];
      if (N[c].left < 0 || as_leaf[c]) {
        if (cnt[c] == 0) {
          set_empty_slot(nd, k);
        } else {
          Bvh2Node leaf = N[c];
          leaf.left = leaf.right = -1;
          leaf.first = first[c];
          leaf.count = cnt[c];
          nd.child[k] = leaf_ref(leaf);
        }
      } else {
        const int32_t slot = (int32_t)wnodes.size();
        wnodes.emplace_back();
        wdepth.push_back(wdepth[wi] + 1);
        wide_depth = std::max(wide_depth, wdepth.back());
        nd.child[k] = node_base + (uint32_t)slot;
        work.push_back({c, slot});
      }
    }
    wnodes[wi] = nd;
  }
}

void collapse_wide(const gi_bvh &b, const gi_tri *tris, std::vector<WideNode> &wnodes,
                   std::vector<IsectTri> &leaf_tris, int32_t &wide_depth, int32_t width) {
  leaf_order_tris(b, tris, leaf_tris);
  collapse_wide_nodes(
      b, [](const Bvh2Node &n) { return kLeafBit | ((uint32_t)n.count << 27) | (uint32_t)n.first; }, 0u, wnodes,
      wide_depth, kCollapseMergeLimit, width);
}

// ------------------------------------------------------------- Halton ----
namespace {
struct PrimeTable {
  std::vector<int32_t> p;
  PrimeTable() {
    for (int32_t c = 2; (int)p.size() < 1024; ++c) {
      bool ok = true;
      for (int32_t q : p) {
        if (q * q > c) break;
        if (c % q == 0) {
          ok = false;
          break;
        }
      }
      if (ok) p.push_back(c);
    }
  }
};
const PrimeTable &prime_table() {
  static PrimeTable t;
  return t;
}
// PCG32 (O'Neill 2014): srandom / random / boundedrand.
struct Pcg {
  uint64_t state = 0, inc = 1;
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  Pcg(uint64_t initstate, uint64_t initseq) {
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
}  // namespace

int32_t prime_of(int32_t dim) { return prime_table().p.at(dim); }

void halton_permutation(int32_t dim, uint32_t seed, uint16_t *perm) {
  const int32_t p = prime_of(dim);
  for (int32_t i = 0; i < p; ++i) perm[i] = (uint16_t)i;
  Pcg r(((uint64_t)seed << 32) ^ (0x9E3779B97F4A7C15ULL * (uint64_t)(dim + 1)), (uint64_t)(dim + 1));
  for (int32_t i = p - 1; i >= 1; --i) std::swap(perm[i], perm[r.bounded((uint32_t)(i + 1))]);
}

void halton_tables(uint32_t seed, int32_t ndims, std::vector<HaltonDim> &dims, std::vector<uint16_t> &chunks) {
  dims.resize(ndims);
  chunks.clear();
  std::vector<uint16_t> perm;
  for (int32_t d = 0; d < ndims; ++d) {
    const uint32_t p = (uint32_t)prime_of(d);
    perm.resize(p);
    halton_permutation(d, seed, perm.data());
    uint32_t m = 1, P = p;
    while (P * p <= 1024u) {
      P *= p;
      ++m;
    }
    HaltonDim &h = dims[d];
    h.prime = p;
    h.chunk_off = (uint32_t)chunks.size();
    h.P = P;
    h.perm0 = perm[0];
    h.magic = (uint64_t)(~0ULL / P) + 1ULL;  // exact floor(n / P) = mulhi(n, magic) for n < 2^32
    h.kmax = 0;
    for (uint64_t pk = 1; pk < ((uint64_t)1 << 32); pk *= P) ++h.kmax;
    h.pad = 0;
    chunks.resize(chunks.size() + P);
    for (uint32_t c = 0; c < P; ++c) {
      uint32_t v = 0, x = c;
      for (uint32_t i = 0; i < m; ++i) {  // least significant digit first
        v = v * p + perm[x % p];
        x /= p;
      }
      chunks[h.chunk_off + c] = (uint16_t)v;
    }
  }
}

// Scrambled radical inverse as the exact rational A/B (see gi_device.h).
float halton_host(uint64_t n, int32_t dim, uint32_t seed) {
  const int32_t p = prime_of(dim);
  std::vector<uint16_t> perm(p);
  halton_permutation(dim, seed, perm.data());
  uint64_t rev = 0, pk = 1;
  while (n > 0) {
    uint64_t q = n / (uint64_t)p;
    rev = rev * (uint64_t)p + perm[n - q * (uint64_t)p];
    pk *= (uint64_t)p;
    n = q;
  }
  const uint64_t A = rev * (uint64_t)(p - 1) + perm[0];
  const uint64_t B = pk * (uint64_t)(p - 1);
  float f = (float)((double)A / (double)B);
  return f >= 1.0f ? 0x1.fffffep-1f : f;
}

}  // namespace gi
