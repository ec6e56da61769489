// Host-only check of the binary -> 4-wide collapse (collapse_wide_nodes):
// every primitive sits in exactly one leaf, every child box contains the
// triangles below it, references and the reported depth are consistent,
// merged leaves stay within the merge limit, and no node fills more child
// slots than the width.  Trees come from the SAH
// builder and from a median-split builder with one triangle per leaf (so
// the collapse's leaf merging is exercised).  Prints one line per case and
// exits non-zero on the first violation.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <numeric>
#include <random>
#include <vector>

#include "gi_internal.h"

using namespace gi;

static std::vector<gi_tri> make_tris(int n, unsigned seed) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> U(0, 100), S(0.01f, 1.5f);
  std::vector<gi_tri> t(n);
  for (int i = 0; i < n; ++i) {
    std::memset(&t[i], 0, sizeof(gi_tri));
    const float cx = U(rng), cy = U(rng) * (i % 3 == 0 ? 0.05f : 1.0f), cz = U(rng), s = S(rng);
    t[i].a[0] = cx; t[i].a[1] = cy; t[i].a[2] = cz;
    t[i].b[0] = cx + s; t[i].b[1] = cy + 0.3f * s; t[i].b[2] = cz;
    t[i].c[0] = cx; t[i].c[1] = cy + s; t[i].c[2] = cz + 0.5f * s;
    if (i % 97 == 0 && i > 0) t[i] = t[i - 1];  // duplicate triangles
  }
  return t;
}

// median split on the longest centroid axis down to one triangle per leaf
static void median_tree(const std::vector<gi_tri> &t, gi_bvh &b) {
  const int n = (int)t.size();
  b.perm.resize(n);
  std::iota(b.perm.begin(), b.perm.end(), 0u);
  b.nodes.assign(1, Bvh2Node{});
  std::function<void(int, int, int)> rec = [&](int node, int lo, int hi) {
    float bl[3] = {1e30f, 1e30f, 1e30f}, bh[3] = {-1e30f, -1e30f, -1e30f};
    float cl[3] = {1e30f, 1e30f, 1e30f}, ch[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = lo; i < hi; ++i) {
      const gi_tri &x = t[b.perm[i]];
      for (int a = 0; a < 3; ++a) {
        bl[a] = std::min(bl[a], std::min(x.a[a], std::min(x.b[a], x.c[a])));
        bh[a] = std::max(bh[a], std::max(x.a[a], std::max(x.b[a], x.c[a])));
        const float c = (x.a[a] + x.b[a] + x.c[a]) / 3.0f;
        cl[a] = std::min(cl[a], c);
        ch[a] = std::max(ch[a], c);
      }
    }
    Bvh2Node &nd = b.nodes[node];
    std::memcpy(nd.lo, bl, 12);
    std::memcpy(nd.hi, bh, 12);
    pad_box(nd.lo, nd.hi);
    if (hi - lo == 1) {
      nd.left = nd.right = -1;
      nd.first = lo;
      nd.count = 1;
      return;
    }
    int ax = 0;
    for (int a = 1; a < 3; ++a)
      if (ch[a] - cl[a] > ch[ax] - cl[ax]) ax = a;
    const int mid = (lo + hi) / 2;
    std::nth_element(b.perm.begin() + lo, b.perm.begin() + mid, b.perm.begin() + hi, [&](uint32_t p, uint32_t q) {
      return t[p].a[ax] + t[p].b[ax] + t[p].c[ax] < t[q].a[ax] + t[q].b[ax] + t[q].c[ax];
    });
    const int l = (int)b.nodes.size();
    b.nodes.push_back(Bvh2Node{});
    b.nodes.push_back(Bvh2Node{});
    b.nodes[node].left = l;
    b.nodes[node].right = l + 1;
    b.nodes[node].first = b.nodes[node].count = 0;
    rec(l, lo, mid);
    rec(l + 1, mid, hi);
  };
  rec(0, 0, n);
}

static int check(const char *what, const std::vector<gi_tri> &t, const gi_bvh &b, int merge_limit, int width) {
  const int n = (int)t.size();
  std::vector<WideNode> w;
  int32_t depth = -1;
  collapse_wide_nodes(
      b, [](const Bvh2Node &nd) { return kLeafBit | ((uint32_t)nd.count << 27) | (uint32_t)nd.first; }, 0u, w, depth,
      merge_limit, width);
  for (const WideNode &nd : w)
    for (int k = width; k < 4; ++k)
      if (nd.child[k] != kEmptyRef) {
        std::printf("FAIL %s: slot %d used at width %d\n", what, k, width);
        return 1;
      }
  std::vector<int> seen(n, 0);
  int maxd = 0, leaves = 0, big = 0;
  bool ok = true;
  // returns the primitives below wide node m (leaf-order positions)
  std::function<void(uint32_t, int, std::vector<int> &)> walk = [&](uint32_t m, int d, std::vector<int> &below) {
    if (m >= w.size()) {
      std::printf("FAIL %s: node reference %u of %zu\n", what, m, w.size());
      ok = false;
      return;
    }
    maxd = std::max(maxd, d);
    const WideNode &nd = w[m];
    for (int k = 0; k < 4 && ok; ++k) {
      const uint32_t c = nd.child[k];
      if (c == kEmptyRef) continue;
      std::vector<int> sub;
      if (c & kLeafBit) {
        const uint32_t cnt = (c >> 27) & 15u, first = c & kMaxLeafFirst;
        if (cnt == 0 || first + cnt > (uint32_t)n) {
          std::printf("FAIL %s: leaf %u+%u\n", what, first, cnt);
          ok = false;
          return;
        }
        ++leaves;
        if ((int)cnt > std::max(merge_limit, 4)) big++;
        for (uint32_t j = 0; j < cnt; ++j) sub.push_back((int)(first + j));
      } else {
        walk(c, d + 1, sub);
      }
      const float lo[3] = {nd.lox[k], nd.loy[k], nd.loz[k]}, hi[3] = {nd.hix[k], nd.hiy[k], nd.hiz[k]};
      for (int p : sub) {
        const gi_tri &x = t[b.perm[p]];
        const float *v[3] = {x.a, x.b, x.c};
        for (int q = 0; q < 3; ++q)
          for (int a = 0; a < 3; ++a)
            if (!(v[q][a] >= lo[a] && v[q][a] <= hi[a])) {
              std::printf("FAIL %s: triangle %u outside its box\n", what, b.perm[p]);
              ok = false;
              return;
            }
      }
      below.insert(below.end(), sub.begin(), sub.end());
    }
  };
  std::vector<int> all;
  walk(0, 0, all);
  if (!ok) return 1;
  for (int p : all) seen[p]++;
  for (int i = 0; i < n; ++i)
    if (seen[i] != 1) {
      std::printf("FAIL %s: primitive %d in %d leaves\n", what, i, seen[i]);
      return 1;
    }
  if (maxd != depth) {
    std::printf("FAIL %s: depth %d reported %d\n", what, maxd, depth);
    return 1;
  }
  if (big) {
    std::printf("FAIL %s: %d leaves above the merge limit\n", what, big);
    return 1;
  }
  std::printf("ok %s n=%d merge=%d width=%d wide=%zu leaves=%d depth=%d\n", what, n, merge_limit, width, w.size(),
              leaves, depth);
  return 0;
}

// The threaded SAH build gives the same nodes and permutation as the
// sequential one.
static int check_parallel(int n) {
  const std::vector<gi_tri> t = make_tris(n, 7u);
  gi_bvh a, b;
  auto t0 = std::chrono::steady_clock::now();
  build_bvh2(t.data(), n, 4, a, 1);
  auto t1 = std::chrono::steady_clock::now();
  build_bvh2(t.data(), n, 4, b, 0);
  auto t2 = std::chrono::steady_clock::now();
  if (a.nodes.size() != b.nodes.size() || a.perm != b.perm || a.max_depth != b.max_depth ||
      std::memcmp(a.nodes.data(), b.nodes.data(), a.nodes.size() * sizeof(Bvh2Node)) != 0) {
    std::printf("FAIL parallel build n=%d differs\n", n);
    return 1;
  }
  std::printf("ok parallel n=%d nodes=%zu sequential_ms=%.1f threaded_ms=%.1f\n", n, a.nodes.size(),
              std::chrono::duration<double, std::milli>(t1 - t0).count(),
              std::chrono::duration<double, std::milli>(t2 - t1).count());
  return 0;
}

int main() {
  if (check_parallel(300000)) return 1;
  for (int n : {1, 2, 3, 5, 17, 1000, 20000}) {
    const std::vector<gi_tri> t = make_tris(n, (unsigned)n);
    gi_bvh sah;
    build_bvh2(t.data(), n, 4, sah);
    gi_bvh med;
    median_tree(t, med);
    const int cases[5][2] = {{1, 4}, {4, 4}, {8, 4}, {4, 2}, {4, 3}};  // (merge limit, width)
    for (const auto &c : cases) {
      if (check("sah", t, sah, c[0], c[1])) return 1;
      if (check("median", t, med, c[0], c[1])) return 1;
    }
  }
  return 0;
}
