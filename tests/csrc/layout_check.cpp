// The record checks of the integration binding, compiled against two
// stand-ins for rta's records: 4-byte-aligned vec2 texture coordinates (the
// records cast straight through) and an 8-byte-aligned float2 (simple_triangle
// padded to 104 bytes: copied field by field).  Prints the path taken and
// checks that the uploaded bytes equal gi_tri records built field by field.
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/gi_hip.h"

namespace four {
struct vec2f { float x, y; };
struct vec3f { float x, y, z; };
struct simple_triangle { vec3f a, b, c, na, nb, nc; vec2f ta, tb, tc; int material_index; };
struct triangle_intersection { float t, beta, gamma; unsigned int ref; };
}  // namespace four
namespace eight {
struct alignas(8) float2 { float x, y; };
struct vec3f { float x, y, z; };
struct simple_triangle { vec3f a, b, c, na, nb, nc; float2 ta, tb, tc; int material_index; };
struct triangle_intersection { float t, beta, gamma; unsigned int ref; };
}  // namespace eight

static std::vector<gi_tri> g_uploaded;
static const char *g_path = "";

template <class rta_tri, class rta_hit>
int upload_tris(const rta_tri *t, size_t n) {
  static_assert(offsetof(rta_tri, a) == offsetof(gi_tri, a) && offsetof(rta_tri, b) == offsetof(gi_tri, b) &&
                offsetof(rta_tri, c) == offsetof(gi_tri, c) && offsetof(rta_tri, na) == offsetof(gi_tri, na) &&
                offsetof(rta_tri, nb) == offsetof(gi_tri, nb) && offsetof(rta_tri, nc) == offsetof(gi_tri, nc) &&
                offsetof(rta_tri, ta) == offsetof(gi_tri, ta) && offsetof(rta_tri, tb) == offsetof(gi_tri, tb) &&
                offsetof(rta_tri, tc) == offsetof(gi_tri, tc) &&
                offsetof(rta_tri, material_index) == offsetof(gi_tri, material_index),
                "simple_triangle fields are not at gi_tri's offsets");
  static_assert(sizeof(rta_hit) == sizeof(gi_hit) && offsetof(rta_hit, t) == offsetof(gi_hit, t) &&
                offsetof(rta_hit, beta) == offsetof(gi_hit, beta) &&
                offsetof(rta_hit, gamma) == offsetof(gi_hit, gamma) && offsetof(rta_hit, ref) == offsetof(gi_hit, ref),
                "triangle_intersection is not gi_hit's layout");
  if constexpr (sizeof(rta_tri) == sizeof(gi_tri)) {
    g_path = "cast";
    const gi_tri *g = reinterpret_cast<const gi_tri *>(t);
    g_uploaded.assign(g, g + n);
  } else {
    g_path = "copy";
    g_uploaded.resize(n);
    for (size_t i = 0; i < n; ++i) std::memcpy(&g_uploaded[i], &t[i], sizeof(gi_tri));
  }
  return 0;
}

template <class T>
static bool check(const char *name, size_t expect_size, const char *expect_path) {
  T tris[3];
  std::memset(tris, 0xAB, sizeof(tris));  // padding bytes carry garbage
  for (int i = 0; i < 3; ++i) {
    float base = 10.0f * (float)i;
    tris[i].a = {base + 1, base + 2, base + 3};
    tris[i].b = {base + 4, base + 5, base + 6};
    tris[i].c = {base + 7, base + 8, base + 9};
    tris[i].na = {0, 1, 0};
    tris[i].nb = {0, 0, 1};
    tris[i].nc = {1, 0, 0};
    tris[i].ta = {0.25f, base};
    tris[i].tb = {0.5f, base};
    tris[i].tc = {0.75f, base};
    tris[i].material_index = 7 + i;
  }
  upload_tris<T, four::triangle_intersection>(tris, 3);
  bool ok = sizeof(T) == expect_size && std::strcmp(g_path, expect_path) == 0 && g_uploaded.size() == 3;
  for (int i = 0; i < 3 && ok; ++i) {
    const gi_tri &g = g_uploaded[(size_t)i];
    float base = 10.0f * (float)i;
    ok = g.a[0] == base + 1 && g.c[2] == base + 9 && g.nc[0] == 1 && g.ta[0] == 0.25f && g.tc[1] == base &&
         g.material_index == 7 + i;
  }
  std::printf("%s: sizeof %zu, %s, %s\n", name, sizeof(T), g_path, ok ? "ok" : "FAIL");
  return ok;
}

int main() {
  static_assert(sizeof(gi_tri) == 100 && sizeof(gi_hit) == 16, "boundary records");
  bool ok = check<four::simple_triangle>("vec2f", 100, "cast");
  ok = check<eight::simple_triangle>("float2", 104, "copy") && ok;
  return ok ? 0 : 1;
}
