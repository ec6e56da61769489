// Scene ingestion: Wavefront OBJ / MTL into the triangle and material
// tables, and the .pbrdf principled-parameter files.
//
// Reference behaviour being followed:
//   add_objfile_to_flat_tri_list (src/hl-main.cpp:277-350): triangles in
//     group order with positions, vertex normals and material index; the
//     texture coordinates are kept only when the material has a diffuse
//     texture, else (0,0); a fallback material "gi/fallback" with diffuse
//     and specular (1,0,0) is registered before the first file's own
//     materials; a group without a material uses the fallback; a material
//     is registered once per name (first use), with diffuse Kd, specular
//     Ks and alpha d from the MTL file.
//   convert_and_upload_materials (src/material-conversion.cpp:59-151): a
//     material whose name has a file <dir>/<name>.pbrdf gets principled
//     parameters from it; others stay Lambert.
//   PrincipledBRDFParameters::readFrom (src/principledParameters.h:64-125):
//     whitespace-separated "identifier value..." pairs, unknown tokens
//     skipped (so the reference files' "transmission0.0" sets nothing).
//
// The OBJ reader itself is an external library of the reference; its
// conventions are this file's own and stated here:
//   - vertex references i, i/j, i//k, i/j/k, 1-based, negative = relative to
//     the end of the list so far;
//   - polygons are split into a fan (v0, v[i], v[i+1]);
//   - a face without normals gets its geometric normal normalize((b-a)x(c-a))
//     at all three corners;
//   - an optional 4x4 row-major transform M maps positions as M (x,y,z,1)
//     (divided by the resulting w) and normals by the inverse transpose of
//     M's upper 3x3, renormalised;
//   - "usemtl" of a name no MTL file defines uses the fallback material;
//   - MTL keys used: newmtl, Kd, Ks, d, Tr (= 1 - d), map_Kd; the rest are
//     ignored.  map_Kd's file (its last token, relative to the MTL file's
//     directory) joins the list of the scene's textures in order of first
//     use, and the material's gi_material.texture is its index there (the
//     image is decoded by gi_image_read, the texture uploaded by the
//     caller, as convert_texture does, src/material-conversion.cpp:18-41).
// Unlike the reference (which loops forever on a malformed number in a
// .pbrdf file and leaves unset principled fields uninitialised) a parse
// error is reported, and unset fields take PrincipledBRDFParameters'
// default-constructor values (opacity 1, transmission 1, others 0) with
// color defaulting to the material's diffuse colour.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gi_hip.h"

struct gi_scene_file {
  std::vector<gi_tri> tris;
  std::vector<gi_material> mats;
  std::vector<std::string> names;
  std::vector<bool> has_diffuse_texture;
  std::vector<std::string> textures;     // distinct map_Kd files (resolved), gi_material.texture indexes them
  std::map<std::string, int32_t> index;  // registered material name -> index
  std::string error;
};

namespace {

struct MtlEntry {
  float kd[3] = {0.f, 0.f, 0.f}, ks[3] = {0.f, 0.f, 0.f};
  float alpha = 1.f;
  bool tex_d = false;
  std::string tex_path;  // map_Kd file, resolved against the MTL's directory
};

gi_material lambert_material(const float kd[3], const float ks[3], float alpha) {
  gi_material m;
  std::memset(&m, 0, sizeof(m));
  for (int i = 0; i < 3; ++i) {
    m.diffuse_color[i] = kd[i];
    m.specular_color[i] = ks[i];
    m.color[i] = kd[i];
  }
  m.alpha = alpha;
  m.principled = 0;
  m.opacity = 1.f;
  m.transmission = 1.f;
  m.texture = -1;
  return m;
}

std::string dir_of(const std::string &path) {
  const size_t s = path.find_last_of('/');
  return s == std::string::npos ? std::string(".") : path.substr(0, s);
}

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) ++a;
  while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) --b;
  return s.substr(a, b - a);
}

bool parse_floats(std::istringstream &in, float *v, int n) {
  for (int i = 0; i < n; ++i)
    if (!(in >> v[i])) return false;
  return true;
}

bool read_mtl(const std::string &path, std::map<std::string, MtlEntry> &out, std::string &err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open material library " + path;
    return false;
  }
  std::string line, cur;
  int ln = 0;
  while (std::getline(f, line)) {
    ++ln;
    std::istringstream in(line);
    std::string key;
    if (!(in >> key) || key[0] == '#') continue;
    if (key == "newmtl") {
      std::string rest;
      std::getline(in, rest);
      cur = trim(rest);
      out[cur] = MtlEntry{};
      continue;
    }
    if (cur.empty()) continue;
    MtlEntry &m = out[cur];
    bool ok = true;
    if (key == "Kd") ok = parse_floats(in, m.kd, 3);
    else if (key == "Ks") ok = parse_floats(in, m.ks, 3);
    else if (key == "d") ok = parse_floats(in, &m.alpha, 1);
    else if (key == "Tr") {
      float tr;
      ok = parse_floats(in, &tr, 1);
      if (ok) m.alpha = 1.f - tr;
    } else if (key == "map_Kd") {
      // the file is the last token (options such as -bm come before it);
      // relative paths are relative to the MTL file
      m.tex_d = true;
      std::string tok, last;
      while (in >> tok) last = tok;
      if (!last.empty()) m.tex_path = (last[0] == '/') ? last : dir_of(path) + "/" + last;
    }
    if (!ok) {
      err = path + ":" + std::to_string(ln) + ": bad number after " + key;
      return false;
    }
  }
  return true;
}

// One corner of a face: indices into v / vt / vn (-1 = absent).
struct Corner {
  long v, t, n;
};

bool parse_ref(const std::string &tok, long nv, long nt, long nn, Corner &c) {
  long idx[3] = {0, 0, 0};
  bool have[3] = {false, false, false};
  size_t p = 0;
  for (int part = 0; part < 3 && p <= tok.size(); ++part) {
    size_t q = tok.find('/', p);
    std::string s = tok.substr(p, q == std::string::npos ? std::string::npos : q - p);
    if (!s.empty()) {
      char *end = nullptr;
      errno = 0;
      long x = std::strtol(s.c_str(), &end, 10);
      if (errno || *end != '\0' || x == 0) return false;
      idx[part] = x;
      have[part] = true;
    }
    if (q == std::string::npos) break;
    p = q + 1;
  }
  if (!have[0]) return false;
  const long count[3] = {nv, nt, nn};
  long out[3] = {-1, -1, -1};
  for (int k = 0; k < 3; ++k) {
    if (!have[k]) continue;
    const long x = idx[k] > 0 ? idx[k] - 1 : count[k] + idx[k];
    if (x < 0 || x >= count[k]) return false;
    out[k] = x;
  }
  c = Corner{out[0], out[1], out[2]};
  return true;
}

struct V3 {
  double x, y, z;
};

V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

int32_t register_material(gi_scene_file *f, const std::string &name, const MtlEntry *e) {
  auto it = f->index.find(name);
  if (it != f->index.end()) return it->second;
  const float red[3] = {1.f, 0.f, 0.f};
  const int32_t id = (int32_t)f->mats.size();
  f->mats.push_back(e ? lambert_material(e->kd, e->ks, e->alpha) : lambert_material(red, red, 1.f));
  if (e && e->tex_d && !e->tex_path.empty()) {
    int32_t t = 0;
    while (t < (int32_t)f->textures.size() && f->textures[(size_t)t] != e->tex_path) ++t;
    if (t == (int32_t)f->textures.size()) f->textures.push_back(e->tex_path);
    f->mats.back().texture = t;
  }
  f->names.push_back(name);
  f->has_diffuse_texture.push_back(e ? e->tex_d : false);
  f->index[name] = id;
  return id;
}

int fail(gi_scene_file *f, int code, const std::string &msg) {
  if (f) f->error = msg;
  return code;
}

}  // namespace

extern "C" {

int gi_obj_open(gi_scene_file **out) {
  if (!out) return GI_ERR_INVALID_ARG;
  *out = new (std::nothrow) gi_scene_file();
  return *out ? GI_OK : GI_ERR_OUT_OF_MEMORY;
}

void gi_obj_free(gi_scene_file *f) { delete f; }

const char *gi_obj_error(const gi_scene_file *f) { return f ? f->error.c_str() : "null scene file"; }

int gi_obj_append(gi_scene_file *f, const char *path, const float *trafo) {
  if (!f || !path) return GI_ERR_INVALID_ARG;
  std::ifstream in(path);
  if (!in) return fail(f, GI_ERR_INVALID_ARG, std::string("cannot open ") + path);
  // the fallback material comes first, registered with the first file
  if (f->mats.empty()) register_material(f, "gi/fallback", nullptr);
  const std::string dir = dir_of(path);
  std::map<std::string, MtlEntry> mtl;
  std::vector<V3> v, vn;
  std::vector<float> vt;  // u, v pairs
  int32_t cur = f->index["gi/fallback"];
  bool cur_tex = false;
  // normal matrix: inverse transpose of the upper 3x3 of trafo
  double nm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (trafo) {
    const double a = trafo[0], b = trafo[1], c = trafo[2], d = trafo[4], e = trafo[5], g = trafo[6], h = trafo[8],
                 i = trafo[9], j = trafo[10];
    const double det = a * (e * j - g * i) - b * (d * j - g * h) + c * (d * i - e * h);
    if (det == 0.0) return fail(f, GI_ERR_INVALID_ARG, "singular transform");
    // inverse transpose = cofactor matrix / det
    nm[0] = (e * j - g * i) / det;
    nm[1] = -(d * j - g * h) / det;
    nm[2] = (d * i - e * h) / det;
    nm[3] = -(b * j - c * i) / det;
    nm[4] = (a * j - c * h) / det;
    nm[5] = -(a * i - b * h) / det;
    nm[6] = (b * g - c * e) / det;
    nm[7] = -(a * g - c * d) / det;
    nm[8] = (a * e - b * d) / det;
  }
  std::string line;
  long ln = 0;
  std::vector<Corner> poly;
  while (std::getline(in, line)) {
    ++ln;
    std::istringstream ls(line);
    std::string key;
    if (!(ls >> key) || key[0] == '#') continue;
    auto bad = [&](const std::string &what) {
      return fail(f, GI_ERR_INVALID_ARG, std::string(path) + ":" + std::to_string(ln) + ": " + what);
    };
    if (key == "v") {
      double p[3];
      if (!(ls >> p[0] >> p[1] >> p[2])) return bad("bad vertex");
      if (trafo) {
        const double x = trafo[0] * p[0] + trafo[1] * p[1] + trafo[2] * p[2] + trafo[3];
        const double y = trafo[4] * p[0] + trafo[5] * p[1] + trafo[6] * p[2] + trafo[7];
        const double z = trafo[8] * p[0] + trafo[9] * p[1] + trafo[10] * p[2] + trafo[11];
        const double w = trafo[12] * p[0] + trafo[13] * p[1] + trafo[14] * p[2] + trafo[15];
        if (w == 0.0) return bad("transform maps a vertex to infinity");
        v.push_back(V3{x / w, y / w, z / w});
      } else {
        v.push_back(V3{p[0], p[1], p[2]});
      }
    } else if (key == "vn") {
      double p[3];
      if (!(ls >> p[0] >> p[1] >> p[2])) return bad("bad normal");
      if (trafo) {
        const double x = nm[0] * p[0] + nm[1] * p[1] + nm[2] * p[2];
        const double y = nm[3] * p[0] + nm[4] * p[1] + nm[5] * p[2];
        const double z = nm[6] * p[0] + nm[7] * p[1] + nm[8] * p[2];
        const double l = std::sqrt(x * x + y * y + z * z);
        vn.push_back(l > 0 ? V3{x / l, y / l, z / l} : V3{x, y, z});
      } else {
        vn.push_back(V3{p[0], p[1], p[2]});
      }
    } else if (key == "vt") {
      float uv[2];
      if (!(ls >> uv[0] >> uv[1])) return bad("bad texture coordinate");
      vt.push_back(uv[0]);
      vt.push_back(uv[1]);
    } else if (key == "mtllib") {
      std::string rest;
      std::getline(ls, rest);
      std::istringstream names(rest);
      std::string lib;
      while (names >> lib) {
        std::string err;
        if (!read_mtl(dir + "/" + lib, mtl, err)) return fail(f, GI_ERR_INVALID_ARG, err);
      }
    } else if (key == "usemtl") {
      std::string rest;
      std::getline(ls, rest);
      const std::string name = trim(rest);
      auto it = mtl.find(name);
      if (it == mtl.end()) {
        cur = f->index["gi/fallback"];
        cur_tex = false;
      } else {
        cur = register_material(f, name, &it->second);
        cur_tex = f->has_diffuse_texture[cur];
      }
    } else if (key == "f") {
      poly.clear();
      std::string tok;
      while (ls >> tok) {
        Corner c;
        if (!parse_ref(tok, (long)v.size(), (long)vt.size() / 2, (long)vn.size(), c)) return bad("bad face reference " + tok);
        poly.push_back(c);
      }
      if (poly.size() < 3) return bad("face with fewer than 3 vertices");
      for (size_t k = 1; k + 1 < poly.size(); ++k) {
        const Corner cs[3] = {poly[0], poly[k], poly[k + 1]};
        gi_tri t;
        std::memset(&t, 0, sizeof(t));
        float *P[3] = {t.a, t.b, t.c}, *N[3] = {t.na, t.nb, t.nc}, *T[3] = {t.ta, t.tb, t.tc};
        for (int q = 0; q < 3; ++q) {
          const V3 &p = v[(size_t)cs[q].v];
          P[q][0] = (float)p.x;
          P[q][1] = (float)p.y;
          P[q][2] = (float)p.z;
        }
        const bool have_n = cs[0].n >= 0 && cs[1].n >= 0 && cs[2].n >= 0;
        if (have_n) {
          for (int q = 0; q < 3; ++q) {
            const V3 &n = vn[(size_t)cs[q].n];
            N[q][0] = (float)n.x;
            N[q][1] = (float)n.y;
            N[q][2] = (float)n.z;
          }
        } else {
          const V3 a{P[0][0], P[0][1], P[0][2]}, b{P[1][0], P[1][1], P[1][2]}, c{P[2][0], P[2][1], P[2][2]};
          V3 g = cross(V3{b.x - a.x, b.y - a.y, b.z - a.z}, V3{c.x - a.x, c.y - a.y, c.z - a.z});
          const double l = std::sqrt(g.x * g.x + g.y * g.y + g.z * g.z);
          if (l > 0) g = V3{g.x / l, g.y / l, g.z / l};
          for (int q = 0; q < 3; ++q) {
            N[q][0] = (float)g.x;
            N[q][1] = (float)g.y;
            N[q][2] = (float)g.z;
          }
        }
        if (cur_tex) {
          for (int q = 0; q < 3; ++q)
            if (cs[q].t >= 0) {
              T[q][0] = vt[2 * (size_t)cs[q].t];
              T[q][1] = vt[2 * (size_t)cs[q].t + 1];
            }
        }
        t.material_index = cur;
        f->tris.push_back(t);
      }
    }
    // g, o, s, l, p and anything else: no geometry of a triangle table
  }
  return GI_OK;
}

int gi_pbrdf_read(const char *path, gi_material *m) {
  if (!path || !m) return GI_ERR_INVALID_ARG;
  std::ifstream in(path);
  if (!in) return GI_ERR_INVALID_ARG;
  gi_material r = *m;
  r.principled = 1;
  r.metallic = r.subsurface = r.specular = r.roughness = 0.f;
  r.specular_tint = r.anisotropic = r.sheen = r.sheen_tint = 0.f;
  r.clearcoat = r.clearcoat_gloss = 0.f;
  r.opacity = 1.f;
  r.transmission = 1.f;
  for (int i = 0; i < 3; ++i) r.color[i] = m->diffuse_color[i];
  struct Key {
    const char *name;
    float *dst;
    int n;
  } keys[] = {{"color", r.color, 3},
              {"metallic", &r.metallic, 1},
              {"opacity", &r.opacity, 1},
              {"subsurface", &r.subsurface, 1},
              {"specular", &r.specular, 1},
              {"roughness", &r.roughness, 1},
              {"tintSpecular", &r.specular_tint, 1},
              {"anisotropic", &r.anisotropic, 1},
              {"sheen", &r.sheen, 1},
              {"tintSheen", &r.sheen_tint, 1},
              {"clearcoat", &r.clearcoat, 1},
              {"glossClearcoat", &r.clearcoat_gloss, 1},
              {"transmission", &r.transmission, 1}};
  std::string tok;
  while (in >> tok) {
    for (const Key &k : keys) {
      if (tok != k.name) continue;
      for (int i = 0; i < k.n; ++i)
        if (!(in >> k.dst[i])) return GI_ERR_INVALID_ARG;
      break;
    }
  }
  *m = r;
  return GI_OK;
}

int gi_obj_apply_pbrdf(gi_scene_file *f, const char *dir) {
  if (!f || !dir) return GI_ERR_INVALID_ARG;
  for (size_t i = 0; i < f->mats.size(); ++i) {
    const std::string p = std::string(dir) + "/" + f->names[i] + ".pbrdf";
    std::ifstream probe(p);
    if (!probe) continue;
    probe.close();
    if (gi_pbrdf_read(p.c_str(), &f->mats[i]) != GI_OK) return fail(f, GI_ERR_INVALID_ARG, "bad number in " + p);
  }
  return GI_OK;
}

int32_t gi_obj_texture_count(const gi_scene_file *f) { return f ? (int32_t)f->textures.size() : 0; }

const char *gi_obj_texture_path(const gi_scene_file *f, int32_t i) {
  if (!f || i < 0 || i >= (int32_t)f->textures.size()) return nullptr;
  return f->textures[(size_t)i].c_str();
}

int gi_obj_sizes(const gi_scene_file *f, int64_t *ntri, int32_t *nmat) {
  if (!f || !ntri || !nmat) return GI_ERR_INVALID_ARG;
  *ntri = (int64_t)f->tris.size();
  *nmat = (int32_t)f->mats.size();
  return GI_OK;
}

int gi_obj_export(const gi_scene_file *f, gi_tri *tris, gi_material *mats) {
  if (!f) return GI_ERR_INVALID_ARG;
  if (tris && !f->tris.empty()) std::memcpy(tris, f->tris.data(), f->tris.size() * sizeof(gi_tri));
  if (mats && !f->mats.empty()) std::memcpy(mats, f->mats.data(), f->mats.size() * sizeof(gi_material));
  return GI_OK;
}

const char *gi_obj_material_name(const gi_scene_file *f, int32_t i) {
  if (!f || i < 0 || (size_t)i >= f->names.size()) return nullptr;
  return f->names[(size_t)i].c_str();
}

}  // extern "C"
