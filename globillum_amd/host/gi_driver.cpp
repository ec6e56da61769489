// Thin headless driver, shaped like the reference's main loop
// (src/hl-main.cpp:646-711): load a scene, register and select the
// integrator, compute(), then update() while in_progress(), and write the
// image.  Scenes come as a binary dump of C-ABI records (written by
// globillum_amd.scenes.save_scene):
//   "GISC" u32 version (1 or 2), u64 n_tris, u32 n_mats, u32 n_lights,
//   u32 n_views, gi_tri[n_tris], gi_material[n_mats], gi_light[n_lights],
//   gi_view[n_views], gi_render_opts
// and in version 2 tagged sections until "END ":
//   "SKY " u32 w, u32 h, float rgb[3*w*h]      the sky light's map
//   "TEX " u32 w, u32 h, u8 rgba[4*w*h]        a texture (in id order)
//   "MESH" u64 n, gi_tri[n]                    a mesh (in id order)
//   "INST" u32 n, gi_instance[n]               the instances
// Output by extension: .png / .exr as the reference's save_image writes them
// (gi_write_png / gi_write_exr), anything else PFM (3 floats per pixel,
// bottom row first, as the renderer's row order y*w+x already is).
//
//   gi_driver scene.gisc out.{pfm,png,exr} [--view i] [--spp n] [--depth d]
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "mi355x_pt.h"

namespace {

struct SceneFile {
  std::vector<gi_tri> tris;
  std::vector<gi_material> mats;
  std::vector<gi_light> lights;
  std::vector<gi_view> views;
  gi_render_opts opts{};
  std::vector<float> sky;
  uint32_t sky_w = 0, sky_h = 0;
  std::vector<std::vector<uint8_t>> tex;
  std::vector<std::pair<uint32_t, uint32_t>> tex_wh;
  std::vector<std::vector<gi_tri>> meshes;
  std::vector<gi_instance> inst;
};

template <class T> void read_n(std::ifstream &f, std::vector<T> &v, uint64_t n) {
  v.resize(n);
  if (n) f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(T)));
}

SceneFile load(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  char magic[4];
  uint32_t version, nm, nl, nv;
  uint64_t nt;
  f.read(magic, 4);
  f.read(reinterpret_cast<char *>(&version), 4);
  if (std::memcmp(magic, "GISC", 4) != 0 || (version != 1 && version != 2))
    throw std::runtime_error(path + ": not a GISC v1/v2 file");
  f.read(reinterpret_cast<char *>(&nt), 8);
  f.read(reinterpret_cast<char *>(&nm), 4);
  f.read(reinterpret_cast<char *>(&nl), 4);
  f.read(reinterpret_cast<char *>(&nv), 4);
  SceneFile s;
  read_n(f, s.tris, nt);
  read_n(f, s.mats, nm);
  read_n(f, s.lights, nl);
  read_n(f, s.views, nv);
  f.read(reinterpret_cast<char *>(&s.opts), sizeof(s.opts));
  while (version == 2 && f) {
    char tag[4];
    f.read(tag, 4);
    if (!f || !std::memcmp(tag, "END ", 4)) break;
    uint32_t a = 0, b = 0;
    uint64_t n = 0;
    if (!std::memcmp(tag, "SKY ", 4) || !std::memcmp(tag, "TEX ", 4)) {
      f.read(reinterpret_cast<char *>(&a), 4);
      f.read(reinterpret_cast<char *>(&b), 4);
      if ((uint64_t)a * b > ((uint64_t)1 << 28)) throw std::runtime_error(path + ": image too large");
      if (tag[0] == 'S') {
        s.sky_w = a;
        s.sky_h = b;
        read_n(f, s.sky, 3 * (uint64_t)a * b);
      } else {
        s.tex.emplace_back();
        s.tex_wh.emplace_back(a, b);
        read_n(f, s.tex.back(), 4 * (uint64_t)a * b);
      }
    } else if (!std::memcmp(tag, "MESH", 4)) {
      f.read(reinterpret_cast<char *>(&n), 8);
      if (n > ((uint64_t)1 << 27)) throw std::runtime_error(path + ": mesh too large");
      s.meshes.emplace_back();
      read_n(f, s.meshes.back(), n);
    } else if (!std::memcmp(tag, "INST", 4)) {
      f.read(reinterpret_cast<char *>(&a), 4);
      read_n(f, s.inst, a);
    } else {
      throw std::runtime_error(path + ": unknown section");
    }
  }
  if (!f) throw std::runtime_error(path + ": truncated");
  return s;
}

void write_pfm(const std::string &path, const std::vector<float> &rgb, int w, int h) {
  std::FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fprintf(f, "PF\n%d %d\n-1.0\n", w, h);
  std::fwrite(rgb.data(), sizeof(float), rgb.size(), f);
  std::fclose(f);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    std::cerr << "usage: gi_driver scene.gisc out.pfm [--view i] [--spp n] [--depth d]\n";
    return 2;
  }
  int view = 0, spp = -1, depth = -1;
  for (int i = 3; i + 1 < argc; i += 2) {
    std::string a = argv[i];
    if (a == "--view") view = std::atoi(argv[i + 1]);
    else if (a == "--spp") spp = std::atoi(argv[i + 1]);
    else if (a == "--depth") depth = std::atoi(argv[i + 1]);
    else {
      std::cerr << "unknown option " << a << "\n";
      return 2;
    }
  }
  try {
    SceneFile s = load(argv[1]);
    if (view < 0 || view >= (int)s.views.size()) throw std::runtime_error("view index out of range");
    const gi_view &v = s.views[view];
    gi::scene_desc sd;
    sd.tris = s.tris.data();
    sd.n_tris = (int64_t)s.tris.size();
    sd.mats = s.mats.data();
    sd.n_mats = (int32_t)s.mats.size();
    sd.lights = s.lights.data();
    sd.n_lights = (int32_t)s.lights.size();
    if (!s.sky.empty()) {
      sd.sky_rgb = s.sky.data();
      sd.sky_w = (int32_t)s.sky_w;
      sd.sky_h = (int32_t)s.sky_h;
    }
    std::vector<gi::scene_desc::image> images;
    for (size_t i = 0; i < s.tex.size(); ++i)
      images.push_back({s.tex[i].data(), (int32_t)s.tex_wh[i].first, (int32_t)s.tex_wh[i].second});
    sd.textures = images.data();
    sd.n_textures = (int32_t)images.size();
    std::vector<gi::scene_desc::mesh> meshes;
    for (const auto &m : s.meshes) meshes.push_back({m.data(), (int64_t)m.size()});
    sd.meshes = meshes.data();
    sd.n_meshes = (int32_t)meshes.size();
    sd.instances = s.inst.data();
    sd.n_instances = (int32_t)s.inst.size();
    gi::gi_algorithm::original_scene = &sd;

    gi::mi355x_pt pt(v.w, v.h);
    gi::gi_algorithm::select("mi355x_pt");  // activates: upload + BVH
    gi::camera cam;
    std::memcpy(cam.pos, v.pos, sizeof cam.pos);
    std::memcpy(cam.dir, v.dir, sizeof cam.dir);
    std::memcpy(cam.up, v.up, sizeof cam.up);
    cam.fovy = v.fovy;
    pt.set_camera(cam);
    pt.path_samples(spp > 0 ? spp : v.spp);
    pt.path_length(depth > 0 ? depth : s.opts.max_depth);
    pt.russian_roulette(s.opts.rr_start);
    pt.sampler_seed(s.opts.seed);

    auto t0 = std::chrono::steady_clock::now();
    gi::gi_algorithm::selected->compute();
    int steps = 0;
    while (gi::gi_algorithm::selected->in_progress()) {
      gi::gi_algorithm::selected->update();
      ++steps;
    }
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const std::string out = argv[2];
    auto ends = [&](const char *ext) {
      const size_t n = std::strlen(ext);
      return out.size() >= n && out.compare(out.size() - n, n, ext) == 0;
    };
    if (ends(".png") || ends(".exr")) {
      const int rc = ends(".png") ? gi_write_png(out.c_str(), pt.result().data(), v.w, v.h)
                                  : gi_write_exr(out.c_str(), pt.result().data(), v.w, v.h);
      if (rc != GI_OK) throw std::runtime_error("cannot write " + out);
    } else {
      write_pfm(out, pt.result(), v.w, v.h);
    }
    gi_counters c = pt.counters();
    std::printf("{\"steps\": %d, \"ms\": %.3f, \"rays_closest\": %llu, \"rays_shadow\": %llu, \"samples\": %llu}\n",
                steps, ms, (unsigned long long)c.rays_closest, (unsigned long long)c.rays_shadow,
                (unsigned long long)c.samples);
  } catch (const std::exception &e) {
    std::cerr << "gi_driver: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
