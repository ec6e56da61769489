// Host-side plugin interface of the path-tracing hot path, mirroring the
// reference's gi_algorithm (src/gi_algorithm.h:12-68): integrators register
// by unique name (a duplicate throws std::logic_error), select(name)
// activates the chosen one lazily (an unknown name throws
// std::runtime_error), compute() starts a render and, for progressive
// integrators, update() advances it while in_progress().
//
// The reference's activate() takes an rta::rt_set (libhyb); outside that
// codebase the scene comes in as a gi::scene_desc of plain C-ABI records.
#pragma once

#include <list>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gi_hip.h"

namespace gi {

// The scene an integrator is activated with (stands in for rta::rt_set:
// triangles, the binary BVH in the reference's 2 x float4 layout when one
// was built elsewhere, materials and lights).
struct scene_desc {
  const gi_tri *tris = nullptr;
  int64_t n_tris = 0;
  const gi_material *mats = nullptr;
  int32_t n_mats = 0;
  const gi_light *lights = nullptr;
  int32_t n_lights = 0;
  const float *bvh2_nodes = nullptr;  // optional: nodes[8*n], src/tracers.cu:26-36
  int64_t bvh2_n_nodes = 0;
  const uint32_t *bvh2_indices = nullptr;
  int64_t bvh2_n_indices = 0;
  // optional: the sky light's map (rgb, w*h; load_skylight_data,
  // src/lights.cpp:58-79), RGBA8 textures (gi_material.texture indexes
  // them), and placed meshes (two-level instancing)
  const float *sky_rgb = nullptr;
  int32_t sky_w = 0, sky_h = 0;
  struct image {
    const uint8_t *rgba;
    int32_t w, h;
  };
  const image *textures = nullptr;
  int32_t n_textures = 0;
  struct mesh {
    const gi_tri *tris;
    int64_t n;
  };
  const mesh *meshes = nullptr;
  int32_t n_meshes = 0;
  const gi_instance *instances = nullptr;
  int32_t n_instances = 0;
};

// Camera as the reference extracts it from its lookat matrix
// (src/gpu-pt.cpp:61-67): position, direction, up, vertical field of view.
struct camera {
  float pos[3] = {0, 0, 0}, dir[3] = {0, 0, 1}, up[3] = {0, 1, 0};
  float fovy = 45.f;
};

class gi_algorithm {
 public:
  static const scene_desc *original_scene;
  static gi_algorithm *selected;

 protected:
  static std::vector<gi_algorithm *> &algorithms() {
    static std::vector<gi_algorithm *> a;
    return a;
  }
  bool activated = false;
  std::string name;
  bool verbose = false;

 public:
  explicit gi_algorithm(const std::string &n) : name(n) {
    for (auto *a : algorithms())
      if (a->name == n) throw std::logic_error(std::string("the algorithm ") + n + " is already registered.");
    algorithms().push_back(this);
  }
  virtual ~gi_algorithm() {
    auto &v = algorithms();
    for (size_t i = 0; i < v.size(); ++i)
      if (v[i] == this) {
        v.erase(v.begin() + (long)i);
        break;
      }
    if (selected == this) selected = nullptr;
  }
  gi_algorithm(const gi_algorithm &) = delete;
  gi_algorithm &operator=(const gi_algorithm &) = delete;

  virtual void activate(const scene_desc *) { activated = true; }
  virtual void compute() = 0;
  virtual bool progressive() { return false; }
  virtual bool in_progress() { return false; }
  virtual void update() { compute(); }

  static void select(const std::string &n) {
    for (auto *a : algorithms())
      if (a->name == n) {
        selected = a;
        if (!selected->activated) selected->activate(original_scene);
        return;
      }
    throw std::runtime_error("no gi algorithm called " + n);
  }
  static gi_algorithm *get(const std::string &n) {
    for (auto *a : algorithms())
      if (a->name == n) return a;
    return nullptr;
  }
  static std::list<std::string> list() {
    std::list<std::string> l;
    for (auto *a : algorithms()) l.push_back(a->name);
    return l;
  }
  void debug(bool verb) { verbose = verb; }
  const std::string &algorithm_name() const { return name; }

  virtual void path_samples(int) {}
  virtual void path_length(int) {}
  virtual void light_samples(int) {}
};

inline const scene_desc *gi_algorithm::original_scene = nullptr;
inline gi_algorithm *gi_algorithm::selected = nullptr;

}  // namespace gi
