// mi355x_pt: the MI355X path tracer as a gi_algorithm plugin.  Mirrors
// gpu_pt (src/gpu-pt.cpp:11-73, src/gpu-pt.h:232-255): activate() uploads
// the scene once, compute() restarts the progressive render for the current
// camera, update() traces one more sample per pixel of every path
// (new_path / bounce / trace_further_bounces, src/gpu-pt.h:102-226), and
// in_progress() is true until path_samples() samples were accumulated.  All
// device work goes through the C-ABI of include/gi_hip.h.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "gi_algorithm.h"

namespace gi {

struct hip_error : std::runtime_error {
  int status;
  hip_error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

class mi355x_pt : public gi_algorithm {
 protected:
  int w, h;
  gi_ctx *ctx = nullptr;
  camera cam;
  int samples = 32;   // "pt/passes" default of gpu_pt (src/gpu-pt.cpp:16)
  int length = 4;     // bounces = 4 (src/gpu-pt.cpp:23)
  int rr_start = 0;
  uint32_t seed = 0;
  bool running = false;
  float background[3] = {1.f, 1.f, 1.f};

  void check(int rc, const char *what);

 public:
  mi355x_pt(int w, int h, const std::string &name = "mi355x_pt");
  ~mi355x_pt() override;

  void activate(const scene_desc *scene) override;
  void compute() override;
  void update() override;
  bool progressive() override { return true; }
  bool in_progress() override { return running; }

  void path_samples(int n) override;
  void path_length(int n) override;
  void light_samples(int n) override;

  void set_camera(const camera &c) { cam = c; }
  void russian_roulette(int start) { rr_start = start; }
  void sampler_seed(uint32_t s) { seed = s; }
  // running-mean image, row y*w+x, 3 floats per pixel
  std::vector<float> result() const;
  gi_counters counters() const;
  gi_ctx *context() const { return ctx; }
};

}  // namespace gi
