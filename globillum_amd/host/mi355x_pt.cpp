#include "mi355x_pt.h"

#include <cstring>

namespace gi {

void mi355x_pt::check(int rc, const char *what) {
  if (rc != GI_OK)
    throw hip_error(rc, std::string(name) + ": " + what + ": " + gi_status_string(rc) + " (" +
                            (ctx ? gi_last_error(ctx) : "no context") + ")");
}

mi355x_pt::mi355x_pt(int w_, int h_, const std::string &n) : gi_algorithm(n), w(w_), h(h_) {
  if (w <= 0 || h <= 0) throw std::invalid_argument("mi355x_pt: image size must be positive");
}

mi355x_pt::~mi355x_pt() {
  if (ctx) gi_destroy(ctx);
}

// gpu_pt::activate (src/gpu-pt.cpp:11-47): scene, lights and materials to
// the device, BVH from the rt_set when it has one, else the host builder.
void mi355x_pt::activate(const scene_desc *scene) {
  if (activated) return;
  if (!scene) throw std::invalid_argument("mi355x_pt: activate without a scene");
  check(gi_create(&ctx, 0), "gi_create");
  check(gi_upload_triangles(ctx, scene->tris, scene->n_tris), "upload triangles");
  for (int32_t i = 0; i < scene->n_textures; ++i) {
    int32_t id = -1;
    check(gi_upload_texture(ctx, scene->textures[i].rgba, scene->textures[i].w, scene->textures[i].h, &id),
          "upload texture");
  }
  check(gi_upload_materials(ctx, scene->mats, scene->n_mats), "upload materials");
  check(gi_upload_lights(ctx, scene->lights, scene->n_lights), "upload lights");
  if (scene->sky_rgb) check(gi_upload_sky(ctx, scene->sky_rgb, scene->sky_w, scene->sky_h), "upload sky map");
  if (scene->bvh2_nodes)
    check(gi_upload_bvh2(ctx, scene->bvh2_nodes, scene->bvh2_n_nodes, scene->bvh2_indices, scene->bvh2_n_indices),
          "upload bvh");
  else
    check(gi_build_bvh(ctx, 4), "build bvh");
  if (scene->n_instances > 0) {
    for (int32_t i = 0; i < scene->n_meshes; ++i) {
      int32_t id = -1;
      check(gi_upload_mesh(ctx, scene->meshes[i].tris, scene->meshes[i].n, &id), "upload mesh");
    }
    check(gi_upload_instances(ctx, scene->instances, scene->n_instances), "upload instances");
  }
  gi_algorithm::activate(scene);
}

// gpu_pt::compute (src/gpu-pt.cpp:59-73): restart the progressive render
// for the current camera.
void mi355x_pt::compute() {
  if (!activated) throw std::logic_error(name + ": compute before activate");
  gi_view v;
  std::memset(&v, 0, sizeof(v));
  std::memcpy(v.pos, cam.pos, sizeof v.pos);
  std::memcpy(v.dir, cam.dir, sizeof v.dir);
  std::memcpy(v.up, cam.up, sizeof v.up);
  v.fovy = cam.fovy;
  v.w = w;
  v.h = h;
  v.spp = samples;
  gi_render_opts o;
  std::memset(&o, 0, sizeof(o));
  o.max_depth = length;
  o.rr_start = rr_start;
  o.seed = seed;
  std::memcpy(o.background, background, sizeof o.background);
  check(gi_progress_begin(ctx, &v, 1, &o), "progress begin");
  running = true;
}

// gpu_pt::update (src/gpu-pt.cpp:49-57): one more sample per pixel.
void mi355x_pt::update() {
  if (!running) return;
  int32_t done = 0;
  check(gi_progress_step(ctx, &done), "progress step");
  running = !done;
}

void mi355x_pt::path_samples(int n) {
  if (n <= 0) throw std::invalid_argument("path_samples must be positive");
  samples = n;
}
void mi355x_pt::path_length(int n) {
  if (n <= 0 || n > 32) throw std::invalid_argument("path_length must be in [1, 32]");
  length = n;
}
// One NEE sample per path vertex, as the bounce state machine of
// src/gpu-pt.h:168-214; other counts are not part of that estimator.
void mi355x_pt::light_samples(int n) {
  if (n != 1) throw std::invalid_argument("mi355x_pt takes one light sample per path vertex");
}

std::vector<float> mi355x_pt::result() const {
  std::vector<float> img((size_t)w * h * 3);
  if (ctx && gi_get_image(ctx, 0, img.data()) != GI_OK) throw hip_error(GI_ERR_STATE, name + ": no image yet");
  return img;
}

gi_counters mi355x_pt::counters() const {
  gi_counters c;
  std::memset(&c, 0, sizeof(c));
  if (ctx) gi_get_counters(ctx, &c);
  return c;
}

}  // namespace gi
