/*
 * ORACLE -- test infrastructure.  A CPU restatement of globillum's CPU path
 * tracer (hybrid_pt / cpu_pt) used ONLY as the checker by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  Nothing in the
 * product (globillum_amd/) includes, links or calls it.
 *
 * The reference cannot be built here (it needs the unvendored rta / bbvh /
 * libhyb libraries, guile and nvcc), so there is no oracle/_ref; this file is
 * the restatement, and each function cites the reference file:line it follows.
 */
#ifndef GI_ORACLE_H
#define GI_ORACLE_H

#include <stdint.h>
#include "../include/gi_hip.h" /* data-type layouts only */

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint64_t rays_closest, rays_shadow, vertices, samples;
  uint64_t box_closest, tri_closest; /* child boxes / triangles tested */
  uint64_t box_shadow, tri_shadow;
} or_stats;

typedef struct or_scene or_scene;

/* nodes == NULL: the oracle builds its own object-median binary BVH. */
or_scene *or_scene_create(const gi_tri *tris, int64_t n, const gi_material *mats, int32_t nm,
                          const gi_light *lights, int32_t nl, const float *nodes2f4,
                          int64_t n_nodes, const uint32_t *indices, int64_t n_indices,
                          int32_t max_leaf);
void or_scene_free(or_scene *s);
/* Instancing: a mesh (object-space triangles, own BVH) and the instances
 * placing meshes; the scene's triangles stay the world geometry. */
int32_t or_scene_add_mesh(or_scene *s, const gi_tri *tris, int64_t n, int32_t max_leaf);
/* Textures: RGBA8 w x h (row 0 at t = 1), its mip chain built here;
 * materials name them by gi_material.texture (diffuse texture). */
int32_t or_scene_add_texture(or_scene *s, const uint8_t *rgba, int32_t w, int32_t h);
int32_t or_texture_levels(const or_scene *s, int32_t id);
int64_t or_texture_chain(const or_scene *s, int32_t id, uint8_t *out);
void or_tex_sample(const or_scene *s, int32_t id, float u, float v, float diff, float *rgb);
int32_t or_tex_level(float diff, int32_t w, int32_t h, int32_t levels);
float or_pow22(float x);
int or_scene_set_instances(or_scene *s, const gi_instance *inst, int32_t n);
int64_t or_scene_bvh_nodes(const or_scene *s);
void or_scene_export_bvh(const or_scene *s, float *nodes, uint32_t *indices);
float or_scene_diagonal(const or_scene *s);

/* Tracing in FP32 through the binary BVH (reference traversal rules). */
void or_trace_closest(const or_scene *s, const float *o, const float *d, const float *tmax,
                      int64_t n, gi_hit *out, uint64_t *n_box, uint64_t *n_tri);
void or_trace_any(const or_scene *s, const float *o, const float *d, const float *tmax,
                  int64_t n, uint8_t *occ);
/* Brute-force linear scan, FP32 (same arithmetic as the BVH path). */
void or_brute_closest(const or_scene *s, const float *o, const float *d, const float *tmax,
                      int64_t n, gi_hit *out);
/* FP64 brute force: the tie adjudicator. */
void or_brute_closest_f64(const or_scene *s, const float *o, const float *d, const float *tmax,
                          int64_t n, double *t, double *beta, double *gamma, uint32_t *ref);
/* Single triangle test on explicit vertices (FP32 / FP64). */
int or_intersect_tri(const float *a, const float *b, const float *c, const float *o,
                     const float *d, float *t, float *beta, float *gamma);
int or_intersect_tri_f64(const double *a, const double *b, const double *c, const double *o,
                         const double *d, double *t, double *beta, double *gamma);

/* Halton (src/util.cpp:22-32 and the scrambled extension). */
float or_radical_inverse(uint64_t n, int32_t base);
float or_halton_sample(uint64_t n, int32_t dim, uint32_t seed);
int32_t or_prime(int32_t dim);
void or_halton_perm(int32_t dim, uint32_t seed, uint16_t *perm);

/* Camera rays for one sample of one view (src/raygen.cu:10-118; flags:
 * GI_FLAG_PIXEL_CENTRE for src/dofrays.cu:10-68). */
void or_camera_rays(const gi_view *v, int32_t sample, uint32_t seed, uint64_t index_base, uint32_t flags,
                    float *o, float *d, float *tmax);

/* Tangent frame (stand-in for libhyb make_tangent_frame). */
void or_tangent_frame(const float *n, float *t, float *b);

/* NEE for one shading point (src/arealight-sampler-cpu.cpp:25-119). */
void or_nee(const or_scene *s, const float *P, const float *N, const float *xi, float *dir,
            float *len, float *contrib, int32_t *light);

/* BRDF sample + eval in tangent space (src/material-wrapper.h:15-51).
 * Returns the lobe type; f_sample = eval(wo, wi), f_light = eval(wo, l). */
int or_brdf_sample(const gi_material *m, const float *wo_ts, const float *xi, float *wi_ts,
                   float *pdf);
void or_brdf_eval(const gi_material *m, int type, const float *wo, const float *wi,
                  const float *N, float *f);
int or_brdf_sample_f64(const gi_material *m, const double *wo_ts, const double *xi,
                       double *wi_ts, double *pdf);
void or_brdf_eval_f64(const gi_material *m, int type, const double *wo, const double *wi,
                      const double *N, double *f);

/* The pieces of the GTR2 specular sample, FP64: the halfway vector
 * (tangent space) for (u, v) and slope scales (ax, ay); the Walter sample
 * weight of a reflected pair (L, V) about m; the GTR1 (clearcoat) density. */
void or_gtr2_half_f64(double u, double v, double ax, double ay, double *h);
double or_walter_weight_f64(const double *L, const double *V, const double *m, double ax, double ay);
double or_gtr1_f64(double ndh, double a);
/* gpu_pt's Phong model (GI_FLAG_PHONG): the lobe's powers x^100 and
 * x^(1/101), and one vertex -- kd, ks, the normalised shading normal, the
 * incoming direction, the light direction and the bounce sample in, the
 * NEE weight, the new direction and the throughput factor out; returns 0
 * when the path is absorbed. */
float or_phong_pow100(float x);
float or_phong_root101(float x);
int or_phong_vertex(const float *kd, const float *ks, const float *N, const float *org_dir, const float *sd,
                    const float *xp, float *fnee, float *wi, float *tpmul);
/* Camera basis U, V, W and image-plane half extents (width, height) of a
 * view (src/raygen.cu:17-39). */
void or_camera_basis(const gi_view *v, float *U, float *V, float *W, float *wh);

/* Full renderer.  images: views concatenated, 3 floats per pixel.
 * Renders pixels [p_begin, p_end) of every view (clamped per view; p_end<0 =
 * all) and samples [0, spp).  threads <= 0: all OpenMP threads. */
void or_render(const or_scene *s, const gi_view *views, int32_t n_views,
               const gi_render_opts *opts, float *images, or_stats *st, int32_t threads,
               int64_t p_begin, int64_t p_end);
/* Same in FP64 (closest hits by the FP64 BVH path); images in double. */
void or_render_f64(const or_scene *s, const gi_view *views, int32_t n_views,
                   const gi_render_opts *opts, double *images, int32_t threads);

/* Path-replay gradient of sum(dL_dimage * image) (views concatenated) with
 * respect to (color rgb, roughness, metallic) per material -> d_mats[5*nm]
 * and the emission (rgb) per light -> d_emission[3*nl].  f64: arithmetic in
 * FP64 (the reference gradient), else FP32.  Returns 1 for unsupported
 * options (Russian roulette). */
int or_adjoint(const or_scene *s, const gi_view *views, int32_t n_views, const gi_render_opts *opts,
               const float *dL_dimage, double *d_mats, double *d_emission, int32_t threads, int32_t f64);
/* or_adjoint plus the gradient w.r.t. every texture's level-0 texel values
 * (byte / 255), d_tex = sum over textures of w*h*3 doubles, RGB. */
int or_adjoint_tex(const or_scene *s, const gi_view *views, int32_t n_views, const gi_render_opts *opts,
                   const float *dL_dimage, double *d_mats, double *d_emission, double *d_tex, int32_t threads,
                   int32_t f64);

/* Sky light (gi::sky_light): the map of the scene's single GI_LIGHT_SKY
 * light (scale in col[0]); returns nonzero unless exactly one sky light. */
int or_scene_set_sky(or_scene *s, const float *rgb, int32_t w, int32_t h);
void or_sky_eval(const or_scene *s, const float *d, float *rgb);
void or_sky_sample(const or_scene *s, const float *xi, const float *N, float *dir, float *contrib);
float or_atan2_turns(float y, float x);
float or_cdf_sample(const float *cdf, const float *pdf, int32_t size, float u, float *p);

#ifdef __cplusplus
}
#endif
#endif
