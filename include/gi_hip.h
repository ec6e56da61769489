/*
 * gi_hip.h -- C-ABI of the MI355X (gfx950) path-tracing hot path.
 *
 * This is the drop-in boundary that replaces globillum's GPU path-tracing
 * hot path behind its gi_algorithm / rta tracer interface:
 *
 *   reference piece (file:line)                         replaced by
 *   ---------------------------------------------------------------------------
 *   rta::plugin_create_rt_set upload (src/hl-main.cpp:426)  gi_upload_triangles,
 *                                                          gi_upload_bvh2 / gi_build_bvh
 *   convert_and_upload_materials (src/material-conversion.cpp:59-151)
 *                                                          gi_upload_materials
 *   convert_and_upload_lights / add-rectlight (src/lights.cpp:84-96,
 *     src/arealight-sampler.cu:25-62)                     gi_upload_lights
 *   k::trace_cis_ailabox_indexed_with_alphamaps (src/tracers.cu:38-169),
 *     bbvh-cuda closest-hit tracer (src/hl-main.cpp:375-385)
 *                                                          gi_trace_closest
 *   matching_any_hit_tracer() (src/gpu-pt.cpp:40)          gi_trace_any
 *   k::setup_jittered_shirley / _dof_shirley (src/raygen.cu:10-145),
 *     k::generate_rectlight_sample / generate_arealight_sample
 *     (src/arealight-sampler.cu:110-327),
 *     k::compute_path_contribution_and_bounce (src/gpu-pt-kernels.cu:43-272),
 *     k::reset_data / k::combine_color_samples (src/util-kernels.cu:36-71),
 *     gpu_pt_bouncer::{new_pass,new_path,bounce,trace_further_bounces}
 *     (src/gpu-pt.h:102-226)                             gi_render, gi_progress_step
 *   output_color readback (src/gpu-pt.cpp:53-55)           gi_get_image
 *   gi::halton / generate_halton_pool_on_gpu (src/util.cpp:22-51)
 *                                                          gi_halton_sample
 *
 * Conventions: integer status codes, no exceptions cross the ABI.  Host
 * buffers are caller-owned and copied in or out; device buffers are owned
 * by the context.  A context is externally synchronised (one caller thread)
 * and owns one HIP stream; every call returns after its work completed.
 */
#ifndef GI_HIP_H
#define GI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GI_ABI_VERSION 1

typedef enum {
  GI_OK = 0,
  GI_ERR_INVALID_ARG = 1,   /* null pointer, zero/negative size, bad enum      */
  GI_ERR_NO_DEVICE = 2,     /* no HIP device / HIP runtime error at create     */
  GI_ERR_HIP = 3,           /* HIP runtime error during a call                 */
  GI_ERR_STATE = 4,         /* call order violated (e.g. render before upload) */
  GI_ERR_OUT_OF_MEMORY = 5,
  GI_ERR_LIMIT = 6,         /* size beyond a documented limit; also a traversal
                               stack overflow (impossible for trees this library
                               builds: stacks are sized from the tree's depth) */
  GI_ERR_UNKNOWN_NAME = 7   /* unknown algorithm name (gi_algorithm::select)   */
} gi_status;

/* Triangle record: the field set of rta::cuda::simple_triangle as filled in
 * src/hl-main.cpp:332-348 (positions, vertex normals, texture coordinates,
 * material index).  100 bytes, copied as-is to the device. */
typedef struct {
  float a[3], b[3], c[3];
  float na[3], nb[3], nc[3];
  float ta[2], tb[2], tc[2];
  int32_t material_index;
} gi_tri;

/* Hit record: rta::triangle_intersection {t, beta, gamma, ref}
 * (src/hybrid-pt.cpp:288-289, src/tracers.cu:110-113).
 * Barycentrics: P = (1-beta-gamma)*a + beta*b + gamma*c.
 * Invalid hit: t = FLT_MAX and ref = GI_INVALID_REF. */
typedef struct {
  float t, beta, gamma;
  uint32_t ref;
} gi_hit;
#define GI_INVALID_REF 0xFFFFFFFFu

/* Material: rta::cuda::material_t (src/material.h:142-184) with the
 * PrincipledBRDFParameters (src/principledParameters.h:7-21) inlined.
 * principled == 0 selects the LambertianMaterial path
 * (src/material-wrapper.h:11-28, ALL_MATERIAL_LAMBERT). */
typedef struct {
  float diffuse_color[3];
  float specular_color[3];
  float alpha;
  int32_t principled;
  float metallic, subsurface, specular, roughness;
  float specular_tint, anisotropic, sheen, sheen_tint;
  float clearcoat, clearcoat_gloss, opacity, transmission;
  float color[3];
  int32_t texture;          /* diffuse texture (gi_upload_texture id), -1: none */
} gi_material;

/* Light: gi::light with the rect_light or sky_light member
 * (src/lights.h:94-154).  power is the selection weight; gi_rect_light_power()
 * gives the value the reference computes in add-rectlight
 * (src/lights.cpp:92), gi_sky_light_power() the add-skylight value
 * (src/lights.cpp:98-110).  A GI_LIGHT_SKY light keeps its scale in col[0]
 * (the other fields are unused); its map comes with gi_upload_sky.  At most
 * one sky light per scene. */
typedef struct {
  float center[3], dir[3], up[3], col[3];
  float wh[2];
  float power;
  int32_t type;             /* GI_LIGHT_RECT or GI_LIGHT_SKY */
} gi_light;
#define GI_LIGHT_RECT 0
#define GI_LIGHT_SKY 1

/* One view of a render: camera as given to camera_ray_generator_shirley::
 * setup(pos, dir, up, fovy) (src/gpu-pt.cpp:67) plus the per-view image
 * size and sample count.  fovy is the setup() argument in degrees; the ray
 * generator uses tan(pi*(fovy/2)/180) as the half-height of the image plane
 * (src/raygen.cu:18-19).  aperture > 0 selects the thin-lens generator
 * (src/raygen.cu:61-118). */
typedef struct {
  float pos[3], dir[3], up[3];
  float fovy;
  int32_t w, h, spp;
  float aperture, focus_distance;
} gi_view;

/* Render options.  max_depth is max_path_len (src/gpu-pt.h:215-226): each
 * sample traces max_depth closest-hit rays and up to max_depth shadow rays.
 * rr_start > 0 enables Russian roulette once rr_start segments were traced
 * (survival = max channel of the throughput).  offset_bounce / offset_shadow
 * replace the scene-scale constants 0.3 (src/hybrid-pt.cpp:551) and 0.9
 * (src/arealight-sampler-cpu.cpp:115); <= 0 selects 1e-4 * scene diagonal
 * (the bounding box of the gi_upload_triangles geometry; instances do not
 * change it). */
typedef struct {
  int32_t max_depth;
  int32_t rr_start;
  uint32_t seed;
  uint32_t flags;
  float offset_bounce;
  float offset_shadow;
  float background[3];
  int32_t samples_per_pass; /* samples per pixel kept in flight; 0 = auto    */
} gi_render_opts;

#define GI_FLAG_BATCHED   0x1u  /* all views share one set of queues       */
#define GI_FLAG_PIXEL_CENTRE 0x8u /* primary rays through the pixel centres
                                   instead of jittered positions: with
                                   aperture > 0 the reference's
                                   k::setup_shirley_lens_rays (src/dofrays.cu:
                                   10-68; its unit-disk rejection never
                                   rejects the [-1/2,1/2]^2 lens jitter used
                                   here), with aperture 0 its pixel-centre
                                   pinhole (src/raygen.cu:24-25) */
#define GI_FLAG_PHONG     0x4u  /* gpu_pt's material model for every material
                                   (compute_path_contribution_and_bounce,
                                   src/gpu-pt-kernels.cu:89-199): Lambert with
                                   diffuse_color (times the diffuse texture)
                                   plus a Phong lobe of shininess 100 with
                                   specular_color, both divided by their six-
                                   channel sum when it exceeds 1; the lobe is
                                   chosen by the colour sums (absorption past
                                   them, and below the surface); single-cosine
                                   NEE; no throughput clamp; the next ray starts
                                   offset_bounce along its direction.  Principled
                                   parameters are not used.  Not differentiable
                                   (gi_adjoint refuses it). */
#define GI_FLAG_GRAPH     0x2u  /* bounce loop captured into a hipGraph; the
                                   context keeps the executable graph and
                                   relaunches it for later renders whose
                                   launches take the same arguments (same
                                   view sizes and sample counts, options,
                                   tuning and scene buffers; camera poses,
                                   material and light values may change)   */

/* Telemetry of the last gi_render / gi_adjoint call. */
typedef struct {
  uint64_t rays_closest;    /* closest-hit rays traced (primary + bounce) */
  uint64_t rays_shadow;     /* any-hit rays traced                         */
  uint64_t samples;         /* sum over views of w*h*spp                   */
  uint64_t path_vertices;   /* shaded hits                                 */
  double ms_total;          /* wall time of the call, host clock           */
  double ms_trace_closest;  /* HIP-event time of the closest-hit kernels   */
  double ms_trace_any;
  double ms_shade;
  double ms_graph_build;    /* hipGraph capture + instantiate (GI_FLAG_GRAPH
                               renders that could not relaunch the graph the
                               context keeps), host clock                   */
  int32_t launches_closest; /* kernel launches of the closest-hit kernel   */
  int32_t passes;
} gi_counters;

/* Tunable launch parameters (chosen by the autotuner). 0 = default. */
typedef struct {
  int32_t trace_block;      /* threads per block of the traversal kernels  */
  int32_t trace_blocks_per_cu;
  int32_t shade_block;
  int32_t samples_per_pass;
  int32_t stack_lds;        /* traversal stack entries per lane in LDS: 8, 16 or 32 */
  int32_t target_slots;     /* paths in flight per pass, in units of 2^20: 0..256,
                               0 = 128 (about 30 GB of path state with textures) */
  int32_t node_width;       /* child slots the collapse fills per BVH node: 2, 3 or 4
                               (0 = 4).  A build parameter: a host-built world tree
                               is collapsed again at once (refused while instances
                               are placed); a device-built one at its next build.
                               Mesh and instance trees keep 4. */
  int32_t refill_lanes;     /* compaction threshold of the traversal kernels: a wave
                               takes new rays once this many of its lanes are idle,
                               in chunks of this many rays: 8, 16, 32 or 64 (0 = 32) */
} gi_tuning;

/* Per-material gradient of the adjoint pass. */
typedef struct {
  float d_color[3];
  float d_roughness;
  float d_metallic;
} gi_material_grad;

typedef struct gi_ctx gi_ctx;
typedef struct gi_bvh gi_bvh;

/* ---- version / errors ---------------------------------------------------- */
int gi_abi_version(void);
const char *gi_status_string(int status);

/* ---- context (gi_algorithm::activate, src/gpu-pt.cpp:11-47) --------------- */
int gi_create(gi_ctx **out, int device);
void gi_destroy(gi_ctx *ctx);
const char *gi_last_error(const gi_ctx *ctx);

/* ---- host BVH build (stands in for rta's -b bsah, src/hl-main.cpp:377-379) -
 * Binned-SAH binary BVH over the triangle list.  Host only; no device call. */
int gi_bvh_build(const gi_tri *tris, int64_t n, int32_t max_leaf, gi_bvh **out);
void gi_bvh_free(gi_bvh *bvh);
/* Size of the binary BVH in the reference's 2 x float4 node layout
 * (src/tracers.cu:26-36) and of its triangle index array. */
int gi_bvh_sizes(const gi_bvh *bvh, int64_t *n_nodes, int64_t *n_indices);
/* Export in the reference layout: nodes[8*n_nodes] floats, indices[n_indices]. */
int gi_bvh_export2(const gi_bvh *bvh, float *nodes, uint32_t *indices);

/* ---- scene upload (src/hl-main.cpp:426-428) --------------------------------
 * gi_upload_triangles starts a new scene: it also drops the meshes and
 * instances.  Rendering checks that every material index shading can reach
 * (triangles, instanced meshes, instance overrides) names an uploaded
 * material. */
int gi_upload_triangles(gi_ctx *ctx, const gi_tri *tris, int64_t n);
int gi_upload_materials(gi_ctx *ctx, const gi_material *mats, int32_t n);
int gi_upload_lights(gi_ctx *ctx, const gi_light *lights, int32_t n);
/* Map of the sky light (rgb: w*h*3 floats, row y = 0 at theta = 0), as
 * load_skylight_data reads it (src/lights.cpp:58-65).  Builds the sampling
 * tables of initSkylight (src/lights.h:107-116).  NEE samples it by
 * importance (sky_light::sample, src/lights.h:121-139); a path that leaves
 * the scene after a bounce picks up the map's radiance (evaluateSkyLight,
 * src/hybrid-pt.cpp:175-189, unscaled as there). */
int gi_upload_sky(gi_ctx *ctx, const float *rgb, int32_t w, int32_t h);
/* Build the binned-SAH BVH on the host (subtrees of 2^16 or more triangles
 * near the root on other threads; the same tree as a sequential build),
 * collapse it into 4-wide nodes by surface-area cost (gi_tuning.node_width
 * child slots per node) and upload it. */
int gi_build_bvh(gi_ctx *ctx, int32_t max_leaf);
/* The same tree format built on the device: a linear BVH over 30-bit
 * Morton codes of the triangle centroids (the reference's GPU LBVH,
 * src/main.cpp:218-225), refit level by level, collapsed into 4-wide nodes
 * by the host builder's surface-area-cost dynamic program, with leaves of
 * at most max_leaf (<= 15) triangles.  Boxes are padded as the host builder
 * pads them, so closest-hit results are the same with either tree. */
int gi_build_bvh_device(gi_ctx *ctx, int32_t max_leaf);
#define GI_BVH_LBVH 0 /* radix tree over the Morton keys (gi_build_bvh_device) */
#define GI_BVH_PLOC 1 /* locally-ordered clustering over the Morton order      */
#define GI_BVH_SAH 2  /* binned SAH, the host builder's splits (gi_build_bvh)  */
/* gi_build_bvh_device with a choice of binary tree before the collapse.
 * GI_BVH_PLOC merges mutual nearest neighbours (smallest union surface
 * area) within `radius` (1..64, 0 = 16) positions of the Morton order,
 * round by round (Meister and Bittner 2018).  GI_BVH_SAH takes the host
 * builder's binned-SAH split decisions (32 centroid bins per axis, SAH leaf
 * test, object median from depth 48) level by level on the device: the
 * host tree's shape at device build speed; radius is unused.  Same node
 * format, padding and closest-hit results as the other builders.  The
 * reference builds its GPU trees with rta's LBVH (src/main.cpp:218-225);
 * PLOC and device SAH have no counterpart there. */
int gi_build_bvh_device_ex(gi_ctx *ctx, int32_t max_leaf, int32_t method, int32_t radius);
/* Upload an externally built binary BVH in the reference's 2 x float4 layout
 * (src/tracers.cu:26-36): nodes[8*n_nodes], indices[n_indices]. */
int gi_upload_bvh2(gi_ctx *ctx, const float *nodes, int64_t n_nodes,
                   const uint32_t *indices, int64_t n_indices);

/* ---- tracing primitives (trace_rays of the closest / any-hit tracers) ------
 * orig[3n], dir[3n], max_t[n] are host arrays in the reference's AoS float
 * layout (src/raygen.cu:53-58).  Closest hit is the minimum t, ties broken by
 * the lower triangle id; a triangle counts when 0 < t <= max_t. */
int gi_trace_closest(gi_ctx *ctx, const float *orig, const float *dir,
                     const float *max_t, int64_t n, gi_hit *out);
/* occluded[i] = 1 iff some triangle is hit with 0 < t <= max_t[i]. */
int gi_trace_any(gi_ctx *ctx, const float *orig, const float *dir,
                 const float *max_t, int64_t n, uint8_t *occluded);
/* Same as gi_trace_* on device pointers (orig/dir as float[3n]); runs on the
 * context stream and returns after completion. */
int gi_trace_closest_device(gi_ctx *ctx, const float *d_orig, const float *d_dir,
                            const float *d_max_t, int64_t n, gi_hit *d_out);
int gi_trace_any_device(gi_ctx *ctx, const float *d_orig, const float *d_dir,
                        const float *d_max_t, int64_t n, uint8_t *d_occluded);
/* gi_trace_closest_device without any device allocation: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same rays, same results as gi_trace_closest_device.  Its scratch -- the
 * packed rays, the global part of the traversal stacks and the overflow
 * counter -- is a fixed 32 MiB device array of the library (one per GPU,
 * held until the process ends) instead of context buffers.  Limits: at most
 * 524288 rays per call (more: GI_ERR_LIMIT); the traversal grid is capped
 * to the stack columns the array holds (a tree too deep for one block of
 * columns: GI_ERR_LIMIT).  Calls that overlap on one GPU share the array
 * and race; each call clears what must start at zero. */
int gi_trace_closest_device_noalloc(gi_ctx *ctx, const float *d_orig, const float *d_dir,
                                    const float *d_max_t, int64_t n, gi_hit *d_out);

/* ---- camera rays (setup_jittered_shirley, src/raygen.cu:10-59) ------------
 * Writes the w*h primary rays of sample `sample` of view `v` (host arrays). */
int gi_generate_camera_rays(gi_ctx *ctx, const gi_view *v, int32_t sample,
                            uint32_t seed, float *orig, float *dir, float *max_t);

/* ---- renderer ------------------------------------------------------------- */
int gi_render(gi_ctx *ctx, const gi_view *views, int32_t n_views,
              const gi_render_opts *opts);
/* Running-mean image of view v (float rgb[3*w*h], row y*w+x). */
int gi_get_image(gi_ctx *ctx, int32_t view, float *rgb);
/* Device pointer of the running-mean image of view v: 4 floats per pixel
 * (r, g, b, unused), row y*w+x.  Valid until the next gi_render or
 * gi_progress_begin. */
int gi_get_image_device(gi_ctx *ctx, int32_t view, float **d_rgb);
int gi_get_counters(const gi_ctx *ctx, gi_counters *out);
/* Traversal statistics: per kind of trace (closest hit; any hit), the rays
 * traced, the wide-BVH interior nodes visited (one 128-byte node fetch
 * each), the triangles tested (48 bytes each) and the instance entries.
 * Counted by a separate build of the traversal kernels (LDS stack depth 16)
 * that runs only while statistics are on, so the default kernels pay
 * nothing.  The reference prints per-bounce trace times only
 * (src/gpu-pt.h:185); these counts have no counterpart there. */
typedef struct {
  uint64_t closest_rays, closest_nodes, closest_triangles, closest_instances;
  uint64_t any_rays, any_nodes, any_triangles, any_instances;
} gi_traversal_stats;
/* enable != 0: zero the totals and count every later traversal launch
 * (renders, adjoint replays, gi_trace_closest / gi_trace_any); 0: stop
 * counting (the totals are kept). */
int gi_set_traversal_stats(gi_ctx *ctx, int32_t enable);
/* Totals since the last gi_set_traversal_stats(ctx, 1). */
int gi_get_traversal_stats(gi_ctx *ctx, gi_traversal_stats *out);
/* Consistency check of the traversal's half-precision node copy (no
 * counterpart in the reference, whose traversal reads f32 boxes,
 * src/tracers.cu:70-73): the kernels test each child's box as f16 plane
 * offsets from a per-node f32 origin, rounded outward.  Counts the child
 * slots of the current tree whose decoded box does not enclose the f32
 * box widened by its margins, or whose reference differs; 0 expected. */
int gi_check_half_nodes(gi_ctx *ctx, int64_t *bad_slots);
int gi_set_tuning(gi_ctx *ctx, const gi_tuning *t);
int gi_get_tuning(const gi_ctx *ctx, gi_tuning *t);

/* Progressive mode (gpu_pt::compute / update, src/gpu-pt.cpp:49-73):
 * begin resets the accumulation, each step renders one more sample per
 * pixel of every view; returns GI_OK while samples remain, and sets *done. */
int gi_progress_begin(gi_ctx *ctx, const gi_view *views, int32_t n_views,
                      const gi_render_opts *opts);
int gi_progress_step(gi_ctx *ctx, int32_t *done);

/* ---- adjoint (path-replay backprop; no counterpart in the reference) -------
 * Gradient of sum_i dL_dimage[i] * image[i] over the views of the last
 * gi_render (dL_dimage: 3 floats per pixel, views concatenated) with respect
 * to every material's (color rgb, roughness, metallic) -- for a Lambert
 * material d_color is d/d diffuse_color -- and every light's emission rgb
 * (d_emission[3*n_lights]; the light-selection weight gi_light.power is held
 * fixed).  The paths are replayed with the same Halton indices; samples
 * (lobe choice, cosine direction or GGX half-vector, light choice) are held
 * fixed and the sampling's dependence on metallic and roughness enters as a
 * score-function term.  Not defined with Russian roulette (rr_start > 0). */
int gi_adjoint(gi_ctx *ctx, const float *dL_dimage, gi_material_grad *d_mats,
               float *d_emission);

/* ---- Textures (texture_data, src/material.h:21-137; compute_mipmaps,
 * src/material.cu:235-274) ----------------------------------------------------
 * An RGBA8 image, w x h texels row-major (row 0 is the top, t = 1), and its
 * box-filtered mip chain built on the device: floor(log2(min(w, h))) levels
 * after level 0, each ((W+1)/2) x ((H+1)/2), texels the truncating integer
 * mean of the 2x2 below.  A material names its diffuse texture in
 * gi_material.texture.  Shading then scales the diffuse colour (Lambert) or
 * the base colour (principled) by sample_bilin_lod at the hit's texture
 * coordinates, with the level from the footprint of the path's two ray
 * differentials (computeRayDiffs, src/hybrid-pt.cpp:300-328).  Replaces
 * convert_texture / compute_mipmaps (src/material-conversion.cpp:18-41);
 * the image file is decoded by the caller. */
int gi_upload_texture(gi_ctx *ctx, const uint8_t *rgba, int32_t w, int32_t h, int32_t *tex_id);
/* Levels after level 0, and texels of the whole chain. */
int gi_texture_info(gi_ctx *ctx, int32_t tex_id, int32_t *levels, int64_t *texels);
/* The chain, level 0 first: 4 * texels bytes. */
int gi_download_texture(gi_ctx *ctx, int32_t tex_id, uint8_t *rgba);
/* Gradient of the last gi_adjoint's objective with respect to the level-0
 * texel values of a texture (value = byte / 255): w*h*3 floats, RGB,
 * row-major.  A lookup blends four texels of one level with the bilinear
 * weights; a coarser level's texel is the mean of the (up to) 2x2 below
 * (the integer truncation is not differentiated).  The material colour
 * gradient of a textured material is d/d colour of colour * texel. */
int gi_texture_grad(gi_ctx *ctx, int32_t tex_id, float *d_rgb);

/* ---- Instancing (two-level BVH) --------------------------------------------
 * No counterpart in the reference; needed by the large configuration
 * (thousands of placed copies of one mesh).  A mesh is a triangle table in
 * object space with its own BVH; an instance places a mesh with a 3x4
 * row-major object-to-world transform (affine, invertible) and optionally
 * one material for all its triangles.  The triangles of gi_upload_triangles
 * stay the world geometry.  Rays hit the world and every instance: the
 * closest hit is the minimum (t, id) over all of them, with t the world ray
 * parameter.  Instance triangles have virtual ids n_world + base_i + k
 * (k = index in the mesh; base_i = sum of the mesh sizes of instances
 * before i), as gi_hit.ref and in rendering.  Call gi_upload_instances
 * after the world BVH is built; it builds the meshes' BVHs and the instance
 * tree. */
typedef struct {
  float xform[12];
  int32_t mesh;
  int32_t material;         /* -1: the mesh triangles' own material indices */
} gi_instance;
int gi_upload_mesh(gi_ctx *ctx, const gi_tri *tris, int64_t n, int32_t *mesh_id);
int gi_upload_instances(gi_ctx *ctx, const gi_instance *inst, int32_t n);

/* ---- Halton (gi::halton, src/util.cpp:22-32) ------------------------------
 * Scrambled radical inverse of index n in the prime base of dimension dim,
 * with the digit permutations derived from seed.  Host function. */
float gi_halton_sample(uint64_t n, int32_t dim, uint32_t seed);
/* Unscrambled radical inverse (the reference's gi::halton). */
float gi_radical_inverse(uint64_t n, int32_t base);
/* Prime base of dimension dim, and the permutation of its digits. */
int32_t gi_halton_prime(int32_t dim);
int gi_halton_permutation(int32_t dim, uint32_t seed, uint16_t *perm);

/* Light selection weight of a rect light, as add-rectlight computes it
 * (src/lights.cpp:92). */
float gi_rect_light_power(const gi_light *l);
/* scale * 4 pi (diameter/2)^2, as add-skylight computes it. */
float gi_sky_light_power(float scale, float diameter);

/* ---- Scene ingestion (host functions, no device needed) -------------------
 * Wavefront OBJ / MTL into the triangle and material tables, as
 * add_objfile_to_flat_tri_list builds them (src/hl-main.cpp:277-350):
 * triangles in file order; material 0 is the fallback "gi/fallback"
 * (diffuse = specular = (1,0,0)) registered with the first file, used by
 * faces without a (defined) material; other materials are registered once
 * per name on first use with diffuse Kd, specular Ks and alpha d; texture
 * coordinates are kept only when the material names a diffuse texture.
 * Several files append to one table (the reference's multi-file ftl).
 * trafo (optional, 16 floats row-major) maps positions as M (x,y,z,1) and
 * normals by the inverse transpose of its 3x3 part.  Faces without normals
 * get their geometric normal; polygons are split into fans. */
typedef struct gi_scene_file gi_scene_file;
int gi_obj_open(gi_scene_file **out);
int gi_obj_append(gi_scene_file *f, const char *obj_path, const float *trafo);
/* Principled parameters for every material that has <dir>/<name>.pbrdf
 * (convert_and_upload_materials, src/material-conversion.cpp:59-151). */
int gi_obj_apply_pbrdf(gi_scene_file *f, const char *dir);
int gi_obj_sizes(const gi_scene_file *f, int64_t *n_tris, int32_t *n_mats);
int gi_obj_export(const gi_scene_file *f, gi_tri *tris, gi_material *mats);
const char *gi_obj_material_name(const gi_scene_file *f, int32_t i);
/* The distinct map_Kd files of the loaded materials, in order of first use
 * (resolved against their MTL file's directory); gi_material.texture of an
 * exported material is its index in this list (-1: none). */
int32_t gi_obj_texture_count(const gi_scene_file *f);
const char *gi_obj_texture_path(const gi_scene_file *f, int32_t i);
const char *gi_obj_error(const gi_scene_file *f);
void gi_obj_free(gi_scene_file *f);
/* One .pbrdf file (PrincipledBRDFParameters::readFrom,
 * src/principledParameters.h:64-125) into *m: principled = 1, the named
 * parameters set, the others at the default constructor's values (opacity
 * 1, transmission 1, others 0; color = m's diffuse colour).  A malformed
 * number is GI_ERR_INVALID_ARG. */
int gi_pbrdf_read(const char *path, gi_material *m);

/* ---- Image output (save_image, src/util.cpp:111-150; host functions) -------
 * rgb: w*h*3 floats, row-major, row 0 first as gi_get_image returns them.
 * Pixel (x, y) is written at (w-1-x, h-1-y), as the reference's writer
 * flips both axes.  PNG: 8-bit RGB, clamp(int(255*v), 0, 255) per channel
 * (no gamma).  EXR: RGBA half, A = 1, uncompressed scanlines, float to half
 * rounded to nearest even. */
int gi_write_png(const char *path, const float *rgb, int32_t w, int32_t h);
int gi_write_exr(const char *path, const float *rgb, int32_t w, int32_t h);
/* Texture images for map_Kd (the reference decodes them with rta's image
 * library, external): PNG (grey, grey+alpha, RGB, RGBA at 8/16 bits,
 * palette at 1/2/4/8 bits with tRNS; not interlaced) and binary PPM (P6).
 * gi_image_size gives w and h; gi_image_read writes w*h RGBA8 texels, row 0
 * the file's top row (as gi_upload_texture takes them). */
int gi_image_size(const char *path, int32_t *w, int32_t *h);
int gi_image_read(const char *path, uint8_t *rgba);

#ifdef __cplusplus
}
#endif
#endif /* GI_HIP_H */
