/*
 * pt_api.h -- C ABI of the MI355X path tracer (libpt_hip.so).
 *
 * This is the drop-in boundary for the reference's per-sample render path:
 * everything the reference hands to OptiX 6 (scene graph, acceleration
 * structures, camera variables, the per-frame launch and the output-buffer
 * maps) crosses this interface as plain structs, pointers and sizes.  No C++,
 * no HIP and no torch types appear in these signatures.
 *
 * Ownership: every input is deep-copied; the context owns all device memory.
 * Errors: every call returns a pt_status; nothing throws or exits across the
 *         ABI.  pt_last_error() gives a human-readable message.
 * Threading: a context must not be used from two threads at once; distinct
 *         contexts are independent.  pt_render is asynchronous on the
 *         context's stream; pt_sync and pt_read_* synchronise.
 *
 * Reference semantics: integrator defaults reproduce the reference
 * (programs/raygen.cu:66-123) unless a PT_FLAG_* switch says otherwise.
 */
#ifndef PT_API_H
#define PT_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_API_VERSION 10

/* ------------------------------------------------------------------ status */
typedef int32_t pt_status;
#define PT_OK 0
#define PT_ERR_INVALID_ARG (-1)  /* bad pointer, size or enum value          */
#define PT_ERR_DEVICE (-2)       /* HIP runtime error (message in last_error) */
#define PT_ERR_STATE (-3)        /* call out of order (e.g. render w/o accel) */
#define PT_ERR_OUT_OF_MEMORY (-4)
#define PT_ERR_UNSUPPORTED (-5)

/* ------------------------------------------------------------ enumerations */
/* Axis of an axis-aligned rectangle (programs/common.hpp:26 AXIS). */
#define PT_AXIS_X 0
#define PT_AXIS_Y 1
#define PT_AXIS_Z 2

/* Primitive kinds of a geometry group (the BLAS an instance refers to). */
#define PT_GEOM_SPHERES 0   /* programs/hitables/sphere.cu              */
#define PT_GEOM_RECTS 1     /* programs/hitables/aarect.cu              */
#define PT_GEOM_BOXES 2     /* programs/hitables/box.cu                 */
#define PT_GEOM_MESH 3      /* programs/hitables/triangle.cu, mesh.hpp  */
#define PT_GEOM_VOLUMES 4   /* programs/hitables/volume_{sphere,box}.cu */
#define PT_GEOM_CYLINDERS 5 /* programs/hitables/cylinder.cu            */
#define PT_GEOM_MOVING_SPHERES 6 /* programs/hitables/moving_sphere.cu  */

/* Materials (host_includes/materials.hpp). */
#define PT_MAT_LAMBERTIAN 0        /* tex[0]=albedo                            */
#define PT_MAT_METAL 1             /* tex[0]=albedo, param[0]=fuzz             */
#define PT_MAT_DIELECTRIC 2        /* tex[0]=base, tex[1]=extinction,
                                      param[0]=ior, param[1]=density           */
#define PT_MAT_DIFFUSE_LIGHT 3     /* tex[0]=emission                          */
#define PT_MAT_ASHIKHMIN_SHIRLEY 4 /* tex[0]=diffuse, tex[1]=specular,
                                      param[0]=nu, param[1]=nv (already >= 1)  */
#define PT_MAT_TORRANCE_SPARROW 5  /* tex[0]=Rs, param[0]=alpha_u,
                                      param[1]=alpha_v (already roughness^2)   */
#define PT_MAT_OREN_NAYAR 6        /* tex[0]=albedo, param[0]=A, param[1]=B    */
#define PT_MAT_ISOTROPIC 7         /* tex[0]=albedo; phase function 1/(4 pi)   */
#define PT_MAT_NORMAL 8            /* param[0] != 0: shading normal, else the
                                      geometric one, as 0.5 N + 0.5           */
#define PT_MAT_COUNT 9

/* Textures (host_includes/textures.hpp). */
#define PT_TEX_CONSTANT 0  /* rgb                                           */
#define PT_TEX_IMAGE 1     /* width x height float4 texels, bilinear, repeat */
/* checkered_texture.cu: the `width` texture where sin(10x) sin(10y) sin(10z)
 * of the world hit point is negative, else the `height` texture (neither
 * may be a checker) */
#define PT_TEX_CHECKER 2
/* noise_texture.cu: rgb[0] = scale; texels = 1792 floats: 256 gradient
 * vectors (x, y, z, 0), then perm_x, perm_y, perm_z (256 each, as floats) */
#define PT_TEX_NOISE 3
/* gradient_texture.cu: texels = 12 floats, colours A, B, C as (r, g, b, 0) */
#define PT_TEX_GRADIENT 4
#define PT_NOISE_FLOATS 1792
/* vector.cu: child texture i of `width` (texels: the child indices as
 * floats), i = the hit's face index (pt_mesh_desc.face_indices; 0 on other
 * primitives); black when i is out of range.  Children are not vectors. */
#define PT_TEX_VECTOR 5

/* Lights (host_includes/pdfs.hpp + Light_Sampler, host_common.hpp:65). */
#define PT_LIGHT_RECT 0    /* axis, a0, a1, b0, b1, k                   */
#define PT_LIGHT_SPHERE 1  /* center, radius                            */
#define PT_LIGHT_ENV 2     /* importance-sampled environment (miss texture) */

/* Miss programs (host_includes/programs.hpp:32 Miss_Programs). */
#define PT_MISS_GRADIENT 0
#define PT_MISS_CONSTANT 1
#define PT_MISS_IMAGE 2    /* LDR lat-long background (miss.cu:51)      */
#define PT_MISS_HDR 3      /* environment map (miss.cu:64)              */

/* Render flags. */
#define PT_FLAG_MATERIAL_SORT 0x1u   /* sort shading by material key         */
#define PT_FLAG_KEEP_PATHS 0x2u      /* keep last frame's per-path radiance  */
#define PT_FLAG_ENV_MIS 0x4u         /* env NEE + MIS, no miss clamp (new)   */
#define PT_FLAG_COUNT_TRAVERSAL 0x8u /* count node/primitive fetches         */
#define PT_FLAG_RADIANCE_CACHE 0x10u /* learned radiance cache (new): training
                                        paths + queries, see pt_cache_*     */
#define PT_FLAG_GRAPH 0x20u          /* replay each full batch from one
                                        captured hipGraph (no per-kernel
                                        timing in the stats)                */
#define PT_FLAG_TUNED 0x40u          /* use pt_autotune's batch size, sort
                                        and graph choice when it matches     */
#define PT_FLAG_DIRECT 0x80u         /* launch directly even where the tuned
                                        shape replays graphs: per-kernel
                                        times in the stats (extend_ms,
                                        shade_ms)                            */

/* ------------------------------------------------------------ scene input */
typedef struct pt_sphere_desc {
  float center[3];
  float radius;
  int32_t material;
} pt_sphere_desc;

typedef struct pt_rect_desc {
  int32_t axis; /* PT_AXIS_* */
  int32_t flip; /* non-zero: normal points along -axis */
  float a0, a1, b0, b1, k;
  int32_t material;
} pt_rect_desc;

typedef struct pt_box_desc {
  float pmin[3];
  float pmax[3];
  int32_t material;
} pt_box_desc;

/* Constant-density medium (hitables.hpp:175-209 Volumetric_Sphere,
 * :316-349 Volumetric_Box): a path entering the boundary scatters after a
 * free flight of -ln(xi) / density, if that ends inside. */
#define PT_VOLUME_SPHERE 0 /* a = center, b[0] = radius */
#define PT_VOLUME_BOX 1    /* a = pmin, b = pmax         */
typedef struct pt_volume_desc {
  int32_t shape; /* PT_VOLUME_* */
  float a[3];
  float b[3];
  float density;
  int32_t material;
} pt_volume_desc;

/* Open cylinder along +Y (hitables.hpp:411-454, cylinder.cu fixed to its
 * intent): the lateral surface of radius `radius` from origin.y to
 * origin.y + length, outward radial normal, no uv. */
typedef struct pt_cylinder_desc {
  float origin[3];
  float length;
  float radius;
  int32_t material;
} pt_cylinder_desc;

/* Sphere whose center moves linearly from center0 at time0 to center1 at
 * time1 (hitables.hpp:131-172, moving_sphere.cu); paths sample their time
 * in [camera time0, time1]; ray queries see it at time 0. */
typedef struct pt_moving_sphere_desc {
  float center0[3];
  float center1[3];
  float radius;
  float time0, time1;
  int32_t material;
} pt_moving_sphere_desc;

typedef struct pt_mesh_desc {
  uint32_t num_vertices;
  const float* vertices; /* 3 * num_vertices                                */
  const float* normals;  /* optional, 3 * num_vertices (NULL: geometric N)  */
  const float* uvs;      /* optional, 2 * num_vertices (NULL: u = v = 0)    */
  uint32_t num_triangles;
  const uint32_t* indices;       /* 3 * num_triangles                       */
  const int32_t* face_materials; /* optional per-face material (NULL: use
                                    `material` for every face)               */
  int32_t material;
  const int32_t* face_indices;   /* optional per-face index that a
                                    PT_TEX_VECTOR texture selects by (the
                                    reference's material buffer; NULL: 0)    */
} pt_mesh_desc;

/* One instance = one group of primitives of a single kind (a bottom-level
 * acceleration structure) placed in the world by an affine transform.  The
 * matrices are row-major 3x4: p_world = M * (p_obj, 1).  Instances with the
 * same (kind, first, count) share one bottom-level structure. */
typedef struct pt_instance_desc {
  int32_t kind;   /* PT_GEOM_* */
  uint32_t first; /* first primitive (spheres/rects/boxes) or mesh index */
  uint32_t count; /* primitive count (ignored for PT_GEOM_MESH)          */
  int32_t has_transform;
  float obj_to_world[12];
  float world_to_obj[12];
} pt_instance_desc;

typedef struct pt_material_desc {
  int32_t type; /* PT_MAT_* */
  int32_t tex[2];
  float param[4];
} pt_material_desc;

typedef struct pt_texture_desc {
  int32_t type; /* PT_TEX_* */
  float rgb[3];
  uint32_t width, height;
  const float* texels; /* image: width * height * 4 floats, row 0 = v in
                        * [0, 1/h); noise and gradient: see PT_TEX_* */
} pt_texture_desc;

typedef struct pt_light_desc {
  int32_t type; /* PT_LIGHT_* */
  int32_t axis; /* rect only */
  float a0, a1, b0, b1, k;
  float center[3]; /* sphere only */
  float radius;
  float emission[3]; /* Light_Emissions entry (programs.hpp:26) */
} pt_light_desc;

typedef struct pt_miss_desc {
  int32_t type; /* PT_MISS_* */
  float color1[3];
  float color2[3];
  int32_t texture;   /* PT_MISS_IMAGE / PT_MISS_HDR: texture index */
  int32_t spherical; /* PT_MISS_HDR: 1 spherical, 0 cylindrical   */
} pt_miss_desc;

typedef struct pt_scene_desc {
  uint32_t num_spheres;
  const pt_sphere_desc* spheres;
  uint32_t num_rects;
  const pt_rect_desc* rects;
  uint32_t num_boxes;
  const pt_box_desc* boxes;
  uint32_t num_meshes;
  const pt_mesh_desc* meshes;
  uint32_t num_instances;
  const pt_instance_desc* instances;
  uint32_t num_materials;
  const pt_material_desc* materials;
  uint32_t num_textures;
  const pt_texture_desc* textures;
  uint32_t num_lights;
  const pt_light_desc* lights;
  pt_miss_desc miss;
  uint32_t num_volumes;
  const pt_volume_desc* volumes;
  uint32_t num_cylinders;
  const pt_cylinder_desc* cylinders;
  uint32_t num_moving_spheres;
  const pt_moving_sphere_desc* moving_spheres;
} pt_scene_desc;

/* Camera variables exactly as Camera::set uploads them (camera.hpp:40-51). */
typedef struct pt_camera {
  float lower_left[3];
  float horizontal[3];
  float vertical[3];
  float origin[3];
  float u[3], v[3], w[3];
  float lens_radius;
  float time0, time1;
} pt_camera;

typedef struct pt_render_params {
  uint32_t width, height;   /* launch dimensions (main.cpp:33)             */
  uint32_t max_depth;       /* maxDepth (main.cpp:63)                      */
  int32_t russian;          /* russian (main.cpp:62)                       */
  uint32_t rr_min_depth;    /* RR applies when depth >= this; reference 11 */
  uint32_t frames_per_batch;/* frames traced together; 0 = automatic: as
                               many as 2^29 paths or half of the device's
                               memory (252 B of path state each) holds,
                               whichever is less                           */
  uint32_t flags;           /* PT_FLAG_*                                   */
  uint32_t reserved;
} pt_render_params;

typedef struct pt_render_stats {
  uint64_t paths;           /* camera paths traced                          */
  uint64_t extension_rays;  /* closest-hit queries                          */
  uint64_t shadow_rays;     /* occlusion queries                            */
  uint64_t node_fetches;    /* only with PT_FLAG_COUNT_TRAVERSAL            */
  uint64_t prim_fetches;    /* only with PT_FLAG_COUNT_TRAVERSAL            */
  uint64_t instance_fetches;/* only with PT_FLAG_COUNT_TRAVERSAL            */
  uint64_t shadow_node_fetches;     /* the same three counts for the     */
  uint64_t shadow_prim_fetches;     /* occlusion queries (only with      */
  uint64_t shadow_instance_fetches; /* PT_FLAG_COUNT_TRAVERSAL)          */
  uint64_t exhausted_paths; /* paths still alive after max_depth bounces     */
  uint64_t cache_queries;   /* radiance cache, last batch: paths ended by a query */
  uint64_t cache_train_vertices; /* training vertices recorded              */
  double cache_loss;        /* loss of the last training step              */
  uint32_t cache_steps;     /* training steps of the last batch            */
  double gpu_ms;            /* render wall time on the stream (events)      */
  double extend_ms;         /* summed closest-hit kernel time               */
  double shade_ms;          /* summed shading kernel time (hit and miss
                               programs, shadow rays, continuation)        */
  uint32_t extend_launches;
  uint32_t shade_launches;
  uint64_t nan_paths;       /* paths whose color() had a NaN channel, which
                               de_nan zeroes (raygen.cu:126-134)           */
  double cache_judge;       /* radiance cache, last batch: the judge's error
                               ratio (-1: not judged, NaN: outputs not finite) */
  uint32_t cache_fallbacks; /* batches of this call rendered again by the
                               plain tracer because the judge failed them  */
  /* radiance cache, whole call (HIP events around the network kernels):  */
  double cache_train_ms;    /* training steps (forward, backward, weight
                               gradients, Adam)                            */
  double cache_infer_ms;    /* inference over the queries and the judge's rows */
  uint64_t cache_train_samples;  /* samples the training steps processed   */
  uint64_t cache_infer_samples;  /* samples inference processed            */
  uint32_t cache_plain_batches;  /* batches of this call rendered without
                                    queries because the network had not
                                    passed the judge (new, or failed last)  */
} pt_render_stats;

typedef struct pt_accel_stats {
  uint32_t num_blas;
  uint32_t num_instances;
  uint64_t num_primitives;
  uint64_t num_nodes;       /* 128-byte BVH4 nodes (all levels)             */
  uint64_t bytes;           /* device bytes of nodes + primitive records    */
  double build_ms;
} pt_accel_stats;

/* Ray and hit records of the query entry points. */
typedef struct pt_ray {
  float org[3];
  float tmin;
  float dir[3]; /* not normalised; t is in units of |dir| */
  float tmax;
} pt_ray;

typedef struct pt_hit {
  float t;      /* +inf on a miss                                 */
  int32_t inst; /* instance index, -1 on a miss                   */
  int32_t prim; /* primitive index of that kind (sphere/rect/box
                   index in the scene arrays, or triangle index in
                   the mesh); -1 on a miss                         */
  uint32_t reserved;
} pt_hit;

/* ------------------------------------------------------------ entry points */
/* Context::create() (host_includes/host_common.hpp:24). */
pt_status pt_create(int32_t device, struct pt_context** out);
/* Context destruction; frees all device memory. */
void pt_destroy(struct pt_context* ctx);
/* Message of the last failing call on ctx (or of a failed pt_create when
 * ctx is NULL).  Valid until the next call on ctx. */
const char* pt_last_error(const struct pt_context* ctx);

/* Scene graph assembly (hitables.hpp:518 addElementsTo, transforms.hpp:400
 * addAndTransform, programs.hpp:17-92 light table and miss program).
 * Deep copy; invalidates the acceleration structure. */
pt_status pt_set_scene(struct pt_context* ctx, const pt_scene_desc* scene);

/* createAcceleration("Trbvh") + the warm-up launch renderFrame(ctx,0,0)
 * (scenes.hpp:31-32, main.cpp:99): GPU LBVH build of every bottom-level
 * structure and the top level. */
pt_status pt_build_accel(struct pt_context* ctx, pt_accel_stats* out_or_null);

/* Camera::set (camera.hpp:40-51). */
pt_status pt_set_camera(struct pt_context* ctx, const pt_camera* cam);

/* renderFrame / Context::launch(0, W, H) for frames
 * [frame_begin, frame_begin + frame_count) (main.cpp:26-39, :264-265).
 * frame_begin == 0 clears the accumulation (raygen.cu:154).  Asynchronous;
 * stats (if given) are filled after an internal synchronisation. */
pt_status pt_render(struct pt_context* ctx, const pt_render_params* params,
                    uint32_t frame_begin, uint32_t frame_count,
                    pt_render_stats* out_or_null);

/* No counterpart in the reference (it renders one scene per OptiX context;
 * this replaces its per-scene launch loop, main.cpp:26-39,264-265):
 * frames [0, frame_counts[i]) of every ctxs[i], i < n <= 256, in one
 * wavefront.  The (scene, frame) work items are cut into sub-batches of at
 * most the automatic batch size of ctxs[0] (see frames_per_batch); within a
 * sub-batch all paths share one path pool and
 * one launch per kernel and bounce, so many small or uneven scenes fill the
 * GPU together.  Every context needs a built scene and a camera, all on one
 * device; ctxs[0] runs the batch.  Each context's accumulation restarts at
 * frame 0 and ends equal, bit for bit, to pt_render(ctxs[i], &params[i],
 * 0, frame_counts[i]).  The radiance cache is not available in a batch.
 * Batch-wide flags are taken from params[0].flags alone (the other scenes'
 * values of these bits are ignored): PT_FLAG_MATERIAL_SORT,
 * PT_FLAG_COUNT_TRAVERSAL, PT_FLAG_GRAPH (the whole batch -- clears,
 * sub-batches, bounce loops, accumulations -- captured once into a graph
 * and replayed while contexts, parameters, frame counts, scenes and
 * buffers are unchanged) and PT_FLAG_TUNED (the shape pt_autotune_batch
 * picked for this batch's contexts and parameters, if any).  Per-scene
 * flags: PT_FLAG_ENV_MIS, max_depth, Russian roulette, resolution.  stats
 * (if given) cover the whole batch; this call synchronises when they are
 * asked for, else it is asynchronous on ctxs[0]'s stream. */
pt_status pt_render_batch(struct pt_context* const* ctxs, const pt_render_params* params,
                          const uint32_t* frame_counts, uint32_t n, pt_render_stats* stats_or_null);

pt_status pt_sync(struct pt_context* ctx);

/* accBuffer->map() (image_save.hpp:11): W*H*3 doubles, row 0 = top of the
 * image; *frames = number of accumulated frames. */
pt_status pt_read_accum(struct pt_context* ctx, double* rgb, uint32_t* frames);

/* Save_PNG / Save_HDR (image_save.hpp:7-73) of the accumulation average:
 * PT_IMAGE_PNG writes 8-bit RGB, 255.99 * clamp(sqrt(c), 0, 1) truncated;
 * PT_IMAGE_HDR writes Radiance RGBE of c / (1 + c); rows top first. */
#define PT_IMAGE_PNG 0
#define PT_IMAGE_HDR 1
pt_status pt_save_image(struct pt_context* ctx, const char* path, int32_t format);

/* displayBuffer->map() (main.cpp:269): W*H*4 bytes, make_Color of the
 * accumulation (raygen.cu:136-146). */
pt_status pt_read_display(struct pt_context* ctx, uint8_t* rgba);

/* Per-path result of the last rendered frame (needs PT_FLAG_KEEP_PATHS):
 * rgb = W*H*3 floats of color() before de_nan, indexed by launch index
 * (x + W*y); signature = W*H hashes of the path's discrete decisions;
 * bounces = W*H bounce counts.  Any pointer may be NULL. */
pt_status pt_read_paths(struct pt_context* ctx, float* rgb,
                        uint32_t* signature, uint32_t* bounces);

/* rtTrace of a radiance ray (raygen.cu:77): closest hit in (tmin, tmax),
 * ties broken by (instance, primitive).  Host arrays. */
pt_status pt_trace_closest(struct pt_context* ctx, const pt_ray* rays,
                           uint64_t n, pt_hit* hits);
/* Same on device-resident arrays, asynchronous on the context stream. */
pt_status pt_trace_closest_device(struct pt_context* ctx, const pt_ray* d_rays,
                                  uint64_t n, pt_hit* d_hits);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): pt_trace_closest_device without any
 * allocation.  Its traversal-stack spill lives in a fixed per-device
 * scratch array of the library (48 MiB + 256 B, held by every GPU the
 * library runs on until the process ends); the launch uses at most 384
 * blocks.  n <= PT_NOALLOC_MAX_RAYS, else PT_ERR_INVALID_ARG.  Asynchronous
 * on the context stream; if d_overflow is not null, the number of
 * traversal-stack overflows of this call (0 for a valid result) is copied
 * there (device memory) in stream order.  Calls on one GPU share its
 * scratch: calls that overlap on one GPU (different streams or threads)
 * race; serialise them. */
#define PT_NOALLOC_MAX_RAYS (1ull << 31)
pt_status pt_trace_closest_device_noalloc(struct pt_context* ctx, const pt_ray* d_rays, uint64_t n,
                                          pt_hit* d_hits, uint32_t* d_overflow_or_null);
/* Counted variant: per-ray node, primitive and instance fetch counts
 * (3 uint32 per ray), host arrays. */
pt_status pt_trace_closest_counted(struct pt_context* ctx, const pt_ray* rays,
                                   uint64_t n, pt_hit* hits, uint32_t* counts);

/* Shadow rtTrace (light_sample.cuh:61-72) under the deterministic rule:
 * lit (0) iff the closest hit in (tmin, tmax) is a diffuse light or there is
 * no hit; occluded (1) otherwise.  Host arrays. */
pt_status pt_trace_occluded(struct pt_context* ctx, const pt_ray* rays,
                            uint64_t n, uint8_t* occluded);

/* Copies of the built acceleration structure for bit-exact checks.
 * Pass NULL buffers to query sizes.  nodes: 32 uint32 per 128-byte node;
 * prim_order: sorted primitive order of every bottom-level structure
 * concatenated; morton: the 30-bit codes in sorted order. */
pt_status pt_read_accel(struct pt_context* ctx, uint32_t* nodes,
                        uint64_t* num_nodes, uint32_t* prim_order,
                        uint32_t* morton, uint64_t* num_prims,
                        uint32_t* blas_info, uint32_t* num_blas);

/* Copies of the environment importance-sampling tables built for an
 * environment light (no reference counterpart): cond = height x width
 * inclusive prefix sums of 64-bit fixed-point texel weights per row, marg =
 * prefix sums of the row totals; width = height = total = 0 without an
 * environment light.  Any pointer may be NULL. */
pt_status pt_read_env_tables(struct pt_context* ctx, uint64_t* cond, uint64_t* marg,
                             uint32_t* width, uint32_t* height, uint64_t* total);

/* ------------------------------------------------------------ autotuning
 * No counterpart in the reference.  pt_autotune renders `frames` frames of
 * the current scene under candidate launch shapes (paths in flight 1, 4,
 * 16, 64, 256 and 512 Mi, at most the automatic batch size, in whole
 * frames; material sort on/off, batch graphs on/off; a warm-up render and
 * the faster of two timed ones each), keeps the fastest, and
 * pt_render uses it for renders with PT_FLAG_TUNED and the same resolution,
 * depth and flags.  Results are identical under every shape. */
typedef struct pt_tune_result {
  uint32_t frames_per_batch;  /* pt_autotune only (0 from pt_autotune_batch) */
  int32_t material_sort;
  int32_t graph;
  uint32_t candidates;
  double ms_per_frame;        /* the fastest candidate's time per frame     */
  uint64_t paths_per_batch;   /* paths in flight per batch                  */
} pt_tune_result;
pt_status pt_autotune(struct pt_context* ctx, const pt_render_params* params, uint32_t frames,
                      pt_tune_result* out_or_null);

/* No counterpart in the reference.  The same for a batch of scenes: renders
 * the batch (pt_render_batch semantics, with the given frame counts, which
 * may be fewer than the production render's) under 8, 16, 32, 64, 128 and
 * 256 Mi paths per sub-batch (at most the automatic batch size) x material
 * sort on/off (when the batch has more than
 * one shading program) x graph replay on/off, twice each, and keeps the
 * fastest second run in ctxs[0] for pt_render_batch calls whose params[0]
 * has PT_FLAG_TUNED and whose contexts and parameters (other than the
 * shape's own flags and frames_per_batch) are the same.  Every context's
 * accumulation is left holding the tuning render. */
pt_status pt_autotune_batch(struct pt_context* const* ctxs, const pt_render_params* params,
                            const uint32_t* frame_counts, uint32_t n, pt_tune_result* out_or_null);

/* ------------------------------------------------- learned radiance cache
 * No counterpart in the reference.  A 96 -> 64 -> 64 -> 64 -> 64 -> 3 MLP
 * with ReLU hidden layers (PT_CACHE_PARAMS parameters in the order W1 (64 x
 * 96, row-major), b1, W2, b2, W3, b3, W4, b4, W5 (3 x 64), b5), FP32 master
 * weights, BF16 MFMA forward and backward with FP32 accumulation, Adam
 * (beta1 0.9, beta2 0.999, eps 1e-8).  Features are BF16 bit patterns,
 * PT_CACHE_INPUTS per sample; outputs and targets are PT_CACHE_OUTPUTS
 * floats per sample: outputs y stand for radiance r = expm1(min(y, 20)),
 * targets are radiance.  Loss: mean over samples, summed over the outputs,
 * of the relative L2 (r - t)^2 / (r^2 + 0.01) with the denominator held
 * constant (Mueller et al. 2021): its minimiser over noisy targets of one
 * input is their mean. */
#define PT_CACHE_INPUTS 96
#define PT_CACHE_OUTPUTS 3
#define PT_CACHE_PARAMS 18883
/* (Re)initialise the network: He-uniform weights from mt19937(seed), zero
 * biases, zero Adam state. */
pt_status pt_cache_init(struct pt_context* ctx, uint32_t seed);
/* Copy the FP32 parameters out (out != NULL) and/or replace them (in != NULL). */
pt_status pt_cache_params(struct pt_context* ctx, float* out, const float* in);
/* Network outputs for n samples (host arrays). */
pt_status pt_cache_infer(struct pt_context* ctx, const uint16_t* features, uint64_t n, float* out);
/* One Adam step on one batch (host arrays); loss and the summed gradient
 * (PT_CACHE_PARAMS floats) are returned when the pointers are non-NULL. */
pt_status pt_cache_train(struct pt_context* ctx, const uint16_t* features, const float* targets,
                         uint64_t n, float lr, float* loss, float* grad);

/* Render-loop use of the cache (PT_FLAG_RADIANCE_CACHE): paths not chosen
 * for training end at bounce `query_depth` (0 = the camera ray's hit) with
 * L += T * expm1(min(network, 20)) (negative and NaN answers count 0); a
 * `train_fraction` of the paths (hashed on pixel and frame) are traced to
 * the end by the plain tracer and record every vertex; after each batch of
 * frames the network takes `steps` Adam steps of up to `batch` vertices
 * each (steps x batch vertices spread evenly over the batch's records,
 * each used once; when the batch recorded fewer, the steps cycle through
 * all of them) with learning rate `lr`, is judged
 * (pt_cache_judge_configure), then answers the batch's queries.
 * Defaults: 2, 0.03, 16384, 4, 1e-3. */
pt_status pt_cache_configure(struct pt_context* ctx, uint32_t query_depth, float train_fraction,
                             uint32_t batch, uint32_t steps, float lr);
/* The per-batch judge of the cache's error.  The plain tracer's training
 * paths give, at `rows` of the batch's vertices (halfway between the ones
 * trained on), an unbiased one-sample radiance t; with the network's answer
 * c and the throughput T there,
 *   ratio = sum T^2 (c - t)^2 / sum T^2 t^2,
 * the squared error the answers would add to the image over that of ending
 * the paths without one.  A batch whose ratio is >= max_ratio, or whose
 * judged outputs are not finite, is rendered again by the plain tracer (its
 * accumulation is then the plain render's, bit for bit), and the batches
 * after it render plain (no queries; the training paths still train and
 * the network is judged after each) until the network passes again; a new
 * network (pt_cache_init) starts held the same way.  A network with
 * non-finite outputs is re-initialised from the pt_cache_init seed.
 * Defaults: max_ratio 1, rows 16384 (32 .. 2^20). */
pt_status pt_cache_judge_configure(struct pt_context* ctx, float max_ratio, uint32_t rows);
/* Records of the last render batch, for checks: which = 0 training
 * vertices (values: L before the vertex, T at it, target radiance),
 * 1 queries (values: T, 0, 0, 0, network output).  9 floats and one slot
 * per record, at most `max`; *count = records the batch produced. */
pt_status pt_cache_read_records(struct pt_context* ctx, int32_t which, uint32_t max, uint16_t* features,
                                float* values, uint32_t* slots, uint32_t* count);

/* Version of this ABI (PT_API_VERSION). */
int32_t pt_api_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PT_API_H */
