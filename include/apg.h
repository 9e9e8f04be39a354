/*
 * apg.h -- C ABI of the MI355X (gfx950) ray-tracing back end for alpine.
 *
 * This is the drop-in boundary for alpine's two plug-in interfaces
 *   Accelerator   (reference: src/accelerators/accelerator.h:13-39)
 *   LightSampler  (reference: src/light_samplers/light_sampler.h:12-25)
 * Every entry point below names the reference member it replaces.  The per-ray
 * reference calls (intersect(const Ray&), intersectAny(const Ray&, float)) map
 * to the batched calls with n = 1; the render path submits whole queues.
 *
 * Conventions
 *  - Plain pointers and sizes only; no C++ or torch types.
 *  - Every call returns an int status: APG_OK (0) or a negative APG_E_* code.
 *    apg_last_error(ctx) returns a message for the most recent failure on the
 *    context (apg_last_error(NULL): the most recent failure of apg_create).
 *  - "device" pointers are HBM allocations on the context's GPU; "host"
 *    pointers are ordinary process memory.  `stream` is a hipStream_t (NULL =
 *    the default stream).  Batched calls are asynchronous on `stream`.
 *  - Calls that launch kernels may be issued from several host threads;
 *    the library serialises their kernels on the GPU in call order.
 *  - Triangle indices in results are *global*: meshes are numbered in append
 *    order and their triangles concatenated (global = prim_offset(shape) +
 *    mesh-local primId, the reference's Bvh layout, bvh.cpp:83-106).
 */
#ifndef APG_H
#define APG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define APG_OK 0
#define APG_E_INVALID (-1)     /* bad argument */
#define APG_E_NO_DEVICE (-2)   /* no usable GPU: behaves like wide_bvh_dummy.cpp */
#define APG_E_HIP (-3)         /* HIP runtime error */
#define APG_E_STATE (-4)       /* call out of order (e.g. intersect before build) */
#define APG_E_CAPACITY (-5)    /* size beyond a documented limit */
#define APG_E_OVERFLOW (-6)    /* traversal stack overflow reported by a kernel: a ray needed more
                                * stack entries than the launch allowed and its result is not
                                * reliable (the reference leaves this unchecked, bvh.cpp:230-231) */

#define APG_INVALID_ID 0xFFFFFFFFu /* Accelerator::INVALID_SHAPE_ID, accelerator.h:38 */

typedef struct apg_context apg_context;

/* Rays in HBM, structure of arrays, n entries each (28 B per ray).
 * tmax may be NULL: closest-hit then uses FLT_MAX (bvh.cpp:230-233). */
typedef struct apg_rays {
    const float* ox; const float* oy; const float* oz;
    const float* dx; const float* dy; const float* dz;
    const float* tmax;
} apg_rays;

/* Closest-hit results in HBM (16 B per ray): prim = APG_INVALID_ID on a miss
 * (t = FLT_MAX, b0 = b1 = 0), otherwise the global triangle index, the hit
 * distance and the barycentrics weighting vertices 1 and 2 (the reference
 * Intersection fields primId, t, barycentric; intersection.h:8-16). */
typedef struct apg_hits {
    uint32_t* prim; float* t; float* b0; float* b1;
} apg_hits;

/* Which builder and which tree shape scene_build produces. */
typedef struct apg_build_params {
    uint32_t leaf_threshold;  /* binned-SAH leaf threshold (reference Bvh: 8); 0 = default 1 */
    uint32_t builder;         /* APG_BUILDER_GPU (default) or APG_BUILDER_HOST */
    uint32_t roots;           /* APG_ROOTS_SCENE: one tree over all shapes (the reference's
                               * updateScene); APG_ROOTS_PER_SHAPE: an independent tree per shape,
                               * built together, wide node s the root of shape s (GPU builder;
                               * rays then name their tree, apg_intersect_roots) */
    uint32_t flags;           /* APG_BUILD_NO_STATS: skip the host-side tree statistics and
                               * leaf keys (per-frame rebuilds) */
    uint32_t reserved[4];
} apg_build_params;
#define APG_BUILDER_GPU 0u
#define APG_BUILDER_HOST 1u
#define APG_ROOTS_SCENE 0u
#define APG_ROOTS_PER_SHAPE 1u
#define APG_BUILD_NO_STATS 1u

typedef struct apg_scene_stats {
    uint32_t triangles;
    uint32_t shapes;
    uint32_t binary_nodes;     /* nodes of the binary binned-SAH tree */
    uint32_t binary_leaves;
    uint32_t wide_nodes;       /* 64 B quantised 8-wide nodes */
    uint32_t max_depth_binary;
    double sah_cost;           /* binary tree, C_trav = C_isect = 1, / SA(root) */
    double sah_cost_leaf;      /* binary tree, leaf term only */
    double build_ms;           /* last scene_build, wall clock */
    uint64_t device_bytes;     /* nodes + triangles + index maps */
    uint32_t wide_depth;       /* levels of 8-wide nodes on the longest root-to-leaf path; a
                                * traversal holds at most one stack entry per level, so builds
                                * deeper than APG_TRACE_STACK_MAX fail with APG_E_CAPACITY */
    uint32_t reserved;
} apg_scene_stats;

/* Traversal stack entries per ray (the reference's STACK_SIZE, bvh_common.h:15). */
#define APG_TRACE_STACK_MAX 64u

/* Traversal launch shape (filled with defaults by apg_trace_config_default;
 * the autotuner in alpine_amd/autotune.py writes these). */
typedef struct apg_trace_config {
    uint32_t block_size;       /* threads per workgroup (multiple of 64) */
    uint32_t blocks_per_cu;    /* persistent grid = CUs * blocks_per_cu */
    uint32_t fetch_threshold;  /* idle lanes a wave waits for before handing out rays */
    uint32_t stack_lds;        /* traversal stack entries per lane kept in LDS: 8 or 16 */
    uint32_t tri_batch;        /* queued triangle tests per wave that start a round: 1..64 */
    uint32_t triangle_test;    /* APG_TRI_REFERENCE (default): the reference's Moller-Trumbore,
                                * bit-exact hits (bvh_common.cpp:66-103); APG_TRI_WATERTIGHT: the
                                * watertight test of Woop, Benthin and Wald (JCGT 2013) on the input
                                * vertices, no ray passes between triangles sharing an edge */
    uint32_t stack_limit;      /* 0: APG_TRACE_STACK_MAX entries; 1..63 caps the stack (test hook:
                                * a scene deeper than the cap makes traversals report
                                * APG_E_OVERFLOW, as a tree deeper than the stack would) */
    uint32_t reserved;
} apg_trace_config;
#define APG_TRI_REFERENCE 0u
#define APG_TRI_WATERTIGHT 1u

/* ---- lifetime (Alpine::resetScene creates the accelerator, alpine.cpp:158-176) ---- */
int apg_create(apg_context** ctx, int device);
int apg_destroy(apg_context* ctx);
const char* apg_last_error(const apg_context* ctx);

/* ---- geometry (Accelerator::appendMesh / getVertexBuffer / updateShape / updateScene,
 *      accelerator.h:19-31) ---- */
/* Copies nv vertices (xyz, host) and nt triangles (3 x uint32, host).  A NaN or infinite
 * coordinate, or an index >= nv, is rejected with APG_E_INVALID (the reference only asserts
 * its capacity, bvh.cpp:89,101, and would build boxes of NaN). */
int apg_mesh_append(apg_context* ctx, const float* xyz, uint32_t nv, const uint32_t* idx,
                    uint32_t nt, uint32_t* shape_id);
/* Mutable host view of the library's copy of the shape's vertices
 * (Accelerator::getVertexBuffer).  Write it, then call apg_mesh_update. */
int apg_mesh_vertices(apg_context* ctx, uint32_t shape_id, float** xyz, uint32_t* nv);
/* updateShape: the vertex copy is checked for NaN / infinite coordinates (APG_E_INVALID). */
int apg_mesh_update(apg_context* ctx, uint32_t shape_id);
int apg_mesh_prim_offset(const apg_context* ctx, uint32_t shape_id, uint32_t* offset);
/* Rigid transform of vertices [first, first + count) of a shape, applied on the GPU to the
 * library's device copy (no upload); m is 3 x 4 row-major (host).  The next
 * apg_scene_build rebuilds from the device copy; apg_mesh_vertices downloads it first.
 * No reference counterpart: it replaces writing through getVertexBuffer + updateShape
 * (mesh.cpp:29-55) for per-frame animation without a host round trip. */
int apg_mesh_transform(apg_context* ctx, uint32_t shape_id, const float* m, uint32_t first,
                       uint32_t count, void* stream);
/* Shading attributes of a mesh: the Mesh::Data fields normals, tangents, uvs and their index
 * buffers (shapes/mesh.h:16-33) that Mesh::getIntersectionAttributes interpolates at a hit
 * (shapes/mesh.cpp:57-154).  The render path shades with them; traversal ignores them.
 * All arrays are host memory and are copied. */
typedef struct apg_mesh_shading {
    const float* normals;        /* nn x 3, or NULL: the geometric normal shades (mesh.cpp:147-150) */
    uint32_t nn;
    const uint32_t* normal_idx;  /* nt x 3 indices into normals, or NULL: the position indices */
    const float* tangents;       /* nn x 4 (xyz, handedness w), or NULL: no tangent space */
    const float* uvs;            /* nuv x 2, or NULL: uv = (0, 0) */
    uint32_t nuv;
    const uint32_t* uv_idx;      /* nt x 3 indices into uvs, or NULL: the position indices */
} apg_mesh_shading;
int apg_mesh_set_shading(apg_context* ctx, uint32_t shape_id, const apg_mesh_shading* shading);
/* Morph targets of a mesh (Mesh::Data::Target, shapes/mesh.h:27-32): ntargets x nv x 3 position
 * deltas, and ntargets x nn x 3 normal and tangent deltas (NULL: zero).  The shape's current
 * vertices become the base the position deltas are added to (Mesh::Data::vertices).
 * ntargets = 0 removes them.  At most APG_MAX_MORPH_TARGETS targets. */
#define APG_MAX_MORPH_TARGETS 64u
int apg_mesh_set_targets(apg_context* ctx, uint32_t shape_id, uint32_t ntargets, const float* dpos,
                         const float* dnormals, const float* dtangents);
/* Mesh::update (shapes/mesh.cpp:29-55) on the GPU: vertex i = base_i + sum_j dpos_j,i * w_j
 * written into the library's device copy (no upload), and the shading normals and tangents use
 * the same weights (mesh.cpp:88-124).  weights: ntargets floats (host), the caller's
 * std::lerp of the two keyframes (Animation::update, animation.cpp:10-39).  The next
 * apg_scene_build rebuilds from the morphed vertices. */
int apg_mesh_morph(apg_context* ctx, uint32_t shape_id, const float* weights, void* stream);
/* Appending a sphere is rejected like the reference Bvh (bvh.cpp:108-112). */
int apg_sphere_append(apg_context* ctx, const float* xyzr, uint32_t n);
int apg_scene_build(apg_context* ctx, const apg_build_params* params); /* updateScene */
int apg_scene_get_stats(const apg_context* ctx, apg_scene_stats* stats);

/* ---- traversal (Accelerator::intersect / intersectAny, accelerator.h:33-35) ---- */
int apg_trace_config_default(const apg_context* ctx, apg_trace_config* cfg);
int apg_set_trace_config(apg_context* ctx, const apg_trace_config* cfg);
/* Closest hit for n rays (device SoA in, device SoA out). */
int apg_intersect(apg_context* ctx, const apg_rays* rays, uint32_t n, const apg_hits* hits,
                  void* stream);
/* apg_intersect without any allocation: no reference counterpart.  For callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels and results as apg_intersect, but the ray cursor and the traversal stack's spill
 * live in one file-scope device array per GPU (24 MiB + 256 B, set up by apg_create) instead of
 * context buffers, and the stack holds the LDS top (cfg.stack_lds, 8 for the watertight test)
 * plus 8 spilled entries: a scene whose 8-wide tree has more levels than that plus one is
 * refused with APG_E_CAPACITY (apg_scene_stats.wide_depth; 17 levels with the default
 * configuration).  Every call on a GPU shares that GPU's array, so calls that overlap on one
 * GPU (other streams, other contexts, other threads) race: the caller serialises them.  It does
 * not wait for other traversals and may be captured into a graph; stack_limit is ignored. */
int apg_intersect_noalloc(apg_context* ctx, const apg_rays* rays, uint32_t n, const apg_hits* hits,
                          void* stream);
/* Any hit with t < tmax[i] (FLT_MAX when tmax is NULL); occluded[i] = 0 or 1. */
int apg_occluded(apg_context* ctx, const apg_rays* rays, uint32_t n, uint8_t* occluded,
                 void* stream);
/* Wait for the context's traversals issued so far and report whether any of them overflowed its
 * stack since the last report: APG_OK, or APG_E_OVERFLOW (the flag is then cleared).  The calls
 * that wait for the GPU themselves (apg_intersect_host, apg_occluded_host, apg_intersect_counted,
 * apg_render_read, apg_render_read_aovs) report an overflow the same way.  No reference
 * counterpart: the reference's traversal asserts in debug builds only (bvh.cpp:230-231). */
int apg_trace_status(apg_context* ctx);
/* Host-memory, synchronous variants for per-ray callers: Accelerator::intersect(const Ray&) and
 * intersectAny(const Ray&, float tFar) (accelerator.h:33-35), called one ray at a time from up to
 * 256 tile threads (alpine.cpp:276-285).  rays6: n rays in alpine's Ray layout (org xyz, dir xyz;
 * ray.h:6-10), host memory; results in host memory as apg_intersect / apg_occluded write them.
 * Thread-safe: each concurrently calling thread stages through its own stream and pinned buffers. */
int apg_intersect_host(apg_context* ctx, const float* rays6, uint32_t n, uint32_t* prim, float* t,
                       float* b0, float* b1);
int apg_occluded_host(apg_context* ctx, const float* rays6, const float* tmax, uint32_t n,
                      uint8_t* occluded);
/* Instrumented closest/any-hit run: same results, plus total 64 B node
 * fetches and 36 B triangle fetches over all rays (host uint64 out).  Used for
 * the byte model of the roofline. */
/* As apg_intersect / apg_occluded in a per-shape forest (APG_ROOTS_PER_SHAPE): ray i is traced
 * in the tree of shape roots[i] (device array).  No reference counterpart: the reference keeps
 * one scene per accelerator; many independent scenes share one launch this way. */
int apg_intersect_roots(apg_context* ctx, const apg_rays* rays, const uint32_t* roots, uint32_t n,
                        const apg_hits* hits, void* stream);
int apg_occluded_roots(apg_context* ctx, const apg_rays* rays, const uint32_t* roots, uint32_t n,
                       uint8_t* occluded, void* stream);
int apg_intersect_counted(apg_context* ctx, const apg_rays* rays, uint32_t n,
                          const apg_hits* hits, int any, uint8_t* occluded,
                          uint64_t* node_fetches, uint64_t* tri_fetches, void* stream);
/* The triangle data the kernels read, in tree order: 9 floats per triangle
 * (v0, e0 = v1 - v0, e1 = v2 - v0; bvh_common.cpp:105-125) and the global
 * index of each.  Host copies out, for tests. */
int apg_scene_triangles(const apg_context* ctx, float* tri9, uint32_t* ids, uint32_t n);
/* The binary tree's leaf sets: for each global triangle, the smallest global
 * triangle index in its leaf (canonical form; host out). */
int apg_scene_leaf_keys(const apg_context* ctx, uint32_t* keys, uint32_t n);
/* Raw 64 B wide nodes (host out), for structural tests. */
int apg_scene_wide_nodes(const apg_context* ctx, void* nodes64, uint32_t n);

/* ---- shading data (Material, Light, LightSampler; materials/material.h:12-29,
 *      lights/light.h:13-46, light_samplers/light_sampler.h:12-25) ---- */
#define APG_MAT_DIELECTRIC 0u  /* Lambert + GGX coat F0 0.04 when |alpha|^2 < 1 (dielectric.cpp) */
#define APG_MAT_CONDUCTOR 1u   /* GGX, F0 = base colour (conductor.cpp) */
typedef struct apg_material {
    uint32_t kind;
    float base_color[3];
    float alpha[2];
    uint32_t base_color_tex;   /* 0: base_color; k: texture k - 1 (Material::getBaseColor) */
    uint32_t normal_tex;       /* 0: none; k: tangent-space normals of texture k - 1 (getNormal) */
} apg_material;

/* Textures (Texture4f, texture.h:11-42): texture k is sizes[2k] x sizes[2k+1] float4 texels,
 * row-major, linear values (the loader's conversions already applied, gltf_loader.cpp:417-460);
 * texels holds all textures back to back (host).  Sampling takes the nearest texel of uv
 * wrapped to [0, 1)^2.  ntex = 0 removes them. */
int apg_textures_set(apg_context* ctx, const float* texels, const uint32_t* sizes, uint32_t ntex);

#define APG_LIGHT_DISK 0u      /* disk_light.cpp */
#define APG_LIGHT_POINT 1u     /* point_light.cpp */
#define APG_LIGHT_QUAD 2u      /* one-sided parallelogram p0 + a (p1 - p0) + b (p2 - p0), radiance = color */
#define APG_LIGHT_TRIANGLE 3u  /* one-sided triangle p0 p1 p2, radiance = color */
#define APG_LIGHT_POINT_INTENSITY 4u /* point light of intensity = color (point_light.cpp:8-13, the
                                      * glTF punctual lights, gltf_loader.cpp:460-472) */
typedef struct apg_light {
    uint32_t kind;
    float p0[3], p1[3], p2[3];
    float normal[3];           /* disk */
    float color[3];            /* disk/point: colour (x power); quad/triangle: radiance */
    float power;               /* disk/point */
    float radius;              /* disk */
} apg_light;

#define APG_LS_UNIFORM 0u      /* uniform_light_sampler.cpp */
#define APG_LS_POWER 1u        /* power_light_sampler.cpp + alias_table.cpp */
#define APG_LS_BVH 2u          /* bvh_light_sampler.cpp (5-bin light BVH) */

/* Material of every global triangle (Mesh per-primitive material, mesh.cpp:76-79). */
int apg_materials_set(apg_context* ctx, const apg_material* mats, uint32_t nmats,
                      const uint32_t* tri_material, uint32_t ntris);
/* Lights and the sampler over them (Alpine::buildLightSampler, alpine.cpp:238-250). */
int apg_lights_set(apg_context* ctx, const apg_light* lights, uint32_t n, uint32_t sampler);
/* LightSampler::sample for n queries (device): u[n], hit/ns as n x 3 floats; light = -1 when
 * none (the reference's empty optional). */
int apg_light_sampler_sample(apg_context* ctx, const float* u, const float* hit,
                             const float* ns, uint32_t n, int32_t* light, float* pdf,
                             void* stream);
/* The LightSampler constructors (uniform_light_sampler.cpp, power_light_sampler.cpp:6-20,
 * bvh_light_sampler.cpp:167-240) over lights given by what they read of each Light: its power
 * (Light::getPower, 3 floats) and bounds (Light::getBoundingBox, lo xyz then hi xyz), host arrays.
 * Enough for sampling; a render needs the lights' geometry (apg_lights_set), which this replaces. */
int apg_light_sampler_set(apg_context* ctx, const float* power, const float* bounds, uint32_t n,
                          uint32_t sampler);
/* LightSampler::sample (light_sampler.h:24) for n queries in host memory, synchronous and
 * thread-safe like apg_intersect_host; light = -1 for the empty optional. */
int apg_light_sampler_sample_host(apg_context* ctx, const float* u, const float* hit,
                                  const float* ns, uint32_t n, int32_t* light, float* pdf);
/* The probability sample() returns `light` at (hit, ns): no reference counterpart; MIS with a
 * third strategy needs it. */
int apg_light_sampler_pdf(apg_context* ctx, const int32_t* light, const float* hit,
                          const float* ns, uint32_t n, float* pdf, void* stream);

/* ---- wavefront path tracer (Alpine::render, alpine.cpp:271-432) ---- */
typedef struct apg_camera {
    float eye[3], target[3], up[3];
    float fovy, aspect;        /* Camera::setLookAt, camera.cpp:13-29 */
} apg_camera;

typedef struct apg_render_params {
    uint32_t spp;              /* samples per pixel in this call */
    uint32_t max_depth;        /* path vertices (alpine.cpp:314) */
    uint32_t rr_depth;         /* 0: no Russian roulette; else from this depth on */
    uint32_t fp64_accum;       /* also accumulate in double */
    float background[3];       /* radiance on a miss (alpine.cpp:318-324) */
    uint32_t guide;            /* 0: plain; 1: guided + online training; 2: guided, frozen */
    uint32_t train_steps;      /* guide = 1: after each pass, one epoch over its radiance samples
                                * in this many Adam minibatch steps; 0 = 4 */
    uint32_t guide_warmup;     /* passes traced before the proposal is mixed in (guide = 1) */
    float guide_frac;          /* probability of sampling the proposal; 0 = 0.5 */
    uint32_t graph;            /* 1: capture a pass into a HIP graph once and replay it (plain or
                                * guide = 2); captured again when the scene, frame or params change */
    uint32_t train_samples;    /* guide = 1: radiance samples kept per pass for training, spread
                                * over the frame; 0 = 262144 */
    uint32_t guide_depth;      /* guide = 2: path vertices below this depth use the proposal, deeper
                                * ones BSDF sampling only (no inference); 0 = all */
    uint32_t aovs;             /* 1: also accumulate the first hit's base colour and shading normal
                                * (the albedo / normal render targets, alpine.cpp:333-337) */
    uint32_t count_fetches;    /* 1: trace with the instrumented traversal and accumulate its node /
                                * triangle fetch counts (apg_render_fetch_counts); no counterpart */
} apg_render_params;

/* Allocate per-pixel samplers (seeded with pixel index + seed, alpine.cpp:259-269) and clear
 * the accumulation (Alpine::resetAccumulation). */
int apg_render_reset(apg_context* ctx, uint32_t width, uint32_t height, uint32_t seed);
/* Launch shape of the wavefront stages (no reference counterpart; the autotuner
 * alpine_amd/autotune.py sets it): persistent 256-thread workgroups per CU of the shade,
 * shadow-visibility and resolve stages, each 1..64; 0 keeps that stage's default
 * (8, 4, 4).  Applies from the next apg_render call; a captured pass is captured again. */
typedef struct apg_render_launch
{
    uint32_t shade_blocks_per_cu;
    uint32_t shadow_blocks_per_cu;
    uint32_t resolve_blocks_per_cu;
    uint32_t reserved;
} apg_render_launch;
int apg_render_set_launch(apg_context* ctx, const apg_render_launch* launch);
int apg_render_get_launch(const apg_context* ctx, apg_render_launch* launch);
/* As apg_render_reset with `slots` samples of every pixel traced side by side in each pass
 * (width x height x slots paths), which fills the GPU on small frames.  Slot 0 of each pixel
 * keeps the reference's per-pixel random sequence; the other slots are seeded from a hash.
 * apg_render's spp must then be a multiple of slots.  No reference counterpart. */
int apg_render_reset_slots(apg_context* ctx, uint32_t width, uint32_t height, uint32_t seed,
                           uint32_t slots);
int apg_render_set_camera(apg_context* ctx, const apg_camera* cam);
/* Tiled frames (no reference counterpart): the frame is a grid of ntiles tiles of
 * tile_width x tile_height pixels, row-major; tile t is seen through cams[t] and, with a
 * per-shape forest, traced in tree roots[t] (host arrays; roots may be NULL = tree 0).
 * ntiles = 0 returns to the single camera.  Many independent scenes render in one pass. */
int apg_render_set_tiles(apg_context* ctx, const apg_camera* cams, const uint32_t* roots,
                         uint32_t ntiles, uint32_t tile_width, uint32_t tile_height);
/* Trace params->spp samples per pixel into the accumulation (asynchronous on stream). */
int apg_render(apg_context* ctx, const apg_render_params* params, void* stream);
/* Copy the accumulated radiance sum (width x height x 3, host) and the sample count.
 * rgb64 (optional) receives the double accumulation when fp64_accum was set. */
int apg_render_read(apg_context* ctx, float* rgb, double* rgb64, uint32_t* total_spp);
/* The albedo and normal render-target sums (width x height x 3 each, host; either may be NULL)
 * of the passes traced with params.aovs = 1 since the last reset (denoiser::RenderTarget,
 * alpine.cpp:259-269, 333-337). */
int apg_render_read_aovs(apg_context* ctx, float* albedo, float* normal);
/* Environment map seen by rays that leave the scene (Alpine::loadEnvironmentMap,
 * alpine.cpp:252-257; EnvironmentMap::sample, environment_map.cpp:15-37): width x height RGB
 * floats, row-major (host, copied).  rgb = NULL returns to the background colour. */
int apg_render_set_environment(apg_context* ctx, const float* rgb, uint32_t width, uint32_t height);
/* Alpine::postProcess without denoising (alpine.cpp:446-483): mean colour, ACES fit when
 * tonemap != 0, clamp, sRGB, bytes.  frame: width x height x 3 bytes, device memory. */
int apg_render_postprocess(apg_context* ctx, int tonemap, uint8_t* frame, void* stream);
/* Ray and path counters accumulated since apg_render_reset (host out): closest-hit rays,
 * shadow rays, path vertices shaded. */
int apg_render_counters(const apg_context* ctx, uint64_t* closest_rays, uint64_t* shadow_rays,
                        uint64_t* vertices);
/* Traversal fetch counts accumulated since apg_render_reset by passes with count_fetches set
 * (host out): 64-byte node records and triangles fetched by closest-hit and by shadow rays.
 * No reference counterpart (the byte model of the render's roofline). */
int apg_render_fetch_counts(const apg_context* ctx, uint64_t* closest_nodes, uint64_t* closest_tris,
                            uint64_t* shadow_nodes, uint64_t* shadow_tris);
/* Guided rendering: training samples of the last pass and the last training loss (host). */
int apg_render_guide_stats(const apg_context* ctx, uint32_t* train_samples, float* loss);

/* ---- animated frames as one HIP graph (no reference counterpart: the reference's frame is
 * Alpine::updateScene then render, alpine.cpp:203-212 and 271-288, host-driven) ----
 * One frame: apg_mesh_transform of each listed vertex range, apg_scene_build with `build` (GPU
 * builder), apg_render_reset_slots at the current size and slots with `seed`, and apg_render
 * with `render`.  The first frame runs eagerly through those entry points; the next one is
 * captured into a HIP graph, which later frames replay (the matrices and the seed may change
 * every frame; a change of anything else the graph baked in, such as the transform list's
 * ranges, the parameters, camera, tiles, materials or launch shapes, makes that frame eager and
 * the next one a new capture).  A replayed rebuild reads nothing back: it runs the last eager
 * build's binary and wide level counts plus a margin; apg_frame_status reports a replay whose
 * tree needed more levels, and the frame after it is built eagerly.  apg_scene_get_stats keeps
 * the last eager build's figures.  Not with guide = 1 or the watertight triangle test. */
typedef struct apg_frame_transform {
    uint32_t shape_id;
    uint32_t first, count;     /* vertex range of the shape */
    float m[12];               /* row-major 3 x 4, as apg_mesh_transform */
} apg_frame_transform;
int apg_frame_render(apg_context* ctx, const apg_frame_transform* xf, uint32_t nxf,
                     const apg_build_params* build, const apg_render_params* render, uint32_t seed,
                     void* stream);
/* Waits for the last replayed frame: APG_OK, or APG_E_CAPACITY when its rebuild ran out of
 * levels (that frame's image shows the previous tree). */
int apg_frame_status(apg_context* ctx);

/* ---- learned path guiding (no reference counterpart) ----
 * An MLP (96 -> 4 x 64 -> 48, BF16 MFMA, FP32 accumulation) maps a shading point to an
 * 8-lobe von Mises-Fisher mixture over directions.  Features: 96 floats per query
 * (position normalised to the scene bounds with 12-frequency sin/cos encoding, shading
 * normal, outgoing direction; alpine_amd/csrc/guide_core.h).  Outputs: 48 floats per query
 * (40 used: per lobe logit, log kappa, mean direction; the forward entry points leave
 * floats 40..47 of each row unwritten). */
int apg_guide_param_count(void);
/* Deterministic initialisation and the Adam learning rate. */
int apg_guide_init(apg_context* ctx, uint32_t seed, float lr);
/* FP32 master parameters (host, apg_guide_param_count() floats): write != 0 sets them. */
int apg_guide_params(apg_context* ctx, float* params, int write);
/* Forward pass: feat n x 96 (device) -> out n x 48 (device). */
int apg_guide_forward(apg_context* ctx, const float* feat, uint32_t n, float* out, void* stream);
/* Forward pass with the encoding fused in: q n x 9 (device; position normalised to the scene
 * bounds, shading normal, outgoing direction) -> out n x 48 (device). */
int apg_guide_infer(apg_context* ctx, const float* q, uint32_t n, float* out, void* stream);
/* One training step on n samples (device): minimise sum_i w_i (-log q(dir_i | feat_i)) /
 * sum_i w_i.  apply = 0 computes the gradient only.  loss (host, optional) receives the
 * weighted mean negative log-likelihood; grad_host (optional) the gradient vector. */
int apg_guide_train_step(apg_context* ctx, const float* feat, const float* dir,
                         const float* weight, uint32_t n, int apply, float* loss,
                         float* grad_host, void* stream);
/* The same step from queries (n x 9 device: position normalised to the scene bounds, shading
 * normal, outgoing direction), with the encoding, forward pass, loss head and backward pass
 * through the layers fused into one kernel (the renderer's training path). */
int apg_guide_train_queries(apg_context* ctx, const float* q, const float* dir,
                            const float* weight, uint32_t n, int apply, float* loss,
                            float* grad_host, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* APG_H */
