/*
 * apg_debug.h -- test hooks of libapg.so (not part of the drop-in boundary).
 *
 * They run the host build of the library's own traversal steps
 * (alpine_amd/csrc/traverse_core.h) over a host-built scene, so CPU-only tests
 * can check the wide-node layout and the traversal logic against the oracle
 * without a GPU.  The product path (apg.h) never calls them.
 */
#ifndef APG_DEBUG_H
#define APG_DEBUG_H

#include <stdint.h>

#include "apg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct apg_debug_scene apg_debug_scene;

int apg_debug_host_build(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                         uint32_t leaf_threshold, apg_debug_scene** out);
int apg_debug_scene_info(const apg_debug_scene* s, uint32_t* n_nodes, uint32_t* n_tris,
                         double* sah_cost, double* sah_cost_leaf, uint32_t* binary_nodes,
                         uint32_t* max_depth);
int apg_debug_scene_copy(const apg_debug_scene* s, void* nodes64, float* tri9, uint32_t* ids,
                         uint32_t* leaf_keys);
/* org3/dir3: n x 3 floats (host).  tmax may be NULL (FLT_MAX).  max_stack
 * (optional) receives the deepest stack the traversals used. */
int apg_debug_host_trace(const apg_debug_scene* s, const float* org3, const float* dir3,
                         const float* tmax, uint32_t n, int any, uint32_t* prim, float* t,
                         float* b0, float* b1, uint8_t* occluded, uint32_t* max_stack);
/* As apg_debug_host_trace with the traversal stack capped at stack_limit entries (1 ..
 * APG_TRACE_STACK_MAX), pushes onto a full stack dropped as trace_kernel drops them;
 * overflowed receives the number of rays that dropped one. */
int apg_debug_host_trace_limited(const apg_debug_scene* s, const float* org3, const float* dir3,
                                 const float* tmax, uint32_t n, int any, uint32_t stack_limit,
                                 uint32_t* prim, float* t, float* b0, float* b1, uint8_t* occluded,
                                 uint32_t* overflowed);
/* Streaming read of `bytes` (a multiple of 16) of device memory at buf by `blocks` workgroups of
 * 256 threads (16-byte non-temporal loads), asynchronous on stream; sink (device, 4 bytes) is
 * written only if the data's XOR is a fixed constant.  The achievable-bandwidth measurement of
 * the roofline. */
int apg_debug_stream_read(const void* buf, uint64_t bytes, uint32_t* sink, uint32_t blocks, void* stream);
/* Per-ray callers measured: nthreads host threads, each over a contiguous share of the n rays
 * (alpine's Ray layout, 6 floats each, host), call apg_intersect_host (any = 0; prim and t
 * receive the results) or apg_occluded_host (any = 1; occluded) with rays_per_call rays per
 * call, as alpine's tile tasks call Accelerator::intersect per ray (alpine.cpp:276-285);
 * seconds receives the wall time from the threads' common start to the last one's end. */
int apg_debug_host_rate(apg_context* ctx, const float* rays6, uint32_t n, uint32_t nthreads,
                        uint32_t rays_per_call, int any, uint32_t* prim, float* t, uint8_t* occluded,
                        double* seconds);
/* Render-state device allocations of ctx (render.hip) fail with APG_E_HIP after `after` more
 * have succeeded; after < 0 restores normal allocation.  For tests of the failure paths. */
int apg_debug_fail_allocations(apg_context* ctx, int after);
/* Replayed frame graphs (apg_frame_render) captured from now on run `cut` binary levels fewer
 * than the last eager build needed (0 restores the margin): their rebuild stops short, every
 * kernel still ends normally, and apg_frame_status reports it.  For tests of that path. */
int apg_debug_frame_level_cut(apg_context* ctx, uint32_t cut);
/* Frames apg_frame_render ran eagerly, graphs it captured and frames it replayed. */
int apg_debug_frame_counts(const apg_context* ctx, uint32_t* eager, uint32_t* captures, uint32_t* replays);
/* Levels of 8-wide nodes of the scene's deepest root-to-leaf path. */
int apg_debug_scene_depth(const apg_debug_scene* s, uint32_t* wide_depth);
void apg_debug_scene_free(apg_debug_scene* s);

/* Host build of the shading steps (alpine_amd/csrc/shade_core.h), n queries each.
 * Material: sample (ok, est, wi, pdf) at (wo, u) and eval (f, pdf) at (wo, wi_eval). */
int apg_debug_material(const apg_material* m, const float* wo, const float* u,
                       const float* wi_eval, uint32_t n, uint8_t* ok, float* est, float* wi,
                       float* pdf, float* f_eval, float* pdf_eval);
/* Light sampler over `lights`: sample (light, pdf) at (u, hit, ns) and the pdf extension. */
int apg_debug_light_sampler(const apg_light* lights, uint32_t nl, uint32_t sampler,
                            const float* u, const float* hit, const float* ns, uint32_t n,
                            int32_t* light, float* pdf, float* pdf_ext);
/* Shading attributes of one mesh of nt triangles (attributes as apg_mesh_set_shading takes them;
 * ntargets normal / tangent deltas, nn x 3 per target, NULL = zero, with their weights), one
 * material for every triangle and its textures as apg_textures_set takes them.  Per query
 * (prim, b = (b1, b2), geometric normal ng): ns, ss, ts, uv and the base colour there. */
int apg_debug_shading(const uint32_t* idx, uint32_t nt, uint32_t nv, const apg_mesh_shading* sh,
                      uint32_t ntargets, const float* dnormals, const float* dtangents,
                      const float* weights, const apg_material* mat, const float* texels,
                      const uint32_t* sizes, uint32_t ntex, const uint32_t* prim, const float* b,
                      const float* ng, uint32_t n, float* ns, float* ss, float* ts, float* uv,
                      float* base_color);
/* Brute-force closest hits over ntri triangles given by their vertices (v0, v1, v2; 9 floats
 * each) with the kernels' triangle tests: watertight = 0 the reference's Moller-Trumbore on
 * (v0, v1 - v0, v2 - v0), 1 the watertight test.  tmax may be NULL (FLT_MAX). */
int apg_debug_triangles(const float* v9, uint32_t ntri, const float* org3, const float* dir3,
                        const float* tmax, uint32_t n, int watertight, uint32_t* prim, float* t,
                        float* b0, float* b1);
/* Environment map lookups over a width x height RGB map for n directions. */
int apg_debug_environment(const float* rgb, uint32_t w, uint32_t h, const float* dirs, uint32_t n,
                          float* out);
/* One light: sample (ok, le, wi, dist, pdf) at (u2, hit) and pdf (qpdf, qdist) at (hit, q). */
int apg_debug_light(const apg_light* light, const float* u2, const float* hit, const float* q,
                    uint32_t n, uint8_t* ok, float* le, float* wi, float* dist, float* pdf,
                    float* qpdf, float* qdist);

#ifdef __cplusplus
}
#endif

#endif
