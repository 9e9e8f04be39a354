// C ABI: context lifetime, geometry ingest and scene build orchestration.
//
// Reference behaviour mirrored here:
//  - appendMesh copies the caller's vertices and indices and assigns the next
//    shape id (bvh.cpp:83-106); the per-triangle data (v0, e0, e1, bbox) is
//    derived exactly as Primitive::updateVertices (bvh_common.cpp:105-125).
//  - getVertexBuffer hands out the library's own vertex copy; updateShape
//    marks it for re-derivation; updateScene rebuilds the whole tree
//    (bvh.cpp:114-142).
//  - appendSphere is rejected like the reference Bvh (bvh.cpp:108-112).
#include "apg_internal.h"

#include <chrono>
#include <cmath>
#include <cstring>

namespace apg {

int build_scene_gpu(Context* ctx, uint32_t leaf_threshold, bool per_shape, bool stats, BinaryTree& bt);
int gather_vertex_triangles(Context* ctx);

Context::~Context()
{
    (void)hipSetDevice(device);
    if (frame) frame_graph_free(frame);  // waits for its replays, which read what follows
    if (scene.nodes) (void)hipFree(scene.nodes);
    if (scene.tris) (void)hipFree(scene.tris);
    if (scene.tri_ids) (void)hipFree(scene.tri_ids);
    if (scene.tris_v) (void)hipFree(scene.tris_v);
    if (counters) (void)hipFree(counters);
    if (d_xyz) (void)hipFree(d_xyz);
    if (d_tri) (void)hipFree(d_tri);
    if (spill) (void)hipFree(spill);
    if (readback) (void)hipHostFree(readback);
    if (overflow_host) (void)hipHostFree(overflow_host);
    for (Shape& sh : shapes)
    {
        if (sh.d_base) (void)hipFree(sh.d_base);
        if (sh.d_tpos) (void)hipFree(sh.d_tpos);
    }
    for (HostStaging* s : staging_all)
    {
        if (s->stream) (void)hipStreamDestroy(s->stream);
        if (s->h_in) (void)hipHostFree(s->h_in);
        if (s->h_out) (void)hipHostFree(s->h_out);
        if (s->trace_words) (void)hipFree(s->trace_words);
        delete s;
    }
    if (last_launch) (void)hipEventDestroy(last_launch);
    if (noalloc_launch) (void)hipEventDestroy(noalloc_launch);
    if (render) render_state_free(render);
    if (guide) guide_free(guide);
    if (build_cache) build_cache_free(build_cache);
}

static void free_scene(Context* ctx)
{
    if (ctx->scene.nodes) (void)hipFree(ctx->scene.nodes);
    if (ctx->scene.tris) (void)hipFree(ctx->scene.tris);
    if (ctx->scene.tri_ids) (void)hipFree(ctx->scene.tri_ids);
    if (ctx->scene.tris_v) (void)hipFree(ctx->scene.tris_v);
    ctx->scene = DeviceScene{};
}

static int upload_scene(Context* ctx, const std::vector<ApgWideNode>& nodes,
                        const std::vector<float>& tri9, const std::vector<uint32_t>& ids)
{
    free_scene(ctx);
    DeviceScene& s = ctx->scene;
    s.node_count = uint32_t(nodes.size());
    s.tri_count = uint32_t(ids.size());
    if (s.node_count == 0) return APG_OK;
    APG_HIP_CHECK(ctx, hipMalloc(&s.nodes, nodes.size() * sizeof(ApgWideNode)));
    APG_HIP_CHECK(ctx, hipMalloc(&s.tris, tri9.size() * sizeof(float)));
    APG_HIP_CHECK(ctx, hipMalloc(&s.tri_ids, ids.size() * sizeof(uint32_t)));
    APG_HIP_CHECK(ctx, hipMemcpy(s.nodes, nodes.data(), nodes.size() * sizeof(ApgWideNode), hipMemcpyHostToDevice));
    APG_HIP_CHECK(ctx, hipMemcpy(s.tris, tri9.data(), tri9.size() * sizeof(float), hipMemcpyHostToDevice));
    APG_HIP_CHECK(ctx, hipMemcpy(s.tri_ids, ids.data(), ids.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    return APG_OK;
}

int Context::sync_geometry_to_device(hipStream_t s)
{
    if (!geom_dirty && d_xyz) return APG_OK;
    size_t nv = 0;
    for (const Shape& sh : shapes) nv += sh.xyz.size() / 3;
    if (nv != d_nv || total_tris != d_nt || !d_xyz)
    {
        if (d_xyz) (void)hipFree(d_xyz);
        if (d_tri) (void)hipFree(d_tri);
        d_xyz = nullptr;
        d_tri = nullptr;
        APG_HIP_CHECK(this, hipMalloc(&d_xyz, std::max<size_t>(1, nv) * 3 * sizeof(float)));
        APG_HIP_CHECK(this, hipMalloc(&d_tri, std::max<size_t>(1, total_tris) * 3 * sizeof(uint32_t)));
        d_nv = nv;
        d_nt = total_tris;
    }
    std::vector<uint32_t> tri;
    tri.reserve(size_t(total_tris) * 3);
    size_t vbase = 0;
    for (const Shape& sh : shapes)
    {
        if (!sh.xyz.empty())
            APG_HIP_CHECK(this, hipMemcpyAsync(d_xyz + 3 * vbase, sh.xyz.data(), sh.xyz.size() * sizeof(float),
                                               hipMemcpyHostToDevice, s));
        for (uint32_t v : sh.idx) tri.push_back(uint32_t(v + vbase));
        vbase += sh.xyz.size() / 3;
    }
    if (!tri.empty())
        APG_HIP_CHECK(this, hipMemcpyAsync(d_tri, tri.data(), tri.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    APG_HIP_CHECK(this, hipStreamSynchronize(s));
    geom_dirty = false;
    host_stale = false;
    return APG_OK;
}

int Context::sync_geometry_to_host()
{
    if (!host_stale) return APG_OK;
    APG_HIP_CHECK(this, hipDeviceSynchronize());
    size_t vbase = 0;
    for (Shape& sh : shapes)
    {
        if (!sh.xyz.empty())
            APG_HIP_CHECK(this, hipMemcpy(sh.xyz.data(), d_xyz + 3 * vbase, sh.xyz.size() * sizeof(float),
                                          hipMemcpyDeviceToHost));
        vbase += sh.xyz.size() / 3;
    }
    host_stale = false;
    return APG_OK;
}

} // namespace apg

namespace {
thread_local std::string g_create_error;
thread_local std::string g_error_copy;  // what apg_last_error hands out on this thread

// Accelerator::appendMesh has no check (bvh.cpp:89,101 assert capacity only): a NaN or
// infinite coordinate would give the builder NaN boxes and bins.
bool all_finite(const float* v, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}
}

extern "C" {

int apg_create(apg_context** out, int device)
{
    if (!out) return APG_E_INVALID;
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
    {
        g_create_error = "ERROR: GPU accelerator is not available (no HIP device)";
        return APG_E_NO_DEVICE;
    }
    if (device < 0 || device >= count)
    {
        g_create_error = "ERROR: device index out of range";
        return APG_E_INVALID;
    }
    e = hipSetDevice(device);
    if (e != hipSuccess)
    {
        g_create_error = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return APG_E_HIP;
    }
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess)
    {
        g_create_error = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
        return APG_E_HIP;
    }
    std::string err;
    if (int rc = apg::noalloc_scratch_init(device, err))
    {
        g_create_error = err;
        return rc;
    }
    hipEvent_t ev = nullptr;
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess)
    {
        g_create_error = std::string("hipEventCreateWithFlags: ") + hipGetErrorString(e);
        return APG_E_HIP;
    }
    auto* ctx = new apg_context;
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount;
    ctx->noalloc_launch = ev;
    apg_trace_config_default(ctx, &ctx->cfg);
    *out = ctx;
    return APG_OK;
}

int apg_destroy(apg_context* ctx)
{
    delete ctx;
    return APG_OK;
}

const char* apg_last_error(const apg_context* ctx)
{
    if (!ctx) return g_create_error.c_str();
    std::lock_guard<std::mutex> lock(ctx->error_mutex);
    g_error_copy = ctx->last_error;
    return g_error_copy.c_str();
}

int apg_mesh_append(apg_context* ctx, const float* xyz, uint32_t nv, const uint32_t* idx,
                    uint32_t nt, uint32_t* shape_id)
{
    if (!ctx) return APG_E_INVALID;
    if (shape_id) *shape_id = APG_INVALID_ID;
    if ((nv && !xyz) || (nt && !idx)) return ctx->fail(APG_E_INVALID, "null geometry pointer");
    if (uint64_t(ctx->total_tris) + nt > 0x7fffffffull)
        return ctx->fail(APG_E_CAPACITY, "scene exceeds 2^31 - 1 triangles");
    if (int rc = ctx->sync_geometry_to_host()) return rc;
    for (uint64_t i = 0; i < uint64_t(nt) * 3; ++i)
        if (idx[i] >= nv) return ctx->fail(APG_E_INVALID, "triangle index out of range");
    if (!all_finite(xyz, size_t(nv) * 3)) return ctx->fail(APG_E_INVALID, "vertex coordinate is NaN or infinite");
    apg::Shape sh;
    sh.xyz.assign(xyz, xyz + size_t(nv) * 3);
    sh.idx.assign(idx, idx + size_t(nt) * 3);
    sh.prim_offset = ctx->total_tris;
    ctx->total_tris += nt;
    ctx->shapes.push_back(std::move(sh));
    ctx->geom_dirty = true;
    if (shape_id) *shape_id = uint32_t(ctx->shapes.size() - 1);
    return APG_OK;
}

int apg_mesh_vertices(apg_context* ctx, uint32_t shape_id, float** xyz, uint32_t* nv)
{
    if (!ctx || !xyz) return APG_E_INVALID;
    if (shape_id >= ctx->shapes.size()) return ctx->fail(APG_E_INVALID, "unknown shape id");
    if (int rc = ctx->sync_geometry_to_host()) return rc;
    *xyz = ctx->shapes[shape_id].xyz.data();
    if (nv) *nv = uint32_t(ctx->shapes[shape_id].xyz.size() / 3);
    return APG_OK;
}

int apg_mesh_update(apg_context* ctx, uint32_t shape_id)
{
    if (!ctx) return APG_E_INVALID;
    if (shape_id >= ctx->shapes.size()) return ctx->fail(APG_E_INVALID, "unknown shape id");
    if (int rc = ctx->sync_geometry_to_host()) return rc;
    const std::vector<float>& v = ctx->shapes[shape_id].xyz;
    if (!all_finite(v.data(), v.size())) return ctx->fail(APG_E_INVALID, "vertex coordinate is NaN or infinite");
    // Triangle data is re-derived from the vertex copy at the next build.
    ctx->geom_dirty = true;
    return APG_OK;
}

int apg_mesh_prim_offset(const apg_context* ctx, uint32_t shape_id, uint32_t* offset)
{
    if (!ctx || !offset) return APG_E_INVALID;
    if (shape_id >= ctx->shapes.size()) return APG_E_INVALID;
    *offset = ctx->shapes[shape_id].prim_offset;
    return APG_OK;
}

int apg_sphere_append(apg_context* ctx, const float*, uint32_t)
{
    if (!ctx) return APG_E_INVALID;
    return ctx->fail(APG_E_INVALID, "ERROR: Sphere intersection has not been implemented yet.");
}

int apg_scene_build(apg_context* ctx, const apg_build_params* params)
{
    if (!ctx) return APG_E_INVALID;
    APG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    uint32_t leaf_threshold = params && params->leaf_threshold ? params->leaf_threshold : 1u;
    uint32_t builder = params ? params->builder : APG_BUILDER_GPU;
    const uint32_t roots = params ? params->roots : APG_ROOTS_SCENE;
    const bool stats = !(params && (params->flags & APG_BUILD_NO_STATS));
    if (leaf_threshold > 255) return ctx->fail(APG_E_INVALID, "leaf_threshold must be <= 255");
    if (roots != APG_ROOTS_SCENE && roots != APG_ROOTS_PER_SHAPE)
        return ctx->fail(APG_E_INVALID, "roots must be APG_ROOTS_SCENE or APG_ROOTS_PER_SHAPE");
    if (roots == APG_ROOTS_PER_SHAPE)
    {
        if (builder != APG_BUILDER_GPU) return ctx->fail(APG_E_INVALID, "per-shape trees need the GPU builder");
        for (const auto& sh : ctx->shapes)
            if (sh.idx.empty()) return ctx->fail(APG_E_INVALID, "per-shape trees need every shape non-empty");
    }

    // Wait for kernels still reading the old scene.
    if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
    if (ctx->noalloc_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->noalloc_launch));
    int rc = APG_OK;
    ctx->built = false;  // until this build has succeeded

    std::vector<apg::TriPrep> tris;
    if (builder == APG_BUILDER_HOST)
    {
        if ((rc = ctx->sync_geometry_to_host())) return rc;
        apg::prepare_triangles(*ctx, tris);
    }
    apg::BinaryTree bt;
    std::vector<ApgWideNode> nodes;
    std::vector<float> tri9;
    std::vector<uint32_t> ids;
    if (builder == APG_BUILDER_HOST)
    {
        apg::build_binary_host(tris, leaf_threshold, bt);
        const uint32_t depth = apg::collapse_to_wide(bt, tris, nodes, tri9, ids);
        rc = apg::upload_scene(ctx, nodes, tri9, ids);
        ctx->scene.wide_depth = depth;
    }
    else if (builder == APG_BUILDER_GPU)
    {
        rc = apg::build_scene_gpu(ctx, leaf_threshold, roots == APG_ROOTS_PER_SHAPE, stats, bt);
    }
    else
        return ctx->fail(APG_E_INVALID, "unknown builder");
    if (rc) return rc;
    // One stack entry per wide level at most (a lane pushes the rest of a node's hit children
    // once per descent), so a tree no deeper than the stack can never overflow it.
    if (const uint32_t depth = ctx->scene.wide_depth; depth > APG_TRACE_STACK_MAX)
    {
        apg::free_scene(ctx);
        return ctx->fail(APG_E_CAPACITY, "the 8-wide tree has " + std::to_string(depth) +
                                             " levels, more than the traversal stack's " +
                                             std::to_string(APG_TRACE_STACK_MAX) + " entries");
    }

    ctx->leaf_keys = std::move(bt.leaf_keys);
    for (int a = 0; a < 3; ++a)
    {
        ctx->scene_lo[a] = bt.scene_lo[a];
        ctx->scene_hi[a] = bt.scene_hi[a];
    }
    apg_scene_stats& st = ctx->stats;
    st = apg_scene_stats{};
    st.triangles = ctx->total_tris;
    st.shapes = uint32_t(ctx->shapes.size());
    st.binary_nodes = bt.sah_nodes;
    st.binary_leaves = bt.sah_leaves;
    st.wide_nodes = ctx->scene.node_count;
    st.max_depth_binary = bt.max_depth;
    st.wide_depth = ctx->scene.wide_depth;
    st.sah_cost = bt.sah_cost;
    st.sah_cost_leaf = bt.sah_cost_leaf;
    // nodes; reference-form triangles and ids (40 B); vertex-form triangles (36 B)
    st.device_bytes = uint64_t(ctx->scene.node_count) * 64 + uint64_t(ctx->scene.tri_count) * 76;
    if ((rc = ctx->rebuild_shading())) return rc;
    if (ctx->cfg.triangle_test == APG_TRI_WATERTIGHT && (rc = apg::gather_vertex_triangles(ctx))) return rc;
    st.build_ms = bt.build_ms > 0.0
        ? bt.build_ms
        : std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->built = true;
    return APG_OK;
}

int apg_scene_get_stats(const apg_context* ctx, apg_scene_stats* stats)
{
    if (!ctx || !stats) return APG_E_INVALID;
    *stats = ctx->stats;
    return APG_OK;
}

int apg_trace_config_default(const apg_context* ctx, apg_trace_config* cfg)
{
    if (!cfg) return APG_E_INVALID;
    (void)ctx;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->block_size = 256;
    cfg->blocks_per_cu = 6;
    cfg->fetch_threshold = 16;
    cfg->stack_lds = 8;
    cfg->tri_batch = 48;
    return APG_OK;
}

int apg_set_trace_config(apg_context* ctx, const apg_trace_config* cfg)
{
    if (!ctx || !cfg) return APG_E_INVALID;
    if (cfg->block_size == 0 || cfg->block_size % 64 || cfg->block_size > 256)
        return ctx->fail(APG_E_INVALID, "block_size must be 64, 128, 192 or 256");
    if (cfg->blocks_per_cu == 0 || cfg->blocks_per_cu > 64)
        return ctx->fail(APG_E_INVALID, "blocks_per_cu must be in 1..64");
    if (cfg->fetch_threshold == 0 || cfg->fetch_threshold > 64)
        return ctx->fail(APG_E_INVALID, "fetch_threshold must be in 1..64");
    if (cfg->stack_lds != 8 && cfg->stack_lds != 16)
        return ctx->fail(APG_E_INVALID, "stack_lds must be 8 or 16");
    if (cfg->tri_batch == 0 || cfg->tri_batch > 64)
        return ctx->fail(APG_E_INVALID, "tri_batch must be in 1..64");
    if (cfg->triangle_test > APG_TRI_WATERTIGHT)
        return ctx->fail(APG_E_INVALID, "triangle_test must be APG_TRI_REFERENCE or APG_TRI_WATERTIGHT");
    if (cfg->stack_limit > APG_TRACE_STACK_MAX)
        return ctx->fail(APG_E_INVALID, "stack_limit must be at most APG_TRACE_STACK_MAX");
    // the watertight test reads vertex-form triangles, gathered once per build when first needed
    if (cfg->triangle_test == APG_TRI_WATERTIGHT && ctx->built && !ctx->scene.tris_v_valid)
    {
        if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
        if (int rc = apg::gather_vertex_triangles(ctx)) return rc;
    }
    ctx->cfg = *cfg;
    return APG_OK;
}

int apg_scene_triangles(const apg_context* ctx, float* tri9, uint32_t* ids, uint32_t n)
{
    if (!ctx || !tri9 || !ids) return APG_E_INVALID;
    if (!ctx->built || n != ctx->scene.tri_count) return APG_E_STATE;
    if (n == 0) return APG_OK;
    if (hipMemcpy(tri9, ctx->scene.tris, size_t(n) * 9 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return APG_E_HIP;
    if (hipMemcpy(ids, ctx->scene.tri_ids, size_t(n) * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
        return APG_E_HIP;
    return APG_OK;
}

int apg_scene_leaf_keys(const apg_context* ctx, uint32_t* keys, uint32_t n)
{
    if (!ctx || !keys) return APG_E_INVALID;
    if (!ctx->built || n != ctx->leaf_keys.size()) return APG_E_STATE;
    std::memcpy(keys, ctx->leaf_keys.data(), size_t(n) * sizeof(uint32_t));
    return APG_OK;
}

int apg_scene_wide_nodes(const apg_context* ctx, void* nodes64, uint32_t n)
{
    if (!ctx || !nodes64) return APG_E_INVALID;
    if (!ctx->built || n != ctx->scene.node_count) return APG_E_STATE;
    if (n == 0) return APG_OK;
    if (hipMemcpy(nodes64, ctx->scene.nodes, size_t(n) * 64, hipMemcpyDeviceToHost) != hipSuccess)
        return APG_E_HIP;
    return APG_OK;
}

} // extern "C"
