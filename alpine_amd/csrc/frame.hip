// Frames of an animated scene as one HIP graph: the rigid transforms of listed vertex ranges,
// the GPU rebuild of every tree and the render of the frame's samples, captured once and
// replayed frame after frame (the reference's per-frame Alpine::updateScene then render,
// alpine.cpp:203-212 and 271-288, with no host work between the stages).
//
// A replay cannot read anything back, so its rebuild runs a fixed number of binary and wide
// levels (the last eager build's, plus a margin) and leaves a status word on the device; a
// replay whose tree needed more levels is reported by apg_frame_status, and the next frame is
// built eagerly and captured again.  The transform matrices and the frame's seed live in a
// device buffer the host refills before each replay; everything else the graph baked in is
// compared against the current state before each replay (a change means an eager frame and a
// new capture).
#include "apg_internal.h"

#include <cstring>
#include <vector>

namespace apg {

int build_scene_gpu_enqueue(Context* ctx, uint32_t leaf_threshold, bool per_shape, uint32_t extra, uint32_t cut,
                            hipStream_t s, uint32_t** status);
int render_reset_enqueue(Context* ctx, const uint32_t* seed_dev, hipStream_t s);
int render_tri_ng_enqueue(Context* ctx, hipStream_t s);
void render_frame_key(Context* ctx, const apg_render_params* prm, std::vector<char>& key);
void render_set_total_spp(Context* ctx, uint32_t spp);
void render_size(Context* ctx, uint32_t& W, uint32_t& H, uint32_t& S);

// Levels a replayed build runs beyond the last eager build's (binary and wide alike): an
// animated scene's trees change depth by a level or two between frames.
constexpr uint32_t FRAME_EXTRA_LEVELS = 3;

struct FrameGraph
{
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;     // the last replay's end
    std::vector<char> key;         // what the captured graph baked in
    float* h_args = nullptr;       // pinned: 12 floats per transform, then the seed word
    float* d_args = nullptr;       // device copy the graph's kernels read
    uint32_t cap = 0;              // transforms the argument buffers hold
    uint32_t* status = nullptr;    // device: the replayed build's status word
    bool replayed = false;         // the last frame was a replay, its status not yet read
    bool eager = true;             // the next frame is built eagerly (and then captured)
    uint32_t level_cut = 0;        // test hook: binary levels a capture leaves out
    uint32_t captures = 0, replays = 0, eagers = 0;
};

void frame_graph_free(FrameGraph* f)
{
    if (!f) return;
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    if (f->exec) (void)hipGraphExecDestroy(f->exec);
    if (f->done) (void)hipEventDestroy(f->done);
    if (f->stream) (void)hipStreamDestroy(f->stream);
    if (f->h_args) (void)hipHostFree(f->h_args);
    if (f->d_args) (void)hipFree(f->d_args);
    delete f;
}

namespace {

// v' = M v for vertices [0, n) of xyz, M a row-major 3 x 4 matrix in device memory.
__global__ void k_transform_dev(float* xyz, uint32_t n, const float* m)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* v = xyz + 3 * size_t(i);
    const float x = v[0], y = v[1], z = v[2];
    for (int r = 0; r < 3; ++r) v[r] = m[4 * r] * x + m[4 * r + 1] * y + m[4 * r + 2] * z + m[4 * r + 3];
}

template <typename T>
void put(std::vector<char>& key, const T& v)
{
    key.insert(key.end(), reinterpret_cast<const char*>(&v), reinterpret_cast<const char*>(&v) + sizeof(T));
}

// Wait for the last replay and read its build status; a failed one makes the next frame eager.
int take_replay_status(Context* ctx, FrameGraph& F)
{
    if (!F.replayed) return APG_OK;
    F.replayed = false;
    APG_HIP_CHECK(ctx, hipEventSynchronize(F.done));
    uint32_t st = 0;
    APG_HIP_CHECK(ctx, hipMemcpy(&st, F.status, sizeof(st), hipMemcpyDeviceToHost));
    if (st == 0u) return APG_OK;
    F.eager = true;
    return ctx->fail(APG_E_CAPACITY, std::string("the replayed frame's rebuild ran out of ") +
                                         ((st & 1u) ? "binary levels (the frame rendered the previous tree)"
                                                    : "wide levels or capacity (the frame's tree was incomplete)") +
                                         "; the next frame is rebuilt eagerly");
}

} // namespace
} // namespace apg

using namespace apg;

extern "C" int apg_frame_render(apg_context* ctx, const apg_frame_transform* xf, uint32_t nxf,
                                const apg_build_params* build, const apg_render_params* render, uint32_t seed,
                                void* stream_)
{
    if (!ctx || !render || (nxf && !xf)) return APG_E_INVALID;
    APG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const uint32_t builder = build ? build->builder : APG_BUILDER_GPU;
    if (builder != APG_BUILDER_GPU) return ctx->fail(APG_E_INVALID, "frame graphs rebuild with the GPU builder");
    if (render->guide == 1) return ctx->fail(APG_E_INVALID, "frame graphs cannot train the guide (guide = 1)");
    if (ctx->cfg.triangle_test != APG_TRI_REFERENCE)
        return ctx->fail(APG_E_INVALID, "frame graphs use the reference triangle test");
    uint32_t W = 0, H = 0, S = 0;
    render_size(ctx, W, H, S);
    if (!W) return ctx->fail(APG_E_STATE, "apg_render_reset has not been called (the frame's size and slots)");
    for (uint32_t i = 0; i < nxf; ++i)
    {
        if (xf[i].shape_id >= ctx->shapes.size()) return ctx->fail(APG_E_INVALID, "unknown shape id");
        const uint64_t nv = ctx->shapes[xf[i].shape_id].xyz.size() / 3;
        if (uint64_t(xf[i].first) + xf[i].count > nv) return ctx->fail(APG_E_INVALID, "vertex range out of bounds");
    }
    if (!ctx->frame) ctx->frame = new FrameGraph;
    FrameGraph& F = *ctx->frame;
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    apg_build_params bp{};
    if (build) bp = *build;
    const bool per_shape = bp.roots == APG_ROOTS_PER_SHAPE;
    const uint32_t thr = bp.leaf_threshold ? bp.leaf_threshold : 1u;

    // the previous replay's build must have finished its tree (a failed one makes this frame
    // eager; apg_frame_status reports it to callers that ask after each frame)
    (void)take_replay_status(ctx, F);

    // what a captured frame bakes in
    std::vector<char> key;
    put(key, nxf);
    for (uint32_t i = 0; i < nxf; ++i)
    {
        put(key, xf[i].shape_id);
        put(key, xf[i].first);
        put(key, xf[i].count);
    }
    put(key, bp);
    put(key, *render);
    put(key, ctx->scene);
    put(key, ctx->cfg);
    put(key, ctx->d_xyz);
    put(key, ctx->d_tri);
    put(key, ctx->total_tris);
    put(key, ctx->build_cache);
    render_frame_key(ctx, render, key);

    // host-side vertex edits (apg_mesh_update) reach the device through the eager path only
    if (F.eager || !ctx->built || ctx->geom_dirty || (F.exec && key != F.key))
    {
        // An eager frame: the same stages through the public entry points (their host
        // read-backs included), with the scene buffers at the capacity any rebuild needs.
        if (F.exec)
        {
            APG_HIP_CHECK(ctx, hipStreamSynchronize(F.stream));
            (void)hipGraphExecDestroy(F.exec);
            F.exec = nullptr;
        }
        ctx->scene_full_capacity = true;
        for (uint32_t i = 0; i < nxf; ++i)
            if (int rc = apg_mesh_transform(ctx, xf[i].shape_id, xf[i].m, xf[i].first, xf[i].count, stream_)) return rc;
        if (stream) APG_HIP_CHECK(ctx, hipStreamSynchronize(stream));
        if (int rc = apg_scene_build(ctx, &bp)) return rc;
        if (int rc = apg_render_reset_slots(ctx, W, H, seed, S)) return rc;
        if (int rc = apg_render(ctx, render, stream_)) return rc;
        F.eager = false;  // the next frame is captured against the state this one left
        F.eagers++;
        return APG_OK;
    }

    if (!F.exec)
    {
        if (!F.stream) APG_HIP_CHECK(ctx, hipStreamCreateWithFlags(&F.stream, hipStreamNonBlocking));
        if (!F.done) APG_HIP_CHECK(ctx, hipEventCreateWithFlags(&F.done, hipEventDisableTiming));
        if (F.cap < nxf || !F.d_args)
        {
            if (F.h_args) (void)hipHostFree(F.h_args);
            if (F.d_args) (void)hipFree(F.d_args);
            F.h_args = F.d_args = nullptr;
            F.cap = 0;
            const size_t words = 12 * size_t(nxf) + 4;
            APG_HIP_CHECK(ctx, hipHostMalloc(&F.h_args, words * sizeof(float)));
            APG_HIP_CHECK(ctx, hipMalloc(&F.d_args, words * sizeof(float)));
            F.cap = nxf;
        }
        // vertex offsets of the transformed shapes in the device geometry
        std::vector<size_t> vbase(ctx->shapes.size(), 0);
        for (size_t k = 1; k < ctx->shapes.size(); ++k) vbase[k] = vbase[k - 1] + ctx->shapes[k - 1].xyz.size() / 3;
        hipGraph_t graph = nullptr;
        APG_HIP_CHECK(ctx, hipStreamBeginCapture(F.stream, hipStreamCaptureModeRelaxed));
        int rc = APG_OK;
        for (uint32_t i = 0; i < nxf && rc == APG_OK; ++i)
            if (xf[i].count)
                k_transform_dev<<<(xf[i].count + 255) / 256, 256, 0, F.stream>>>(
                    ctx->d_xyz + 3 * (vbase[xf[i].shape_id] + xf[i].first), xf[i].count, F.d_args + 12 * size_t(i));
        if (rc == APG_OK)
            rc = build_scene_gpu_enqueue(ctx, thr, per_shape, F.level_cut ? 0u : FRAME_EXTRA_LEVELS, F.level_cut, F.stream,
                                         &F.status);
        if (rc == APG_OK) rc = render_tri_ng_enqueue(ctx, F.stream);
        if (rc == APG_OK)
            rc = render_reset_enqueue(ctx, reinterpret_cast<const uint32_t*>(F.d_args + 12 * size_t(nxf)), F.stream);
        if (rc == APG_OK)
        {
            apg_render_params rp = *render;
            rp.graph = 0;  // the passes are enqueued into this capture
            ctx->capturing = true;
            rc = apg_render(ctx, &rp, F.stream);
            ctx->capturing = false;
        }
        const hipError_t e = hipStreamEndCapture(F.stream, &graph);
        if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
        if (e != hipSuccess) return ctx->fail(APG_E_HIP, std::string("frame capture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&F.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess)
        {
            F.exec = nullptr;
            return ctx->fail(APG_E_HIP, std::string("frame graph: ") + hipGetErrorString(ei));
        }
        F.key = key;
        F.captures++;
    }

    // this frame's matrices and seed (the previous replay has finished reading them)
    for (uint32_t i = 0; i < nxf; ++i) std::memcpy(F.h_args + 12 * size_t(i), xf[i].m, 12 * sizeof(float));
    std::memcpy(F.h_args + 12 * size_t(nxf), &seed, sizeof(seed));
    std::lock_guard<std::mutex> lock(ctx->launch_mutex);
    if (!ctx->last_launch) APG_HIP_CHECK(ctx, hipEventCreateWithFlags(&ctx->last_launch, hipEventDisableTiming));
    // after the caller's stream and the context's earlier work
    APG_HIP_CHECK(ctx, hipEventRecord(F.done, stream));
    APG_HIP_CHECK(ctx, hipStreamWaitEvent(F.stream, F.done, 0));
    APG_HIP_CHECK(ctx, hipStreamWaitEvent(F.stream, ctx->last_launch, 0));
    APG_HIP_CHECK(ctx, hipMemcpyAsync(F.d_args, F.h_args, (12 * size_t(nxf) + 1) * sizeof(float), hipMemcpyHostToDevice,
                                      F.stream));
    APG_HIP_CHECK(ctx, hipGraphLaunch(F.exec, F.stream));
    APG_HIP_CHECK(ctx, hipEventRecord(F.done, F.stream));
    APG_HIP_CHECK(ctx, hipEventRecord(ctx->last_launch, F.stream));
    APG_HIP_CHECK(ctx, hipStreamWaitEvent(stream, F.done, 0));
    F.replayed = true;
    F.replays++;
    ctx->host_stale = true;  // the transforms moved the device vertices
    render_set_total_spp(ctx, render->spp);
    return APG_OK;
}

extern "C" int apg_frame_status(apg_context* ctx)
{
    if (!ctx) return APG_E_INVALID;
    if (!ctx->frame) return APG_OK;
    APG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return take_replay_status(ctx, *ctx->frame);
}

extern "C" int apg_debug_frame_level_cut(apg_context* ctx, uint32_t cut)
{
    if (!ctx) return APG_E_INVALID;
    if (!ctx->frame) ctx->frame = new FrameGraph;
    ctx->frame->level_cut = cut;
    return APG_OK;
}

extern "C" int apg_debug_frame_counts(const apg_context* ctx, uint32_t* eager, uint32_t* captures, uint32_t* replays)
{
    if (!ctx || !eager || !captures || !replays) return APG_E_INVALID;
    const FrameGraph* F = ctx->frame;
    *eager = F ? F->eagers : 0u;
    *captures = F ? F->captures : 0u;
    *replays = F ? F->replays : 0u;
    return APG_OK;
}
