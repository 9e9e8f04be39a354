// Internal state of an apg_context (not part of the C ABI).
#pragma once

#include "../../include/apg.h"
#include "wide_node.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

namespace apg {

struct Shape
{
    std::vector<float> xyz;      // the library's copy (Accelerator::getVertexBuffer)
    std::vector<uint32_t> idx;   // 3 per triangle
    uint32_t prim_offset = 0;
    // shading attributes (Mesh::Data, shapes/mesh.h:16-33); empty = absent
    std::vector<float> normals;      // 3 per normal vertex
    std::vector<uint32_t> nidx;      // 3 per triangle (empty: idx)
    std::vector<float> tangents;     // 4 per normal vertex (xyz, handedness)
    std::vector<float> uvs;          // 2 per uv vertex
    std::vector<uint32_t> uvidx;     // 3 per triangle (empty: idx)
    // morph targets (Mesh::Data::Target): per target nv x 3 position deltas and
    // nn x 3 normal / tangent deltas; the base the position deltas are added to
    uint32_t ntargets = 0;
    std::vector<float> tpos, tnrm, ttan;
    std::vector<float> weights;      // current morph weights (ntargets)
    float* d_base = nullptr;         // device: Mesh::Data::vertices
    float* d_tpos = nullptr;         // device: position deltas
};

// Binary tree produced by either builder (host mirror for stats and collapse).
struct BinNode
{
    float lo[3], hi[3];
    uint32_t left = 0, right = 0;  // inner: children
    uint32_t first = 0, count = 0; // leaf: range in the ordered triangle list (count > 0)
};

struct DeviceScene
{
    ApgWideNode* nodes = nullptr;
    float* tris = nullptr;      // 9 floats per triangle, tree order
    uint32_t* tri_ids = nullptr; // tree order -> global triangle index
    uint32_t node_count = 0;
    uint32_t tri_count = 0;
    uint32_t roots = 1;          // wide nodes 0 .. roots-1 are tree roots
    uint32_t wide_depth = 0;     // wide-node levels of the deepest root-to-leaf path
    float* tris_v = nullptr;     // 9 floats per triangle, tree order: the input vertices v0, v1, v2
                                 // (watertight test; gathered on first use after a build)
    bool tris_v_valid = false;   // tris_v holds this build's triangles
    // allocated sizes (the GPU builder reuses the buffers while they are large enough)
    uint32_t node_cap = 0, tri_cap = 0, tris_v_cap = 0;
};

struct BuildCache;  // build.hip: the GPU builder's device scratch, kept across builds
void build_cache_free(BuildCache* c);
struct FrameGraph;  // frame.hip: a captured frame (transforms, rebuild, render)
void frame_graph_free(FrameGraph* f);

// A stream with device and pinned host buffers for the host-memory entry points
// (per-ray callers); the context keeps a pool, one per concurrently calling thread.
struct HostStaging
{
    hipStream_t stream = nullptr;
    size_t cap = 0;            // items
    float* h_in = nullptr;     // pinned: 7 floats per item (rays, SoA)
    float* h_out = nullptr;    // pinned: 4 floats per item (hits, SoA), then the overflow word
    float* dev_in = nullptr;   // the same two buffers as the kernel addresses them
    float* dev_out = nullptr;
    // device: this staging's own traversal counters (32 words) and stack spill, so batches on
    // different stagings need no ordering between them
    uint64_t* trace_words = nullptr;
    uint32_t trace_threads = 0;  // persistent threads the spill holds
};

// Pending host-memory calls of the per-ray entry points (host_calls.cpp), per kind (closest,
// any hit): the calls waiting for a batch and the batches in flight.
struct HostCoalescer
{
    std::mutex mutex;                     // guards both members
    std::vector<void*> pending[2];        // calls not yet taken into a batch, oldest first
    uint32_t leaders[2] = { 0, 0 };       // batches running or handed to a waiting caller
};

struct RenderState;  // render.hip
void render_state_free(RenderState* r);
struct GuideNet;     // guide.hip
void guide_free(GuideNet* g);

struct Context
{
    int device = 0;
    int cu_count = 0;
    // the most recent failure; per-ray entry points fail from many threads at once
    mutable std::mutex error_mutex;
    std::string last_error;
    std::vector<Shape> shapes;
    uint32_t total_tris = 0;

    // device copy of all vertices (meshes concatenated) and global-index triangles;
    // geom_dirty: host copy changed (upload at next build or transform);
    // host_stale: device transforms changed the device copy (download before host access)
    float* d_xyz = nullptr;
    uint32_t* d_tri = nullptr;
    size_t d_nv = 0, d_nt = 0;
    bool geom_dirty = true;
    bool host_stale = false;
    int sync_geometry_to_device(hipStream_t s);
    int sync_geometry_to_host();

    // scene
    bool built = false;
    DeviceScene scene;
    apg_scene_stats stats{};
    float scene_lo[3] = { 0, 0, 0 }, scene_hi[3] = { 1, 1, 1 };
    std::vector<uint32_t> leaf_keys;  // global triangle -> smallest index in its leaf

    // traversal launch state
    apg_trace_config cfg{};
    apg_render_launch rlaunch{ 8u, 4u, 4u, 0u };  // wavefront stage grids (blocks per CU)
    std::mutex launch_mutex;
    hipEvent_t last_launch = nullptr;
    hipEvent_t noalloc_launch = nullptr;  // the last apg_intersect_noalloc outside graph capture
    uint32_t* counters = nullptr;     // [0] ray cursor, [1] overflow flag, [2] waves done
    uint64_t* spill = nullptr;        // per persistent thread stack spill
    size_t spill_threads = 0;
    uint32_t* readback = nullptr;     // pinned host words the GPU builder reads level sizes into
    BuildCache* build_cache = nullptr;
    uint32_t* overflow_host = nullptr; // pinned host word the overflow flag is read back into
    // After the context's traversals so far have finished (the caller waited for them):
    // APG_E_OVERFLOW if one of them overflowed its stack since the last check, and the
    // flag is cleared; else APG_OK.
    int take_overflow();

    // shading data and the path tracer (render.hip)
    RenderState* render = nullptr;
    GuideNet* guide = nullptr;
    uint64_t attr_version = 0;   // bumped when a shape's shading attributes or targets change
    bool weights_dirty = false;  // morph weights changed since the render pools were filled
    int fail_allocs_after = -1;  // test hook: render-state allocations fail after this many more

    // host-memory entry points (host_calls.cpp)
    std::mutex staging_mutex;
    std::vector<HostStaging*> staging_free, staging_all;
    HostStaging* take_staging(size_t n, int& rc);
    void give_staging(HostStaging* s);
    HostCoalescer host_batches;

    // scene buffers sized for any rebuild of the current triangles (graph-replayed frames)
    bool scene_full_capacity = false;
    bool capturing = false;  // apg_render enqueues into a frame graph being captured
    FrameGraph* frame = nullptr;

    ~Context();
    int rebuild_shading();  // per-triangle normals after a scene build (render.hip)
    int fail(int code, const std::string& msg)
    {
        std::lock_guard<std::mutex> lock(error_mutex);
        last_error = msg;
        return code;
    }
};

// Triangle data in reference form (bvh_common.cpp:105-125): v0, e0, e1 and
// the vertex bounding box.
struct TriPrep
{
    float v0[3], e0[3], e1[3];
    float lo[3], hi[3];
};

void prepare_triangles(const Context& ctx, std::vector<TriPrep>& out);

// Host binned-SAH builder (same split rule as the reference, bvh_common.cpp:185-343).
// Produces the binary tree over `order` (ordered -> global index), then
// splits leaves above APG_MAX_LEAF into balanced sub-trees.
struct BinaryTree
{
    std::vector<BinNode> nodes;   // nodes[0] is the root
    std::vector<uint32_t> order;  // tree order -> global triangle index
    uint32_t sah_leaves = 0;
    uint32_t sah_nodes = 0;
    uint32_t max_depth = 0;
    double sah_cost = 0.0, sah_cost_leaf = 0.0;
    double build_ms = 0.0;  // device build time (GPU builder), excluding statistics
    float scene_lo[3] = { 0, 0, 0 }, scene_hi[3] = { 0, 0, 0 };
    std::vector<uint32_t> leaf_keys;
};
void build_binary_host(const std::vector<TriPrep>& tris, uint32_t leaf_threshold,
                       BinaryTree& out);

// Collapse a binary tree into 64 B wide nodes (reference collapse policy,
// wide_bvh.cpp:200-243) with octant-ordered slots and quantised boxes.
// Returns the number of wide-node levels (0 for an empty tree).
uint32_t collapse_to_wide(const BinaryTree& bt, const std::vector<TriPrep>& tris,
                      std::vector<ApgWideNode>& nodes, std::vector<float>& tri9,
                      std::vector<uint32_t>& tri_ids);

// Scene-wide quantisation slack (see wide_node.h).
float quant_slack(const std::vector<TriPrep>& tris);

} // namespace apg

struct apg_context : apg::Context {};

#define APG_HIP_CHECK(ctx, expr)                                                   \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess)                                                      \
            return (ctx)->fail(APG_E_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

namespace apg {
// The device address of apg_intersect_noalloc's scratch on `device` (current), looked up once
// per device (traverse.hip).
int noalloc_scratch_init(int device, std::string& err);
int ensure_trace_capacity(Context* ctx, uint32_t n);
int launch_trace_raw(Context* ctx, const apg_rays& rays, uint32_t n, const uint32_t* n_dev,
                     const apg_hits* hits, bool any, uint8_t* occ, unsigned long long* counts,
                     hipStream_t stream, const uint32_t* roots = nullptr);
}
