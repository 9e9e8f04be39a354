// Host build of the traversal steps over a host-built scene (test hooks,
// include/apg_debug.h).  The loop below is the per-lane body of trace_kernel
// (traverse.hip) with an unbounded vector as the stack.
#include "../../include/apg_debug.h"
#include "apg_internal.h"
#include "traverse_core.h"

#include <cstring>

struct apg_debug_scene
{
    std::vector<ApgWideNode> nodes;
    std::vector<float> tri9;
    std::vector<uint32_t> ids;
    apg::BinaryTree bt;
};

extern "C" {

int apg_debug_host_build(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                         uint32_t leaf_threshold, apg_debug_scene** out)
{
    if (!out || (nv && !xyz) || (nt && !idx)) return APG_E_INVALID;
    for (uint64_t i = 0; i < uint64_t(nt) * 3; ++i)
        if (idx[i] >= nv) return APG_E_INVALID;
    apg::Context tmp;
    apg::Shape sh;
    sh.xyz.assign(xyz, xyz + size_t(nv) * 3);
    sh.idx.assign(idx, idx + size_t(nt) * 3);
    tmp.shapes.push_back(std::move(sh));
    tmp.total_tris = nt;
    std::vector<apg::TriPrep> tris;
    apg::prepare_triangles(tmp, tris);
    auto* s = new apg_debug_scene;
    apg::build_binary_host(tris, leaf_threshold ? leaf_threshold : 4u, s->bt);
    apg::collapse_to_wide(s->bt, tris, s->nodes, s->tri9, s->ids);
    *out = s;
    return APG_OK;
}

int apg_debug_scene_info(const apg_debug_scene* s, uint32_t* n_nodes, uint32_t* n_tris,
                         double* sah_cost, double* sah_cost_leaf, uint32_t* binary_nodes,
                         uint32_t* max_depth)
{
    if (!s) return APG_E_INVALID;
    if (n_nodes) *n_nodes = uint32_t(s->nodes.size());
    if (n_tris) *n_tris = uint32_t(s->ids.size());
    if (sah_cost) *sah_cost = s->bt.sah_cost;
    if (sah_cost_leaf) *sah_cost_leaf = s->bt.sah_cost_leaf;
    if (binary_nodes) *binary_nodes = s->bt.sah_nodes;
    if (max_depth) *max_depth = s->bt.max_depth;
    return APG_OK;
}

int apg_debug_scene_copy(const apg_debug_scene* s, void* nodes64, float* tri9, uint32_t* ids,
                         uint32_t* leaf_keys)
{
    if (!s) return APG_E_INVALID;
    if (nodes64 && !s->nodes.empty()) std::memcpy(nodes64, s->nodes.data(), s->nodes.size() * 64);
    if (tri9 && !s->tri9.empty()) std::memcpy(tri9, s->tri9.data(), s->tri9.size() * sizeof(float));
    if (ids && !s->ids.empty()) std::memcpy(ids, s->ids.data(), s->ids.size() * sizeof(uint32_t));
    if (leaf_keys && !s->bt.leaf_keys.empty())
        std::memcpy(leaf_keys, s->bt.leaf_keys.data(), s->bt.leaf_keys.size() * sizeof(uint32_t));
    return APG_OK;
}

int apg_debug_host_trace(const apg_debug_scene* s, const float* org3, const float* dir3,
                         const float* tmaxp, uint32_t n, int any, uint32_t* prim, float* tout,
                         float* b0out, float* b1out, uint8_t* occ, uint32_t* max_stack)
{
    using namespace apg;
    if (!s || !org3 || !dir3) return APG_E_INVALID;
    if (any ? !occ : (!prim || !tout || !b0out || !b1out)) return APG_E_INVALID;
    uint32_t deepest = 0;
    std::vector<uint64_t> stack;
    for (uint32_t i = 0; i < n; ++i)
    {
        const float ox = org3[3 * i], oy = org3[3 * i + 1], oz = org3[3 * i + 2];
        const float dx = dir3[3 * i], dy = dir3[3 * i + 1], dz = dir3[3 * i + 2];
        float tmax = tmaxp ? tmaxp[i] : FLT_MAX;
        RayBox rb = make_raybox(ox, oy, oz, dx, dy, dz);
        uint32_t g_base = 0, g_bits = s->nodes.empty() ? 0u : root_group_bits(rb.oct);
        uint32_t t_base = 0, t_mask = 0, best = 0xFFFFFFFFu;
        float hb0 = 0, hb1 = 0;
        bool found = false;
        stack.clear();
        for (;;)
        {
            if ((g_bits >> 8) != 0u)
            {
                uint32_t hits = 0;
                uint32_t node = group_next_node(g_base, g_bits, rb.oct, hits);
                if (hits) stack.push_back((uint64_t((hits << 8) | (g_bits & 0xffu)) << 32) | g_base);
                if (stack.size() > deepest) deepest = uint32_t(stack.size());
                uint32_t w[16];
                std::memcpy(w, &s->nodes[node], 64);
                node_visit(w, rb, tmax, g_base, g_bits, t_base, t_mask);
            }
            while (t_mask)
            {
                uint32_t k = bit_ffs(t_mask) - 1u;
                t_mask &= t_mask - 1u;
                float t, b0, b1;
                if (!tri_hit(&s->tri9[size_t(t_base + k) * 9], ox, oy, oz, dx, dy, dz, tmax, t, b0, b1))
                    continue;
                if (any)
                {
                    found = true;
                    t_mask = 0;
                    g_bits = 0;
                    stack.clear();
                    break;
                }
                tmax = t;
                best = t_base + k;
                hb0 = b0;
                hb1 = b1;
            }
            if ((g_bits >> 8) == 0u)
            {
                if (stack.empty()) break;
                uint64_t e = stack.back();
                stack.pop_back();
                g_base = uint32_t(e);
                g_bits = uint32_t(e >> 32);
            }
        }
        if (any)
            occ[i] = found ? 1 : 0;
        else
        {
            bool hit = best != 0xFFFFFFFFu;
            prim[i] = hit ? s->ids[best] : 0xFFFFFFFFu;
            tout[i] = hit ? tmax : FLT_MAX;
            b0out[i] = hit ? hb0 : 0.0f;
            b1out[i] = hit ? hb1 : 0.0f;
        }
    }
    if (max_stack) *max_stack = deepest;
    return APG_OK;
}

void apg_debug_scene_free(apg_debug_scene* s) { delete s; }

} // extern "C"
