// Host build of the traversal steps over a host-built scene (test hooks,
// include/apg_debug.h).  The loop below is the per-lane body of trace_kernel
// (traverse.hip) with an unbounded vector as the stack.
#include "../../include/apg_debug.h"
#include "apg_internal.h"
#include "traverse_core.h"
#include "render_internal.h"

#include <cmath>
#include <cstring>

struct apg_debug_scene
{
    std::vector<ApgWideNode> nodes;
    std::vector<float> tri9;
    std::vector<uint32_t> ids;
    apg::BinaryTree bt;
    uint32_t wide_depth = 0;
};

extern "C" {

int apg_debug_host_build(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                         uint32_t leaf_threshold, apg_debug_scene** out)
{
    if (!out || (nv && !xyz) || (nt && !idx)) return APG_E_INVALID;
    for (uint64_t i = 0; i < uint64_t(nt) * 3; ++i)
        if (idx[i] >= nv) return APG_E_INVALID;
    for (uint64_t i = 0; i < uint64_t(nv) * 3; ++i)
        if (!std::isfinite(xyz[i])) return APG_E_INVALID;  // as apg_mesh_append
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
    s->wide_depth = apg::collapse_to_wide(s->bt, tris, s->nodes, s->tri9, s->ids);
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

int apg_debug_scene_depth(const apg_debug_scene* s, uint32_t* wide_depth)
{
    if (!s || !wide_depth) return APG_E_INVALID;
    *wide_depth = s->wide_depth;
    return APG_OK;
}

} // extern "C"

namespace {

// The per-lane loop of trace_kernel.  stack_limit = 0: unbounded; otherwise a push onto a
// full stack is dropped as the kernel drops it, and the ray is counted in *overflowed.
int host_trace(const apg_debug_scene* s, const float* org3, const float* dir3, const float* tmaxp, uint32_t n,
               int any, uint32_t stack_limit, uint32_t* prim, float* tout, float* b0out, float* b1out, uint8_t* occ,
               uint32_t* max_stack, uint32_t* overflowed)
{
    using namespace apg;
    if (!s || !org3 || !dir3) return APG_E_INVALID;
    if (any ? !occ : (!prim || !tout || !b0out || !b1out)) return APG_E_INVALID;
    uint32_t deepest = 0, dropped_rays = 0;
    std::vector<uint64_t> stack;
    for (uint32_t i = 0; i < n; ++i)
    {
        const float ox = org3[3 * i], oy = org3[3 * i + 1], oz = org3[3 * i + 2];
        const float dx = dir3[3 * i], dy = dir3[3 * i + 1], dz = dir3[3 * i + 2];
        float tmax = tmaxp ? tmaxp[i] : FLT_MAX;
        RayBox rb = make_raybox(ox, oy, oz, dx, dy, dz);
        uint32_t g_base = 0, g_bits = s->nodes.empty() ? 0u : root_group_bits(rb.oct);
        uint32_t t_base = 0, t_mask = 0, t_cmask = 0, best = 0xFFFFFFFFu;
        float hb0 = 0, hb1 = 0;
        bool found = false, dropped = false;
        stack.clear();
        for (;;)
        {
            if ((g_bits >> 8) != 0u)
            {
                uint32_t hits = 0;
                uint32_t node = group_next_node(g_base, g_bits, rb.oct, hits);
                if (hits && stack_limit && stack.size() >= stack_limit)
                {
                    dropped = true;
                    hits = 0;
                }
                if (hits) stack.push_back((uint64_t((hits << 8) | (g_bits & 0xffu)) << 32) | g_base);
                if (stack.size() > deepest) deepest = uint32_t(stack.size());
                uint32_t w[16];
                std::memcpy(w, &s->nodes[node], 64);
                node_visit(w, rb, tmax, g_base, g_bits, t_base, t_mask, t_cmask);
            }
            while (t_mask)
            {
                const uint32_t tri = tri_index(t_base, t_cmask, bit_ffs(t_mask) - 1u);
                t_mask &= t_mask - 1u;
                float t, b0, b1;
                if (!tri_hit(&s->tri9[size_t(tri) * 9], ox, oy, oz, dx, dy, dz, tmax, t, b0, b1))
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
                best = tri;
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
        dropped_rays += dropped ? 1u : 0u;
    }
    if (max_stack) *max_stack = deepest;
    if (overflowed) *overflowed = dropped_rays;
    return APG_OK;
}

} // namespace

extern "C" {

int apg_debug_host_trace(const apg_debug_scene* s, const float* org3, const float* dir3,
                         const float* tmaxp, uint32_t n, int any, uint32_t* prim, float* tout,
                         float* b0out, float* b1out, uint8_t* occ, uint32_t* max_stack)
{
    return host_trace(s, org3, dir3, tmaxp, n, any, 0u, prim, tout, b0out, b1out, occ, max_stack, nullptr);
}

int apg_debug_host_trace_limited(const apg_debug_scene* s, const float* org3, const float* dir3,
                                 const float* tmaxp, uint32_t n, int any, uint32_t stack_limit,
                                 uint32_t* prim, float* tout, float* b0out, float* b1out, uint8_t* occ,
                                 uint32_t* overflowed)
{
    if (stack_limit == 0 || stack_limit > APG_TRACE_STACK_MAX) return APG_E_INVALID;
    return host_trace(s, org3, dir3, tmaxp, n, any, stack_limit, prim, tout, b0out, b1out, occ, nullptr,
                      overflowed);
}

void apg_debug_scene_free(apg_debug_scene* s) { delete s; }

} // extern "C"

namespace {
// Reads n16 16-byte words, four loads in flight per thread per step; the XOR of everything
// read decides a store that never happens for real data, so the loads cannot be dropped.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_stream_read(const u32x4* __restrict__ p, uint64_t n16, uint32_t* sink)
{
    uint32_t acc = 0;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride)
    {
        const u32x4 a = __builtin_nontemporal_load(p + i);
        const u32x4 b = __builtin_nontemporal_load(p + i + stride);
        const u32x4 c = __builtin_nontemporal_load(p + i + 2 * stride);
        const u32x4 d = __builtin_nontemporal_load(p + i + 3 * stride);
        acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^ d.z ^ d.w;
    }
    for (; i < n16; i += stride)
    {
        const u32x4 a = __builtin_nontemporal_load(p + i);
        acc ^= a.x ^ a.y ^ a.z ^ a.w;
    }
    if (acc == 0x9e3779b9u) sink[0] = acc;
}
} // namespace

extern "C" {

int apg_debug_stream_read(const void* buf, uint64_t bytes, uint32_t* sink, uint32_t blocks, void* stream)
{
    if (!buf || !sink || bytes % 16 || blocks == 0) return APG_E_INVALID;
    k_stream_read<<<blocks, 256, 0, static_cast<hipStream_t>(stream)>>>(static_cast<const u32x4*>(buf), bytes / 16,
                                                                       sink);
    return hipGetLastError() == hipSuccess ? APG_OK : APG_E_HIP;
}

int apg_debug_fail_allocations(apg_context* ctx, int after)
{
    if (!ctx) return APG_E_INVALID;
    ctx->fail_allocs_after = after < 0 ? -1 : after;
    return APG_OK;
}

int apg_debug_material(const apg_material* mat, const float* wo, const float* u,
                       const float* wi_eval, uint32_t n, uint8_t* ok, float* est, float* wi,
                       float* pdf, float* f_eval, float* pdf_eval)
{
    using namespace apg;
    if (!mat) return APG_E_INVALID;
    const DevMaterial m = make_material(mat->kind, mat->base_color, mat->alpha);
    for (uint32_t i = 0; i < n; ++i)
    {
        f3 e{}, w{};
        float p = 0.0f;
        const bool s = mat_sample(m, ld3(wo + 3 * i), f2{ u[2 * i], u[2 * i + 1] }, e, w, p);
        ok[i] = s ? 1 : 0;
        est[3 * i] = s ? e.x : 0.0f; est[3 * i + 1] = s ? e.y : 0.0f; est[3 * i + 2] = s ? e.z : 0.0f;
        wi[3 * i] = s ? w.x : 0.0f; wi[3 * i + 1] = s ? w.y : 0.0f; wi[3 * i + 2] = s ? w.z : 0.0f;
        pdf[i] = s ? p : 0.0f;
        f3 f;
        float fp;
        mat_eval(m, ld3(wo + 3 * i), ld3(wi_eval + 3 * i), f, fp);
        f_eval[3 * i] = f.x; f_eval[3 * i + 1] = f.y; f_eval[3 * i + 2] = f.z;
        pdf_eval[i] = fp;
    }
    return APG_OK;
}

int apg_debug_light_sampler(const apg_light* lights, uint32_t nl, uint32_t sampler,
                            const float* u, const float* hit, const float* ns, uint32_t n,
                            int32_t* light, float* pdf, float* pdf_ext)
{
    using namespace apg;
    std::vector<DevLight> dl(nl);
    for (uint32_t i = 0; i < nl; ++i) dl[i] = prepare_light(lights[i]);
    std::vector<float> ap, aq;
    std::vector<uint32_t> aa;
    std::vector<DevLightNode> nodes;
    std::vector<int32_t> leaf;
    if (sampler == LS_POWER && nl)
    {
        std::vector<float> w(nl);
        for (uint32_t i = 0; i < nl; ++i) w[i] = length(ld3(dl[i].power));
        build_alias(w, ap, aq, aa);
    }
    std::vector<DevLightPair> pairs;
    std::vector<int32_t> parent;
    int32_t root_light = -1;
    if (sampler == LS_BVH && nl)
    {
        build_light_bvh(dl, nodes, leaf);
        pack_light_bvh(nodes, pairs, parent, root_light);
    }
    LightSamplerView v{ sampler, nl, dl.data(), ap.data(), aq.data(), aa.data(), pairs.data(), parent.data(),
                        leaf.data(), root_light };
    for (uint32_t i = 0; i < n; ++i)
    {
        float p = 0.0f;
        light[i] = light_sampler_sample(v, u[i], ld3(hit + 3 * i), ld3(ns + 3 * i), p);
        pdf[i] = light[i] >= 0 ? p : 0.0f;
        pdf_ext[i] = light[i] >= 0 ? light_sampler_pdf(v, light[i], ld3(hit + 3 * i), ld3(ns + 3 * i)) : 0.0f;
    }
    return APG_OK;
}

int apg_debug_triangles(const float* v9, uint32_t ntri, const float* org3, const float* dir3,
                        const float* tmax, uint32_t n, int watertight, uint32_t* prim, float* t,
                        float* b0, float* b1)
{
    using namespace apg;
    if ((ntri && !v9) || (n && (!org3 || !dir3 || !prim || !t || !b0 || !b1))) return APG_E_INVALID;
    // reference form (v0, e0 = v1 - v0, e1 = v2 - v0), bvh_common.cpp:105-125
    std::vector<float> ref(size_t(ntri) * 9);
    for (uint32_t i = 0; i < ntri; ++i)
        for (int a = 0; a < 3; ++a)
        {
            const float* v = v9 + 9 * size_t(i);
            ref[9 * size_t(i) + a] = v[a];
            ref[9 * size_t(i) + 3 + a] = v[3 + a] - v[a];
            ref[9 * size_t(i) + 6 + a] = v[6 + a] - v[a];
        }
    for (uint32_t r = 0; r < n; ++r)
    {
        const float* o = org3 + 3 * size_t(r);
        const float* d = dir3 + 3 * size_t(r);
        float best = tmax ? tmax[r] : FLT_MAX;
        prim[r] = 0xFFFFFFFFu;
        t[r] = FLT_MAX;
        b0[r] = b1[r] = 0.0f;
        const WtRay w = make_wtray(o[0], o[1], o[2], d[0], d[1], d[2]);
        for (uint32_t i = 0; i < ntri; ++i)
        {
            float th, u, v;
            const bool h = watertight ? tri_hit_watertight(v9 + 9 * size_t(i), w, best, th, u, v)
                                      : tri_hit(ref.data() + 9 * size_t(i), o[0], o[1], o[2], d[0], d[1], d[2],
                                                best, th, u, v);
            if (h)
            {
                best = th;
                prim[r] = i;
                t[r] = th;
                b0[r] = u;
                b1[r] = v;
            }
        }
    }
    return APG_OK;
}

int apg_debug_shading(const uint32_t* idx, uint32_t nt, uint32_t nv, const apg_mesh_shading* sh,
                      uint32_t ntargets, const float* dnormals, const float* dtangents,
                      const float* weights, const apg_material* mat, const float* texels,
                      const uint32_t* sizes, uint32_t ntex, const uint32_t* prim, const float* b,
                      const float* ng, uint32_t n, float* ns, float* ss, float* ts, float* uv,
                      float* base_color)
{
    using namespace apg;
    if (!idx || !sh || !mat || (ntargets && !weights)) return APG_E_INVALID;
    Shape s;
    s.idx.assign(idx, idx + size_t(nt) * 3);
    s.xyz.assign(size_t(nv) * 3, 0.0f);
    const uint32_t nn = sh->normals ? sh->nn : 0u;
    if (nn)
    {
        s.normals.assign(sh->normals, sh->normals + size_t(nn) * 3);
        if (sh->normal_idx) s.nidx.assign(sh->normal_idx, sh->normal_idx + size_t(nt) * 3);
        if (sh->tangents) s.tangents.assign(sh->tangents, sh->tangents + size_t(nn) * 4);
    }
    if (sh->uvs)
    {
        s.uvs.assign(sh->uvs, sh->uvs + size_t(sh->nuv) * 2);
        if (sh->uv_idx) s.uvidx.assign(sh->uv_idx, sh->uv_idx + size_t(nt) * 3);
    }
    s.ntargets = ntargets;
    s.tnrm.assign(size_t(ntargets) * nn * 3, 0.0f);
    s.ttan.assign(size_t(ntargets) * nn * 3, 0.0f);
    if (dnormals) std::memcpy(s.tnrm.data(), dnormals, s.tnrm.size() * sizeof(float));
    if (dtangents) std::memcpy(s.ttan.data(), dtangents, s.ttan.size() * sizeof(float));
    s.weights.assign(weights, weights + ntargets);
    std::vector<Shape> shapes(1, s);
    AttrPools P;
    gather_attributes(shapes, nt, P);
    std::vector<DevTexture> tx(ntex);
    size_t total = 0;
    for (uint32_t k = 0; k < ntex; ++k)
    {
        tx[k] = DevTexture{ uint32_t(total), sizes[2 * k], sizes[2 * k + 1], 0u };
        total += size_t(sizes[2 * k]) * sizes[2 * k + 1];
    }
    std::vector<float4> tl(total);
    for (size_t i = 0; i < total; ++i)
        tl[i] = make_float4(texels[4 * i], texels[4 * i + 1], texels[4 * i + 2], texels[4 * i + 3]);
    AttrView V{};
    if (P.any)
    {
        V.tri = P.tri.data();
        V.normals = P.normals.data();
        V.tangents = P.any_tangents ? P.tangents.data() : nullptr;
        V.uvs = P.uvs.data();
        V.morphs = P.morphs.data();
        V.weights = P.weights.data();
        V.tnormals = P.tnormals.data();
        V.ttangents = P.ttangents.data();
    }
    V.texs = tx.data();
    V.texels = tl.data();
    DevMaterial m = make_material(mat->kind, mat->base_color, mat->alpha);
    m.tex_bc = mat->base_color_tex;
    m.tex_n = mat->normal_tex;
    if (m.tex_bc > ntex || m.tex_n > ntex) return APG_E_INVALID;
    for (uint32_t i = 0; i < n; ++i)
    {
        if (prim[i] >= nt) return APG_E_INVALID;
        f3 nsv, ssv, tsv;
        f2 u;
        shading_attributes(V, prim[i], b[2 * i], b[2 * i + 1], ld3(ng + 3 * i), m, nsv, u);
        get_basis(nsv, ssv, tsv);
        const f3 c = material_base_color(V, m, u);
        ns[3 * i] = nsv.x; ns[3 * i + 1] = nsv.y; ns[3 * i + 2] = nsv.z;
        ss[3 * i] = ssv.x; ss[3 * i + 1] = ssv.y; ss[3 * i + 2] = ssv.z;
        ts[3 * i] = tsv.x; ts[3 * i + 1] = tsv.y; ts[3 * i + 2] = tsv.z;
        uv[2 * i] = u.x; uv[2 * i + 1] = u.y;
        base_color[3 * i] = c.x; base_color[3 * i + 1] = c.y; base_color[3 * i + 2] = c.z;
    }
    return APG_OK;
}

int apg_debug_environment(const float* rgb, uint32_t w, uint32_t h, const float* dirs, uint32_t n,
                          float* out)
{
    using namespace apg;
    if (!rgb || w == 0 || h == 0) return APG_E_INVALID;
    std::vector<float4> t(size_t(w) * h);
    for (size_t i = 0; i < t.size(); ++i) t[i] = make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], 0.0f);
    for (uint32_t i = 0; i < n; ++i)
    {
        const f3 e = environment_sample(t.data(), w, h, ld3(dirs + 3 * i));
        out[3 * i] = e.x; out[3 * i + 1] = e.y; out[3 * i + 2] = e.z;
    }
    return APG_OK;
}

int apg_debug_light(const apg_light* light, const float* u2, const float* hit, const float* q,
                    uint32_t n, uint8_t* ok, float* le, float* wi, float* dist, float* pdf,
                    float* qpdf, float* qdist)
{
    using namespace apg;
    const DevLight L = prepare_light(*light);
    for (uint32_t i = 0; i < n; ++i)
    {
        f3 l{}, w{};
        float d = 0.0f, p = 0.0f;
        const bool s = light_sample(L, f2{ u2[2 * i], u2[2 * i + 1] }, ld3(hit + 3 * i), l, w, d, p);
        ok[i] = s ? 1 : 0;
        le[3 * i] = l.x; le[3 * i + 1] = l.y; le[3 * i + 2] = l.z;
        wi[3 * i] = w.x; wi[3 * i + 1] = w.y; wi[3 * i + 2] = w.z;
        dist[i] = s ? d : 0.0f;
        pdf[i] = s ? p : 0.0f;
        float qd = 0.0f;
        qpdf[i] = light_pdf(L, ld3(hit + 3 * i), ld3(q + 3 * i), qd);
        qdist[i] = qd;
    }
    return APG_OK;
}

} // extern "C"
