// Host-side preparation of the shading tables: light records, the alias
// table of the power sampler and the light BVH.
//
// The light BVH build follows the reference exactly (bvh_light_sampler.cpp:
// 54-240): node power = sum of |power|, bbox = merge, one light per leaf,
// clusters of <= 2 lights split by index, otherwise 5 bins per axis with cost
// (sum power x SA) x kr and strict-< argmin.  It is built once per
// apg_lights_set (10 K lights take milliseconds) and sampled on the GPU.
#include "render_internal.h"

#include <cstring>
#include <numeric>

namespace apg {

namespace {
struct HBox
{
    f3 lo{ FLT_MAX, FLT_MAX, FLT_MAX }, hi{ -FLT_MAX, -FLT_MAX, -FLT_MAX };
    void merge(f3 l, f3 h)
    {
        lo = mk(min_s(l.x, lo.x), min_s(l.y, lo.y), min_s(l.z, lo.z));
        hi = mk(max_s(h.x, hi.x), max_s(h.y, hi.y), max_s(h.z, hi.z));
    }
    // merge(binAccum.bbox, bin.bbox): accumulated box first
    void merge_first(f3 l, f3 h)
    {
        lo = mk(min_s(lo.x, l.x), min_s(lo.y, l.y), min_s(lo.z, l.z));
        hi = mk(max_s(hi.x, h.x), max_s(hi.y, h.y), max_s(hi.z, h.z));
    }
};

inline float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

struct LightBvhBuilder
{
    const std::vector<DevLight>& lights;
    std::vector<DevLightNode>& nodes;
    std::vector<int32_t>& leaf_of;

    int32_t build(const std::vector<int32_t>& ids, int32_t parent)
    {
        const int32_t me = int32_t(nodes.size());
        nodes.push_back(DevLightNode{});
        float power = 0.0f;
        HBox box;
        for (int32_t i : ids)
        {
            const DevLight& L = lights[i];
            power += length(ld3(L.power));
            // merge(bbox, lightNode->bbox): the light's box first
            const f3 llo = ld3(L.lo), lhi = ld3(L.hi);
            box.lo = mk(min_s(llo.x, box.lo.x), min_s(llo.y, box.lo.y), min_s(llo.z, box.lo.z));
            box.hi = mk(max_s(lhi.x, box.hi.x), max_s(lhi.y, box.hi.y), max_s(lhi.z, box.hi.z));
        }
        {
            DevLightNode& nd = nodes[me];
            nd.lo[0] = box.lo.x; nd.lo[1] = box.lo.y; nd.lo[2] = box.lo.z;
            nd.hi[0] = box.hi.x; nd.hi[1] = box.hi.y; nd.hi[2] = box.hi.z;
            nd.power = power;
            nd.child0 = nd.child1 = -1;
            nd.light = -1;
            nd.parent = parent;
        }
        if (ids.size() == 1)
        {
            nodes[me].light = ids[0];
            leaf_of[ids[0]] = me;
            return me;
        }
        const f3 diag = box.hi - box.lo;
        bool have = false;
        int sdim = 0, sidx = 0;
        float smn = 0.0f, smx = 0.0f;
        if (ids.size() > 2)
        {
            HBox cb;
            for (int32_t i : ids)
            {
                const f3 c = (ld3(lights[i].lo) + ld3(lights[i].hi)) * 0.5f;
                cb.merge_first(c, c);
            }
            if (maxc(cb.hi - cb.lo) != 0.0f)
            {
                constexpr int SPL = 4, BC = SPL + 1;
                have = true;
                float min_cost = FLT_MAX;
                for (int d = 0; d < 3; ++d)
                {
                    const float mn = comp(cb.lo, d), mx = comp(cb.hi, d);
                    if (mn == mx) continue;
                    const float kr = maxc(diag) / max_s(comp(diag, d), 0.001f);
                    float bp[BC] = {};
                    HBox bb[BC];
                    for (int32_t i : ids)
                    {
                        const DevLight& L = lights[i];
                        const float c = comp((ld3(L.lo) + ld3(L.hi)) * 0.5f, d);
                        uint8_t bi = static_cast<uint8_t>(float(BC) * (c - mn) / (mx - mn));
                        if (bi > BC - 1) bi = BC - 1;
                        bp[bi] += length(ld3(L.power));
                        bb[bi].merge_first(ld3(L.lo), ld3(L.hi));
                    }
                    float costs[SPL] = {};
                    float ap = 0.0f;
                    HBox acc;
                    for (int s = 0; s < SPL; ++s)
                    {
                        ap += bp[s];
                        acc.merge_first(bb[s].lo, bb[s].hi);
                        const f3 l3 = acc.hi - acc.lo;
                        const float sa = acc.lo.x < FLT_MAX ? (l3.x * l3.y + l3.y * l3.z + l3.z * l3.x) * 2.0f : 0.0f;
                        costs[s] += ap * sa;
                    }
                    ap = 0.0f;
                    acc = HBox();
                    for (int s = SPL - 1; s >= 0; --s)
                    {
                        ap += bp[s + 1];
                        acc.merge_first(bb[s + 1].lo, bb[s + 1].hi);
                        const f3 l3 = acc.hi - acc.lo;
                        const float sa = acc.lo.x < FLT_MAX ? (l3.x * l3.y + l3.y * l3.z + l3.z * l3.x) * 2.0f : 0.0f;
                        costs[s] += ap * sa;
                    }
                    for (int s = 0; s < SPL; ++s)
                    {
                        costs[s] *= kr;
                        if (costs[s] < min_cost)
                        {
                            min_cost = costs[s];
                            sdim = d;
                            sidx = s;
                            smn = mn;
                            smx = mx;
                        }
                    }
                }
            }
        }
        std::vector<int32_t> sub[2];
        if (have)
        {
            for (int32_t i : ids)
            {
                const float c = comp((ld3(lights[i].lo) + ld3(lights[i].hi)) * 0.5f, sdim);
                uint8_t bi = static_cast<uint8_t>(float(5) * (c - smn) / (smx - smn));
                if (bi > 4) bi = 4;
                sub[bi <= sidx ? 0 : 1].push_back(i);
            }
        }
        else
        {
            const size_t mid = ids.size() / 2;
            for (size_t i = 0; i < ids.size(); ++i) sub[i < mid ? 0 : 1].push_back(ids[i]);
        }
        const int32_t c0 = build(sub[0], me);
        const int32_t c1 = build(sub[1], me);
        nodes[me].child0 = c0;
        nodes[me].child1 = c1;
        return me;
    }
};
} // namespace

DevLight prepare_light(const apg_light& s)
{
    DevLight L;
    std::memset(&L, 0, sizeof(L));
    L.kind = s.kind;
    const f3 col = ld3(s.color);
    const f3 pos = ld3(s.p0);
    auto put = [](float* d, f3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; };
    put(L.pos, pos);
    if (s.kind == LIGHT_DISK)
    {
        // disk_light.cpp:9-28
        const f3 power = col * s.power;
        const f3 n = ld3(s.normal);
        f3 b, t;
        get_basis(n, b, t);
        L.radius = s.radius;
        L.area = L.radius * L.radius * kPi;
        put(L.power, power);
        put(L.n, n);
        put(L.b, b);
        put(L.t, t);
        put(L.le, power / (L.area * kPi));
        const f3 c[4] = { pos + b * L.radius, pos - b * L.radius, pos + t * L.radius, pos - t * L.radius };
        HBox box;
        for (const f3& p : c) box.merge_first(p, p);
        put(L.lo, box.lo);
        put(L.hi, box.hi);
    }
    else if (s.kind == LIGHT_POINT)
    {
        // point_light.cpp:15-20
        const f3 power = col * s.power;
        put(L.power, power);
        put(L.le, power / (4.0f * kPi));
        put(L.lo, pos);
        put(L.hi, pos);
    }
    else if (s.kind == APG_LIGHT_POINT_INTENSITY)
    {
        // point_light.cpp:8-13: given its intensity (glTF punctual lights)
        L.kind = LIGHT_POINT;
        put(L.le, col);
        put(L.power, col * 4.0f * kPi);
        put(L.lo, pos);
        put(L.hi, pos);
    }
    else
    {
        const f3 e0 = ld3(s.p1) - pos, e1 = ld3(s.p2) - pos;
        const f3 c = cross(e0, e1);
        const float cl = length(c);
        put(L.e0, e0);
        put(L.e1, e1);
        put(L.n, c / cl);
        L.area = s.kind == LIGHT_QUAD ? cl : 0.5f * cl;
        put(L.le, col);
        put(L.power, col * (L.area * kPi));
        const f3 pts[4] = { pos, pos + e0, pos + e1, pos + e0 + e1 };
        const int np = s.kind == LIGHT_QUAD ? 4 : 3;
        HBox box;
        for (int i = 0; i < np; ++i) box.merge_first(pts[i], pts[i]);
        put(L.lo, box.lo);
        put(L.hi, box.hi);
    }
    return L;
}

void build_light_bvh(const std::vector<DevLight>& lights, std::vector<DevLightNode>& nodes,
                     std::vector<int32_t>& leaf_of)
{
    nodes.clear();
    leaf_of.assign(lights.size(), -1);
    if (lights.empty()) return;
    nodes.reserve(2 * lights.size());
    std::vector<int32_t> ids(lights.size());
    std::iota(ids.begin(), ids.end(), 0);
    LightBvhBuilder b{ lights, nodes, leaf_of };
    b.build(ids, -1);
}

// The device form of a built light BVH: one pair record per inner node with both children's
// centre, half diagonal and power, a parent-slot array for the pdf walk, and the light of a
// root that is itself a leaf.
void pack_light_bvh(const std::vector<DevLightNode>& nodes, std::vector<DevLightPair>& pairs,
                    std::vector<int32_t>& parent_slot, int32_t& root_light)
{
    pairs.assign(nodes.empty() ? 0 : nodes.size(), DevLightPair{});
    parent_slot.assign(nodes.size(), -1);
    root_light = -1;
    auto child = [&](int32_t id) {
        const DevLightNode& nd = nodes[size_t(id)];
        const f3 lo = ld3(nd.lo), hi = ld3(nd.hi);
        const f3 c = (lo + hi) * 0.5f;
        DevLightChild r{};
        r.c[0] = c.x; r.c[1] = c.y; r.c[2] = c.z;
        r.br = 0.5f * length(hi - lo);
        r.power = nd.power;
        r.light = nd.light;
        const bool inner = nd.child0 >= 0 || nd.child1 >= 0;
        r.node = nd.light >= 0 && !inner ? -1 : (nd.child0 >= 0 && nd.child1 >= 0 ? id : -2);
        if (r.node == -2) r.light = -1;
        return r;
    };
    if (nodes.empty()) return;
    if (nodes[0].child0 < 0 && nodes[0].child1 < 0)
    {
        pairs[0].ch[0] = child(0);
        root_light = nodes[0].light;
        return;
    }
    for (size_t i = 0; i < nodes.size(); ++i)
    {
        const DevLightNode& nd = nodes[i];
        if (nd.child0 < 0 || nd.child1 < 0) continue;
        pairs[i].ch[0] = child(nd.child0);
        pairs[i].ch[1] = child(nd.child1);
        parent_slot[size_t(nd.child0)] = int32_t(2 * i);
        parent_slot[size_t(nd.child1)] = int32_t(2 * i + 1);
    }
}

// AliasTable (alias_table.cpp:8-63)
void build_alias(const std::vector<float>& w, std::vector<float>& p, std::vector<float>& q,
                 std::vector<uint32_t>& alias)
{
    const size_t n = w.size();
    p.assign(n, 0.0f);
    q.assign(n, 0.0f);
    alias.assign(n, 0u);
    std::vector<float> sp(n);
    const float sum = static_cast<float>(std::accumulate(w.begin(), w.end(), 0.0));
    for (size_t i = 0; i < n; ++i)
    {
        p[i] = w[i] / sum;
        sp[i] = p[i] * float(n);
    }
    std::vector<uint32_t> small, large;
    for (uint32_t i = 0; i < n; ++i) (sp[i] < 1.0f ? small : large).push_back(i);
    while (!small.empty() && !large.empty())
    {
        const uint32_t si = small.back(), li = large.back();
        small.pop_back();
        large.pop_back();
        q[si] = sp[si];
        alias[si] = li;
        sp[li] += sp[si] - 1.0f;
        (sp[li] < 1.0f ? small : large).push_back(li);
    }
    for (auto* work : { &small, &large })
        while (!work->empty())
        {
            q[work->back()] = 1.0f;
            alias[work->back()] = UINT32_MAX;
            work->pop_back();
        }
}

void camera_setup(const apg_camera& c, DevCamera& out)
{
    // camera.cpp:13-29, 86-96
    const f3 eye = ld3(c.eye), tgt = ld3(c.target), up = ld3(c.up);
    out.film_h = 1.0f * std::tan(0.5f * c.fovy);
    out.film_w = out.film_h * c.aspect;
    const f3 z = normalize(tgt - eye);
    const f3 x = normalize(cross(z, up));
    const f3 y = cross(x, z);
    auto put = [](float* d, f3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; };
    put(out.eye, eye);
    put(out.vx, x);
    put(out.vy, y);
    put(out.vz, z);
}

void gather_attributes(const std::vector<Shape>& shapes, uint32_t total_tris, AttrPools& P)
{
    P = AttrPools{};
    for (const Shape& sh : shapes) P.any = P.any || !sh.normals.empty() || !sh.uvs.empty();
    if (!P.any) return;
    P.tri.assign(total_tris, DevTriAttr{});
    for (const Shape& sh : shapes)
    {
        const uint32_t nt = uint32_t(sh.idx.size() / 3);
        const uint32_t n_base = uint32_t(P.normals.size()), uv_base = uint32_t(P.uvs.size() / 2);
        const uint32_t nn = uint32_t(sh.normals.size() / 3);
        uint32_t flags = 0, morph = 0;
        if (nn)
        {
            flags |= TA_NORMALS;
            for (uint32_t k = 0; k < nn; ++k)
                P.normals.push_back(make_float4(sh.normals[3 * k], sh.normals[3 * k + 1], sh.normals[3 * k + 2], 0.0f));
            P.tangents.resize(P.normals.size(), make_float4(0, 0, 0, 0));
            if (!sh.tangents.empty())
            {
                flags |= TA_TANGENTS;
                P.any_tangents = true;
                for (uint32_t k = 0; k < nn; ++k)
                    P.tangents[n_base + k] = make_float4(sh.tangents[4 * k], sh.tangents[4 * k + 1],
                                                         sh.tangents[4 * k + 2], sh.tangents[4 * k + 3]);
            }
            if (sh.ntargets)
            {
                flags |= TA_MORPH;
                morph = uint32_t(P.morphs.size());
                DevMorph M{};
                M.ntargets = sh.ntargets;
                M.w_off = uint32_t(P.weights.size());
                M.n_base = n_base;
                M.nn = nn;
                M.d_off = uint32_t(P.tnormals.size());
                P.morphs.push_back(M);
                P.weights.insert(P.weights.end(), sh.weights.begin(), sh.weights.end());
                for (size_t k = 0; k < size_t(sh.ntargets) * nn; ++k)
                {
                    P.tnormals.push_back(make_float4(sh.tnrm[3 * k], sh.tnrm[3 * k + 1], sh.tnrm[3 * k + 2], 0.0f));
                    P.ttangents.push_back(make_float4(sh.ttan[3 * k], sh.ttan[3 * k + 1], sh.ttan[3 * k + 2], 0.0f));
                }
            }
        }
        if (!sh.uvs.empty())
        {
            flags |= TA_UVS;
            P.uvs.insert(P.uvs.end(), sh.uvs.begin(), sh.uvs.end());
        }
        const std::vector<uint32_t>& ni = sh.nidx.empty() ? sh.idx : sh.nidx;
        const std::vector<uint32_t>& ui = sh.uvidx.empty() ? sh.idx : sh.uvidx;
        for (uint32_t t = 0; t < nt; ++t)
        {
            DevTriAttr& a = P.tri[sh.prim_offset + t];
            a.flags = flags;
            a.morph = morph;
            for (int k = 0; k < 3; ++k)
            {
                a.n[k] = nn ? n_base + ni[3 * t + k] : 0u;
                a.uv[k] = (flags & TA_UVS) ? uv_base + ui[3 * t + k] : 0u;
            }
        }
    }
    if (!P.any_tangents) P.tangents.clear();
}

void gather_weights(const std::vector<Shape>& shapes, std::vector<float>& out)
{
    out.clear();
    for (const Shape& sh : shapes)
        if (sh.ntargets && !sh.normals.empty()) out.insert(out.end(), sh.weights.begin(), sh.weights.end());
}

} // namespace apg
