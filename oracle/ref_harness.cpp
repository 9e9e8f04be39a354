// ORACLE -- test infrastructure only.
//
// Thin extern "C" harness around the reference's own binary BVH, compiled
// from the unmodified sources under the reference tree
// (src/accelerators/bvh.cpp, src/accelerators/bvh_common.cpp) by
// oracle/Makefile into oracle/_ref/libalpine_ref.so.  It is used to write the
// golden vectors under tests/golden/ and, on the GPU host, as the timed CPU
// baseline.  Nothing in the product links it.
#include <accelerators/bvh.h>
#include <accelerators/bvh_common.h>
#include <intersection.h>
#include <ray.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <cstring>
#include <memory>
#include <memory_resource>
#include <thread>
#include <vector>

namespace {
struct RefBvh
{
    std::vector<std::byte> arena;
    std::unique_ptr<alpine::Bvh> bvh;
};

size_t arenaBytes(uint32_t nt) { return std::max<size_t>(size_t(64) << 20, size_t(nt) * 320); }

std::vector<alpine::float3> toVerts(const float* xyz, uint32_t nv)
{
    std::vector<alpine::float3> v(nv);
    for (uint32_t i = 0; i < nv; ++i) v[i] = alpine::float3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    return v;
}

std::vector<alpine::uint3> toTris(const uint32_t* idx, uint32_t nt)
{
    std::vector<alpine::uint3> t(nt);
    for (uint32_t i = 0; i < nt; ++i) t[i] = alpine::uint3(idx[3 * i], idx[3 * i + 1], idx[3 * i + 2]);
    return t;
}

template <typename F>
void parallelFor(uint32_t n, uint32_t nthreads, F&& f)
{
    if (nthreads <= 1 || n < 1024)
    {
        f(0u, n);
        return;
    }
    std::vector<std::thread> th;
    uint32_t chunk = (n + nthreads - 1) / nthreads;
    for (uint32_t k = 0; k < nthreads; ++k)
    {
        uint32_t b = std::min(n, k * chunk), e = std::min(n, (k + 1) * chunk);
        th.emplace_back([&f, b, e]() { f(b, e); });
    }
    for (auto& t : th) t.join();
}
}

extern "C" {

void* ref_bvh_create(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt)
{
    auto* r = new RefBvh;
    r->arena.resize(arenaBytes(nt));
    r->bvh = std::make_unique<alpine::Bvh>(std::span<std::byte>(r->arena.data(), r->arena.size()));
    static const int tag = 0;
    r->bvh->appendMesh(toVerts(xyz, nv), toTris(idx, nt), &tag);
    r->bvh->updateScene();
    return r;
}

void ref_bvh_destroy(void* h) { delete static_cast<RefBvh*>(h); }

void ref_bvh_intersect(void* h, const float* org, const float* dir, uint32_t n,
                       uint32_t* prim, float* t, float* b0, float* b1, uint32_t nthreads)
{
    const auto* r = static_cast<RefBvh*>(h);
    parallelFor(n, nthreads, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i)
        {
            alpine::Ray ray;
            ray.org = alpine::float3(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
            ray.dir = alpine::float3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
            auto isect = r->bvh->intersect(ray);
            if (isect.has_value())
            {
                prim[i] = isect->primId;
                t[i] = isect->t;
                b0[i] = isect->barycentric[0];
                b1[i] = isect->barycentric[1];
            }
            else
            {
                prim[i] = 0xFFFFFFFFu;
                t[i] = FLT_MAX;
                b0[i] = 0.0f;
                b1[i] = 0.0f;
            }
        }
    });
}

void ref_bvh_occluded(void* h, const float* org, const float* dir, const float* tfar,
                      uint32_t n, uint8_t* occ, uint32_t nthreads)
{
    const auto* r = static_cast<RefBvh*>(h);
    parallelFor(n, nthreads, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i)
        {
            alpine::Ray ray;
            ray.org = alpine::float3(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
            ray.dir = alpine::float3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
            occ[i] = r->bvh->intersectAny(ray, tfar ? tfar[i] : FLT_MAX) ? 1 : 0;
        }
    });
}

// Runs the reference builder (buildBvh) directly and reports the tree:
// SAH cost (C_trav = C_isect = 1, and leaf-only), node/leaf counts, depth and
// canonical leaf keys (smallest original triangle index in each leaf).
int ref_tree_stats(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                   uint32_t leafThreshold, double* sahCost, double* sahLeaf,
                   uint32_t* nodeCount, uint32_t* leafCount, uint32_t* maxDepth,
                   uint32_t* leafKeys)
{
    alpine::BvhShape shape{ toVerts(xyz, nv), toTris(idx, nt), 0 };
    std::vector<alpine::Primitive> prims(nt);
    for (uint32_t i = 0; i < nt; ++i)
    {
        prims[i].primId = i;
        prims[i].updateVertices(shape);
    }
    std::vector<std::byte> arena(arenaBytes(nt));
    std::pmr::monotonic_buffer_resource res(arena.data(), arena.size(), std::pmr::null_memory_resource());
    std::vector<alpine::Primitive> ordered;
    const alpine::BuildNode* root = alpine::buildBvh(prims, ordered, uint8_t(leafThreshold), &res);

    double inner = 0.0, leaf = 0.0;
    uint32_t nodes = 0, leaves = 0, depthMax = 0;
    struct Item { const alpine::BuildNode* n; uint32_t d; };
    std::vector<Item> stack{ { root, 0 } };
    while (!stack.empty())
    {
        Item it = stack.back();
        stack.pop_back();
        nodes++;
        depthMax = std::max(depthMax, it.d);
        double sa = double(it.n->bbox.computeSurfaceArea());
        if (it.n->isLeaf())
        {
            leaves++;
            leaf += sa * double(it.n->primitiveCount);
            uint32_t mn = 0xFFFFFFFFu;
            for (uint32_t k = 0; k < it.n->primitiveCount; ++k)
                mn = std::min(mn, ordered[it.n->offset + k].primId);
            for (uint32_t k = 0; k < it.n->primitiveCount; ++k)
                leafKeys[ordered[it.n->offset + k].primId] = mn;
        }
        else
        {
            inner += sa;
            stack.push_back({ it.n->children[1], it.d + 1 });
            stack.push_back({ it.n->children[0], it.d + 1 });
        }
    }
    double rootSa = double(root->bbox.computeSurfaceArea());
    *sahCost = (inner + leaf) / rootSa;
    *sahLeaf = leaf / rootSa;
    *nodeCount = nodes;
    *leafCount = leaves;
    *maxDepth = depthMax;
    return 0;
}
}
