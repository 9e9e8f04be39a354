// The `gpu` Accelerator and LightSampler for alpine (see gpu_accelerator.h).
//
// Reference conventions kept: no exceptions; failures print an "ERROR: ..."
// line and return the interface's sentinels (INVALID_SHAPE_ID, nullptr, the
// empty optional, false), as alpine's unavailable accelerators do
// (src/accelerators/wide_bvh_dummy.cpp:9-53).  Intersection::primId is
// mesh-local and shapePtr the appendMesh `ptr` (bvh.cpp:83-106); ng is
// normalize(cross(v1 - v0, v2 - v0)) of the vertices at the last build, the
// reference's Primitive::updateVertices (bvh_common.cpp:105-125).
#include "gpu_accelerator.h"

#include <lights/light.h>
#include <utils/bounding_box.h>

#include <algorithm>
#include <cstdio>

namespace alpine {
static_assert(sizeof(float3) == 3 * sizeof(float), "float3 must be three packed floats");
static_assert(sizeof(uint3) == 3 * sizeof(uint32_t), "uint3 must be three packed uint32");
static_assert(sizeof(Ray) == 6 * sizeof(float), "Ray must be org xyz, dir xyz");

GpuAccelerator::GpuAccelerator(int device)
{
    if (apg_create(&mCtx, device) != APG_OK)
    {
        printf("%s\n", apg_last_error(nullptr));
        mCtx = nullptr;
    }
}

GpuAccelerator::~GpuAccelerator()
{
    if (mCtx) apg_destroy(mCtx);
}

uint32_t
GpuAccelerator::appendMesh(
    const std::vector<float3>& vertices,
    const std::vector<uint3>& prims,
    const void* ptr)
{
    if (!mCtx) return INVALID_SHAPE_ID;
    uint32_t id = INVALID_SHAPE_ID, offset = 0;
    const float* xyz = vertices.empty() ? nullptr : &vertices[0].x;
    const uint32_t* idx = prims.empty() ? nullptr : &prims[0].x;
    if (apg_mesh_append(mCtx, xyz, uint32_t(vertices.size()), idx, uint32_t(prims.size()), &id) != APG_OK ||
        apg_mesh_prim_offset(mCtx, id, &offset) != APG_OK)
    {
        printf("ERROR: %s\n", apg_last_error(mCtx));
        return INVALID_SHAPE_ID;
    }
    mShapePtrs.push_back(ptr);
    mPrimOffsets.push_back(offset);
    mPrims.push_back(prims);
    return id;
}

void
GpuAccelerator::appendSphere(const std::vector<float4>& vertices, const void* ptr)
{
    if (!mCtx) return;
    if (apg_sphere_append(mCtx, vertices.empty() ? nullptr : &vertices[0].x, uint32_t(vertices.size())) != APG_OK)
        printf("%s", apg_last_error(mCtx));
}

void*
GpuAccelerator::getVertexBuffer(uint32_t shapeId)
{
    if (!mCtx) return nullptr;
    float* xyz = nullptr;
    if (apg_mesh_vertices(mCtx, shapeId, &xyz, nullptr) != APG_OK) return nullptr;
    return xyz;
}

void
GpuAccelerator::updateShape(uint32_t shapeId)
{
    if (mCtx) apg_mesh_update(mCtx, shapeId);
}

void
GpuAccelerator::updateScene()
{
    if (!mCtx) return;
    apg_build_params params{};
    params.builder = APG_BUILDER_GPU;
    params.flags = APG_BUILD_NO_STATS;
    if (apg_scene_build(mCtx, &params) != APG_OK)
    {
        printf("ERROR: %s\n", apg_last_error(mCtx));
        return;
    }
    mNg.clear();
    for (uint32_t s = 0; s < mPrims.size(); ++s)
    {
        const auto* v = static_cast<const float3*>(getVertexBuffer(s));
        for (const uint3& p : mPrims[s])
        {
            float3 e0 = v[p[1]] - v[p[0]];
            float3 e1 = v[p[2]] - v[p[0]];
            mNg.push_back(normalize(cross(e0, e1)));
        }
    }
}

Intersection
GpuAccelerator::makeIntersection(uint32_t prim, float t, float b0, float b1) const
{
    const auto it = std::upper_bound(mPrimOffsets.begin(), mPrimOffsets.end(), prim);
    const size_t shape = size_t(it - mPrimOffsets.begin()) - 1;
    Intersection isect;
    isect.shapePtr = mShapePtrs[shape];
    isect.primId = prim - mPrimOffsets[shape];
    isect.ng = mNg[prim];
    isect.barycentric = float2(b0, b1);
    isect.t = t;
    return isect;
}

std::optional<Intersection>
GpuAccelerator::intersect(const Ray& ray) const
{
    std::optional<Intersection> out;
    intersect(&ray, 1, &out);
    return out;
}

bool
GpuAccelerator::intersectAny(const Ray& ray, float tFar) const
{
    bool out = false;
    intersectAny(&ray, &tFar, 1, &out);
    return out;
}

void
GpuAccelerator::intersect(const Ray* rays, uint32_t n, std::optional<Intersection>* out) const
{
    for (uint32_t i = 0; i < n; ++i) out[i].reset();
    if (!mCtx || n == 0) return;
    std::vector<uint32_t> prim(n);
    std::vector<float> t(n), b0(n), b1(n);
    if (apg_intersect_host(mCtx, reinterpret_cast<const float*>(rays), n, prim.data(), t.data(), b0.data(),
                           b1.data()) != APG_OK)
    {
        printf("ERROR: %s\n", apg_last_error(mCtx));
        return;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (prim[i] != APG_INVALID_ID) out[i] = makeIntersection(prim[i], t[i], b0[i], b1[i]);
}

void
GpuAccelerator::intersectAny(const Ray* rays, const float* tFar, uint32_t n, bool* out) const
{
    for (uint32_t i = 0; i < n; ++i) out[i] = false;
    if (!mCtx || n == 0) return;
    std::vector<uint8_t> occ(n);
    if (apg_occluded_host(mCtx, reinterpret_cast<const float*>(rays), tFar, n, occ.data()) != APG_OK)
    {
        printf("ERROR: %s\n", apg_last_error(mCtx));
        return;
    }
    for (uint32_t i = 0; i < n; ++i) out[i] = occ[i] != 0;
}

GpuLightSampler::GpuLightSampler(const std::vector<std::shared_ptr<Light>>& lights, Kind kind, int device)
    : mLights(lights)
{
    if (apg_create(&mCtx, device) != APG_OK)
    {
        printf("%s\n", apg_last_error(nullptr));
        mCtx = nullptr;
        return;
    }
    std::vector<float> power, bounds;
    for (const auto& light : mLights)
    {
        const float3 p = light->getPower();
        const BoundingBox b = light->getBoundingBox();
        power.insert(power.end(), { p.x, p.y, p.z });
        bounds.insert(bounds.end(), { b.min.x, b.min.y, b.min.z, b.max.x, b.max.y, b.max.z });
    }
    if (apg_light_sampler_set(mCtx, power.data(), bounds.data(), uint32_t(mLights.size()),
                              static_cast<uint32_t>(kind)) != APG_OK)
    {
        printf("ERROR: %s\n", apg_last_error(mCtx));
        apg_destroy(mCtx);
        mCtx = nullptr;
    }
}

GpuLightSampler::~GpuLightSampler()
{
    if (mCtx) apg_destroy(mCtx);
}

std::optional<LightSampler::Sample>
GpuLightSampler::sample(float u, const float3& hit, const float3& ns) const
{
    if (!mCtx || mLights.empty()) return {};
    int32_t light = -1;
    float pdf = 0.0f;
    if (apg_light_sampler_sample_host(mCtx, &u, &hit.x, &ns.x, 1, &light, &pdf) != APG_OK)
    {
        printf("ERROR: %s\n", apg_last_error(mCtx));
        return {};
    }
    if (light < 0) return {};
    Sample s;
    s.light = mLights[size_t(light)].get();
    s.pdf = pdf;
    return s;
}
}
