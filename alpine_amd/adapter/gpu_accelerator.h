// The `gpu` Accelerator and LightSampler for alpine, over the C ABI of
// alpine_amd/libapg.so (include/apg.h).
//
// Drop-in files for alpine's tree: they implement alpine's own plug-in
// interfaces, Accelerator (src/accelerators/accelerator.h:13-39) and
// LightSampler (src/light_samplers/light_sampler.h:12-25), and compile against
// alpine's headers (src/ on the include path, like mesh.cpp's includes).
// alpine's enums gain a Gpu value (include/alpine/alpine.h:18-22, 44-48) that
// Alpine::resetScene / buildLightSampler map to these classes
// (src/alpine.cpp:159-176, 238-250).
#pragma once

#include <accelerators/accelerator.h>
#include <intersection.h>
#include <light_samplers/light_sampler.h>
#include <ray.h>

#include <apg.h>

#include <cstdint>
#include <memory>
#include <optional>
#include <vector>

namespace alpine {
class Light;

class GpuAccelerator final : public Accelerator
{
public:
    explicit GpuAccelerator(int device = 0);
    ~GpuAccelerator() override;
    GpuAccelerator(const GpuAccelerator&) = delete;
    GpuAccelerator& operator=(const GpuAccelerator&) = delete;

    uint32_t appendMesh(
        const std::vector<float3>& vertices,
        const std::vector<uint3>& prims,
        const void* ptr) override;

    void appendSphere(const std::vector<float4>& vertices, const void* ptr) override;

    void* getVertexBuffer(uint32_t shapeId) override;

    void updateShape(uint32_t shapeId) override;

    void updateScene() override;

    std::optional<Intersection> intersect(const Ray& ray) const override;

    bool intersectAny(const Ray& ray, float tFar) const override;

    // Batched forms: n rays in one staged launch, results as the per-ray calls return them.
    void intersect(const Ray* rays, uint32_t n, std::optional<Intersection>* out) const;
    void intersectAny(const Ray* rays, const float* tFar, uint32_t n, bool* out) const;

    // The library context (the GPU render path, apg_render, runs on it).
    apg_context* context() const { return mCtx; }

private:
    Intersection makeIntersection(uint32_t prim, float t, float b0, float b1) const;

    apg_context* mCtx = nullptr;
    std::vector<const void*> mShapePtrs;     // shape id -> Shape*, echoed in Intersection::shapePtr
    std::vector<uint32_t> mPrimOffsets;      // shape id -> its first global triangle
    std::vector<std::vector<uint3>> mPrims;  // shape id -> triangles, for the geometric normals
    std::vector<float3> mNg;                 // global triangle -> geometric normal at the last build
};

class GpuLightSampler final : public LightSampler
{
public:
    enum class Kind : uint32_t { Uniform = APG_LS_UNIFORM, Power = APG_LS_POWER, Bvh = APG_LS_BVH };

    GpuLightSampler(const std::vector<std::shared_ptr<Light>>& lights, Kind kind = Kind::Bvh,
                    int device = 0);
    ~GpuLightSampler() override;
    GpuLightSampler(const GpuLightSampler&) = delete;
    GpuLightSampler& operator=(const GpuLightSampler&) = delete;

    std::optional<Sample> sample(float u, const float3& hit, const float3& ns) const override;

private:
    apg_context* mCtx = nullptr;
    std::vector<std::shared_ptr<Light>> mLights;
};
}
