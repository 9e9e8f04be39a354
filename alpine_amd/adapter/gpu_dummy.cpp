// The `gpu` Accelerator and LightSampler for builds without the MI355X library
// (ALPINE_ENABLE_HIP off): the same classes as gpu_accelerator.cpp, each call
// printing nothing more than the constructor's ERROR line and returning the
// interface's sentinels, as alpine's own dummy accelerators do
// (src/accelerators/wide_bvh_dummy.cpp:9-53, embree_dummy.cpp).
#include "gpu_accelerator.h"

#include <cstdio>

namespace alpine {

GpuAccelerator::GpuAccelerator(int)
{
    printf("ERROR: GPU accelerator is not available in this build.\n");
}

GpuAccelerator::~GpuAccelerator() = default;

uint32_t GpuAccelerator::appendMesh(const std::vector<float3>&, const std::vector<uint3>&, const void*)
{
    return INVALID_SHAPE_ID;
}

void GpuAccelerator::appendSphere(const std::vector<float4>&, const void*) {}

void* GpuAccelerator::getVertexBuffer(uint32_t)
{
    return nullptr;
}

void GpuAccelerator::updateShape(uint32_t) {}

void GpuAccelerator::updateScene() {}

std::optional<Intersection> GpuAccelerator::intersect(const Ray&) const
{
    return {};
}

bool GpuAccelerator::intersectAny(const Ray&, float) const
{
    return false;
}

void GpuAccelerator::intersect(const Ray*, uint32_t n, std::optional<Intersection>* out) const
{
    for (uint32_t i = 0; i < n; ++i) out[i].reset();
}

void GpuAccelerator::intersectAny(const Ray*, const float*, uint32_t n, bool* out) const
{
    for (uint32_t i = 0; i < n; ++i) out[i] = false;
}

GpuLightSampler::GpuLightSampler(const std::vector<std::shared_ptr<Light>>& lights, Kind, int)
    : mLights(lights)
{
    printf("ERROR: GPU light sampler is not available in this build.\n");
}

GpuLightSampler::~GpuLightSampler() = default;

std::optional<LightSampler::Sample> GpuLightSampler::sample(float, const float3&, const float3&) const
{
    return {};
}

} // namespace alpine
