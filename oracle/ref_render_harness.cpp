// ORACLE -- test infrastructure only.
//
// extern "C" harness around the reference's own render-path translation
// units, compiled from the unmodified sources by oracle/Makefile (target
// `ref`) into oracle/_ref/libalpine_ref_render.so:
//   sampler.cpp, camera.cpp, utils/util.cpp, utils/alias_table.cpp,
//   materials/{lambertian,microfacet,dielectric,conductor}.cpp,
//   lights/{disk_light,point_light}.cpp,
//   light_samplers/{uniform,power,bvh}_light_sampler.cpp.
// tests/golden/gen_golden_render.py uses it to write input/output vectors.
#include <camera.h>
#include <lights/disk_light.h>
#include <lights/point_light.h>
#include <light_samplers/bvh_light_sampler.h>
#include <light_samplers/power_light_sampler.h>
#include <light_samplers/uniform_light_sampler.h>
#include <materials/conductor.h>
#include <materials/dielectric.h>
#include <ray.h>
#include <sampler.h>
#include <shapes/shape.h>

#include <cstdint>
#include <memory>
#include <vector>

using namespace alpine;

namespace {
std::unique_ptr<Material> makeMaterial(uint32_t kind, const float* bc, const float* alpha)
{
    float2 a(alpha[0], alpha[1]);
    float3 c(bc[0], bc[1], bc[2]);
    if (kind == 1) return std::make_unique<Conductor>(a, c, nullptr, nullptr);
    return std::make_unique<Dielectric>(a, c, nullptr, nullptr);
}

// spec: kind (0 disk, 1 point), then 11 floats: pos xyz, normal xyz, color xyz, power, radius
std::vector<std::shared_ptr<Light>> makeLights(const float* spec, uint32_t n)
{
    std::vector<std::shared_ptr<Light>> out;
    for (uint32_t i = 0; i < n; ++i)
    {
        const float* s = spec + 12 * i;
        float3 pos(s[1], s[2], s[3]), nrm(s[4], s[5], s[6]), col(s[7], s[8], s[9]);
        if (s[0] == 0.0f) out.push_back(std::make_shared<DiskLight>(s[10], col, pos, nrm, s[11]));
        else out.push_back(std::make_shared<PointLight>(s[10], col, pos));
    }
    return out;
}
}

extern "C" {

void ref_sampler_sequence(uint32_t seed, uint32_t n, float* out)
{
    Sampler s;
    s.reset(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = s.get1D();
}

void ref_camera_rays(const float* eye, const float* target, const float* up, float fovy,
                     float aspect, const float* xy, uint32_t n, float* org, float* dir)
{
    Camera cam;
    cam.setLookAt(eye, target, up, fovy, aspect);
    for (uint32_t i = 0; i < n; ++i)
    {
        Ray r = cam.generateRay(xy[2 * i], xy[2 * i + 1]);
        for (int a = 0; a < 3; ++a) { org[3 * i + a] = r.org[a]; dir[3 * i + a] = r.dir[a]; }
    }
}

// Material::sample for n (wo, u) pairs; ok[i] = 0 when no sample.
void ref_material_sample(uint32_t kind, const float* bc, const float* alpha, const float* wo,
                         const float* u, uint32_t n, uint8_t* ok, float* est, float* wi, float* pdf)
{
    auto m = makeMaterial(kind, bc, alpha);
    IntersectionAttributes attr;
    for (uint32_t i = 0; i < n; ++i)
    {
        auto s = m->sample(float3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), float2(u[2 * i], u[2 * i + 1]), attr);
        ok[i] = s.has_value();
        if (!s) { pdf[i] = 0; for (int a = 0; a < 3; ++a) { est[3 * i + a] = 0; wi[3 * i + a] = 0; } continue; }
        for (int a = 0; a < 3; ++a) { est[3 * i + a] = s->estimator[a]; wi[3 * i + a] = s->wi[a]; }
        pdf[i] = s->pdf;
    }
}

void ref_material_eval(uint32_t kind, const float* bc, const float* alpha, const float* wo,
                       const float* wi, uint32_t n, float* f, float* pdf)
{
    auto m = makeMaterial(kind, bc, alpha);
    IntersectionAttributes attr;
    for (uint32_t i = 0; i < n; ++i)
    {
        float3 o(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), w(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        float3 b = m->computeBsdf(o, w, attr);
        for (int a = 0; a < 3; ++a) f[3 * i + a] = b[a];
        pdf[i] = m->computePdf(o, w);
    }
}

// LightSampler::sample (kind 0 uniform, 1 power, 2 bvh) for n (u, hit, ns); idx = -1 if none.
void ref_light_sampler(uint32_t kind, const float* spec, uint32_t nl, const float* u,
                       const float* hit, const float* ns, uint32_t n, int32_t* idx, float* pdf)
{
    auto lights = makeLights(spec, nl);
    std::unique_ptr<LightSampler> ls;
    if (kind == 0) ls = std::make_unique<UniformLightSampler>(lights);
    else if (kind == 1) ls = std::make_unique<PowerLightSampler>(lights);
    else ls = std::make_unique<BvhLightSampler>(lights);
    for (uint32_t i = 0; i < n; ++i)
    {
        auto s = ls->sample(u[i], float3(hit[3 * i], hit[3 * i + 1], hit[3 * i + 2]),
                            float3(ns[3 * i], ns[3 * i + 1], ns[3 * i + 2]));
        idx[i] = -1;
        pdf[i] = 0.0f;
        if (!s) continue;
        for (uint32_t k = 0; k < nl; ++k)
            if (lights[k].get() == s->light) idx[i] = int32_t(k);
        pdf[i] = s->pdf;
    }
}

// Light::sample and Light::computePdf of one light for n queries.
void ref_light_sample(const float* spec, const float* u, const float* hit, uint32_t n,
                      uint8_t* ok, float* le, float* wi, float* dist, float* pdf)
{
    auto L = makeLights(spec, 1)[0];
    for (uint32_t i = 0; i < n; ++i)
    {
        auto s = L->sample(float2(u[2 * i], u[2 * i + 1]), float3(hit[3 * i], hit[3 * i + 1], hit[3 * i + 2]));
        ok[i] = s.has_value();
        if (!s) { dist[i] = pdf[i] = 0; for (int a = 0; a < 3; ++a) le[3 * i + a] = wi[3 * i + a] = 0; continue; }
        for (int a = 0; a < 3; ++a) { le[3 * i + a] = s->emittedRadiance[a]; wi[3 * i + a] = s->wiWorld[a]; }
        dist[i] = s->distance;
        pdf[i] = s->pdf;
    }
}

void ref_light_pdf(const float* spec, const float* hit, const float* wi, uint32_t n, float* pdf,
                   float* dist)
{
    auto L = makeLights(spec, 1)[0];
    for (uint32_t i = 0; i < n; ++i)
    {
        auto [p, d] = L->computePdf(float3(hit[3 * i], hit[3 * i + 1], hit[3 * i + 2]),
                                    float3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]));
        pdf[i] = p;
        dist[i] = d;
    }
}
}
