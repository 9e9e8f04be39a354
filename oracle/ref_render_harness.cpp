// ORACLE -- test infrastructure only.
//
// extern "C" harness around the reference's own render-path translation
// units, compiled from the unmodified sources by oracle/Makefile (target
// `ref`) into oracle/_ref/libalpine_ref_render.so:
//   sampler.cpp, camera.cpp, utils/util.cpp, utils/alias_table.cpp,
//   materials/{lambertian,microfacet,dielectric,conductor}.cpp,
//   lights/{disk_light,point_light}.cpp,
//   light_samplers/{uniform,power,bvh}_light_sampler.cpp,
//   shapes/mesh.cpp, environment_map.cpp, animation.cpp.
// tests/golden/gen_golden_render.py uses it to write input/output vectors.
#include <accelerators/accelerator.h>
#include <animation.h>
#include <camera.h>
#include <environment_map.h>
#include <intersection.h>
#include <lights/disk_light.h>
#include <lights/point_light.h>
#include <light_samplers/bvh_light_sampler.h>
#include <light_samplers/power_light_sampler.h>
#include <light_samplers/uniform_light_sampler.h>
#include <materials/conductor.h>
#include <materials/dielectric.h>
#include <math/matrix.h>
#include <ray.h>
#include <sampler.h>
#include <shapes/mesh.h>
#include <shapes/shape.h>
#include <texture.h>

#include <cstdint>
#include <memory>
#include <vector>

using namespace alpine;

namespace {
// The accelerator side of Mesh::appendTo / Mesh::update: owns the vertex buffer the mesh writes.
class VertexBufferAccelerator final : public Accelerator
{
public:
    std::vector<float3> vertices;
    uint32_t appendMesh(const std::vector<float3>& v, const std::vector<uint3>&, const void*) override
    {
        vertices = v;
        return 0;
    }
    void appendSphere(const std::vector<float4>&, const void*) override {}
    void* getVertexBuffer(uint32_t) override { return vertices.data(); }
    void updateShape(uint32_t) override {}
    void updateScene() override {}
    std::optional<Intersection> intersect(const Ray&) const override { return {}; }
    bool intersectAny(const Ray&, float) const override { return false; }
};

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

// Mesh::getIntersectionAttributes after Animation::update at `time` (when keys are given), for
// a mesh whose triangles all use one material (kind, base colour, alpha, optional base-colour
// and normal textures).  Optional inputs may be null; index buffers null = the position indices.
// Outputs: the accelerator's vertex buffer after the update (nv x 3), and per query (prim, b1, b2)
// ns, ss, ts, uv and the material's base colour at uv.
void ref_mesh_attributes(const float* verts, uint32_t nv, const uint32_t* prims, uint32_t nt,
                         const float* normals, uint32_t nn, const uint32_t* normal_prims,
                         const float* tangents, const float* uvs, uint32_t nuv,
                         const uint32_t* uv_prims, uint32_t ntargets, const float* tpos,
                         const float* tnrm, const float* ttan, uint32_t nkeys,
                         const float* key_times, const float* key_weights, float time,
                         uint32_t mat_kind, const float* bc, const float* alpha, const float* bc_tex,
                         uint32_t bw, uint32_t bh, const float* n_tex, uint32_t nw, uint32_t nh,
                         const uint32_t* q_prim, const float* q_b, uint32_t nq, float* out_verts,
                         float* ns, float* ss, float* ts, float* uv, float* base_color)
{
    Mesh::Data d;
    for (uint32_t i = 0; i < nv; ++i) d.vertices.emplace_back(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    for (uint32_t i = 0; i < nt; ++i) d.prims.emplace_back(prims[3 * i], prims[3 * i + 1], prims[3 * i + 2]);
    if (normals)
    {
        for (uint32_t i = 0; i < nn; ++i) d.normals.emplace_back(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
        if (normal_prims)
            for (uint32_t i = 0; i < nt; ++i)
                d.normalPrims.emplace_back(normal_prims[3 * i], normal_prims[3 * i + 1], normal_prims[3 * i + 2]);
        if (tangents)
            for (uint32_t i = 0; i < nn; ++i)
                d.tangents.emplace_back(tangents[4 * i], tangents[4 * i + 1], tangents[4 * i + 2], tangents[4 * i + 3]);
    }
    if (uvs)
    {
        for (uint32_t i = 0; i < nuv; ++i) d.uvs.emplace_back(uvs[2 * i], uvs[2 * i + 1]);
        if (uv_prims)
            for (uint32_t i = 0; i < nt; ++i)
                d.uvPrims.emplace_back(uv_prims[3 * i], uv_prims[3 * i + 1], uv_prims[3 * i + 2]);
    }
    for (uint32_t j = 0; j < ntargets; ++j)
    {
        Mesh::Data::Target t;
        for (uint32_t i = 0; i < nv; ++i)
        {
            const float* p = tpos + 3 * (size_t(j) * nv + i);
            t.vertices.emplace_back(p[0], p[1], p[2]);
        }
        for (uint32_t i = 0; normals && i < nn; ++i)
        {
            const float* p = tnrm + 3 * (size_t(j) * nn + i);
            t.normals.emplace_back(p[0], p[1], p[2]);
            const float* q = ttan + 3 * (size_t(j) * nn + i);
            t.tangents.emplace_back(q[0], q[1], q[2]);
        }
        d.targets.push_back(std::move(t));
    }
    auto tex = [](const float* texels, uint32_t w, uint32_t h) -> std::shared_ptr<Texture4f> {
        if (!texels) return nullptr;
        std::vector<float4> v;
        for (size_t i = 0; i < size_t(w) * h; ++i)
            v.emplace_back(texels[4 * i], texels[4 * i + 1], texels[4 * i + 2], texels[4 * i + 3]);
        return std::make_shared<Texture4f>(w, h, std::move(v));
    };
    float2 a(alpha[0], alpha[1]);
    float3 c(bc[0], bc[1], bc[2]);
    std::shared_ptr<Material> mat;
    if (mat_kind == 1) mat = std::make_shared<Conductor>(a, c, tex(bc_tex, bw, bh), tex(n_tex, nw, nh));
    else mat = std::make_shared<Dielectric>(a, c, tex(bc_tex, bw, bh), tex(n_tex, nw, nh));
    d.materials.assign(nt, mat);

    VertexBufferAccelerator acc;
    auto mesh = std::make_shared<Mesh>(std::move(d), ntargets);
    mesh->appendTo(&acc);
    if (ntargets && nkeys >= 2)
    {
        Animation::Channel ch;
        for (uint32_t k = 0; k < nkeys; ++k)
            ch.morphKeyframes.push_back({ key_times[k], std::vector<float>(key_weights + size_t(k) * ntargets,
                                                                           key_weights + size_t(k + 1) * ntargets) });
        ch.shape = mesh;
        std::vector<Animation::Channel> chs;
        chs.push_back(std::move(ch));
        Animation anim(std::move(chs));
        anim.update(&acc, time);
    }
    for (size_t i = 0; i < size_t(nv) * 3; ++i) out_verts[i] = acc.vertices.empty() ? verts[i] : acc.vertices[i / 3][i % 3];
    for (uint32_t i = 0; i < nq; ++i)
    {
        Intersection isect;
        isect.primId = q_prim[i];
        isect.barycentric = float2(q_b[2 * i], q_b[2 * i + 1]);
        isect.ng = float3(0.0f, 0.0f, 1.0f);
        IntersectionAttributes at = mesh->getIntersectionAttributes(isect);
        float3 b = at.material->getBaseColor(at.uv);
        for (int k = 0; k < 3; ++k)
        {
            ns[3 * i + k] = at.ns[k];
            ss[3 * i + k] = at.ss[k];
            ts[3 * i + k] = at.ts[k];
            base_color[3 * i + k] = b[k];
        }
        uv[2 * i] = at.uv.x;
        uv[2 * i + 1] = at.uv.y;
    }
}

// math/matrix.h float4x4: prod = mul(m0, m1) (16 floats, column-major), mv[i] = mul(m0, v[i]).
void ref_mat4(const float* m0, const float* m1, const float* v, uint32_t n, float* prod, float* mv)
{
    float4x4 a, b;
    for (uint32_t i = 0; i < 16; ++i)
    {
        a[i] = m0[i];
        b[i] = m1[i];
    }
    float4x4 p = mul(a, b);
    for (uint32_t i = 0; i < 16; ++i) prod[i] = p[i];
    for (uint32_t i = 0; i < n; ++i)
    {
        float4 r = mul(a, float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]));
        for (int k = 0; k < 4; ++k) mv[4 * i + k] = r[k];
    }
}

// EnvironmentMap::sample for n directions over a width x height RGB map.
void ref_environment_sample(const float* rgb, uint32_t w, uint32_t h, const float* dirs, uint32_t n,
                            float* out)
{
    std::vector<float3> data;
    for (size_t i = 0; i < size_t(w) * h; ++i) data.emplace_back(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    EnvironmentMap env(std::move(data), w, h);
    for (uint32_t i = 0; i < n; ++i)
    {
        float3 e = env.sample(float3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        for (int k = 0; k < 3; ++k) out[3 * i + k] = e[k];
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
