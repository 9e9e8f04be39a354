// Shading steps of the wavefront path tracer, shared by the gfx950 kernels
// (render.hip) and their host build (test hooks in debug_host.hip).
//
// Every function restates the reference's arithmetic in the same operand
// order (files cited per function); the translation units that include this
// header are compiled with -ffp-contract=off.  cosf/sinf/powf/tanf are the
// platform's (ocml on the GPU, libm on the host), so results agree with the
// reference to a few ulps rather than bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#define SHD __host__ __device__ __forceinline__

namespace apg {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kOneMinusEps = 1.0f - FLT_EPSILON;  // util.h:8
constexpr float kRayOffset = 0.001f;                 // alpine_config.h:7

struct f3
{
    float x, y, z;
};
SHD f3 mk(float x, float y, float z) { return f3{ x, y, z }; }
SHD f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
SHD f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
SHD f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
SHD f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
SHD f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
SHD f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
SHD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SHD float length(f3 a) { return sqrtf(dot(a, a)); }
SHD f3 cross(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
SHD f3 normalize(f3 a) { return a / length(a); }
SHD float min_s(float a, float b) { return (b < a) ? b : a; }
SHD float max_s(float a, float b) { return (a < b) ? b : a; }
SHD float clamp_s(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
SHD float maxc(f3 v) { return max_s(v.x, max_s(v.y, v.z)); }
SHD f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

// ---- sampler (sampler.cpp:7-42) ----
SHD void sampler_reset(uint32_t s[4], uint32_t seed)
{
    for (uint32_t i = 1; i <= 4; ++i) s[i - 1] = seed = 1812433253u * (seed ^ (seed >> 30)) + i;
}
SHD uint32_t sampler_next(uint32_t s[4])
{
    uint32_t t = s[3];
    const uint32_t a = s[0];
    s[3] = s[2];
    s[2] = s[1];
    s[1] = a;
    t ^= t << 11;
    t ^= t >> 8;
    return s[0] = t ^ a ^ (a >> 19);
}
SHD float get1d(uint32_t s[4])
{
    const uint32_t r = sampler_next(s);
    const double v = double(r) / double(UINT32_MAX);
    return min_s(float(v), kOneMinusEps);
}
struct f2
{
    float x, y;
};
SHD f2 get2d(uint32_t s[4])
{
    f2 r;
    r.x = get1d(s);
    r.y = get1d(s);
    return r;
}

// ---- util.cpp:6-51, util.h, vector.h:214-222 ----
SHD f2 concentric_disk(f2 u)
{
    const f2 o{ 2.0f * u.x - 1.0f, 2.0f * u.y - 1.0f };
    if (o.x == 0 && o.y == 0) return f2{ 0.0f, 0.0f };
    float theta, r;
    if (fabsf(o.x) > fabsf(o.y))
    {
        r = o.x;
        theta = 0.25f * kPi * (o.y / o.x);
    }
    else
    {
        r = o.y;
        theta = 0.5f * kPi * (1.0f - 0.5f * (o.x / o.y));
    }
    return f2{ r * cosf(theta), r * sinf(theta) };
}
SHD f3 cosine_hemisphere(f2 u, float& pdf)
{
    const f2 d = concentric_disk(u);
    const float z = sqrtf(max_s(0.0f, 1.0f - d.x * d.x - d.y * d.y));
    pdf = z / kPi;
    return normalize(mk(d.x, d.y, z));
}
SHD float cos2t(f3 w) { return w.z * w.z; }
SHD float sin2t(f3 w) { return max_s(1.0f - cos2t(w), 0.0f); }
SHD float sint(f3 w) { return sqrtf(sin2t(w)); }
SHD float tan2t(f3 w) { return sin2t(w) / cos2t(w); }
SHD float cosphi(f3 w) { const float st = sint(w); return st == 0.0f ? 1.0f : clamp_s(w.x / st, -1.0f, 1.0f); }
SHD float sinphi(f3 w) { const float st = sint(w); return st == 0.0f ? 0.0f : clamp_s(w.y / st, -1.0f, 1.0f); }
SHD f3 reflect(f3 v, f3 n) { return -v + n * 2.0f * dot(v, n); }
SHD bool same_hemi(f3 a, f3 b) { return a.z * b.z > 0.0f; }
SHD float power_heuristic(float a, float b) { return (a * a) / (a * a + b * b); }
SHD f3 schlick(f3 f0, f3 wi, f3 wh)
{
    const float f = powf(max_s(0.0f, 1.0f - dot(wi, wh)), 5.0f);
    return mk(f, f, f) + f0 * (1.0f - f);
}
SHD void get_basis(f3 v, f3& b1, f3& b2)
{
    b1 = fabsf(v.x) > fabsf(v.y) ? normalize(mk(-v.z, 0.0f, v.x)) : normalize(mk(0.0f, v.z, -v.y));
    b2 = cross(v, b1);
}
SHD f3 to_local(f3 v, f3 s, f3 t, f3 n) { return mk(dot(s, v), dot(t, v), dot(n, v)); }
SHD f3 to_world(f3 v, f3 s, f3 t, f3 n) { return s * v.x + t * v.y + n * v.z; }

// ---- materials (microfacet.cpp, lambertian.cpp, dielectric.cpp, conductor.cpp) ----
enum : uint32_t { MAT_DIELECTRIC = 0, MAT_CONDUCTOR = 1 };

struct DevMaterial
{
    float bc[3];
    float ax, ay;   // clamped >= 0.001 (microfacet.h:9-12)
    float f0;       // < 0: conductor (F0 = base colour)
    uint32_t kind;
    uint32_t nbsdf; // dielectric: 1 = Lambert only, 2 = Lambert + GGX coat
    uint32_t tex_bc, tex_n;  // base colour / normal texture: 0 = none, else texture index + 1
};

SHD DevMaterial make_material(uint32_t kind, const float bc[3], const float alpha[2])
{
    DevMaterial m;
    m.bc[0] = bc[0]; m.bc[1] = bc[1]; m.bc[2] = bc[2];
    m.tex_bc = 0;
    m.tex_n = 0;
    m.kind = kind;
    m.ax = max_s(0.001f, alpha[0]);
    m.ay = max_s(0.001f, alpha[1]);
    if (kind == MAT_CONDUCTOR)
    {
        m.f0 = -1.0f;
        m.nbsdf = 1;
    }
    else
    {
        const float refl = 0.5f;  // dielectric.cpp:10-11
        m.f0 = 0.16f * refl * refl;
        m.nbsdf = (alpha[0] * alpha[0] + alpha[1] * alpha[1] < 1.0f) ? 2u : 1u;
    }
    return m;
}

SHD float mf_d(const DevMaterial& m, f3 wh)
{
    const float cos4 = cos2t(wh) * cos2t(wh);
    const float c2p = cosphi(wh) * cosphi(wh);
    const float s2p = sinphi(wh) * sinphi(wh);
    const float term = 1.0f + tan2t(wh) * (c2p / (m.ax * m.ax) + s2p / (m.ay * m.ay));
    const float invD = kPi * m.ax * m.ay * cos4 * term * term;
    return invD != 0.0f ? 1.0f / invD : 0.0f;
}
SHD float mf_lambda(const DevMaterial& m, f3 v)
{
    const float a = sqrtf(cosphi(v) * cosphi(v) * m.ax * m.ax + sinphi(v) * sinphi(v) * m.ay * m.ay);
    return (-1.0f + sqrtf(1.0f + a * a * tan2t(v))) / 2.0f;
}
SHD f3 mf_f0(const DevMaterial& m) { return m.f0 < 0.0f ? ld3(m.bc) : mk(m.f0, m.f0, m.f0); }

SHD bool mf_sample(const DevMaterial& m, f3 wo, f2 u, f3& est, f3& wi, float& pdf)
{
    const f3 woh = normalize(mk(m.ax * wo.x, m.ay * wo.y, wo.z));
    f2 t = concentric_disk(u);
    const float s = 0.5f * (1.0f + woh.z);
    t.y = (1.0f - s) * sqrtf(1.0f - t.x * t.x) + s * t.y;
    const f3 nhh = mk(t.x, t.y, sqrtf(max_s(0.0f, 1.0f - t.x * t.x - t.y * t.y)));
    f3 b1, b2;
    get_basis(woh, b1, b2);
    const f3 nh = b1 * nhh.x + b2 * nhh.y + woh * nhh.z;
    f3 wh = normalize(mk(m.ax * nh.x, m.ay * nh.y, max_s(0.0f, nh.z)));
    if (!same_hemi(wo, wh)) wh = -wh;
    wi = reflect(wo, wh);
    if (!same_hemi(wo, wi)) return false;
    const f3 f = schlick(mf_f0(m), wi, wh);
    const float g1 = 1.0f / (1.0f + mf_lambda(m, wo));
    const float g2 = 1.0f / (1.0f + mf_lambda(m, wo) + mf_lambda(m, wi));
    const float d = mf_d(m, wh);
    est = f * g2 / g1;
    pdf = d / (4.0f * dot(wo, wh));
    return true;
}
SHD f3 mf_bsdf(const DevMaterial& m, f3 wo, f3 wi)
{
    const float co = fabsf(wo.z), ci = fabsf(wi.z);
    if (co == 0.0f || ci == 0.0f) return mk(0, 0, 0);
    f3 wh = wo + wi;
    const float l = length(wh);
    if (l == 0.0f) return mk(0, 0, 0);
    wh = wh / l;
    const f3 f = schlick(mf_f0(m), wi, wh);
    const float d = mf_d(m, wh);
    const float g2 = 1.0f / (1.0f + mf_lambda(m, wo) + mf_lambda(m, wi));
    return f * d * g2 / (4.0f * co * ci);
}
SHD float mf_pdf(const DevMaterial& m, f3 wo, f3 wi)
{
    f3 wh = wo + wi;
    const float l = length(wh);
    if (l == 0.0f) return 0.0f;
    wh = wh / l;
    return mf_d(m, wh) / (4.0f * dot(wo, wh));
}

SHD bool bsdf_i_sample(const DevMaterial& m, uint32_t i, f3 wo, f2 u, f3& est, f3& wi, float& pdf)
{
    if (i == 0)
    {
        wi = cosine_hemisphere(u, pdf);
        if (pdf == 0.0f) return false;
        est = ld3(m.bc);
        return true;
    }
    return mf_sample(m, wo, u, est, wi, pdf);
}
SHD f3 bsdf_i_eval(const DevMaterial& m, uint32_t i, f3 wo, f3 wi)
{
    return i == 0 ? ld3(m.bc) / kPi : mf_bsdf(m, wo, wi);
}
SHD float bsdf_i_pdf(const DevMaterial& m, uint32_t i, f3 wo, f3 wi)
{
    return i == 0 ? max_s(wi.z, 0.0f) / kPi : mf_pdf(m, wo, wi);
}

// Material::sample (conductor.cpp:19-25, dielectric.cpp:29-67)
SHD bool mat_sample(const DevMaterial& m, f3 wo, f2 u, f3& est, f3& wi, float& pdf)
{
    if (m.kind == MAT_CONDUCTOR) return mf_sample(m, wo, u, est, wi, pdf);
    const uint8_t idx = static_cast<uint8_t>(static_cast<uint32_t>(u.x * float(m.nbsdf)));
    u.x = u.x * float(m.nbsdf) - float(idx);
    if (!bsdf_i_sample(m, idx, wo, u, est, wi, pdf)) return false;
    const float c = wi.z;
    if (c <= 0.0f) return false;
    f3 f = est * pdf / c;
    for (uint32_t i = 0; i < m.nbsdf; ++i)
    {
        if (i != idx)
        {
            pdf += bsdf_i_pdf(m, i, wo, wi);
            f = f + bsdf_i_eval(m, i, wo, wi);
        }
    }
    pdf /= float(m.nbsdf);
    est = f * c / pdf;
    return true;
}
// Material::computeBsdf and computePdf
SHD void mat_eval(const DevMaterial& m, f3 wo, f3 wi, f3& f, float& pdf)
{
    if (m.kind == MAT_CONDUCTOR)
    {
        f = mf_bsdf(m, wo, wi);
        pdf = mf_pdf(m, wo, wi);
        return;
    }
    f3 s = mk(0, 0, 0);
    float ps = 0.0f;
    for (uint32_t i = 0; i < m.nbsdf; ++i)
    {
        s = s + bsdf_i_eval(m, i, wo, wi);
        ps += bsdf_i_pdf(m, i, wo, wi);
    }
    f = s;
    pdf = ps / float(m.nbsdf);
}

// ---- textures (texture.h:11-42), the environment map (environment_map.cpp:15-37) and
//      shading attributes (Mesh::getIntersectionAttributes, shapes/mesh.cpp:57-154) ----
struct DevTexture
{
    uint32_t offset, w, h, pad;  // texels [offset, offset + w * h) of the texel pool, row-major
};

// Texture::sample: the nearest texel of uv wrapped to [0, 1)^2 (texture.h:19-35).
SHD float4 texture_sample(const DevTexture& t, const float4* texels, f2 uv)
{
    const float u = uv.x - floorf(uv.x);
    const float v = uv.y - floorf(uv.y);
    uint32_t x = static_cast<uint32_t>(u * float(t.w));
    uint32_t y = static_cast<uint32_t>(v * float(t.h));
    x = t.w - 1u < x ? t.w - 1u : x;
    y = t.h - 1u < y ? t.h - 1u : y;
    return texels[t.offset + x + y * t.w];
}

// EnvironmentMap::sample: latitude-longitude lookup of the nearest texel.
SHD f3 environment_sample(const float4* texels, uint32_t w, uint32_t h, f3 dir)
{
    float u = atan2f(dir.z, dir.x) / (2.0f * kPi) + 0.5f;
    float v = acosf(clamp_s(dir.y, -1.0f, 1.0f)) / kPi;
    u = u - floorf(u);
    v = clamp_s(v, 0.0f, 1.0f);
    uint32_t x = static_cast<uint32_t>(u * float(w));
    uint32_t y = static_cast<uint32_t>(v * float(h));
    x = w - 1u < x ? w - 1u : x;
    y = h - 1u < y ? h - 1u : y;
    const float4 e = texels[y * w + x];
    return mk(e.x, e.y, e.z);
}

enum : uint32_t { TA_NORMALS = 1u, TA_TANGENTS = 2u, TA_UVS = 4u, TA_MORPH = 8u };

// Per global triangle: indices into the global normal / tangent and uv pools.
struct DevTriAttr
{
    uint32_t n[3];
    uint32_t flags;   // TA_*
    uint32_t uv[3];
    uint32_t morph;   // TA_MORPH: index of the shape's DevMorph
};

// Morph targets of one shape (Mesh::Data::Target, shapes/mesh.h:27-32): target j's normal and
// tangent deltas of the shape's normal vertex k are tnormals / ttangents[d_off + j * nn + k].
struct DevMorph
{
    uint32_t ntargets, w_off, n_base, nn, d_off, pad[3];
};

struct AttrView
{
    const DevTriAttr* tri;     // nullptr: every triangle shades with ns = ng, uv = (0, 0)
    const float4* normals;     // xyz
    const float4* tangents;    // xyz, w = handedness
    const float* uvs;          // 2 per uv vertex
    const DevMorph* morphs;
    const float* weights;
    const float4* tnormals;
    const float4* ttangents;
    const DevTexture* texs;
    const float4* texels;
};

SHD f3 xyz4(float4 v) { return mk(v.x, v.y, v.z); }
SHD f3 interp3(f3 a, f3 b, f3 c, float b0, float b1, float b2) { return a * b0 + b * b1 + c * b2; }

// Material::getBaseColor (dielectric.cpp:97-101, conductor.cpp:45-49).
SHD f3 material_base_color(const AttrView& V, const DevMaterial& m, f2 uv)
{
    if (m.tex_bc == 0u) return ld3(m.bc);
    return xyz4(texture_sample(V.texs[m.tex_bc - 1u], V.texels, uv));
}

// Mesh::getIntersectionAttributes: the shading normal ns and uv of a hit on global triangle
// `prim` with barycentrics (b1, b2) weighting vertices 1 and 2.  The basis is getBasis(ns).
SHD void shading_attributes(const AttrView& V, uint32_t prim, float b1, float b2, f3 ng,
                            const DevMaterial& m, f3& ns, f2& uv)
{
    ns = ng;
    uv = f2{ 0.0f, 0.0f };
    if (!V.tri) return;
    const DevTriAttr a = V.tri[prim];
    const float b0 = 1.0f - b1 - b2;
    if (a.flags & TA_UVS)
    {
        const float* u0 = V.uvs + 2 * size_t(a.uv[0]);
        const float* u1 = V.uvs + 2 * size_t(a.uv[1]);
        const float* u2 = V.uvs + 2 * size_t(a.uv[2]);
        uv.x = u0[0] * b0 + u1[0] * b1 + u2[0] * b2;
        uv.y = u0[1] * b0 + u1[1] * b1 + u2[1] * b2;
    }
    if (!(a.flags & TA_NORMALS)) return;
    const f3 n0 = xyz4(V.normals[a.n[0]]), n1 = xyz4(V.normals[a.n[1]]), n2 = xyz4(V.normals[a.n[2]]);
    ns = interp3(n0, n1, n2, b0, b1, b2);
    DevMorph M{};
    uint32_t k[3] = { 0, 0, 0 };
    if (a.flags & TA_MORPH)
    {
        // mesh.cpp:88-98: add each target's interpolated normal delta times its weight
        M = V.morphs[a.morph];
        for (int i = 0; i < 3; ++i) k[i] = a.n[i] - M.n_base;
        for (uint32_t j = 0; j < M.ntargets; ++j)
        {
            const float4* d = V.tnormals + M.d_off + size_t(j) * M.nn;
            ns = ns + interp3(xyz4(d[k[0]]), xyz4(d[k[1]]), xyz4(d[k[2]]), b0, b1, b2) * V.weights[M.w_off + j];
        }
        ns = normalize(ns);
    }
    if (!(a.flags & TA_TANGENTS)) return;
    // tangent space (mesh.cpp:104-144)
    f3 t[3], bt[3];
    float sgn[3];
    const f3 nb[3] = { n0, n1, n2 };
    for (int i = 0; i < 3; ++i)
    {
        const float4 tv = V.tangents[a.n[i]];
        t[i] = mk(tv.x, tv.y, tv.z);
        sgn[i] = tv.w;
    }
    if (a.flags & TA_MORPH)
    {
        for (int i = 0; i < 3; ++i)
        {
            for (uint32_t j = 0; j < M.ntargets; ++j)
                t[i] = t[i] + xyz4(V.ttangents[M.d_off + size_t(j) * M.nn + k[i]]) * V.weights[M.w_off + j];
            t[i] = normalize(t[i]);
        }
    }
    for (int i = 0; i < 3; ++i) bt[i] = normalize(cross(nb[i], t[i]) * sgn[i]);
    const f3 tan = interp3(t[0], t[1], t[2], b0, b1, b2);
    const f3 bitan = interp3(bt[0], bt[1], bt[2], b0, b1, b2);
    const f3 v = m.tex_n ? xyz4(texture_sample(V.texs[m.tex_n - 1u], V.texels, uv)) : mk(0.0f, 0.0f, 1.0f);
    ns = normalize(bitan * v.x + tan * v.y + ns * v.z);
}

// ---- lights (disk_light.cpp, point_light.cpp; quad and triangle area lights) ----
enum : uint32_t { LIGHT_DISK = 0, LIGHT_POINT = 1, LIGHT_QUAD = 2, LIGHT_TRIANGLE = 3 };

struct DevLight
{
    float pos[3], n[3], b[3], t[3];  // disk: centre, normal, basis; quad/tri: p0, normal
    float e0[3], e1[3];              // quad/tri edges
    float le[3];                     // emitted radiance (area lights) or intensity (point)
    float power[3];
    float lo[3], hi[3];
    float radius, area;
    uint32_t kind, pad;
};

SHD bool light_is_delta(const DevLight& L) { return L.kind == LIGHT_POINT; }

// Light::sample (disk_light.cpp:30-53, point_light.cpp:25-38)
SHD bool light_sample(const DevLight& L, f2 u, f3 hit, f3& le, f3& wi, float& dist, float& pdf)
{
    if (L.kind == LIGHT_POINT)
    {
        f3 w = ld3(L.pos) - hit;
        const float d2 = dot(w, w);
        if (d2 == 0.0f) return false;
        const float d = sqrtf(d2);
        wi = w / d;
        le = ld3(L.le) / d2;
        dist = d;
        pdf = 1.0f;
        return true;
    }
    f3 p;
    if (L.kind == LIGHT_DISK)
    {
        const f2 ds = concentric_disk(u);
        p = ld3(L.pos) + (ld3(L.b) * ds.x + ld3(L.t) * ds.y) * L.radius;
    }
    else if (L.kind == LIGHT_QUAD)
        p = ld3(L.pos) + ld3(L.e0) * u.x + ld3(L.e1) * u.y;
    else
    {
        const float su = sqrtf(u.x);
        p = ld3(L.pos) + ld3(L.e0) * (1.0f - su) + ld3(L.e1) * (u.y * su);
    }
    f3 w = p - hit;
    const float d = length(w);
    if (d == 0.0f) return false;
    w = w / d;
    const float c = max_s(0.0f, dot(-w, ld3(L.n)));
    if (c == 0.0f) return false;
    pdf = (d * d) / (L.area * c);
    le = ld3(L.le);
    wi = w;
    dist = d;
    return true;
}

// Light::computePdf (disk_light.cpp:55-81); quad/triangle: plane hit + inside test
SHD float light_pdf(const DevLight& L, f3 hit, f3 wi, float& dist)
{
    dist = 0.0f;
    if (L.kind == LIGHT_POINT) return 0.0f;
    const float c = max_s(0.0f, dot(-wi, ld3(L.n)));
    if (c == 0.0f) return 0.0f;
    const float d = dot(hit - ld3(L.pos), ld3(L.n)) / c;
    if (d < 0.0f) return 0.0f;
    const f3 p = hit + wi * d;
    bool inside;
    if (L.kind == LIGHT_DISK)
    {
        const f3 q = p - ld3(L.pos);
        inside = dot(q, q) < L.radius * L.radius;
    }
    else
    {
        const f3 q = p - ld3(L.pos), e0 = ld3(L.e0), e1 = ld3(L.e1);
        const float e00 = dot(e0, e0), e01 = dot(e0, e1), e11 = dot(e1, e1);
        const float q0 = dot(q, e0), q1 = dot(q, e1);
        const float det = e00 * e11 - e01 * e01;
        const float a = (q0 * e11 - q1 * e01) / det;
        const float b = (q1 * e00 - q0 * e01) / det;
        inside = L.kind == LIGHT_QUAD ? (a >= 0.0f && a <= 1.0f && b >= 0.0f && b <= 1.0f)
                                      : (a >= 0.0f && b >= 0.0f && a + b <= 1.0f);
    }
    if (!inside) return 0.0f;
    dist = d;
    return (d * d) / (L.area * c);
}

// ---- light samplers (uniform_light_sampler.cpp, power_light_sampler.cpp + alias_table.cpp,
//      bvh_light_sampler.cpp) ----
enum : uint32_t { LS_UNIFORM = 0, LS_POWER = 1, LS_BVH = 2 };

// Light BVH node as the host builder produces it (bvh_light_sampler.cpp:167-240).
struct DevLightNode
{
    float lo[3], hi[3];
    float power;
    int32_t child0, child1;  // -1 for a leaf
    int32_t light;           // leaf: light index
    int32_t parent;
    int32_t pad;
};

// What LightNode::importance reads of one node, with the box reduced ahead of time to
// its centre (lo + hi) * 0.5 and half diagonal 0.5 * |hi - lo| (the same float
// expressions importance() evaluates, so the values are identical).
struct DevLightChild
{
    float c[3];
    float br;
    float power;
    int32_t node;   // >= 0: an inner node (index of its pair); -1: a leaf; -2: an inner node
                    // without two children (sampling ends there without a light)
    int32_t light;  // leaf: light index, else -1
};

// The two children of inner node i, read together in one 64-byte record at every step of
// the descent (the importance of both children decides the branch).
struct __attribute__((aligned(64))) DevLightPair
{
    DevLightChild ch[2];
    int32_t pad[2];
};

struct LightSamplerView
{
    uint32_t kind;
    uint32_t n;
    const DevLight* lights;
    const float* alias_p;
    const float* alias_q;
    const uint32_t* alias_a;
    const DevLightPair* pairs;     // by inner node; a single-light tree keeps its leaf in pairs[0].ch[0]
    const int32_t* parent_slot;    // by node: 2 * parent + side, -1 for the root
    const int32_t* leaf_of_light;
    int32_t root_light;            // the light when the root is a leaf, else -1
};

// LightNode::importance (bvh_light_sampler.cpp:24-51)
SHD float light_importance(const DevLightChild& nd, f3 p, f3 n)
{
    f3 w = ld3(nd.c) - p;
    const float d2 = dot(w, w);
    float imp = nd.power / max_s(d2, nd.br);
    const float br = nd.br;
    const float br2 = br * br;
    if (d2 > 0.0f && d2 > br2)
    {
        const float sinB2 = br2 / d2;
        const float sinB = sqrtf(sinB2);
        const float cosB = sqrtf(1.0f - sinB2);
        w = normalize(w);
        const float cosI = fabsf(dot(w, n));
        const float sinI = sqrtf(1 - cosI * cosI);
        imp *= cosI <= cosB ? cosI * cosB + sinI * sinB : 1.0f;
    }
    return imp;
}

// BvhLightSampler::sample (bvh_light_sampler.cpp:267-319); returns the light index or -1.
SHD int32_t light_sampler_sample(const LightSamplerView& S, float u, f3 hit, f3 ns, float& pdf)
{
    if (S.n == 0) return -1;
    if (S.kind == LS_UNIFORM)
    {
        uint32_t i = static_cast<uint32_t>(u * float(S.n));
        if (i > S.n - 1) i = S.n - 1;
        pdf = 1.0f / float(S.n);
        return int32_t(i);
    }
    if (S.kind == LS_POWER)
    {
        uint32_t bin = static_cast<uint32_t>(u * float(S.n));
        if (bin > S.n - 1) bin = S.n - 1;
        const float up = min_s(u * float(S.n) - float(bin), kOneMinusEps);
        const uint32_t idx = up < S.alias_q[bin] ? bin : S.alias_a[bin];
        pdf = S.alias_p[idx];
        return int32_t(idx);
    }
    if (S.root_light >= 0)
    {
        if (!(light_importance(S.pairs[0].ch[0], hit, ns) > 0.0f)) return -1;
        pdf = 1.0f;
        return S.root_light;
    }
    int32_t node = 0;
    float p = 1.0f;
    for (;;)
    {
        const DevLightPair& P = S.pairs[node];
        const float w0 = light_importance(P.ch[0], hit, ns);
        const float w1 = light_importance(P.ch[1], hit, ns);
        if (w0 == 0.0f && w1 == 0.0f) return -1;
        const float up = (w0 + w1) * u;
        const int idx = up < w0 ? 0 : 1;
        const float wc = idx == 0 ? w0 : w1;
        p *= wc / (w0 + w1);
        u = idx == 0 ? min_s(up / w0, kOneMinusEps) : min_s((up - w0) / w1, kOneMinusEps);
        const DevLightChild& C = P.ch[idx];
        if (C.node >= 0)
        {
            node = C.node;
            continue;
        }
        // a leaf is returned when its importance (wc, the same evaluation) is positive
        if (C.light >= 0 && wc > 0.0f)
        {
            pdf = p;
            return C.light;
        }
        return -1;
    }
}

// Probability that light_sampler_sample picks `light` (the LightSampler::pdf extension).
SHD float light_sampler_pdf(const LightSamplerView& S, int32_t light, f3 hit, f3 ns)
{
    if (S.n == 0 || light < 0 || uint32_t(light) >= S.n) return 0.0f;
    if (S.kind == LS_UNIFORM) return 1.0f / float(S.n);
    if (S.kind == LS_POWER) return S.alias_p[light];
    if (S.root_light >= 0)
        return light == S.root_light && light_importance(S.pairs[0].ch[0], hit, ns) > 0.0f ? 1.0f : 0.0f;
    int32_t ps = S.parent_slot[S.leaf_of_light[light]];
    if (!(light_importance(S.pairs[ps >> 1].ch[ps & 1], hit, ns) > 0.0f)) return 0.0f;
    float p = 1.0f;
    while (ps >= 0)
    {
        const DevLightPair& P = S.pairs[ps >> 1];
        const float w0 = light_importance(P.ch[0], hit, ns);
        const float w1 = light_importance(P.ch[1], hit, ns);
        if (w0 == 0.0f && w1 == 0.0f) return 0.0f;
        p *= ((ps & 1) == 0 ? w0 : w1) / (w0 + w1);
        ps = S.parent_slot[ps >> 1];
    }
    return p;
}

// ---- camera (camera.cpp:13-29, 73-96) ----
struct DevCamera
{
    float eye[3];
    float vx[3], vy[3], vz[3];
    float film_w, film_h;
};

SHD void camera_ray(const DevCamera& c, float x, float y, f3& o, f3& d)
{
    const float fx = (x - 0.5f) * c.film_w;
    const float fy = (0.5f - y) * c.film_h;
    o = ld3(c.eye);
    d = normalize(ld3(c.vx) * fx + ld3(c.vy) * fy + ld3(c.vz) * 1.0f);
}

} // namespace apg
