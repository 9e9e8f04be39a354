// Path-guiding proposal: input encoding and the 8-lobe von Mises-Fisher
// mixture head, shared by the gfx950 kernels (guide.hip, render.hip) and their
// host build (test hooks).  No counterpart in the reference.
//
// Network: 96 inputs -> 64 -> 64 -> 64 -> 64 -> 48 outputs (40 used), ReLU
// between layers, BF16 weights and activations, FP32 accumulation.
// Inputs: p = hit position normalised to the scene bounds (3), sin/cos of
// 2^k pi p for k = 0..11 on each axis (72), shading normal (3), outgoing
// direction (3), zero padding (15).
// Outputs per lobe i (5 each): mixture logit, log kappa, mean direction (3,
// normalised).  q(w) = sum_i softmax(a)_i kappa_i / (2 pi (1 - e^{-2 kappa_i}))
// exp(kappa_i (mu_i . w - 1)).
#pragma once

#include "shade_core.h"

namespace apg {

constexpr int GUIDE_IN = 96;
constexpr int GUIDE_HID = 64;
constexpr int GUIDE_OUT = 48;
constexpr int GUIDE_LOBES = 8;
constexpr int GUIDE_FREQS = 12;
constexpr int GUIDE_LAYERS = 5;
// weight offsets (N-major: W[n][k]) and bias offsets in the parameter vector
constexpr int GUIDE_W_OFF[GUIDE_LAYERS + 1] = { 0, 6144, 10240, 14336, 18432, 21504 };
constexpr int GUIDE_B_OFF[GUIDE_LAYERS + 1] = { 21504, 21568, 21632, 21696, 21760, 21808 };
constexpr int GUIDE_PARAMS = 21808;
constexpr int GUIDE_K[GUIDE_LAYERS] = { GUIDE_IN, GUIDE_HID, GUIDE_HID, GUIDE_HID, GUIDE_HID };
constexpr int GUIDE_N[GUIDE_LAYERS] = { GUIDE_HID, GUIDE_HID, GUIDE_HID, GUIDE_HID, GUIDE_OUT };
constexpr float GUIDE_LOGK_MIN = -3.0f, GUIDE_LOGK_MAX = 7.0f;
// the same tables as functions, usable with run-time layer indices in device code
SHD constexpr int guide_k(int l) { return l == 0 ? GUIDE_IN : GUIDE_HID; }
SHD constexpr int guide_n(int l) { return l == GUIDE_LAYERS - 1 ? GUIDE_OUT : GUIDE_HID; }
SHD constexpr int guide_w_off(int l) { return l == 0 ? 0 : 6144 + 4096 * (l - 1); }
SHD constexpr int guide_b_off(int l) { return 21504 + 64 * l; }

SHD f3 load3c(const float* a, uint32_t i)
{
    return mk(a[3 * size_t(i)], a[3 * size_t(i) + 1], a[3 * size_t(i) + 2]);
}

// sin(2 pi x) for x in [0, 1).
SHD float sin_rev(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sinf(x);
#else
    return sinf(6.28318530717958647692f * x);
#endif
}

// Feature f (0..95) of a query.  sin and cos of 2^k pi p are evaluated as
// sin(2 pi frac(2^(k-1) p + c / 4)), c = 0 / 1, so the phase is exact in
// float at every frequency.
SHD float guide_feature(int f, const float (&p)[3], const float (&n)[3], const float (&wo)[3])
{
    if (f < 3) return f == 0 ? p[0] : (f == 1 ? p[1] : p[2]);
    if (f < 75)
    {
        const int g = f - 3, a = g / 24, k = (g % 24) >> 1;
        const float pa = a == 0 ? p[0] : (a == 1 ? p[1] : p[2]);
        float x = ldexpf(pa, k - 1) + ((g & 1) ? 0.25f : 0.0f);
        x -= floorf(x);
        return sin_rev(x);
    }
    if (f < 78) return f == 75 ? n[0] : (f == 76 ? n[1] : n[2]);
    if (f < 81) return f == 78 ? wo[0] : (f == 79 ? wo[1] : wo[2]);
    return 0.0f;
}

// Encoding of one query into 96 floats.
SHD void guide_encode(f3 p01, f3 n, f3 wo, float* f)
{
    const float p[3] = { p01.x, p01.y, p01.z }, nn[3] = { n.x, n.y, n.z }, w[3] = { wo.x, wo.y, wo.z };
    for (int i = 0; i < GUIDE_IN; ++i) f[i] = guide_feature(i, p, nn, w);
}

struct VmfMix
{
    float pi[GUIDE_LOBES];
    float kappa[GUIDE_LOBES];
    f3 mu[GUIDE_LOBES];
};

SHD float clampk(float x) { return x < GUIDE_LOGK_MIN ? GUIDE_LOGK_MIN : (x > GUIDE_LOGK_MAX ? GUIDE_LOGK_MAX : x); }

SHD VmfMix guide_decode(const float* o)
{
    VmfMix m;
    float amax = -FLT_MAX;
#pragma unroll
    for (int i = 0; i < GUIDE_LOBES; ++i) amax = fmaxf(amax, o[5 * i]);
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i < GUIDE_LOBES; ++i)
    {
        m.pi[i] = expf(o[5 * i] - amax);
        sum += m.pi[i];
    }
#pragma unroll
    for (int i = 0; i < GUIDE_LOBES; ++i)
    {
        m.pi[i] /= sum;
        m.kappa[i] = expf(clampk(o[5 * i + 1]));
        f3 v = mk(o[5 * i + 2], o[5 * i + 3], o[5 * i + 4]);
        const float l = length(v);
        m.mu[i] = l > 1e-6f ? v / l : mk(0.0f, 0.0f, 1.0f);
    }
    return m;
}

// log of the normalised vMF density kappa / (2 pi (1 - e^{-2 kappa})) e^{kappa (c - 1)}
SHD float vmf_log(float kappa, float c)
{
    return logf(kappa) - logf(2.0f * kPi) - log1pf(-expf(-2.0f * kappa)) + kappa * (c - 1.0f);
}

SHD float guide_pdf(const VmfMix& m, f3 w)
{
    float q = 0.0f;
    // unrolled: a loop index into the mixture's arrays would move them to per-lane memory
#pragma unroll
    for (int i = 0; i < GUIDE_LOBES; ++i) q += m.pi[i] * expf(vmf_log(m.kappa[i], dot(m.mu[i], w)));
    return q;
}

// Sample a direction: lobe by u0, then the vMF of that lobe (Jakob, "Numerically stable
// sampling of the von Mises-Fisher distribution on S^2").
SHD f3 guide_sample(const VmfMix& m, float u0, float u1, float u2)
{
    // the first lobe whose running weight passes u0 (the last if none does), picked by
    // selects over the unrolled lobes: an index into the mixture's arrays would move them
    // from registers to per-lane memory
    float k = m.kappa[GUIDE_LOBES - 1];
    float mx = m.mu[GUIDE_LOBES - 1].x, my = m.mu[GUIDE_LOBES - 1].y, mz = m.mu[GUIDE_LOBES - 1].z;
    float acc = 0.0f;
    bool taken = false;
#pragma unroll
    for (int i = 0; i < GUIDE_LOBES; ++i)
    {
        acc += m.pi[i];
        const bool take = !taken && u0 < acc;
        k = take ? m.kappa[i] : k;
        mx = take ? m.mu[i].x : mx;
        my = take ? m.mu[i].y : my;
        mz = take ? m.mu[i].z : mz;
        taken = taken || take;
    }
    const f3 mu = mk(mx, my, mz);
    const float w = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * k)) / k;
    const float s = sqrtf(max_s(0.0f, 1.0f - w * w));
    const float phi = 2.0f * kPi * u2;
    f3 b1, b2;
    get_basis(mu, b1, b2);
    return normalize(b1 * (s * cosf(phi)) + b2 * (s * sinf(phi)) + mu * w);
}

// Gradient of -log q(w) with respect to the 40 outputs, scaled by `scale`.
SHD float guide_nll_grad(const float* o, f3 w, float scale, float* g)
{
    const VmfMix m = guide_decode(o);
    float lv[GUIDE_LOBES], lmax = -FLT_MAX;
    for (int i = 0; i < GUIDE_LOBES; ++i)
    {
        lv[i] = logf(max_s(m.pi[i], 1e-30f)) + vmf_log(m.kappa[i], dot(m.mu[i], w));
        lmax = fmaxf(lmax, lv[i]);
    }
    float sum = 0.0f;
    for (int i = 0; i < GUIDE_LOBES; ++i) sum += expf(lv[i] - lmax);
    const float logq = lmax + logf(sum);
    for (int i = 0; i < GUIDE_LOBES; ++i)
    {
        const float r = expf(lv[i] - logq);  // responsibility
        const float k = m.kappa[i];
        const float c = dot(m.mu[i], w);
        g[5 * i] = scale * (m.pi[i] - r);
        // d log vMF / d kappa = 1/k - 2 e^{-2k} / (1 - e^{-2k}) + (c - 1); kappa = exp(o)
        const float e2 = expf(-2.0f * k);
        const float dlk = 1.0f / k - 2.0f * e2 / (1.0f - e2) + (c - 1.0f);
        const float ok = o[5 * i + 1];
        const bool inside = ok > GUIDE_LOGK_MIN && ok < GUIDE_LOGK_MAX;
        g[5 * i + 1] = inside ? scale * (-r * dlk * k) : 0.0f;
        // d/d mu = -r k w; through mu = v / |v|
        const f3 v = mk(o[5 * i + 2], o[5 * i + 3], o[5 * i + 4]);
        const float l = length(v);
        if (l > 1e-6f)
        {
            const f3 dmu = w * (-r * k);
            const f3 dv = (dmu - m.mu[i] * dot(m.mu[i], dmu)) / l;
            g[5 * i + 2] = scale * dv.x;
            g[5 * i + 3] = scale * dv.y;
            g[5 * i + 4] = scale * dv.z;
        }
        else
            g[5 * i + 2] = g[5 * i + 3] = g[5 * i + 4] = 0.0f;
    }
    for (int i = 5 * GUIDE_LOBES; i < GUIDE_OUT; ++i) g[i] = 0.0f;
    return -logq;
}

// True density of the material's BSDF sampler (mat_sample) for a direction wi
// (local frame).  The reference's Microfacet::computePdf returns D / (4 wo.wh)
// (microfacet.cpp:123-142), which is not the density of its VNDF sampler; a
// one-sample mixture with another strategy needs the real one:
// G1(wo) D(wh) / (4 |wo.z|).
SHD float mat_true_pdf(const DevMaterial& m, f3 wo, f3 wi)
{
    auto mf_true = [&](f3 o, f3 i) -> float {
        if (!same_hemi(o, i)) return 0.0f;
        f3 wh = o + i;
        const float l = length(wh);
        if (l == 0.0f || o.z == 0.0f) return 0.0f;
        wh = wh / l;
        if (!same_hemi(o, wh)) wh = -wh;
        const float g1 = 1.0f / (1.0f + mf_lambda(m, o));
        return g1 * mf_d(m, wh) / (4.0f * fabsf(o.z));
    };
    if (m.kind == MAT_CONDUCTOR) return mf_true(wo, wi);
    float p = max_s(wi.z, 0.0f) / kPi;
    if (m.nbsdf == 2) p += mf_true(wo, wi);
    return p / float(m.nbsdf);
}

} // namespace apg
