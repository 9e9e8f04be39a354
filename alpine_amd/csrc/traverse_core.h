// Per-ray traversal steps shared by the gfx950 kernel and its host build.
//
// The kernel (traverse.hip) runs these functions one ray per lane; the host
// build of the same functions (apg_debug_host_trace) lets tests replay a ray's
// traversal on the CPU against the oracle without a GPU.
#pragma once

#include "wide_node.h"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#define APG_HD __host__ __device__ __forceinline__

namespace apg {

APG_HD uint32_t bit_ffs(uint32_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return uint32_t(__ffs(x));
#else
    return uint32_t(__builtin_ffs(int(x)));
#endif
}

APG_HD uint32_t bit_popc(uint32_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return uint32_t(__popc(x));
#else
    return uint32_t(__builtin_popcount(x));
#endif
}

APG_HD float fast_rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

APG_HD float fma_f(float a, float b, float c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fmaf(a, b, c);
#else
    return std::fma(a, b, c);
#endif
}

APG_HD uint32_t plane6(const uint32_t (&q)[9], uint32_t slot, uint32_t j)
{
    const uint32_t bit = 36u * slot + 6u * j;
    const uint32_t w = bit >> 5, s = bit & 31u;
    uint32_t v;
    if (s > 26u)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        v = __builtin_amdgcn_alignbit(q[w + 1], q[w], s);
#else
        v = uint32_t(((uint64_t(q[w + 1]) << 32) | q[w]) >> s);
#endif
    }
    else
        v = q[w] >> s;
    return v & 63u;
}

// Map an 8-bit slot mask to traversal order: slot s -> bit s ^ oct.
APG_HD uint32_t octant_permute(uint32_t x, uint32_t oct)
{
    if (oct & 1u) x = ((x & 0x55u) << 1) | ((x >> 1) & 0x55u);
    if (oct & 2u) x = ((x & 0x33u) << 2) | ((x >> 2) & 0x33u);
    if (oct & 4u) x = ((x & 0x0fu) << 4) | ((x >> 4) & 0x0fu);
    return x;
}

struct RayBox
{
    float ox, oy, oz;
    float ix, iy, iz;  // reciprocal directions (zero components nudged to +-1e-18)
    uint32_t oct;      // bit a set = negative along axis a
};

APG_HD RayBox make_raybox(float ox, float oy, float oz, float dx, float dy, float dz)
{
    const float tiny = 1e-18f;
    RayBox r;
    r.ox = ox; r.oy = oy; r.oz = oz;
    const float sx = fabsf(dx) < tiny ? copysignf(tiny, dx) : dx;
    const float sy = fabsf(dy) < tiny ? copysignf(tiny, dy) : dy;
    const float sz = fabsf(dz) < tiny ? copysignf(tiny, dz) : dz;
    r.ix = fast_rcp(sx);
    r.iy = fast_rcp(sy);
    r.iz = fast_rcp(sz);
    r.oct = (r.ix < 0.0f ? 1u : 0u) | (r.iy < 0.0f ? 2u : 0u) | (r.iz < 0.0f ? 4u : 0u);
    return r;
}

// The root as a node group holding one inner child in slot 0.
APG_HD uint32_t root_group_bits(uint32_t oct) { return 1u | (1u << (8u + oct)); }

// Pop the next child node index from a node group; returns the remaining hit
// bits (traversal order) through `rest`.
APG_HD uint32_t group_next_node(uint32_t g_base, uint32_t g_bits, uint32_t oct, uint32_t& rest)
{
    uint32_t hits = g_bits >> 8;
    const uint32_t pos = bit_ffs(hits) - 1u;
    const uint32_t slot = pos ^ oct;
    rest = hits & (hits - 1u);
    const uint32_t gim = g_bits & 0xffu;
    return g_base + bit_popc(gim & ((1u << slot) - 1u));
}

// Test the 8 children of one node (16 dwords w[0..15]) against the ray with
// current closest distance tmax.  Produces the new node group and triangle
// group.
APG_HD void node_visit(const uint32_t* w, const RayBox& r, float tmax, uint32_t& g_base,
                       uint32_t& g_bits, uint32_t& t_base, uint32_t& t_mask)
{
    float orgx, orgy, orgz;
#if defined(__HIP_DEVICE_COMPILE__)
    orgx = __uint_as_float(w[0]); orgy = __uint_as_float(w[1]); orgz = __uint_as_float(w[2]);
#else
    __builtin_memcpy(&orgx, &w[0], 4); __builtin_memcpy(&orgy, &w[1], 4); __builtin_memcpy(&orgz, &w[2], 4);
#endif
    const uint32_t ei = w[3];
    float stx, sty, stz;
    {
        uint32_t bx_ = (ei & 0xffu) << 23, by_ = ((ei >> 8) & 0xffu) << 23, bz_ = ((ei >> 16) & 0xffu) << 23;
#if defined(__HIP_DEVICE_COMPILE__)
        stx = __uint_as_float(bx_); sty = __uint_as_float(by_); stz = __uint_as_float(bz_);
#else
        __builtin_memcpy(&stx, &bx_, 4); __builtin_memcpy(&sty, &by_, 4); __builtin_memcpy(&stz, &bz_, 4);
#endif
    }
    const uint32_t nim = ei >> 24;
    const uint32_t child_base = w[4], tri_base = w[5], meta = w[6];
    const uint32_t q[9] = { w[7], w[8], w[9], w[10], w[11], w[12], w[13], w[14], w[15] };

    const float ax = stx * r.ix, ay = sty * r.iy, az = stz * r.iz;
    const float bx = (orgx - r.ox) * r.ix, by = (orgy - r.oy) * r.iy, bz = (orgz - r.oz) * r.iz;
    const bool px = r.ix >= 0.0f, py = r.iy >= 0.0f, pz = r.iz >= 0.0f;

    uint32_t hitmask = 0;
#pragma unroll
    for (uint32_t s = 0; s < APG_WIDE; ++s)
    {
        const uint32_t lx = plane6(q, s, 0), ly = plane6(q, s, 1), lz = plane6(q, s, 2);
        const uint32_t hx = plane6(q, s, 3), hy = plane6(q, s, 4), hz = plane6(q, s, 5);
        const float tnx = fma_f(float(px ? lx : hx), ax, bx);
        const float tfx = fma_f(float(px ? hx : lx), ax, bx);
        const float tny = fma_f(float(py ? ly : hy), ay, by);
        const float tfy = fma_f(float(py ? hy : ly), ay, by);
        const float tnz = fma_f(float(pz ? lz : hz), az, bz);
        const float tfz = fma_f(float(pz ? hz : lz), az, bz);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
        const float tf = fminf(fminf(tfx, tfy), tfz);
        if (tn <= tf && tn < tmax) hitmask |= 1u << s;
    }

    const uint32_t inner = hitmask & nim;
    g_base = child_base;
    g_bits = nim | (octant_permute(inner, r.oct) << 8);

    const uint32_t leafs = (meta >> 16) & ~nim & 0xffu;
    uint32_t tm = 0, off = 0;
#pragma unroll
    for (uint32_t s = 0; s < APG_WIDE; ++s)
    {
        if ((leafs >> s) & 1u)
        {
            const uint32_t cnt = ((meta >> (2u * s)) & 3u) + 1u;
            if ((hitmask >> s) & 1u) tm |= ((1u << cnt) - 1u) << off;
            off += cnt;
        }
    }
    t_base = tri_base;
    t_mask = tm;
}

// Moller-Trumbore exactly as the reference (bvh_common.cpp:66-103), combined
// with the closest/any acceptance t < tmax (bvh.cpp:287-307).  The caller's
// translation unit is compiled with -ffp-contract=off.
APG_HD bool tri_hit(const float* tp, float ox, float oy, float oz, float dx, float dy, float dz,
                    float tmax, float& t_out, float& b0_out, float& b1_out)
{
    const float v0x = tp[0], v0y = tp[1], v0z = tp[2];
    const float e0x = tp[3], e0y = tp[4], e0z = tp[5];
    const float e1x = tp[6], e1y = tp[7], e1z = tp[8];
    const float sx = ox - v0x, sy = oy - v0y, sz = oz - v0z;
    const float s1x = dy * e1z - dz * e1y;
    const float s1y = dz * e1x - dx * e1z;
    const float s1z = dx * e1y - dy * e1x;
    const float s2x = sy * e0z - sz * e0y;
    const float s2y = sz * e0x - sx * e0z;
    const float s2z = sx * e0y - sy * e0x;
    const float det = s1x * e0x + s1y * e0y + s1z * e0z;
    if (fabsf(det) < FLT_EPSILON) return false;
    const float t = (s2x * e1x + s2y * e1y + s2z * e1z) / det;
    if (!(t >= 0.0f) || !(t < tmax)) return false;
    const float b0 = (s1x * sx + s1y * sy + s1z * sz) / det;
    if (b0 < 0.0f || b0 > 1.0f) return false;
    const float b1 = (s2x * dx + s2y * dy + s2z * dz) / det;
    if (b1 < 0.0f || b1 > 1.0f - b0) return false;
    t_out = t;
    b0_out = b0;
    b1_out = b1;
    return true;
}

} // namespace apg
