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

// Select the r-th (0-based) set bit of m; r < popcount(m).
APG_HD uint32_t select_bit(uint32_t m, uint32_t r)
{
    uint32_t pos = 0, c;
    c = bit_popc(m & 0xffffu); if (r >= c) { r -= c; pos += 16u; m >>= 16; }
    c = bit_popc(m & 0xffu);   if (r >= c) { r -= c; pos += 8u;  m >>= 8; }
    c = bit_popc(m & 0xfu);    if (r >= c) { r -= c; pos += 4u;  m >>= 4; }
    c = bit_popc(m & 0x3u);    if (r >= c) { r -= c; pos += 2u;  m >>= 2; }
    if (r >= (m & 1u)) pos += 1u;
    return pos;
}

// Index of the triangle at count-mask bit p of a node (wide_node.h, dw6).
APG_HD uint32_t tri_index(uint32_t tri_base, uint32_t cmask, uint32_t p)
{
    return tri_base + bit_popc(cmask & ((1u << p) - 1u));
}

// Six bits of a 48-bit plane chunk held as a low word and a high half word.
template <uint32_t S>
APG_HD uint32_t chunk_field(uint32_t lo, uint32_t hi)
{
    if (S < 5u) return (lo >> (6u * S)) & 63u;
    if (S == 5u)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_alignbit(hi, lo, 30) & 63u;
#else
        return ((lo >> 30) | (hi << 2)) & 63u;
#endif
    }
    return (hi >> (6u * S - 32u)) & 63u;
}

// Spread the 8 bits of x to bit 0 of the 8 nibbles, then fill each nibble.
APG_HD uint32_t spread_nibbles(uint32_t x)
{
    x = (x | (x << 12)) & 0x000f000fu;
    x = (x | (x << 6)) & 0x03030303u;
    x = (x | (x << 3)) & 0x11111111u;
    return x * 15u;
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

// Slab test of one slot against near/far plane chunks of the three axes.
template <uint32_t S>
APG_HD bool slot_hit(const uint32_t (&nw)[3], const uint32_t (&nh)[3], const uint32_t (&fw)[3],
                     const uint32_t (&fh)[3], const float (&a)[3], const float (&b)[3], float tmax)
{
    const float tnx = fma_f(float(chunk_field<S>(nw[0], nh[0])), a[0], b[0]);
    const float tfx = fma_f(float(chunk_field<S>(fw[0], fh[0])), a[0], b[0]);
    const float tny = fma_f(float(chunk_field<S>(nw[1], nh[1])), a[1], b[1]);
    const float tfy = fma_f(float(chunk_field<S>(fw[1], fh[1])), a[1], b[1]);
    const float tnz = fma_f(float(chunk_field<S>(nw[2], nh[2])), a[2], b[2]);
    const float tfz = fma_f(float(chunk_field<S>(fw[2], fh[2])), a[2], b[2]);
    const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
    const float tf = fminf(fminf(tfx, tfy), tfz);
    return tn <= tf && tn < tmax;
}

// Test the 8 children of one node (16 dwords w[0..15]) against the ray with
// current closest distance tmax.  Produces the new node group and the
// triangle group: t_base, the hit triangles as count-mask bits (t_mask) and
// the node's count mask (t_cmask).
APG_HD void node_visit(const uint32_t* w, const RayBox& r, float tmax, uint32_t& g_base,
                       uint32_t& g_bits, uint32_t& t_base, uint32_t& t_mask, uint32_t& t_cmask)
{
    float org[3];
#if defined(__HIP_DEVICE_COMPILE__)
    org[0] = __uint_as_float(w[0]); org[1] = __uint_as_float(w[1]); org[2] = __uint_as_float(w[2]);
#else
    __builtin_memcpy(org, w, 12);
#endif
    const uint32_t ei = w[3];
    const float o[3] = { r.ox, r.oy, r.oz };
    const float inv[3] = { r.ix, r.iy, r.iz };
    float a[3], b[3];
    uint32_t nw[3], nh[3], fw[3], fh[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        const uint32_t sb = ((ei >> (8 * k)) & 0xffu) << 23;
        float st;
#if defined(__HIP_DEVICE_COMPILE__)
        st = __uint_as_float(sb);
#else
        __builtin_memcpy(&st, &sb, 4);
#endif
        a[k] = st * inv[k];
        b[k] = (org[k] - o[k]) * inv[k];
        // near chunk = lo planes on a positive axis, hi planes on a negative one
        const uint32_t sh = ((r.oct >> k) & 1u) << 4;
        const uint32_t lo = w[7 + 2 * k], hi = w[8 + 2 * k], half = w[13 + k];
        nw[k] = sh ? hi : lo;
        fw[k] = sh ? lo : hi;
        nh[k] = half >> sh;
        fh[k] = half >> (16u - sh);
    }
    uint32_t hitmask = 0;
    hitmask |= slot_hit<0>(nw, nh, fw, fh, a, b, tmax) ? 0x01u : 0u;
    hitmask |= slot_hit<1>(nw, nh, fw, fh, a, b, tmax) ? 0x02u : 0u;
    hitmask |= slot_hit<2>(nw, nh, fw, fh, a, b, tmax) ? 0x04u : 0u;
    hitmask |= slot_hit<3>(nw, nh, fw, fh, a, b, tmax) ? 0x08u : 0u;
    hitmask |= slot_hit<4>(nw, nh, fw, fh, a, b, tmax) ? 0x10u : 0u;
    hitmask |= slot_hit<5>(nw, nh, fw, fh, a, b, tmax) ? 0x20u : 0u;
    hitmask |= slot_hit<6>(nw, nh, fw, fh, a, b, tmax) ? 0x40u : 0u;
    hitmask |= slot_hit<7>(nw, nh, fw, fh, a, b, tmax) ? 0x80u : 0u;

    const uint32_t nim = ei >> 24;
    g_base = w[4];
    g_bits = nim | (octant_permute(hitmask & nim, r.oct) << 8);
    t_base = w[5];
    t_cmask = w[6];
    t_mask = spread_nibbles(hitmask) & t_cmask;
}

// ---- watertight triangle test (Woop, Benthin and Wald, "Watertight Ray/Triangle
//      Intersection", JCGT 2(1), 2013) ----
// The ray is sheared so that it runs along +z from the origin; the three 2-D
// edge functions of a triangle then decide the hit.  An edge shared by two
// triangles yields the same edge-function magnitude with opposite signs in
// both (the products are rounded identically and IEEE subtraction is
// sign-symmetric), and a zero is re-evaluated in double precision, so a ray
// through a shared edge or vertex hits at least one of the triangles.  Edges
// are inclusive and both faces are hit, as in the reference's test.
struct WtRay
{
    float ox, oy, oz;
    float sx, sy, sz;  // shear: d[kx] / d[kz], d[ky] / d[kz], 1 / d[kz]
    uint32_t k;        // kx | ky << 2 | kz << 4
};

APG_HD WtRay make_wtray(float ox, float oy, float oz, float dx, float dy, float dz)
{
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    const uint32_t kz = ax > ay ? (ax > az ? 0u : 2u) : (ay > az ? 1u : 2u);
    uint32_t kx = kz == 2u ? 0u : kz + 1u;
    uint32_t ky = kx == 2u ? 0u : kx + 1u;
    const float dk = kz == 0u ? dx : (kz == 1u ? dy : dz);
    if (dk < 0.0f)
    {
        const uint32_t s = kx;
        kx = ky;
        ky = s;
    }
    const float dkx = kx == 0u ? dx : (kx == 1u ? dy : dz);
    const float dky = ky == 0u ? dx : (ky == 1u ? dy : dz);
    WtRay r;
    r.ox = ox; r.oy = oy; r.oz = oz;
    r.sx = dkx / dk;
    r.sy = dky / dk;
    r.sz = 1.0f / dk;
    r.k = kx | (ky << 2) | (kz << 4);
    return r;
}

APG_HD float pick3(float a, float b, float c, uint32_t k) { return k == 0u ? a : (k == 1u ? b : c); }

// v: the triangle's vertices v0, v1, v2 (9 floats).  On a hit with 0 <= t < tmax returns
// t and the barycentrics weighting v1 and v2 (the reference's convention).
APG_HD bool tri_hit_watertight(const float* v, const WtRay& r, float tmax, float& t_out, float& b0_out,
                               float& b1_out)
{
    const uint32_t kx = r.k & 3u, ky = (r.k >> 2) & 3u, kz = r.k >> 4;
    const float ax = v[0] - r.ox, ay = v[1] - r.oy, az = v[2] - r.oz;
    const float bx = v[3] - r.ox, by = v[4] - r.oy, bz = v[5] - r.oz;
    const float cx = v[6] - r.ox, cy = v[7] - r.oy, cz = v[8] - r.oz;
    const float Akz = pick3(ax, ay, az, kz), Bkz = pick3(bx, by, bz, kz), Ckz = pick3(cx, cy, cz, kz);
    const float Ax = pick3(ax, ay, az, kx) - r.sx * Akz;
    const float Ay = pick3(ax, ay, az, ky) - r.sy * Akz;
    const float Bx = pick3(bx, by, bz, kx) - r.sx * Bkz;
    const float By = pick3(bx, by, bz, ky) - r.sy * Bkz;
    const float Cx = pick3(cx, cy, cz, kx) - r.sx * Ckz;
    const float Cy = pick3(cx, cy, cz, ky) - r.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f)
    {
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T = U * (r.sz * Akz) + V * (r.sz * Bkz) + W * (r.sz * Ckz);
    const float rcp = 1.0f / det;
    const float t = T * rcp;
    if (!(t >= 0.0f) || !(t < tmax)) return false;
    t_out = t;
    b0_out = V * rcp;
    b1_out = W * rcp;
    return true;
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
