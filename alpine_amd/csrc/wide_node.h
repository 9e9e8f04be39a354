// 64-byte quantised 8-wide BVH node shared by the builders (host and GPU) and
// the traversal kernels.
//
// Layout (16 dwords):
//   dw0..2  origin x, y, z (float).  Each origin is a multiple of 2^e of its
//           axis, so origin + q * 2^e is an exact float for q in [0, 63].
//   dw3     bits 0..23: biased exponents ex, ey, ez (e + 127, one byte each);
//           bits 24..31: imask (bit s set = slot s holds an inner node).
//   dw4     child_base: index of the first inner child node; inner children
//           are stored contiguously in slot order.
//   dw5     tri_base: index of the first triangle of the first leaf slot; the
//           leaf slots' triangles follow one another in slot order.
//   dw6     bits 0..15: (count - 1) of each leaf slot, 2 bits per slot
//           (leaves hold 1..4 triangles); bits 16..23: valid-slot mask.
//   dw7..15 288-bit stream of 6-bit quantised planes: slot s, plane j
//           (0..5 = lo x, lo y, lo z, hi x, hi y, hi z) at bit 36 * s + 6 * j.
//           Empty slots carry lo = 63, hi = 0, which no ray can hit.
//
// Quantised boxes are conservative: lo = origin + qlo * 2^e <= child min - d
// and hi = origin + qhi * 2^e >= child max + d for a scene-wide slack
// d = 2^-19 * (largest absolute coordinate), which absorbs the rounding of
// the slab test for rays that start inside the scene bounds.
#pragma once

#include <stdint.h>

#define APG_WIDE 8
#define APG_QBITS 6
#define APG_QMAX 63
#define APG_MAX_LEAF 4
#define APG_NODE_BYTES 64
#define APG_TRI_FLOATS 9

struct alignas(64) ApgWideNode
{
    float origin[3];
    uint32_t exp_imask;
    uint32_t child_base;
    uint32_t tri_base;
    uint32_t meta;
    uint32_t q[9];
};
static_assert(sizeof(ApgWideNode) == 64, "wide node must be 64 bytes");

// Host-side packing helper for the 288-bit plane stream.
static inline void apg_put_plane(uint32_t q[9], uint32_t slot, uint32_t plane, uint32_t v)
{
    uint32_t bit = 36u * slot + 6u * plane;
    uint32_t w = bit >> 5, s = bit & 31u;
    q[w] |= (v & 63u) << s;
    if (s > 26u) q[w + 1] |= (v & 63u) >> (32u - s);
}

static inline uint32_t apg_get_plane(const uint32_t q[9], uint32_t slot, uint32_t plane)
{
    uint32_t bit = 36u * slot + 6u * plane;
    uint32_t w = bit >> 5, s = bit & 31u;
    uint32_t v = q[w] >> s;
    if (s > 26u) v |= q[w + 1] << (32u - s);
    return v & 63u;
}
