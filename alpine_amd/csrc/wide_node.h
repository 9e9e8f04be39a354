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
//   dw6     cmask: triangle count mask.  Nibble s holds (1 << count) - 1 for
//           a leaf slot s (leaves hold 1..4 triangles) and 0 otherwise, so
//           bit 4 s + i stands for triangle i of slot s and the triangle's
//           index is tri_base + popcount(cmask & ((1 << (4 s + i)) - 1)).
//   dw7..15 6-bit quantised planes.  For axis a and side (0 = lo, 1 = hi)
//           the eight slots' planes form a 48-bit chunk (slot s at bit 6 s);
//           its low 32 bits are q[2 a + side] and its high 16 bits are
//           bits 16 side .. 16 side + 15 of q[6 + a].  A ray picks its near
//           and far chunk of an axis with two word selects and two shifts.
//           Empty slots carry lo = 63, hi = 0 and a zero cmask nibble.
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

#if defined(__HIPCC__)
#define APG_NODE_FN __host__ __device__ static inline
#else
#define APG_NODE_FN static inline
#endif

struct alignas(64) ApgWideNode
{
    float origin[3];
    uint32_t exp_imask;
    uint32_t child_base;
    uint32_t tri_base;
    uint32_t cmask;
    uint32_t q[9];
};
static_assert(sizeof(ApgWideNode) == 64, "wide node must be 64 bytes");

// Plane j of slot s: j = 0..2 lo x/y/z, 3..5 hi x/y/z.
APG_NODE_FN void apg_put_plane(uint32_t q[9], uint32_t slot, uint32_t plane, uint32_t v)
{
    const uint32_t a = plane % 3u, side = plane / 3u;
    const uint32_t bit = 6u * slot;
    v &= 63u;
    if (bit + 6u <= 32u)
        q[2u * a + side] |= v << bit;
    else if (bit >= 32u)
        q[6u + a] |= v << (bit - 32u + 16u * side);
    else
    {
        q[2u * a + side] |= (v & 3u) << 30;
        q[6u + a] |= (v >> 2) << (16u * side);
    }
}

APG_NODE_FN uint32_t apg_get_plane(const uint32_t q[9], uint32_t slot, uint32_t plane)
{
    const uint32_t a = plane % 3u, side = plane / 3u;
    const uint32_t bit = 6u * slot;
    const uint32_t lo = q[2u * a + side], hi = q[6u + a] >> (16u * side);
    if (bit + 6u <= 32u) return (lo >> bit) & 63u;
    if (bit >= 32u) return (hi >> (bit - 32u)) & 63u;
    return ((lo >> 30) | (hi << 2)) & 63u;
}

// Count mask nibble of a leaf slot holding `count` (1..4) triangles.
APG_NODE_FN uint32_t apg_leaf_nibble(uint32_t slot, uint32_t count)
{
    return ((1u << count) - 1u) << (4u * slot);
}
