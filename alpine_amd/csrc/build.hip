// GPU scene build for gfx950: triangle preparation, binned-SAH binary tree,
// collapse into 64-byte quantised 8-wide nodes, triangle packing.
//
// Binary tree (breadth first, one pass per level over the triangle slots):
//   bounds   node box and centroid box of every node of the level; min/max
//            are accumulated as order-preserving integer atomics, first in
//            LDS for the node that owns the start of a workgroup's slot chunk
//            (the top levels are one or a few big nodes) and then once per
//            workgroup into HBM, directly in HBM for the other nodes.
//   decide   per node: leaf (N <= threshold, zero centroid extent), forced
//            positional split (an SAH leaf above APG_MAX_LEAF triangles) or SAH
//            candidate; candidates get a histogram slot.
//   bin      9 bins x 3 axes of (box, count) per candidate, LDS-privatised
//            like the bounds pass.
//   sweep    per candidate the reference's forward/backward cost sweep and
//            strict-< argmin (bvh_common.cpp:185-272), leaf test
//            split cost >= N * SA (bvh_common.cpp:274-343); children are
//            allocated contiguously after the level.
//   partition stable partition of each node's slot range: a global exclusive
//            scan of the "below" flags gives every slot its new position.
// A node of at most FIN slots leaves the level loop; after the loop, k_finish builds each
// such subtree in one workgroup's LDS with the same rules (no per-level launches below).
// Min/max are exact and order independent and the sweep is sequential per
// node with the reference's operand order, so with the reference's leaf
// threshold the leaf sets equal the reference builder's.
//
// Wide collapse (breadth first over wide levels): per wide node, the
// reference's greedy opening of the largest-area inner child up to 8
// children (wide_bvh.cpp:200-243), octant slot assignment, quantisation;
// child nodes and triangle ranges are placed with exclusive scans.
#include "apg_internal.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstring>

namespace apg {

namespace {

constexpr uint32_t BLOCK = 256;
constexpr uint32_t ITEMS = 8;               // slots per thread
constexpr uint32_t CHUNK = BLOCK * ITEMS;   // slots per workgroup
constexpr uint32_t NBINS = 9;
constexpr uint32_t NSPLITS = 8;
constexpr uint32_t BIN_VALS = 7;            // lo xyz, hi xyz (ordered uints), count
constexpr uint32_t NODE_BIN_WORDS = 3 * NBINS * BIN_VALS;
constexpr uint32_t RETIRED = 0xFFFFFFFFu;
// Nodes of at most SMALL slots get one wave each for bounds and bins (no global
// atomics); larger nodes are accumulated by chunks of CHUNK slots.
constexpr uint32_t SMALL = CHUNK;
// In k_finish, nodes of at most TINY slots get one thread each: bounds by a serial loop, and
// the SAH sweep evaluates the splits from the slots directly (no histogram).
constexpr uint32_t TINY = 16;
// Nodes of at most FIN slots leave the level loop: one workgroup builds each one's whole
// subtree in LDS (k_finish).
constexpr uint32_t FIN = 1024;
static_assert(FIN >= TINY && FIN <= SMALL && FIN < 0x8000u, "finisher size");

enum : uint32_t { NF_LEAF = 1u, NF_SAH_LEAF = 2u, NF_FORCED = 4u, NF_SPLIT = 8u };

__host__ __device__ __forceinline__ uint32_t ord_of(float f)
{
    uint32_t u;
#if defined(__HIP_DEVICE_COMPILE__)
    u = __float_as_uint(f);
#else
    __builtin_memcpy(&u, &f, 4);
#endif
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__host__ __device__ __forceinline__ float float_of(uint32_t o)
{
    uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    float f;
#if defined(__HIP_DEVICE_COMPILE__)
    f = __uint_as_float(u);
#else
    __builtin_memcpy(&f, &u, 4);
#endif
    return f;
}

__device__ __forceinline__ float min_s(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float max_s(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ uint8_t bin_index(float c, float mn, float mx)
{
    uint8_t b = static_cast<uint8_t>(float(NBINS) * (c - mn) / (mx - mn));
    return b < NBINS - 1 ? b : uint8_t(NBINS - 1);
}

struct BuildState
{
    // triangles (original order)
    float* tri9_src;    // 9 per triangle: v0, e0, e1
    // slots, in two buffers: buffer 0 is (idx, pbox), buffer 1 (idx_tmp, pbox_tmp); level l
    // reads its nodes' slots from buffer l % 2 and partitions them into buffer (l + 1) % 2
    uint32_t* idx;      // slot -> triangle (in the end every slot's, written by k_finish)
    uint32_t* idx_tmp;
    float* pbox;        // 6 per slot: the triangle's lo xyz, hi xyz (k_prep: triangle order)
    float* pbox_tmp;
    uint8_t* flag8;     // per slot of the level's nodes: goes to the left child
    uint16_t* sbin;     // per slot of the level's SAH candidates: bin on axis a in bits 4a..4a+3
    // nodes (capacity 2n)
    uint32_t* nbox;     // 6 ordered uints per node (bounds accumulators)
    uint32_t* ncbox;    // 6 ordered uints per node (centroid bounds)
    uint32_t* nfirst;
    uint32_t* ncount;
    uint32_t* nleft;    // right = left + 1
    uint32_t* nflags;
    uint32_t* nsplit;   // dim | idx << 8
    float* nsplit_mm;   // mn, mx
    uint32_t* nbinslot; // histogram slot or RETIRED
    uint32_t* bins;     // NODE_BIN_WORDS per candidate
    uint32_t* counters; // [0] node count, [1] bin slots, [3] k_finish's unused node ids
    // the level being built: [0] first node, [1] end, [2] levels done, [3] listed large nodes,
    // [4] nodes left to k_finish, [5] k_finish's cursor over them, [6] node ids allocated,
    // [7] binary nodes built (allocated less k_finish's unused reservations), [9] levels that
    // held nodes, [10 + level % 2] nodes k_level_init listed
    uint32_t* lvl;
    uint32_t* large;      // the level's nodes (all of more than FIN slots)
    uint32_t* large_off;  // exclusive scan of their chunk counts (total at [count])
    uint32_t* chunk_below;  // per chunk of the level: slots going to the left child
    uint32_t* chunk_off;    // their exclusive scan over the level's chunks
    uint32_t* fin;        // nodes of at most FIN slots, left to k_finish (count at lvl[4]), with
                          // the buffer of their slots in bit 31
};

// The slot buffer a kernel of the current level reads (k_large_scan has advanced lvl[2]).
__device__ __forceinline__ uint32_t level_buffer(const BuildState& S) { return (S.lvl[2] - 1u) & 1u; }
__device__ __forceinline__ const uint32_t* slots_of(const BuildState& S, uint32_t b) { return b ? S.idx_tmp : S.idx; }
__device__ __forceinline__ const float* boxes_of(const BuildState& S, uint32_t b) { return b ? S.pbox_tmp : S.pbox; }

// ---------------------------------------------------------------------------
// triangle preparation (Primitive::updateVertices, bvh_common.cpp:105-125)
// ---------------------------------------------------------------------------
__global__ void k_prep(const float* __restrict__ xyz, const uint32_t* __restrict__ tri,
                       uint32_t n, float* __restrict__ tri9, float* __restrict__ pbox)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[3][3];
    for (int k = 0; k < 3; ++k)
    {
        const uint32_t vi = tri[3 * size_t(i) + k];
        for (int a = 0; a < 3; ++a) v[k][a] = xyz[3 * size_t(vi) + a];
    }
    float* t = tri9 + 9 * size_t(i);
    float* b = pbox + 6 * size_t(i);
    for (int a = 0; a < 3; ++a)
    {
        t[a] = v[0][a];
        t[3 + a] = v[1][a] - v[0][a];
        t[6 + a] = v[2][a] - v[0][a];
        float lo = FLT_MAX, hi = -FLT_MAX;
        for (int k = 0; k < 3; ++k) { lo = min_s(lo, v[k][a]); hi = max_s(hi, v[k][a]); }
        b[a] = lo;
        b[3 + a] = hi;
    }
}

// Rigid transform (3 x 4 row-major) of a vertex range of the device geometry.
__global__ void k_transform(float* xyz, uint32_t n, float m0, float m1, float m2, float m3, float m4,
                            float m5, float m6, float m7, float m8, float m9, float m10, float m11)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = xyz[3 * size_t(i)], y = xyz[3 * size_t(i) + 1], z = xyz[3 * size_t(i) + 2];
    xyz[3 * size_t(i)] = m0 * x + m1 * y + m2 * z + m3;
    xyz[3 * size_t(i) + 1] = m4 * x + m5 * y + m6 * z + m7;
    xyz[3 * size_t(i) + 2] = m8 * x + m9 * y + m10 * z + m11;
}

// The input vertices v0, v1, v2 of every triangle in tree order, for the watertight test
// (the shared vertices of neighbouring triangles must be bit-identical).
__global__ void k_gather_vertices(const uint32_t* tri_ids, uint32_t n, const uint32_t* tri, const float* xyz,
                                  float* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t g = tri_ids[i];
    float* o = out + 9 * size_t(i);
    for (int k = 0; k < 3; ++k)
    {
        const float* v = xyz + 3 * size_t(tri[3 * size_t(g) + k]);
        o[3 * k] = v[0];
        o[3 * k + 1] = v[1];
        o[3 * k + 2] = v[2];
    }
}

// Slot i holds triangle i (a shape's triangles are the slot range of its tree's root).
__global__ void k_init_slots(uint32_t n, uint32_t* idx)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) idx[i] = i;
}


// ---------------------------------------------------------------------------
// binary levels
// ---------------------------------------------------------------------------
// Every kernel of a level reads the level's node range [lvl[0], lvl[1]) from device memory
// and walks it with a grid-stride loop over a grid fixed for the build, so the host enqueues
// level after level without knowing their widths; k_large_scan moves on to the children.

// Per node of the level: hand the nodes of at most FIN slots to k_finish; clear the others'
// bounds accumulators and list them (their order in the list is arbitrary: what they
// accumulate is order independent, and each node's chunks stay contiguous in chunk order).
__global__ __launch_bounds__(256) void k_level_init(BuildState S)
{
    // the children the previous level allocated; the list count alternates between two words
    const uint32_t begin = S.lvl[1], end = S.counters[0];
    uint32_t* listed = &S.lvl[10 + (S.lvl[2] & 1u)];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = begin + blockIdx.x * blockDim.x + threadIdx.x; k < end; k += stride)
    {
        if (S.ncount[k] <= FIN)
        {
            // built by k_finish after the level loop; every kernel of the level passes it by
            S.nbinslot[k] = RETIRED;
            S.fin[atomicAdd(&S.lvl[4], 1u)] = k | ((S.lvl[2] & 1u) << 31);
            continue;
        }
        // the roots' bounds are accumulated by k_bounds_large; every later node's by the
        // partition that placed its slots
        if (S.lvl[2] == 0u)
            for (int a = 0; a < 3; ++a)
            {
                S.nbox[6 * size_t(k) + a] = ord_of(FLT_MAX);
                S.nbox[6 * size_t(k) + 3 + a] = ord_of(-FLT_MAX);
                S.ncbox[6 * size_t(k) + a] = ord_of(FLT_MAX);
                S.ncbox[6 * size_t(k) + 3 + a] = ord_of(-FLT_MAX);
            }
        const uint32_t j = atomicAdd(listed, 1u);
        S.large[j] = k;
        // the node's histogram (its list slot); forced nodes are split positionally, unbinned
        if (S.nflags[k] & NF_FORCED)
        {
            S.nbinslot[k] = RETIRED;
            continue;
        }
        S.nbinslot[k] = j;
        uint32_t* hb = S.bins + size_t(j) * NODE_BIN_WORDS;
        for (uint32_t e = 0; e < 3 * NBINS; ++e)
        {
            for (int a = 0; a < 3; ++a) { hb[e * BIN_VALS + a] = ord_of(FLT_MAX); hb[e * BIN_VALS + 3 + a] = ord_of(-FLT_MAX); }
            hb[e * BIN_VALS + 6] = 0u;
        }
    }
}

// One workgroup of SCAN1 threads: out[j] = v(0) + ... + v(j - 1) for j < m; returns the total.
constexpr uint32_t SCAN1 = 1024;
constexpr uint32_t SCAN1_ITEMS = 8;
template <typename F>
__device__ uint32_t wg_exclusive_scan(uint32_t m, F v, uint32_t* out)
{
    __shared__ uint32_t sh[SCAN1];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < m; base += SCAN1 * SCAN1_ITEMS)
    {
        const uint32_t j0 = base + threadIdx.x * SCAN1_ITEMS;
        uint32_t x[SCAN1_ITEMS], tot = 0;
        for (uint32_t k = 0; k < SCAN1_ITEMS; ++k)
        {
            x[k] = j0 + k < m ? v(j0 + k) : 0u;
            tot += x[k];
        }
        sh[threadIdx.x] = tot;
        __syncthreads();
        for (uint32_t off = 1; off < SCAN1; off <<= 1)
        {
            const uint32_t y = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
            __syncthreads();
            sh[threadIdx.x] += y;
            __syncthreads();
        }
        uint32_t run = carry + sh[threadIdx.x] - tot;
        for (uint32_t k = 0; k < SCAN1_ITEMS; ++k)
        {
            if (j0 + k < m) out[j0 + k] = run;
            run += x[k];
        }
        carry += sh[SCAN1 - 1];
        __syncthreads();
    }
    return carry;
}

// large_off = exclusive scan of the listed nodes' chunk counts (total at [count]); then the
// level becomes current: [lvl[0], lvl[1]) = the nodes k_level_init took, lvl[3] = listed.
__global__ __launch_bounds__(SCAN1) void k_large_scan(BuildState S)
{
    const uint32_t level = S.lvl[2];
    const uint32_t m = S.lvl[10 + (level & 1u)];
    const uint32_t total = wg_exclusive_scan(m, [&](uint32_t j) { return (S.ncount[S.large[j]] + CHUNK - 1) / CHUNK; },
                                             S.large_off);
    if (threadIdx.x == 0)
    {
        S.large_off[m] = total;
        const uint32_t begin = S.lvl[1], end = S.counters[0];
        S.lvl[0] = begin;
        S.lvl[1] = end;
        S.lvl[2] = level + 1u;
        S.lvl[3] = m;
        if (begin != end) S.lvl[9] += 1u;
        S.lvl[10 + ((level + 1u) & 1u)] = 0u;  // the next level's list count
    }
}

// The node, its list index and the slot range of chunk c of the listed nodes (false past the
// last chunk).
__device__ __forceinline__ bool chunk_range(const BuildState& S, uint32_t c, uint32_t& node, uint32_t& lo,
                                            uint32_t& hi, uint32_t* list_index = nullptr)
{
    const uint32_t m = S.lvl[3];
    if (c >= S.large_off[m]) return false;
    uint32_t a = 0, b = m;  // largest j with large_off[j] <= c
    while (b - a > 1)
    {
        const uint32_t mid = (a + b) >> 1;
        if (S.large_off[mid] <= c) a = mid;
        else b = mid;
    }
    node = S.large[a];
    if (list_index) *list_index = a;
    const uint32_t first = S.nfirst[node], cnt = S.ncount[node];
    lo = first + (c - S.large_off[a]) * CHUNK;
    hi = lo + CHUNK < first + cnt ? lo + CHUNK : first + cnt;
    return true;
}

// Bounds and centroid bounds of a large node's chunk: per-thread running
// min / max as order-preserving uints (-0 < +0), reduced over the wave and the
// workgroup, then one atomic per value into the node's entry.
__global__ __launch_bounds__(BLOCK) void k_bounds_large(BuildState S)
{
    __shared__ uint32_t part[BLOCK / 64u][12];
    if (S.lvl[2] != 1u) return;  // past level 0 the partition accumulates the bounds
    uint32_t node, lo_i, hi_i;
    if (!chunk_range(S, blockIdx.x, node, lo_i, hi_i)) return;
    const uint32_t lo_ord = ord_of(FLT_MAX), hi_ord = ord_of(-FLT_MAX);
    uint32_t r[12];
    for (int a = 0; a < 3; ++a)
    {
        r[a] = lo_ord; r[3 + a] = hi_ord; r[6 + a] = lo_ord; r[9 + a] = hi_ord;
    }
    const float* box = boxes_of(S, level_buffer(S));
    for (uint32_t i = lo_i + threadIdx.x; i < hi_i; i += BLOCK)
    {
        const float* b = box + 6 * size_t(i);
        for (int a = 0; a < 3; ++a)
        {
            const float lo = b[a], hi = b[3 + a], c = (lo + hi) * 0.5f;
            r[a] = min(r[a], ord_of(lo));
            r[3 + a] = max(r[3 + a], ord_of(hi));
            r[6 + a] = min(r[6 + a], ord_of(c));
            r[9 + a] = max(r[9 + a], ord_of(c));
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; ++a)
        {
            r[a] = min(r[a], uint32_t(__shfl_xor(int(r[a]), off)));
            r[3 + a] = max(r[3 + a], uint32_t(__shfl_xor(int(r[3 + a]), off)));
            r[6 + a] = min(r[6 + a], uint32_t(__shfl_xor(int(r[6 + a]), off)));
            r[9 + a] = max(r[9 + a], uint32_t(__shfl_xor(int(r[9 + a]), off)));
        }
    if ((threadIdx.x & 63u) == 0u)
        for (int j = 0; j < 12; ++j) part[threadIdx.x / 64u][j] = r[j];
    __syncthreads();
    if (threadIdx.x < 12)
    {
        const uint32_t j = threadIdx.x;
        uint32_t v = part[0][j];
        for (uint32_t w = 1; w < BLOCK / 64u; ++w) v = (j % 6) < 3 ? min(v, part[w][j]) : max(v, part[w][j]);
        uint32_t* dst = j < 6 ? &S.nbox[6 * size_t(node) + j] : &S.ncbox[6 * size_t(node) + (j - 6)];
        if ((j % 6) < 3) atomicMin(dst, v);
        else atomicMax(dst, v);
    }
}

__device__ __forceinline__ float box_area(const float lo[3], const float hi[3])
{
    const float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
    return (x * y + y * z + z * x) * 2.0f;
}

// 9-bin histograms (box + count) per axis of a large SAH candidate's chunk: eight LDS
// histograms per wave (lane % 8: lanes that share a bin then hit eight addresses, not one),
// merged into the
// node's global histogram.
constexpr uint32_t HCOPIES = 8;  // histogram copies per wave, by lane % HCOPIES
__global__ __launch_bounds__(BLOCK) void k_bin_large(BuildState S)
{
    constexpr uint32_t NH = (BLOCK / 64u) * HCOPIES;
    __shared__ uint32_t hist[NH][NODE_BIN_WORDS];
    uint32_t node, lo_i, hi_i;
    if (!chunk_range(S, blockIdx.x, node, lo_i, hi_i)) return;
    const uint32_t slot = S.nbinslot[node];
    if (slot == RETIRED) return;
    for (uint32_t j = threadIdx.x; j < NH * NODE_BIN_WORDS; j += BLOCK)
    {
        const uint32_t v = (j % NODE_BIN_WORDS) % BIN_VALS;
        (&hist[0][0])[j] = v < 3 ? ord_of(FLT_MAX) : (v < 6 ? ord_of(-FLT_MAX) : 0u);
    }
    float mn[3], mx[3];
    for (int d = 0; d < 3; ++d)
    {
        mn[d] = float_of(S.ncbox[6 * size_t(node) + d]);
        mx[d] = float_of(S.ncbox[6 * size_t(node) + 3 + d]);
    }
    __syncthreads();
    uint32_t* acc = hist[(threadIdx.x / 64u) * HCOPIES + (threadIdx.x % HCOPIES)];
    const float* box = boxes_of(S, level_buffer(S));
    for (uint32_t i = lo_i + threadIdx.x; i < hi_i; i += BLOCK)
    {
        const float* b = box + 6 * size_t(i);
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { lo[a] = b[a]; hi[a] = b[3 + a]; }
        uint32_t bins3 = 0;
        for (int d = 0; d < 3; ++d)
        {
            if (mn[d] == mx[d]) continue;
            const float c = (lo[d] + hi[d]) * 0.5f;
            const uint32_t bi = bin_index(c, mn[d], mx[d]);
            bins3 |= bi << (4 * d);
            uint32_t* e = acc + (d * NBINS + bi) * BIN_VALS;
            for (int a = 0; a < 3; ++a)
            {
                atomicMin(&e[a], ord_of(lo[a]));
                atomicMax(&e[3 + a], ord_of(hi[a]));
            }
            atomicAdd(&e[6], 1u);
        }
        S.sbin[i] = uint16_t(bins3);  // the partition's left test reads the split axis' bin
    }
    __syncthreads();
    uint32_t* g = S.bins + size_t(slot) * NODE_BIN_WORDS;
    for (uint32_t j = threadIdx.x; j < NODE_BIN_WORDS; j += BLOCK)
    {
        const uint32_t v = j % BIN_VALS;
        uint32_t x = hist[0][j];
        for (uint32_t w = 1; w < NH; ++w)
            x = v == 6 ? x + hist[w][j] : (v < 3 ? min(x, hist[w][j]) : max(x, hist[w][j]));
        if (v == 6) { if (x) atomicAdd(&g[j], x); }
        else if (v < 3) atomicMin(&g[j], x);
        else atomicMax(&g[j], x);
    }
}

// Two children of node k (first nb slots left, the rest right), allocated after the level.
__device__ __forceinline__ void alloc_children(BuildState& S, uint32_t k, uint32_t cnt, uint32_t nb, uint32_t child_fl)
{
    const uint32_t l = atomicAdd(&S.counters[0], 2u);
    const uint32_t first = S.nfirst[k];
    S.nleft[k] = l;
    S.ncount[l] = nb;
    S.ncount[l + 1] = cnt - nb;
    S.nfirst[l] = first;
    S.nfirst[l + 1] = first + nb;
    S.nflags[l] = child_fl;
    S.nflags[l + 1] = child_fl;
    // the children's bounds accumulators (k_part_scatter fills them as it moves the slots)
    for (uint32_t c = l; c < l + 2; ++c)
        for (int a = 0; a < 3; ++a)
        {
            S.nbox[6 * size_t(c) + a] = ord_of(FLT_MAX);
            S.nbox[6 * size_t(c) + 3 + a] = ord_of(-FLT_MAX);
            S.ncbox[6 * size_t(c) + a] = ord_of(FLT_MAX);
            S.ncbox[6 * size_t(c) + 3 + a] = ord_of(-FLT_MAX);
        }
}

// A positional split (a forced node, or an SAH leaf of more than APG_MAX_LEAF triangles): the
// first half rounded to whole leaves goes left.
__device__ __forceinline__ void positional_split(BuildState& S, uint32_t k, uint32_t fl, uint32_t cnt)
{
    uint32_t half = ((cnt + 1) / 2 + APG_MAX_LEAF - 1) / APG_MAX_LEAF * APG_MAX_LEAF;
    half = half < cnt - 1 ? half : cnt - 1;
    S.nflags[k] = fl;
    S.nsplit[k] = 0xFFu;
    S.nsplit_mm[2 * size_t(k)] = __uint_as_float(half);
    alloc_children(S, k, cnt, half, NF_FORCED);
}

// The SAH sweep of a node's 3 x 9-bin histograms (hb: per bin lo xyz and hi xyz as ordered
// uints, then a count) by one wave, for the axes in the mask `axes` (those of non-zero centroid
// extent).  Lane d * 9 + b loads bin b of axis d once; inclusive scans over the axis' nine lanes
// give each split s its below box and count (bins 0..s) and its above box and count (bins
// s+1..8).  Min and max are exact, so the boxes are those the reference's sequential sweep
// accumulates, up to the sign of a zero (which changes no area), and so are the costs
// (bn * area(below) + an * area(above), in that order).  Splits whose cost is not below
// FLT_MAX are not candidates; the first (axis, split) of least cost wins, as the sequential
// sweep's strict < picks it.  Every lane returns the winner: key = d * 8 + s (63: none), its
// cost and its below count; `at` is the lane that priced it.
struct SweepBest
{
    float cost;
    uint32_t key, below, at;
};

__device__ SweepBest sweep_bins_wave(const uint32_t* hb, uint32_t axes)
{
    const uint32_t lane = threadIdx.x & 63u;
    const bool in = lane < 3u * NBINS;
    const uint32_t d = lane / NBINS, b = lane % NBINS;
    float blo[3], bhi[3], alo[3], ahi[3];
    uint32_t bn = 0;
    for (int a = 0; a < 3; ++a) { blo[a] = FLT_MAX; bhi[a] = -FLT_MAX; }
    if (in)
    {
        const uint32_t* e = hb + lane * BIN_VALS;
        for (int a = 0; a < 3; ++a) { blo[a] = float_of(e[a]); bhi[a] = float_of(e[3 + a]); }
        bn = e[6];
    }
    for (int a = 0; a < 3; ++a) { alo[a] = blo[a]; ahi[a] = bhi[a]; }
    uint32_t an = bn;
    // below: bins 0..b; above: bins b..8 (shifted to s + 1 afterwards)
    for (uint32_t off = 1; off < NBINS; off <<= 1)
    {
        const int up = int((lane - off) & 63u), dn = int((lane + off) & 63u);
        const bool take_up = b >= off, take_dn = b + off < NBINS;
        float t[12];
        for (int a = 0; a < 3; ++a)
        {
            t[a] = __shfl(blo[a], up); t[3 + a] = __shfl(bhi[a], up);
            t[6 + a] = __shfl(alo[a], dn); t[9 + a] = __shfl(ahi[a], dn);
        }
        const uint32_t tbn = uint32_t(__shfl(int(bn), up)), tan = uint32_t(__shfl(int(an), dn));
        if (take_up)
        {
            for (int a = 0; a < 3; ++a) { blo[a] = min_s(blo[a], t[a]); bhi[a] = max_s(bhi[a], t[3 + a]); }
            bn += tbn;
        }
        if (take_dn)
        {
            for (int a = 0; a < 3; ++a) { alo[a] = min_s(alo[a], t[6 + a]); ahi[a] = max_s(ahi[a], t[9 + a]); }
            an += tan;
        }
    }
    {
        const int nx = int((lane + 1u) & 63u);
        for (int a = 0; a < 3; ++a) { alo[a] = __shfl(alo[a], nx); ahi[a] = __shfl(ahi[a], nx); }
        an = uint32_t(__shfl(int(an), nx));
    }
    SweepBest r{ __builtin_inff(), 63u, 0u, 0u };
    if (in && b < NSPLITS && ((axes >> d) & 1u))
    {
        float c = 0.0f;
        c += float(bn) * box_area(blo, bhi);
        c += float(an) * box_area(alo, ahi);
        if (c < FLT_MAX)
        {
            r.cost = c;
            r.key = d * NSPLITS + b;
        }
    }
    for (int off = 1; off < 32; off <<= 1)
    {
        const float oc = __shfl_xor(r.cost, off);
        const uint32_t ok = uint32_t(__shfl_xor(int(r.key), off));
        if (oc < r.cost || (oc == r.cost && ok < r.key)) { r.cost = oc; r.key = ok; }
    }
    r.cost = __shfl(r.cost, 0);
    r.key = uint32_t(__shfl(int(r.key), 0));
    r.at = r.key == 63u ? 0u : (r.key / NSPLITS) * NBINS + r.key % NSPLITS;
    r.below = uint32_t(__shfl(int(bn), int(r.at)));
    return r;
}

// Sweep, split choice and child allocation of one node of the level, by one wave
// (sweep_bins_wave over the node's global histogram).
__device__ void sweep_node_wave(BuildState& S, uint32_t k, uint32_t lane, uint32_t thr)
{
    const uint32_t cnt = S.ncount[k];
    if (cnt <= FIN) return;  // k_finish's
    // the leaf test (a node of the level loop holds more than FIN > APG_MAX_LEAF triangles, so
    // forced nodes and leaves alike are split positionally)
    uint32_t fl = S.nflags[k];
    bool positional = (fl & NF_FORCED) != 0u;
    if (positional) fl |= NF_SPLIT;
    else
    {
        bool leaf = cnt <= thr;
        if (!leaf)
        {
            float clo[3], chi[3];
            for (int a = 0; a < 3; ++a)
            {
                clo[a] = float_of(S.ncbox[6 * size_t(k) + a]);
                chi[a] = float_of(S.ncbox[6 * size_t(k) + 3 + a]);
            }
            const float diag = max_s(chi[0] - clo[0], max_s(chi[1] - clo[1], chi[2] - clo[2]));
            if (diag == 0.0f) leaf = true;
        }
        if (leaf)
        {
            fl |= NF_SAH_LEAF | NF_FORCED | NF_SPLIT;
            positional = true;
        }
    }
    if (positional)
    {
        if (lane == 0u) positional_split(S, k, fl, cnt);
        return;
    }
    const uint32_t* hb = S.bins + size_t(S.nbinslot[k]) * NODE_BIN_WORDS;
    uint32_t axes = 0;
    for (int a = 0; a < 3; ++a)
        if (float_of(S.ncbox[6 * size_t(k) + a]) != float_of(S.ncbox[6 * size_t(k) + 3 + a])) axes |= 1u << a;
    const SweepBest sb = sweep_bins_wave(hb, axes);
    const float cost = sb.cost;
    const uint32_t best = sb.key, nbelow = sb.below;
    if (lane != 0u) return;
    float nlo[3], nhi[3];
    for (int a = 0; a < 3; ++a)
    {
        nlo[a] = float_of(S.nbox[6 * size_t(k) + a]);
        nhi[a] = float_of(S.nbox[6 * size_t(k) + 3 + a]);
    }
    const float leaf_cost = float(cnt) * box_area(nlo, nhi);
    // no plane below FLT_MAX (surface areas overflowed): a leaf, where the reference's
    // default split would send every primitive to one side and never terminate
    if (best == 63u || cost >= leaf_cost)
    {
        positional_split(S, k, fl | NF_SAH_LEAF | NF_FORCED | NF_SPLIT, cnt);
        return;
    }
    S.nflags[k] = fl | NF_SPLIT;
    S.nsplit[k] = (best / NSPLITS) | ((best % NSPLITS) << 8);
    S.nsplit_mm[2 * size_t(k)] = float_of(S.ncbox[6 * size_t(k) + best / NSPLITS]);
    S.nsplit_mm[2 * size_t(k) + 1] = float_of(S.ncbox[6 * size_t(k) + 3 + best / NSPLITS]);
    alloc_children(S, k, cnt, nbelow, 0u);
}

__global__ __launch_bounds__(256) void k_sweep(BuildState S, uint32_t thr)
{
    const uint32_t begin = S.lvl[0], end = S.lvl[1];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (blockDim.x / 64u);
    for (uint32_t k = begin + (blockIdx.x * blockDim.x + threadIdx.x) / 64u; k < end; k += waves)
        sweep_node_wave(S, k, lane, thr);
}

// ---------------------------------------------------------------------------
// partition of the level's nodes, chunk by chunk (slots of other nodes are not touched)
// ---------------------------------------------------------------------------
// Whether slot i of split node nd goes to the left child: its bin on the split axis (recorded
// by k_bin_large with the same centroid range) at most the split bin, or for a positional split
// its place in the first half.
__device__ __forceinline__ bool goes_left(const BuildState& S, uint32_t nd, uint32_t first, uint32_t i)
{
    const uint32_t sp = S.nsplit[nd];
    if ((sp & 0xFFu) == 0xFFu) return (i - first) < __float_as_uint(S.nsplit_mm[2 * size_t(nd)]);
    const uint32_t d = sp & 0xFFu, si = sp >> 8;
    return ((uint32_t(S.sbin[i]) >> (4 * d)) & 15u) <= si;
}

// Per chunk: each slot's left flag (a byte in flag8) and the chunk's count of them.  Thread t
// takes slots lo + t, lo + t + BLOCK, ... (coalesced).
__global__ __launch_bounds__(BLOCK) void k_part_count(BuildState S)
{
    __shared__ uint32_t part[BLOCK / 64u];
    uint32_t node, lo, hi;
    if (!chunk_range(S, blockIdx.x, node, lo, hi)) return;
    const uint32_t first = S.nfirst[node];
    uint32_t cnt = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += BLOCK)
    {
        const bool l = goes_left(S, node, first, i);
        S.flag8[i] = l ? 1u : 0u;
        cnt += l ? 1u : 0u;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += uint32_t(__shfl_xor(int(cnt), off));
    if ((threadIdx.x & 63u) == 0u) part[threadIdx.x / 64u] = cnt;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t t = 0;
        for (uint32_t w = 0; w < BLOCK / 64u; ++w) t += part[w];
        S.chunk_below[blockIdx.x] = t;
    }
}

// One workgroup: chunk_off = exclusive scan of chunk_below over the level's chunks.
__global__ __launch_bounds__(SCAN1) void k_part_scan(BuildState S)
{
    const uint32_t m = S.large_off[S.lvl[3]];
    wg_exclusive_scan(m, [&](uint32_t c) { return S.chunk_below[c]; }, S.chunk_off);
}

// Per chunk: every slot's place in its node's stable partition (left slots first), in the
// other slot buffer (triangle and box move together).
// Slot lo + q * BLOCK + t is thread t's q-th; its rank among the chunk's left slots counts
// the earlier rounds, the earlier waves of its round and the earlier lanes of its wave.
__global__ __launch_bounds__(BLOCK) void k_part_scatter(BuildState S)
{
    constexpr uint32_t W = BLOCK / 64u;
    __shared__ uint32_t wl[ITEMS][W];
    __shared__ uint32_t part[W][24];
    uint32_t node, lo, hi, j;
    if (!chunk_range(S, blockIdx.x, node, lo, hi, &j)) return;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x / 64u;
    uint64_t masks[ITEMS];
    for (uint32_t q = 0; q < ITEMS; ++q)
    {
        const uint32_t i = lo + q * BLOCK + threadIdx.x;
        masks[q] = __ballot(i < hi && S.flag8[i] != 0u);
        if (lane == 0u) wl[q][wave] = uint32_t(__popcll(masks[q]));
    }
    __syncthreads();
    const uint32_t first = S.nfirst[node], l = S.nleft[node];
    const uint32_t nleft_total = S.ncount[l];
    const uint32_t buf = level_buffer(S);
    const uint32_t* idx_in = slots_of(S, buf);
    const float* box_in = boxes_of(S, buf);
    uint32_t* idx_out = buf ? S.idx : S.idx_tmp;
    float* box_out = buf ? S.pbox : S.pbox_tmp;
    // bounds and centroid bounds of the two children (ordered uints: left 0..11, right 12..23)
    uint32_t r[24];
    for (int c = 0; c < 2; ++c)
        for (int a = 0; a < 3; ++a)
        {
            r[12 * c + a] = ord_of(FLT_MAX); r[12 * c + 3 + a] = ord_of(-FLT_MAX);
            r[12 * c + 6 + a] = ord_of(FLT_MAX); r[12 * c + 9 + a] = ord_of(-FLT_MAX);
        }
    // left slots of the node in its earlier chunks
    uint32_t left = S.chunk_off[blockIdx.x] - S.chunk_off[S.large_off[j]];
    const uint64_t below_lane = (1ull << lane) - 1ull;
    for (uint32_t q = 0; q < ITEMS; ++q)
    {
        const uint32_t i = lo + q * BLOCK + threadIdx.x;
        uint32_t before = left;
        for (uint32_t w = 0; w < wave; ++w) before += wl[q][w];
        before += uint32_t(__popcll(masks[q] & below_lane));
        if (i < hi)
        {
            const bool is_left = ((masks[q] >> lane) & 1ull) != 0ull;
            const uint32_t pos = is_left ? first + before : first + nleft_total + (i - first - before);
            idx_out[pos] = idx_in[i];
            // a box is 6 floats at a multiple of 24 bytes: three 8-byte moves
            const float2* src = reinterpret_cast<const float2*>(box_in + 6 * size_t(i));
            const float2 b0 = src[0], b1 = src[1], b2 = src[2];
            float2* dst = reinterpret_cast<float2*>(box_out + 6 * size_t(pos));
            dst[0] = b0;
            dst[1] = b1;
            dst[2] = b2;
            const float lo3[3] = { b0.x, b0.y, b1.x }, hi3[3] = { b1.y, b2.x, b2.y };
            const uint32_t o = is_left ? 0u : 12u;
            for (int a = 0; a < 3; ++a)
            {
                const uint32_t c = ord_of((lo3[a] + hi3[a]) * 0.5f);
                const uint32_t vlo = ord_of(lo3[a]), vhi = ord_of(hi3[a]);
                // both children's accumulators, the other one unchanged
                r[a] = o ? r[a] : min(r[a], vlo);
                r[12 + a] = o ? min(r[12 + a], vlo) : r[12 + a];
                r[3 + a] = o ? r[3 + a] : max(r[3 + a], vhi);
                r[15 + a] = o ? max(r[15 + a], vhi) : r[15 + a];
                r[6 + a] = o ? r[6 + a] : min(r[6 + a], c);
                r[18 + a] = o ? min(r[18 + a], c) : r[18 + a];
                r[9 + a] = o ? r[9 + a] : max(r[9 + a], c);
                r[21 + a] = o ? max(r[21 + a], c) : r[21 + a];
            }
        }
        for (uint32_t w = 0; w < W; ++w) left += wl[q][w];
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < 24; ++k)
        {
            const uint32_t v = uint32_t(__shfl_xor(int(r[k]), off));
            r[k] = (k % 6) < 3 ? min(r[k], v) : max(r[k], v);
        }
    if (lane == 0u)
        for (int k = 0; k < 24; ++k) part[wave][k] = r[k];
    __syncthreads();
    if (threadIdx.x < 24)
    {
        const uint32_t k = threadIdx.x;
        uint32_t v = part[0][k];
        for (uint32_t w = 1; w < W; ++w) v = (k % 6) < 3 ? min(v, part[w][k]) : max(v, part[w][k]);
        const uint32_t c = l + k / 12, e = k % 12;
        uint32_t* dst = e < 6 ? &S.nbox[6 * size_t(c) + e] : &S.ncbox[6 * size_t(c) + (e - 6)];
        if ((e % 6) < 3) atomicMin(dst, v);
        else atomicMax(dst, v);
    }
}

// Roots: the first slot and slot count of each tree, the level control words and counters.
__global__ void k_init_roots(BuildState S, const uint32_t* rfirst, uint32_t R)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R)
    {
        S.nfirst[r] = rfirst[r];
        S.ncount[r] = rfirst[r + 1] - rfirst[r];
        S.nflags[r] = 0u;
    }
    if (r == 0)
    {
        // level 0 (the roots) is [lvl[1], counters[0]) for k_level_init
        S.lvl[0] = 0u;
        S.lvl[1] = 0u;
        S.lvl[2] = 0u;
        S.lvl[3] = 0u;
        S.lvl[4] = 0u;
        S.lvl[5] = 0u;
        S.lvl[9] = 0u;
        S.lvl[10] = 0u;
        S.lvl[11] = 0u;
        S.counters[0] = R;
        S.counters[1] = 0u;
        S.counters[3] = 0u;
    }
}


// ---------------------------------------------------------------------------
// subtree finisher: every node of at most FIN slots, one workgroup each
// ---------------------------------------------------------------------------
// The workgroup copies its node's triangle boxes into LDS once and builds the whole subtree
// there, breadth first, with the level loop's rules (bounds, leaf tests, forced positional
// splits, 9-bin SAH sweeps with the same operation order, stable partitions):
//   wave nodes  (more than TINY slots) one wave each per round: bounds by a wave reduction,
//               histograms by LDS atomics, the 24 (axis, split) costs one lane each and an
//               argmin over the lanes in (axis, split) order (the sweep's strict-< rule);
//   tiny nodes  (at most TINY slots) three steps per sub-level: a thread per node for bounds
//               and leaf tests, a thread per (node, axis) for the sweep from the slots
//               directly, a thread per node for the choice across axes and the partition.
// Child ids come from one block of 2 * count - 2 ids reserved per subtree (a binary tree over
// count slots has at most that many nodes below its root); ids left unused are counted in
// counters[3].  The subtree's slot order is written back to idx at the end.
constexpr uint32_t FIN_BLOCK = 512;
constexpr uint32_t FIN_WAVES = FIN_BLOCK / 64u;
constexpr uint32_t FIN_QCAP = FIN / (TINY + 1) + 1;  // wave nodes in a round (disjoint, > TINY slots)
constexpr uint32_t FIN_CCAP = FIN / 2;               // tiny SAH candidates in a step (>= 2 slots each)

// A node waiting in LDS: global id, and its first local slot | count << 16 | forced << 31.
struct FinNode
{
    uint32_t node, pack;
};

__device__ __forceinline__ FinNode fin_node(uint32_t node, uint32_t first, uint32_t count, bool forced)
{
    return FinNode{ node, first | (count << 16) | (forced ? 0x80000000u : 0u) };
}

// How a node is split: mode 0 none (a leaf), 1 positional, 2 SAH on axis d at bin si with nb
// slots below and the node's centroid range [mn, mx] on that axis.
struct FinSplit
{
    uint32_t mode, d, si, nb;
    float mn, mx;
};

struct FinShared
{
    float box[FIN][6];                      // triangle boxes, by local triangle
    uint32_t tri[FIN];                      // local triangle -> triangle
    uint16_t lid[FIN];                      // slot -> local triangle
    uint16_t tmp[FIN];                      // partition staging
    FinNode queue[2][FIN_QCAP];             // wave nodes: this round's and the next round's
    FinNode tiny[2][FIN];                   // tiny nodes: this sub-level's and the next one's
    FinNode cand[FIN_CCAP];                 // this sub-level's tiny SAH candidates
    float ccost[FIN_CCAP][3];               // centroid minimum per axis, then the best cost (+inf: none)
    float ccmx[FIN_CCAP][3];                // centroid maximum per axis
    uint16_t cinfo[FIN_CCAP][3];            // split | below count << 8 of that cost
    uint16_t order[FIN_CCAP];               // the candidates by slot count (lanes of a wave run alike loops)
    uint32_t csize[TINY + 1];               // candidates per slot count, then their offsets
    uint32_t hist[FIN_WAVES][NODE_BIN_WORDS];
    uint32_t qn[2], tn[2], cn, next, base, item;
};

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t fin_count(FinNode e) { return (e.pack >> 16) & 0x7FFFu; }
__device__ __forceinline__ uint32_t fin_first(FinNode e) { return e.pack & 0xFFFFu; }
__device__ __forceinline__ bool fin_forced(FinNode e) { return (e.pack >> 31) != 0u; }

__device__ __forceinline__ void fin_push(FinShared& L, uint32_t qnext, uint32_t tnext, FinNode e)
{
    if (fin_count(e) > TINY) L.queue[qnext][atomicAdd(&L.qn[qnext], 1u)] = e;
    else L.tiny[tnext][atomicAdd(&L.tn[tnext], 1u)] = e;
}

// Positional split of a forced node: the first half rounded to whole leaves.
__device__ __forceinline__ uint32_t forced_half(uint32_t cnt)
{
    uint32_t half = ((cnt + 1) / 2 + APG_MAX_LEAF - 1) / APG_MAX_LEAF * APG_MAX_LEAF;
    return half < cnt - 1 ? half : cnt - 1;
}


// Leaf test of the level loop (sweep_node_wave): forced nodes split positionally until they hold at
// most APG_MAX_LEAF triangles; others are SAH leaves at most `thr` triangles or with a zero
// centroid extent.  Returns true when the node is decided without a sweep.
__device__ __forceinline__ bool fin_decide(uint32_t cnt, bool forced, uint32_t thr, const float cmn[3],
                                           const float cmx[3], uint32_t& fl, FinSplit& sp)
{
    sp.mode = 0;
    if (forced)
    {
        fl = NF_FORCED | (cnt <= APG_MAX_LEAF ? NF_LEAF : NF_SPLIT);
        if (cnt > APG_MAX_LEAF) { sp.mode = 1; sp.nb = forced_half(cnt); }
        return true;
    }
    bool leaf = cnt <= thr;
    if (!leaf)
    {
        const float diag = max_s(cmx[0] - cmn[0], max_s(cmx[1] - cmn[1], cmx[2] - cmn[2]));
        if (diag == 0.0f) leaf = true;
    }
    if (!leaf) return false;
    fl = NF_SAH_LEAF | (cnt <= APG_MAX_LEAF ? NF_LEAF : (NF_FORCED | NF_SPLIT));
    if (cnt > APG_MAX_LEAF) { sp.mode = 1; sp.nb = forced_half(cnt); }
    return true;
}

// The choice after a sweep (as in sweep_node_wave): a leaf unless a split costs less than the
// leaf.
__device__ __forceinline__ void fin_choose(uint32_t cnt, bool found, float min_cost, float leaf_cost, uint32_t d,
                                           uint32_t si, uint32_t nb, const float cmn[3], const float cmx[3],
                                           uint32_t& fl, FinSplit& sp)
{
    // no plane below FLT_MAX (surface areas overflowed): a leaf, where the reference's
    // default split would send every primitive to one side and never terminate
    if (!found || min_cost >= leaf_cost)
    {
        fl = NF_SAH_LEAF | (cnt <= APG_MAX_LEAF ? NF_LEAF : (NF_FORCED | NF_SPLIT));
        sp.mode = 0;
        if (cnt > APG_MAX_LEAF) { sp.mode = 1; sp.nb = forced_half(cnt); }
        return;
    }
    fl = NF_SPLIT;
    sp.mode = 2;
    sp.d = d;
    sp.si = si;
    sp.nb = nb;
    sp.mn = d == 0 ? cmn[0] : (d == 1 ? cmn[1] : cmn[2]);
    sp.mx = d == 0 ? cmx[0] : (d == 1 ? cmx[1] : cmx[2]);
}

// Children of a split node: ids from the subtree's block, counts, first slots (global).
__device__ __forceinline__ uint32_t fin_children(BuildState& S, FinShared& L, uint32_t k, uint32_t gfirst,
                                                 uint32_t f, uint32_t cnt, uint32_t nb)
{
    const uint32_t l = L.base + atomicAdd(&L.next, 2u);
    S.nleft[k] = l;
    S.ncount[l] = nb;
    S.ncount[l + 1] = cnt - nb;
    S.nfirst[l] = gfirst + f;
    S.nfirst[l + 1] = gfirst + f + nb;
    return l;
}

__device__ __forceinline__ void fin_unpack_bounds(const uint32_t r[12], float nlo[3], float nhi[3], float cmn[3],
                                                  float cmx[3])
{
    for (int a = 0; a < 3; ++a)
    {
        nlo[a] = float_of(r[a]); nhi[a] = float_of(r[3 + a]);
        cmn[a] = float_of(r[6 + a]); cmx[a] = float_of(r[9 + a]);
    }
}

// ---- wave-level steps of a node of more than TINY slots ----

__device__ __forceinline__ void ord_bounds_init(uint32_t r[12])
{
    const uint32_t lo_ord = ord_of(FLT_MAX), hi_ord = ord_of(-FLT_MAX);
    for (int a = 0; a < 3; ++a)
    {
        r[a] = lo_ord; r[3 + a] = hi_ord; r[6 + a] = lo_ord; r[9 + a] = hi_ord;
    }
}

// Bounds and centroid bounds of slots [f + j0, f + j1), reduced across the wave (every lane
// ends with all 12 ordered values).
__device__ void fin_bounds_wave(const FinShared& L, uint32_t f, uint32_t j0, uint32_t j1, uint32_t r[12])
{
    const uint32_t lane = threadIdx.x & 63u;
    ord_bounds_init(r);
    for (uint32_t j = j0 + lane; j < j1; j += 64u)
    {
        const float* b = L.box[L.lid[f + j]];
        for (int a = 0; a < 3; ++a)
        {
            const float lo = b[a], hi = b[3 + a], c = (lo + hi) * 0.5f;
            r[a] = min(r[a], ord_of(lo));
            r[3 + a] = max(r[3 + a], ord_of(hi));
            r[6 + a] = min(r[6 + a], ord_of(c));
            r[9 + a] = max(r[9 + a], ord_of(c));
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; ++a)
        {
            r[a] = min(r[a], uint32_t(__shfl_xor(int(r[a]), off)));
            r[3 + a] = max(r[3 + a], uint32_t(__shfl_xor(int(r[3 + a]), off)));
            r[6 + a] = min(r[6 + a], uint32_t(__shfl_xor(int(r[6 + a]), off)));
            r[9 + a] = max(r[9 + a], uint32_t(__shfl_xor(int(r[9 + a]), off)));
        }
}

// The 9-bin histograms of the three axes of slots [f + j0, f + j1) in the wave's own LDS block.
__device__ void fin_bin_wave(const FinShared& L, uint32_t f, uint32_t j0, uint32_t j1, uint32_t* hist,
                             const float cmn[3], const float cmx[3])
{
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t j = lane; j < NODE_BIN_WORDS; j += 64u)
    {
        const uint32_t v = j % BIN_VALS;
        hist[j] = v < 3 ? ord_of(FLT_MAX) : (v < 6 ? ord_of(-FLT_MAX) : 0u);
    }
    wave_sync_lds();
    for (uint32_t j = j0 + lane; j < j1; j += 64u)
    {
        const float* b = L.box[L.lid[f + j]];
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { lo[a] = b[a]; hi[a] = b[3 + a]; }
        for (int d = 0; d < 3; ++d)
        {
            if (cmn[d] == cmx[d]) continue;
            const float c = (lo[d] + hi[d]) * 0.5f;
            uint32_t* h = hist + (d * NBINS + bin_index(c, cmn[d], cmx[d])) * BIN_VALS;
            for (int a = 0; a < 3; ++a)
            {
                atomicMin(&h[a], ord_of(lo[a]));
                atomicMax(&h[3 + a], ord_of(hi[a]));
            }
            atomicAdd(&h[6], 1u);
        }
    }
    wave_sync_lds();
}

// The sweep (sweep_bins_wave over the wave's histogram), then the leaf test.
__device__ void fin_sweep_wave(const uint32_t* hist, uint32_t cnt, const float nlo[3], const float nhi[3],
                               const float cmn[3], const float cmx[3], uint32_t& fl, FinSplit& sp)
{
    uint32_t axes = 0;
    for (int a = 0; a < 3; ++a)
        if (cmn[a] != cmx[a]) axes |= 1u << a;
    const SweepBest sb = sweep_bins_wave(hist, axes);
    const float leaf_cost = float(cnt) * box_area(nlo, nhi);
    fin_choose(cnt, sb.key != 63u, sb.cost, leaf_cost, sb.key / NSPLITS, sb.key % NSPLITS, sb.below, cmn, cmx, fl,
               sp);
    // the next use of hist must follow every lane's reads above
    wave_sync_lds();
}

__device__ __forceinline__ bool fin_goes_left(const FinShared& L, uint16_t v, const FinSplit& sp)
{
    const float* b = L.box[v];
    return bin_index((b[sp.d] + b[3 + sp.d]) * 0.5f, sp.mn, sp.mx) <= sp.si;
}

// Stable partition of slots [f + j0, f + j1) of a node into tmp: `left` left slots of the node
// precede j0; the node has sp.nb left slots in all.
__device__ void fin_scatter_wave(FinShared& L, uint32_t f, uint32_t j0, uint32_t j1, const FinSplit& sp,
                                 uint32_t left)
{
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t run = left;
    for (uint32_t b = j0; b < j1; b += 64u)
    {
        const uint32_t j = b + lane;
        bool bl = false;
        uint16_t v = 0;
        if (j < j1)
        {
            v = L.lid[f + j];
            bl = fin_goes_left(L, v, sp);
        }
        const uint64_t m = __ballot(bl);
        const uint32_t before = run + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
        if (j < j1) L.tmp[f + (bl ? before : sp.nb + (j - before))] = v;
        run += uint32_t(__popcll(m));
    }
}

// One node of more than TINY slots, by one wave.
__device__ void fin_wave_node(BuildState& S, FinShared& L, uint32_t thr, uint32_t gfirst, FinNode e,
                              uint32_t qnext, uint32_t* hist)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k = e.node, f = fin_first(e), cnt = fin_count(e);
    uint32_t r[12];
    fin_bounds_wave(L, f, 0, cnt, r);
    if (lane == 0u)
        for (int j = 0; j < 6; ++j) S.nbox[6 * size_t(k) + j] = r[j];
    float nlo[3], nhi[3], cmn[3], cmx[3];
    fin_unpack_bounds(r, nlo, nhi, cmn, cmx);
    uint32_t fl = 0;
    FinSplit sp;
    if (!fin_decide(cnt, fin_forced(e), thr, cmn, cmx, fl, sp))
    {
        fin_bin_wave(L, f, 0, cnt, hist, cmn, cmx);
        fin_sweep_wave(hist, cnt, nlo, nhi, cmn, cmx, fl, sp);
    }
    if (lane == 0u) S.nflags[k] = fl;
    if (sp.mode == 0) return;
    const uint32_t nb = sp.nb;
    if (sp.mode == 2)
    {
        fin_scatter_wave(L, f, 0, cnt, sp, 0u);
        wave_sync_lds();
        for (uint32_t j = lane; j < cnt; j += 64u) L.lid[f + j] = L.tmp[f + j];
        wave_sync_lds();
    }
    if (lane == 0u)
    {
        const uint32_t l = fin_children(S, L, k, gfirst, f, cnt, nb);
        fin_push(L, qnext, 0, fin_node(l, f, nb, sp.mode == 1));
        fin_push(L, qnext, 0, fin_node(l + 1, f + nb, cnt - nb, sp.mode == 1));
    }
}

// Exchange a word with the other lane of the pair (lanes 2k and 2k + 1): DPP quad_perm [1,0,3,2].
__device__ __forceinline__ uint32_t pair_swap(uint32_t v)
{
    return uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float pair_swap(float v) { return __uint_as_float(pair_swap(__float_as_uint(v))); }

// Bounds and centroid bounds (ordered uints) of a tiny node's slots [f, f + cnt) by a lane pair:
// the even lane reduces the first half, the odd lane the second, and both end with the node's.
__device__ void fin_pair_bounds(const FinShared& L, uint32_t f, uint32_t cnt, uint32_t half, uint32_t r[12])
{
    const uint32_t split = (cnt + 1u) / 2u;
    const uint32_t j0 = half ? split : 0u, j1 = half ? cnt : split;
    ord_bounds_init(r);
    for (uint32_t j = j0; j < j1; ++j)
    {
        const float* b = L.box[L.lid[f + j]];
        for (int a = 0; a < 3; ++a)
        {
            const float lo = b[a], hi = b[3 + a], cc = (lo + hi) * 0.5f;
            r[a] = min(r[a], ord_of(lo));
            r[3 + a] = max(r[3 + a], ord_of(hi));
            r[6 + a] = min(r[6 + a], ord_of(cc));
            r[9 + a] = max(r[9 + a], ord_of(cc));
        }
    }
    for (int i = 0; i < 12; ++i)
    {
        const uint32_t o = pair_swap(r[i]);
        r[i] = (i % 6) < 3 ? min(r[i], o) : max(r[i], o);
    }
}

// Sub-level step 2: the best split of one axis of a tiny candidate, from the slots directly
// (each split's below and above boxes and counts; the sweep's strict-< rule within the axis),
// by the lane pair 2k, 2k + 1: each lane bins half of the slots, the pair merges the bins, the
// even lane prices the 8 below boxes and the odd lane the 8 above boxes, and with the halves
// exchanged both lanes hold every cost (0 + below term) + above term.  Both lanes of a pair
// are active together (the callers hand out pairs).
__device__ void fin_tiny_axis(FinShared& L, uint32_t c, uint32_t d, uint32_t half)
{
    const FinNode e = L.cand[c];
    const uint32_t cnt = fin_count(e), f = fin_first(e);
    // the centroid range on axis d, left by step 1 (ccost holds the minimum until this step)
    const float mn = L.ccost[c][d], mx = L.ccmx[c][d];
    float min_cost = FLT_MAX;
    uint32_t best_idx = 0, best_below = 0;
    bool found = false;
    if (mn != mx)
    {
        // the slots' 9 bin boxes and counts (min / max are exact: the boxes equal those the
        // slots give directly, up to the sign of zero, which no area comparison sees)
        float lo_[NBINS][3], hi_[NBINS][3];
        uint32_t nbin[NBINS];
#pragma unroll
        for (uint32_t q = 0; q < NBINS; ++q)
        {
            nbin[q] = 0;
            for (int a = 0; a < 3; ++a) { lo_[q][a] = FLT_MAX; hi_[q][a] = -FLT_MAX; }
        }
        const uint32_t split = (cnt + 1u) / 2u;
        const uint32_t j0 = half ? split : 0u, j1 = half ? cnt : split;
        for (uint32_t j = j0; j < j1; ++j)
        {
            const float* b = L.box[L.lid[f + j]];
            float lo[3], hi[3];
            for (int a = 0; a < 3; ++a) { lo[a] = b[a]; hi[a] = b[3 + a]; }
            const uint32_t bi = bin_index((lo[d] + hi[d]) * 0.5f, mn, mx);
            // compare-and-select updates (two instructions a value; fminf / fmaxf would also
            // quieten their operands every time): the boxes are finite, and which zero a tie
            // between -0 and +0 keeps changes no area
#pragma unroll
            for (uint32_t q = 0; q < NBINS; ++q)
            {
                const bool in = bi == q;
                nbin[q] += in ? 1u : 0u;
                for (int a = 0; a < 3; ++a)
                {
                    lo_[q][a] = (in && lo[a] < lo_[q][a]) ? lo[a] : lo_[q][a];
                    hi_[q][a] = (in && hi_[q][a] < hi[a]) ? hi[a] : hi_[q][a];
                }
            }
        }
        // the pair's bins
#pragma unroll
        for (uint32_t q = 0; q < NBINS; ++q)
        {
            nbin[q] += pair_swap(nbin[q]);
            for (int a = 0; a < 3; ++a)
            {
                lo_[q][a] = min_s(lo_[q][a], pair_swap(lo_[q][a]));
                hi_[q][a] = max_s(hi_[q][a], pair_swap(hi_[q][a]));
            }
        }
        uint32_t bn[NSPLITS];
        {
            uint32_t pn = 0;
#pragma unroll
            for (uint32_t q = 0; q < NSPLITS; ++q) { pn += nbin[q]; bn[q] = pn; }
        }
        // this lane's terms: below (even lane) or above (odd lane)
        float part[NSPLITS];
        float plo[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, phi[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
        if (!half)
        {
#pragma unroll
            for (uint32_t q = 0; q < NSPLITS; ++q)
            {
                for (int a = 0; a < 3; ++a) { plo[a] = min_s(plo[a], lo_[q][a]); phi[a] = max_s(phi[a], hi_[q][a]); }
                part[q] = float(bn[q]) * box_area(plo, phi);
            }
        }
        else
        {
#pragma unroll
            for (int q = NSPLITS - 1; q >= 0; --q)
            {
                for (int a = 0; a < 3; ++a) { plo[a] = min_s(plo[a], lo_[q + 1][a]); phi[a] = max_s(phi[a], hi_[q + 1][a]); }
                part[q] = float(cnt - bn[q]) * box_area(plo, phi);
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < NSPLITS; ++q)
        {
            const float other = pair_swap(part[q]);
            const float below = half ? other : part[q], above = half ? part[q] : other;
            float cost = 0.0f;
            cost += below;
            cost += above;
            if (cost < min_cost)
            {
                min_cost = cost;
                best_idx = q;
                best_below = bn[q];
                found = true;
            }
        }
    }
    if (!half)
    {
        L.ccost[c][d] = found ? min_cost : __builtin_inff();
        L.cinfo[c][d] = uint16_t(best_idx | (best_below << 8));
    }
}

// Sub-level step 3: the choice across axes (in axis order, strict <), the leaf test and the
// stable partition of one tiny candidate, by the lane pair 2k, 2k + 1: each lane reduces the
// bounds of half of the slots (the pair merges them), both make the same choice, and each lane
// places its half's slots (the odd lane after the even lane's left and right ones).
__device__ void fin_tiny_choose(BuildState& S, FinShared& L, uint32_t gfirst, uint32_t c, uint32_t tnext,
                                uint32_t half)
{
    const FinNode e = L.cand[c];
    const uint32_t k = e.node, f = fin_first(e), cnt = fin_count(e);
    const uint32_t split = (cnt + 1u) / 2u;
    const uint32_t j0 = half ? split : 0u, j1 = half ? cnt : split;
    uint32_t r[12];
    fin_pair_bounds(L, f, cnt, half, r);
    float nlo[3], nhi[3], cmn[3], cmx[3];
    fin_unpack_bounds(r, nlo, nhi, cmn, cmx);
    float min_cost = FLT_MAX;
    uint32_t best_dim = 0;
    bool found = false;
    for (uint32_t d = 0; d < 3; ++d)
        if (L.ccost[c][d] < min_cost)
        {
            min_cost = L.ccost[c][d];
            best_dim = d;
            found = true;
        }
    const uint32_t info = L.cinfo[c][best_dim];
    uint32_t fl = 0;
    FinSplit sp;
    fin_choose(cnt, found, min_cost, float(cnt) * box_area(nlo, nhi), best_dim, info & 0xFFu, info >> 8, cmn, cmx,
               fl, sp);
    if (!half) S.nflags[k] = fl;
    if (sp.mode == 0) return;
    const uint32_t nb = sp.nb;
    if (sp.mode == 2)
    {
        uint32_t flags = 0, nl = 0;
        for (uint32_t j = j0; j < j1; ++j)
        {
            const float* b = L.box[L.lid[f + j]];
            const bool left = bin_index((b[sp.d] + b[3 + sp.d]) * 0.5f, sp.mn, sp.mx) <= sp.si;
            flags |= (left ? 1u : 0u) << (j - j0);
            nl += left ? 1u : 0u;
        }
        const uint32_t nl_even = pair_swap(nl);  // the odd lane's view: the even lane's left count
        uint32_t wl = half ? nl_even : 0u, wr = half ? split - nl_even : 0u;
        for (uint32_t j = j0; j < j1; ++j)
        {
            const uint16_t v = L.lid[f + j];
            if ((flags >> (j - j0)) & 1u) L.tmp[f + wl++] = v;
            else L.tmp[f + nb + wr++] = v;
        }
        wave_sync_lds();
        for (uint32_t j = j0; j < j1; ++j) L.lid[f + j] = L.tmp[f + j];
    }
    if (!half)
    {
        const uint32_t l = fin_children(S, L, k, gfirst, f, cnt, nb);
        fin_push(L, 0, tnext, fin_node(l, f, nb, sp.mode == 1));
        fin_push(L, 0, tnext, fin_node(l + 1, f + nb, cnt - nb, sp.mode == 1));
    }
}

// Grid-stride over the nodes the level loop handed over, taken in turn from a device cursor.
// Two workgroups of 512 threads per CU (the LDS holds two): the register budget is set for
// four waves per SIMD.
__global__ __launch_bounds__(FIN_BLOCK) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_finish(BuildState S,
                                                                                                  uint32_t thr)
{
    __shared__ FinShared L;
    const uint32_t tid = threadIdx.x, wave = tid / 64u;
    const uint32_t total = S.lvl[4];
    for (;;)
    {
        if (tid == 0) L.item = atomicAdd(&S.lvl[5], 1u);
        __syncthreads();
        const uint32_t item = L.item;
        if (item >= total) break;
        const uint32_t root = S.fin[item] & 0x7FFFFFFFu, buf = S.fin[item] >> 31;
        const uint32_t gfirst = S.nfirst[root], cnt = S.ncount[root];
        const uint32_t* idx_in = slots_of(S, buf);
        const float* box_in = boxes_of(S, buf);
        const bool forced = (S.nflags[root] & NF_FORCED) != 0u;
        for (uint32_t j = tid; j < cnt; j += FIN_BLOCK)
        {
            L.tri[j] = idx_in[gfirst + j];
            L.lid[j] = uint16_t(j);
            const float* b = box_in + 6 * size_t(gfirst + j);
            for (int a = 0; a < 6; ++a) L.box[j][a] = b[a];
        }
        if (tid == 0)
        {
            L.base = cnt > 1 ? atomicAdd(&S.counters[0], 2 * cnt - 2) : 0u;
            L.next = 0;
            L.qn[0] = L.qn[1] = 0;
            L.tn[0] = L.tn[1] = 0;
            L.cn = 0;
            fin_push(L, 0, 0, fin_node(root, 0, cnt, forced));
        }
        if (tid <= TINY) L.csize[tid] = 0;
        __syncthreads();
        // wave nodes: rounds of one wave per node (their tiny children go to tiny[0])
        uint32_t cur = 0;
        for (;;)
        {
            const uint32_t qn = L.qn[cur];
            if (qn == 0) break;
            for (uint32_t q = wave; q < qn; q += FIN_WAVES)
                fin_wave_node(S, L, thr, gfirst, L.queue[cur][q], cur ^ 1u, L.hist[wave]);
            __syncthreads();
            if (tid == 0) L.qn[cur] = 0;
            cur ^= 1u;
            __syncthreads();
        }
        // tiny nodes: sub-levels of three steps
        uint32_t tc = 0;
        for (;;)
        {
            const uint32_t tn = L.tn[tc];
            if (tn == 0) break;
            // a lane pair per tiny node for its bounds; the even lane does the rest
            for (uint32_t u = tid; u < 2 * tn; u += FIN_BLOCK)
            {
                const FinNode e = L.tiny[tc][u >> 1];
                const uint32_t k = e.node, f = fin_first(e), c = fin_count(e);
                uint32_t r[12];
                fin_pair_bounds(L, f, c, u & 1u, r);
                if (u & 1u) continue;
                for (int j = 0; j < 6; ++j) S.nbox[6 * size_t(k) + j] = r[j];
                float nlo[3], nhi[3], cmn[3], cmx[3];
                fin_unpack_bounds(r, nlo, nhi, cmn, cmx);
                uint32_t fl = 0;
                FinSplit sp;
                if (!fin_decide(c, fin_forced(e), thr, cmn, cmx, fl, sp))
                {
                    const uint32_t j = atomicAdd(&L.cn, 1u);
                    L.cand[j] = e;
                    for (int a = 0; a < 3; ++a) { L.ccost[j][a] = cmn[a]; L.ccmx[j][a] = cmx[a]; }
                    atomicAdd(&L.csize[c], 1u);
                    continue;
                }
                S.nflags[k] = fl;
                if (sp.mode == 1)
                {
                    const uint32_t l = fin_children(S, L, k, gfirst, f, c, sp.nb);
                    fin_push(L, 0, tc ^ 1u, fin_node(l, f, sp.nb, true));
                    fin_push(L, 0, tc ^ 1u, fin_node(l + 1, f + sp.nb, c - sp.nb, true));
                }
            }
            __syncthreads();
            const uint32_t cn = L.cn;
            if (tid == 0)
            {
                uint32_t run = 0;
                for (uint32_t k = 0; k <= TINY; ++k)
                {
                    const uint32_t h = L.csize[k];
                    L.csize[k] = run;
                    run += h;
                }
            }
            __syncthreads();
            for (uint32_t c = tid; c < cn; c += FIN_BLOCK)
                L.order[atomicAdd(&L.csize[fin_count(L.cand[c])], 1u)] = uint16_t(c);
            __syncthreads();
            // lane pairs: both lanes of a pair fall in the same pass (FIN_BLOCK and 6 cn are even)
            for (uint32_t u = tid; u < 6 * cn; u += FIN_BLOCK)
                fin_tiny_axis(L, L.order[(u >> 1) / 3], (u >> 1) % 3, u & 1u);
            __syncthreads();
            for (uint32_t u = tid; u < 2 * cn; u += FIN_BLOCK)
                fin_tiny_choose(S, L, gfirst, L.order[u >> 1], tc ^ 1u, u & 1u);
            __syncthreads();
            if (tid == 0)
            {
                L.tn[tc] = 0;
                L.cn = 0;
            }
            if (tid <= TINY) L.csize[tid] = 0;
            tc ^= 1u;
            __syncthreads();
        }
        for (uint32_t j = tid; j < cnt; j += FIN_BLOCK) S.idx[gfirst + j] = L.tri[L.lid[j]];
        if (tid == 0 && cnt > 1) atomicAdd(&S.counters[3], 2 * cnt - 2 - L.next);
        __syncthreads();
    }
}

// After k_finish: the node ids allocated and the binary nodes built.
__global__ void k_finish_done(BuildState S)
{
    S.lvl[6] = S.counters[0];
    S.lvl[7] = S.counters[0] - S.counters[3];
}

// A graph-replayed build after its fixed number of levels: when the last level still allocated
// children the tree is unfinished, so k_finish gets no subtrees and the wide collapse no level
// (the scene keeps the previous tree); lvl[12] = 1 then.
__global__ void k_build_gate(BuildState S)
{
    const bool done = S.lvl[1] == S.counters[0];
    S.lvl[12] = done ? 0u : 1u;
    if (!done) S.lvl[4] = 0u;
}

// ---------------------------------------------------------------------------
// wide collapse
// ---------------------------------------------------------------------------
// Breadth first over wide levels, like the binary levels: the level's wide nodes
// [wl[0], wl[1]) are read from device memory.  Exclusive scans of the level's per-node counts
// (over that range only, on a fixed grid) give its child and triangle offsets in wide-node
// order (deterministic) without the host knowing the level's width; the leaf triangles are
// copied in one pass after the last level (k_tri_gather).
struct WideState
{
    const uint32_t* nbox;   // ordered uints
    const uint32_t* nfirst;
    const uint32_t* ncount;
    const uint32_t* nleft;
    const uint32_t* nflags;
    const uint32_t* idx;    // tree order -> triangle
    const float* tri9_src;
    uint32_t* wbin;         // wide node -> binary node
    uint32_t* wchild;       // 8 per wide node (binary ids, slot order, RETIRED = empty)
    uint32_t* n_inner;      // per wide node: inner children (0 outside the level)
    uint32_t* n_tris;       // per wide node: triangles of its leaf slots (0 outside the level)
    uint32_t* s_inner;      // exclusive scans over the capacity (wcap + 1)
    uint32_t* s_tris;
    uint32_t* wpart;        // per workgroup of the level scan: sums, then offsets (4 * WSCAN_BLOCKS)
    uint32_t* perm;         // triangle position -> slot (k_tri_gather then copies)
    ApgWideNode* nodes;
    float* tri9;
    uint32_t* tri_ids;
    const float* slack;     // device: the quantisation slack
    uint32_t* wl;           // [0] level begin, [1] level end, [2] triangles placed, [3] levels, [4] error,
                            // [5], [6] the level's inner children and leaf triangles
    uint32_t wcap;
};

__device__ __forceinline__ bool is_leaf_node(const WideState& W, uint32_t b)
{
    return (W.nflags[b] & NF_LEAF) != 0;
}

__device__ __forceinline__ void node_box(const WideState& W, uint32_t b, float lo[3], float hi[3])
{
    for (int a = 0; a < 3; ++a)
    {
        lo[a] = float_of(W.nbox[6 * size_t(b) + a]);
        hi[a] = float_of(W.nbox[6 * size_t(b) + 3 + a]);
    }
}

__device__ void wide_gather(const WideState& W, uint32_t w)
{
    const uint32_t root = W.wbin[w];
    // the children so far, with each one's leaf flag, surface area and box centre read once
    uint32_t ch[APG_WIDE];
    bool leaf[APG_WIDE];
    float sa[APG_WIDE], ctr[APG_WIDE][3];
    uint32_t nch = 0;
    auto take = [&](uint32_t i, uint32_t b) {
        float lo[3], hi[3];
        node_box(W, b, lo, hi);
        ch[i] = b;
        leaf[i] = is_leaf_node(W, b);
        sa[i] = box_area(lo, hi);
        for (int a = 0; a < 3; ++a) ctr[i][a] = 0.5f * (lo[a] + hi[a]);
    };
    if (is_leaf_node(W, root)) take(nch++, root);
    else
    {
        const uint32_t l = W.nleft[root];
        take(nch++, l);
        take(nch++, l + 1);
        // the inner child of largest surface area (the first such) opens into its two children
        while (nch < APG_WIDE)
        {
            int best = -1;
            float best_sa = 0.0f;
            for (uint32_t i = 0; i < nch; ++i)
            {
                if (leaf[i]) continue;
                if (best < 0 || sa[i] > best_sa) { best = int(i); best_sa = sa[i]; }
            }
            if (best < 0) break;
            const uint32_t c = W.nleft[ch[best]];
            take(nch++, c);
            take(uint32_t(best), c + 1);
        }
    }
    // octant slot assignment (greedy over the 8 x 8 cost table)
    float plo[3], phi[3];
    node_box(W, root, plo, phi);
    float pc[3];
    for (int a = 0; a < 3; ++a) pc[a] = 0.5f * (plo[a] + phi[a]);
    float cd[APG_WIDE][3];
    for (uint32_t c = 0; c < nch; ++c)
        for (int a = 0; a < 3; ++a) cd[c][a] = ctr[c][a] - pc[a];
    uint32_t slot_child[APG_WIDE];
    bool slot_leaf[APG_WIDE];
    for (uint32_t s = 0; s < APG_WIDE; ++s) { slot_child[s] = RETIRED; slot_leaf[s] = false; }
    uint32_t done = 0, used = 0;
    for (uint32_t r = 0; r < nch; ++r)
    {
        float bc = FLT_MAX;
        int bcc = -1, bs = -1;
        for (uint32_t c = 0; c < nch; ++c)
        {
            if ((done >> c) & 1u) continue;
            for (uint32_t s = 0; s < APG_WIDE; ++s)
            {
                if ((used >> s) & 1u) continue;
                const float cost = ((s & 1) ? -cd[c][0] : cd[c][0]) + ((s & 2) ? -cd[c][1] : cd[c][1]) +
                                   ((s & 4) ? -cd[c][2] : cd[c][2]);
                if (bcc < 0 || cost < bc) { bc = cost; bcc = int(c); bs = int(s); }
            }
        }
        slot_child[bs] = ch[bcc];
        slot_leaf[bs] = leaf[bcc];
        done |= 1u << bcc;
        used |= 1u << bs;
    }
    uint32_t ni = 0, nt = 0;
    for (uint32_t s = 0; s < APG_WIDE; ++s)
    {
        const uint32_t c = slot_child[s];
        W.wchild[8 * size_t(w) + s] = c;
        if (c == RETIRED) continue;
        if (slot_leaf[s]) nt += W.ncount[c];
        else ni++;
    }
    W.n_inner[w] = ni;
    W.n_tris[w] = nt;
}

__global__ __launch_bounds__(128) void k_wide_gather(WideState W)
{
    const uint32_t begin = W.wl[0], end = W.wl[1];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t w = begin + blockIdx.x * blockDim.x + threadIdx.x; w < end; w += stride) wide_gather(W, w);
}

__device__ uint32_t quantise_axis(float lo_f, float hi_f, float slack, const float* clo,
                                  const float* chi, uint32_t valid, float& origin,
                                  uint32_t* qlo, uint32_t* qhi)
{
    const float lo = lo_f - slack, hi = hi_f + slack;
    const float m = fmaxf(fabsf(lo), fabsf(hi));
    int k = 0;
    (void)frexpf(m > 0.0f ? m : 1.0f, &k);  // m < 2^k
    int e = k - 23;
    const float ext = (hi - lo) * (1.0f + 0x1p-20f);
    while (ldexpf(62.0f, e) < ext) ++e;
    if (e < -126) e = -126;
    for (;;)
    {
        const float step = ldexpf(1.0f, e);
        const float inv = ldexpf(1.0f, -e);
        const float o = floorf(lo * inv) * step;
        if (o + 63.0f * step >= hi)
        {
            origin = o;
            for (uint32_t s = 0; s < APG_WIDE; ++s)
            {
                if (!((valid >> s) & 1u)) { qlo[s] = APG_QMAX; qhi[s] = 0; continue; }
                const float a = floorf(((clo[s] - slack) - o) * inv);
                const float b = ceilf(((chi[s] + slack) - o) * inv);
                qlo[s] = uint32_t(fminf(fmaxf(a, 0.0f), float(APG_QMAX)));
                qhi[s] = uint32_t(fminf(fmaxf(b, 0.0f), float(APG_QMAX)));
            }
            return uint32_t(e + 127);
        }
        ++e;
    }
}


__device__ void wide_emit(const WideState& W, uint32_t w, uint32_t level_end, uint32_t tri_begin, float slack,
                          uint32_t seg_inner, uint32_t seg_tris)
{
    ApgWideNode nd;
    const uint32_t child_base = level_end + seg_inner + W.s_inner[w];
    const uint32_t tri_base = tri_begin + seg_tris + W.s_tris[w];
    uint32_t imask = 0, valid = 0, cmask = 0, inner_rank = 0, toff = 0;
    float clo[3][APG_WIDE], chi[3][APG_WIDE];
    float nlo[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, nhi[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
    for (uint32_t s = 0; s < APG_WIDE; ++s)
    {
        const uint32_t c = W.wchild[8 * size_t(w) + s];
        for (int a = 0; a < 3; ++a) { clo[a][s] = 0.0f; chi[a][s] = 0.0f; }
        if (c == RETIRED) continue;
        valid |= 1u << s;
        float lo[3], hi[3];
        node_box(W, c, lo, hi);
        for (int a = 0; a < 3; ++a)
        {
            clo[a][s] = lo[a];
            chi[a][s] = hi[a];
            nlo[a] = fminf(nlo[a], lo[a]);
            nhi[a] = fmaxf(nhi[a], hi[a]);
        }
        if (is_leaf_node(W, c))
        {
            const uint32_t cnt = W.ncount[c], first = W.nfirst[c];
            cmask |= apg_leaf_nibble(s, cnt);
            for (uint32_t k = 0; k < cnt; ++k) W.perm[tri_base + toff++] = first + k;
        }
        else
        {
            imask |= 1u << s;
            const uint32_t cw = child_base + inner_rank;
            if (cw < W.wcap) W.wbin[cw] = c;
            else W.wl[4] = 1u;  // capacity exceeded: reported, nothing written
            inner_rank++;
        }
    }
    uint32_t eb[3], qlo[3][APG_WIDE], qhi[3][APG_WIDE];
    for (int a = 0; a < 3; ++a)
        eb[a] = quantise_axis(nlo[a], nhi[a], slack, clo[a], chi[a], valid, nd.origin[a], qlo[a], qhi[a]);
    nd.exp_imask = eb[0] | (eb[1] << 8) | (eb[2] << 16) | (imask << 24);
    nd.child_base = child_base;
    nd.tri_base = tri_base;
    nd.cmask = cmask;
    for (int j = 0; j < 9; ++j) nd.q[j] = 0;
    for (uint32_t s = 0; s < APG_WIDE; ++s)
        for (uint32_t a = 0; a < 3; ++a)
        {
            apg_put_plane(nd.q, s, a, qlo[a][s]);
            apg_put_plane(nd.q, s, 3 + a, qhi[a][s]);
        }
    W.nodes[w] = nd;
}

// ---------------------------------------------------------------------------
// exclusive scans of one wide level's counts over [wl[0], wl[1]) only: a grid fixed for the
// build, each workgroup a contiguous segment of the level; the segment sums are scanned by
// one workgroup and k_wide_emit adds its segment's offset.
constexpr uint32_t WSCAN_BLOCKS = 256;

__device__ __forceinline__ void wscan_segment(const WideState& W, uint32_t blk, uint32_t& s0, uint32_t& s1,
                                              uint32_t& seg)
{
    const uint32_t b = W.wl[0], e = W.wl[1];
    seg = (e - b + WSCAN_BLOCKS - 1) / WSCAN_BLOCKS;
    s0 = b + blk * seg;
    s1 = s0 + seg < e ? s0 + seg : e;
    if (s0 > s1) s0 = s1;
}

__global__ __launch_bounds__(BLOCK) void k_wscan_local(WideState W)
{
    __shared__ uint32_t sh[2][BLOCK];
    uint32_t s0, s1, seg;
    wscan_segment(W, blockIdx.x, s0, s1, seg);
    uint32_t carry_i = 0, carry_t = 0;
    for (uint32_t base = s0; base < s1; base += CHUNK)
    {
        const uint32_t j0 = base + threadIdx.x * ITEMS;
        uint32_t vi[ITEMS], vt[ITEMS], ti = 0, tt = 0;
        for (uint32_t k = 0; k < ITEMS; ++k)
        {
            const bool in = j0 + k < s1;
            vi[k] = in ? W.n_inner[j0 + k] : 0u;
            vt[k] = in ? W.n_tris[j0 + k] : 0u;
            ti += vi[k];
            tt += vt[k];
        }
        sh[0][threadIdx.x] = ti;
        sh[1][threadIdx.x] = tt;
        __syncthreads();
        for (uint32_t off = 1; off < BLOCK; off <<= 1)
        {
            const uint32_t xi = threadIdx.x >= off ? sh[0][threadIdx.x - off] : 0u;
            const uint32_t xt = threadIdx.x >= off ? sh[1][threadIdx.x - off] : 0u;
            __syncthreads();
            sh[0][threadIdx.x] += xi;
            sh[1][threadIdx.x] += xt;
            __syncthreads();
        }
        uint32_t ri = carry_i + sh[0][threadIdx.x] - ti, rt = carry_t + sh[1][threadIdx.x] - tt;
        for (uint32_t k = 0; k < ITEMS; ++k)
        {
            if (j0 + k < s1)
            {
                W.s_inner[j0 + k] = ri;
                W.s_tris[j0 + k] = rt;
            }
            ri += vi[k];
            rt += vt[k];
        }
        carry_i += sh[0][BLOCK - 1];
        carry_t += sh[1][BLOCK - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        W.wpart[blockIdx.x] = carry_i;
        W.wpart[WSCAN_BLOCKS + blockIdx.x] = carry_t;
    }
}

// One workgroup of WSCAN_BLOCKS threads: the segments' offsets and the level's totals.
__global__ __launch_bounds__(WSCAN_BLOCKS) void k_wscan_top(WideState W)
{
    __shared__ uint32_t sh[2][WSCAN_BLOCKS];
    const uint32_t t = threadIdx.x;
    const uint32_t ti = W.wpart[t], tt = W.wpart[WSCAN_BLOCKS + t];
    sh[0][t] = ti;
    sh[1][t] = tt;
    __syncthreads();
    for (uint32_t off = 1; off < WSCAN_BLOCKS; off <<= 1)
    {
        const uint32_t xi = t >= off ? sh[0][t - off] : 0u;
        const uint32_t xt = t >= off ? sh[1][t - off] : 0u;
        __syncthreads();
        sh[0][t] += xi;
        sh[1][t] += xt;
        __syncthreads();
    }
    W.wpart[2 * WSCAN_BLOCKS + t] = sh[0][t] - ti;
    W.wpart[3 * WSCAN_BLOCKS + t] = sh[1][t] - tt;
    if (t == WSCAN_BLOCKS - 1)
    {
        W.wl[5] = sh[0][t];
        W.wl[6] = sh[1][t];
    }
}

__global__ __launch_bounds__(128) void k_wide_emit(WideState W)
{
    const uint32_t begin = W.wl[0], end = W.wl[1], tri_begin = W.wl[2];
    const uint32_t seg = (end - begin + WSCAN_BLOCKS - 1) / WSCAN_BLOCKS;
    const float slack = *W.slack;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t w = begin + blockIdx.x * blockDim.x + threadIdx.x; w < end; w += stride)
    {
        const uint32_t blk = (w - begin) / seg;
        wide_emit(W, w, end, tri_begin, slack, W.wpart[2 * WSCAN_BLOCKS + blk], W.wpart[3 * WSCAN_BLOCKS + blk]);
    }
}

// The reordered triangles: position p holds the triangle of slot perm[p] (nothing when a
// graph-replayed build's status word is set).
__global__ void k_tri_gather(WideState W, uint32_t n, const uint32_t* fail)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || (fail && *fail)) return;
    const uint32_t g = W.idx[W.perm[p]];
    const float* src = W.tri9_src + 9 * size_t(g);
    float* dst = W.tri9 + 9 * size_t(p);
    // all nine loads before the first store (the buffers may alias as far as the compiler
    // knows, so interleaved copies would wait on memory once per value)
    float v[9];
    for (int j = 0; j < 9; ++j) v[j] = src[j];
    for (int j = 0; j < 9; ++j) dst[j] = v[j];
    W.tri_ids[p] = g;
}

// A graph-replayed build's status in lvl[13]: 0 a complete tree, 1 binary levels unfinished
// (nothing replaced), 2 wide levels unfinished, 4 wide capacity exceeded.
__global__ void k_build_status(uint32_t* lvl, const uint32_t* wl, uint32_t n)
{
    uint32_t st = lvl[12];
    if (!st)
    {
        if (wl[0] != wl[1] || wl[2] != n) st |= 2u;
        if (wl[4]) st |= 4u;
    }
    lvl[13] = st;
}

// The next wide level is the inner children placed by this one.
__global__ void k_next_wide(WideState W)
{
    const uint32_t begin = W.wl[0], end = W.wl[1];
    if (begin == end) return;  // past the last level: nothing placed, not a level
    const uint32_t inner = W.wl[5], tris = W.wl[6];
    const bool fits = end + inner <= W.wcap;
    W.wl[0] = end;
    W.wl[1] = fits ? end + inner : end;  // beyond the capacity: stop, with the error set
    if (!fits) W.wl[4] = 1u;
    W.wl[2] += tris;
    W.wl[3] += 1u;
}

// Wide level 0: the roots' images (wide node r is the root of tree r); no level at all when a
// graph-replayed build's binary levels did not finish (fail set).
__global__ void k_init_wide(WideState W, uint32_t R, const uint32_t* fail)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    const bool skip = fail && *fail;
    if (r < R && !skip) W.wbin[r] = r;
    if (r == 0)
    {
        W.wl[0] = 0u;
        W.wl[1] = skip ? 0u : R;
        W.wl[2] = 0u;
        W.wl[3] = 0u;
        W.wl[4] = 0u;
    }
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

__global__ void k_scene_slack(const float* pbox, uint32_t n, uint32_t* out)
{
    __shared__ uint32_t sh;
    if (threadIdx.x == 0) sh = 0;
    __syncthreads();
    uint32_t m = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int a = 0; a < 6; ++a) m = max(m, __float_as_uint(fabsf(pbox[6 * size_t(i) + a])));
    atomicMax(&sh, m);
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, sh);
}

// The quantisation slack: 2^-17 x the largest absolute coordinate (1 for an all-zero scene).
__global__ void k_slack_final(const uint32_t* m_bits, float* slack)
{
    float mf = __uint_as_float(*m_bits);
    if (mf == 0.0f) mf = 1.0f;
    *slack = ldexpf(mf, -17);
}


template <typename T>
int dalloc(Context* ctx, T*& p, size_t count)
{
    p = nullptr;
    if (count == 0) count = 1;
    APG_HIP_CHECK(ctx, hipMalloc(&p, count * sizeof(T)));
    return APG_OK;
}

inline uint32_t blocks_for(uint32_t n, uint32_t b) { return n ? (n + b - 1) / b : 1; }

} // namespace

// Device scratch of the GPU builder, kept by the context across builds and reallocated only
// when a build needs more (per-frame rebuilds allocate nothing).
struct BuildCache
{
    uint32_t n_cap = 0, r_cap = 0;
    std::vector<void*> ptrs;
    BuildState S{};
    uint32_t* rfirst = nullptr;      // R + 1: first triangle of each tree
    uint32_t* slack_bits = nullptr;  // largest |coordinate| as float bits
    float* slack = nullptr;
    // wide collapse, capacity n + R + 2 wide nodes
    uint32_t wide_cap = 0;
    uint32_t chunk_cap = 0;  // chunks a level can have (the chunk kernels' grid)
    // the previous build: triangles, trees, binary levels that held nodes, wide levels
    uint32_t last_n = 0, last_r = 0, last_levels = 0, last_wide = 0;
    uint32_t *wbin = nullptr, *wchild = nullptr, *n_inner = nullptr, *n_tris = nullptr;
    uint32_t *s_inner = nullptr, *s_tris = nullptr, *wpart = nullptr, *wl = nullptr;
    uint32_t* rb = nullptr;          // pinned read-back words

    void release()
    {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
        if (rb) (void)hipHostFree(rb);
        rb = nullptr;
        n_cap = r_cap = wide_cap = chunk_cap = 0;
    }
    ~BuildCache() { release(); }

    template <typename T>
    int get(Context* ctx, T*& p, size_t count)
    {
        int rc = dalloc(ctx, p, count);
        if (rc == APG_OK) ptrs.push_back(p);
        return rc;
    }

    int ensure(Context* ctx, uint32_t n, uint32_t R)
    {
        if (n <= n_cap && R <= r_cap) return APG_OK;
        release();
        int rc = ensure_sized(ctx, n, R);
        if (rc) release();
        return rc;
    }

    int ensure_sized(Context* ctx, uint32_t n, uint32_t R)
    {
        const size_t ncap = 2 * size_t(n) + 2 + R;
        // histogram slots: only the level loop's nodes (more than FIN slots) take one, and a
        // level's nodes own disjoint slots
        const size_t bincap = size_t(n) / (FIN + 1) + 2;
        const size_t lcap = size_t(n) / (FIN + 1) + 2;  // nodes of more than FIN slots per level
        // chunks of a level: ceil(count / CHUNK) per listed node
        const size_t ccap = size_t(n) / CHUNK + lcap + 1;
        const size_t wcap = size_t(n) + R + 2;      // >= binary nodes / 2 + 2 + R
        int rc;
        if ((rc = get(ctx, S.pbox, 6 * size_t(n))) || (rc = get(ctx, S.tri9_src, 9 * size_t(n))) ||
            (rc = get(ctx, S.pbox_tmp, 6 * size_t(n))) ||
            (rc = get(ctx, S.idx, n)) || (rc = get(ctx, S.idx_tmp, n)) || (rc = get(ctx, S.flag8, n)) ||
            (rc = get(ctx, S.sbin, n)) ||
            (rc = get(ctx, S.chunk_below, ccap)) || (rc = get(ctx, S.chunk_off, ccap)) ||
            (rc = get(ctx, S.nbox, 6 * ncap)) || (rc = get(ctx, S.ncbox, 6 * ncap)) || (rc = get(ctx, S.nfirst, ncap)) ||
            (rc = get(ctx, S.ncount, ncap)) || (rc = get(ctx, S.nleft, ncap)) || (rc = get(ctx, S.nflags, ncap)) ||
            (rc = get(ctx, S.nsplit, ncap)) || (rc = get(ctx, S.nsplit_mm, 2 * ncap)) ||
            (rc = get(ctx, S.nbinslot, ncap)) || (rc = get(ctx, S.bins, bincap * NODE_BIN_WORDS)) ||
            (rc = get(ctx, S.counters, 16)) ||
            (rc = get(ctx, S.lvl, 16)) || (rc = get(ctx, S.large, lcap)) || (rc = get(ctx, S.large_off, lcap + 1)) ||
            (rc = get(ctx, S.fin, size_t(n) + R)) ||
            (rc = get(ctx, rfirst, size_t(R) + 1)) || (rc = get(ctx, slack_bits, 1)) || (rc = get(ctx, slack, 1)) ||
            (rc = get(ctx, wbin, 2 * wcap + 8)) || (rc = get(ctx, wchild, 8 * wcap)) || (rc = get(ctx, n_inner, wcap)) ||
            (rc = get(ctx, n_tris, wcap)) || (rc = get(ctx, s_inner, wcap + 1)) || (rc = get(ctx, s_tris, wcap + 1)) ||
            (rc = get(ctx, wpart, 4 * WSCAN_BLOCKS)) || (rc = get(ctx, wl, 16)))
            return rc;
        APG_HIP_CHECK(ctx, hipHostMalloc(&rb, 64 * sizeof(uint32_t)));
        n_cap = n;
        r_cap = R;
        wide_cap = uint32_t(wcap);
        chunk_cap = uint32_t(ccap);
        return APG_OK;
    }
};

void build_cache_free(BuildCache* c) { delete c; }

namespace {

// The scene's node and triangle buffers, reused across builds while they are large enough.
int ensure_scene_buffers(Context* ctx, uint32_t nodes, uint32_t tris)
{
    DeviceScene& sc = ctx->scene;
    if (sc.nodes && sc.node_cap >= nodes && sc.tri_cap >= tris) return APG_OK;
    if (sc.nodes) (void)hipFree(sc.nodes);
    if (sc.tris) (void)hipFree(sc.tris);
    if (sc.tri_ids) (void)hipFree(sc.tri_ids);
    sc.nodes = nullptr;
    sc.tris = nullptr;
    sc.tri_ids = nullptr;
    sc.node_cap = sc.tri_cap = 0;
    int rc;
    if ((rc = dalloc(ctx, sc.nodes, nodes)) || (rc = dalloc(ctx, sc.tris, 9 * size_t(tris))) ||
        (rc = dalloc(ctx, sc.tri_ids, tris)))
    {
        if (sc.nodes) (void)hipFree(sc.nodes);
        if (sc.tris) (void)hipFree(sc.tris);
        sc.nodes = nullptr;
        sc.tris = nullptr;
        sc.tri_ids = nullptr;
        return rc;
    }
    sc.node_cap = nodes;
    sc.tri_cap = tris;
    return APG_OK;
}

uint32_t ceil_log2(uint32_t n)
{
    uint32_t l = 0;
    while ((1ull << l) < n) ++l;
    return l;
}

} // namespace

namespace {

// One binary level of the level loop (every kernel reads the level's range on the device).
void enqueue_binary_level(const BuildState& S, uint32_t thr, uint32_t g_node, uint32_t g_chunk, hipStream_t s)
{
    k_level_init<<<g_node, 256, 0, s>>>(S);
    k_large_scan<<<1, SCAN1, 0, s>>>(S);
    k_bounds_large<<<g_chunk, BLOCK, 0, s>>>(S);
    k_bin_large<<<g_chunk, BLOCK, 0, s>>>(S);
    k_sweep<<<g_node, 256, 0, s>>>(S, thr);
    k_part_count<<<g_chunk, BLOCK, 0, s>>>(S);
    k_part_scan<<<1, SCAN1, 0, s>>>(S);
    k_part_scatter<<<g_chunk, BLOCK, 0, s>>>(S);
}

// One wide level of the collapse.
void enqueue_wide_level(const WideState& W, uint32_t g_wide, hipStream_t s)
{
    k_wide_gather<<<g_wide, 128, 0, s>>>(W);
    k_wscan_local<<<WSCAN_BLOCKS, BLOCK, 0, s>>>(W);
    k_wscan_top<<<1, WSCAN_BLOCKS, 0, s>>>(W);
    k_wide_emit<<<g_wide, 128, 0, s>>>(W);
    k_next_wide<<<1, 1, 0, s>>>(W);
}

WideState wide_state(const BuildState& S, const BuildCache& C, const DeviceScene& sc, uint32_t wcap)
{
    WideState W;
    W.nbox = S.nbox; W.nfirst = S.nfirst; W.ncount = S.ncount; W.nleft = S.nleft;
    W.nflags = S.nflags; W.idx = S.idx; W.tri9_src = S.tri9_src;
    W.wbin = C.wbin; W.wchild = C.wchild; W.n_inner = C.n_inner; W.n_tris = C.n_tris;
    W.s_inner = C.s_inner; W.s_tris = C.s_tris;
    W.nodes = sc.nodes; W.tri9 = sc.tris; W.tri_ids = sc.tri_ids;
    W.slack = C.slack; W.wl = C.wl; W.wcap = wcap;
    W.wpart = C.wpart; W.perm = S.idx_tmp;  // idx_tmp is free once the binary tree is built
    return W;
}

uint32_t fin_grid(uint32_t n) { return std::min<uint32_t>(blocks_for(n, 256), 3u * 256u); }
uint32_t node_grid(uint32_t n) { return std::min<uint32_t>(blocks_for(n, 256), 2048u); }

} // namespace

int build_scene_gpu(Context* ctx, uint32_t leaf_threshold, bool per_shape, bool stats, BinaryTree& bt)
{
    const auto t_start = std::chrono::steady_clock::now();
    const uint32_t n = ctx->total_tris;
    hipStream_t s = nullptr;
    // the previous scene's buffers stay allocated for this build to reuse
    DeviceScene& sc = ctx->scene;
    sc.tris_v_valid = false;
    sc.node_count = sc.tri_count = sc.wide_depth = 0;
    sc.roots = 1;
    bt = BinaryTree{};
    if (n == 0) return APG_OK;

    int rc;
    // geometry lives on the device (uploaded when the host copy changed)
    if ((rc = ctx->sync_geometry_to_device(s))) return rc;
    const float* d_xyz = ctx->d_xyz;
    const uint32_t* d_tri = ctx->d_tri;
    const uint32_t R = per_shape ? uint32_t(ctx->shapes.size()) : 1u;
    if (!ctx->build_cache) ctx->build_cache = new BuildCache;
    BuildCache& C = *ctx->build_cache;
    if ((rc = C.ensure(ctx, n, R))) return rc;
    BuildState S = C.S;
    const uint32_t thr = std::max<uint32_t>(1, leaf_threshold);

    k_prep<<<blocks_for(n, 256), 256, 0, s>>>(d_xyz, d_tri, n, S.tri9_src, S.pbox);
    // the quantisation slack from every triangle's box, before the level loop permutes them
    k_fill_u32<<<1, 64, 0, s>>>(C.slack_bits, 1, 0u);
    k_scene_slack<<<256, 256, 0, s>>>(S.pbox, n, C.slack_bits);
    k_slack_final<<<1, 1, 0, s>>>(C.slack_bits, C.slack);
    // roots: one per shape (forest) or one for the scene
    std::vector<uint32_t> rfirst(R + 1);
    {
        uint32_t off = 0;
        for (uint32_t r = 0; r < R; ++r)
        {
            rfirst[r] = off;
            off += per_shape ? uint32_t(ctx->shapes[r].idx.size() / 3) : n;
        }
        rfirst[R] = off;
    }
    APG_HIP_CHECK(ctx, hipMemcpyAsync(C.rfirst, rfirst.data(), 4 * (size_t(R) + 1), hipMemcpyHostToDevice, s));
    k_init_slots<<<blocks_for(n, 256), 256, 0, s>>>(n, S.idx);
    k_init_roots<<<blocks_for(R, 256), 256, 0, s>>>(S, C.rfirst, R);

    // ---- binary levels: enqueued in batches, the level range read back between batches ----
    const uint32_t g_node = node_grid(n);
    const uint32_t g_chunk = C.chunk_cap;
    const uint32_t g_fin = fin_grid(n);
    uint32_t* rb = C.rb;
    auto enqueue_level = [&]() { enqueue_binary_level(S, thr, g_node, g_chunk, s); };
    // the level loop ends where every node holds at most FIN slots, a little deeper than
    // log2(n / FIN) levels; a rebuild of the same triangle count starts with the levels the
    // previous build took
    const bool same = C.last_n == n && C.last_r == R;
    uint32_t enqueued = 0;
    for (;;)
    {
        const uint32_t batch = enqueued == 0 ? (same && C.last_levels ? C.last_levels : ceil_log2(n / FIN + 1) + 3u) : 2u;
        for (uint32_t b = 0; b < batch; ++b) enqueue_level();
        enqueued += batch;
        // done when the last level allocated no children
        APG_HIP_CHECK(ctx, hipMemcpyAsync(rb, S.lvl, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        APG_HIP_CHECK(ctx, hipMemcpyAsync(rb + 4, S.counters, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        APG_HIP_CHECK(ctx, hipStreamSynchronize(s));
        APG_HIP_CHECK(ctx, hipGetLastError());
        if (rb[1] == rb[4]) break;
        if (enqueued > 4096) return ctx->fail(APG_E_CAPACITY, "binary tree deeper than 4096 levels");
    }
    // the subtrees of at most FIN slots
    k_finish<<<g_fin, FIN_BLOCK, 0, s>>>(S, thr);
    k_finish_done<<<1, 1, 0, s>>>(S);
    APG_HIP_CHECK(ctx, hipMemcpyAsync(rb, S.lvl, 10 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    APG_HIP_CHECK(ctx, hipStreamSynchronize(s));
    APG_HIP_CHECK(ctx, hipGetLastError());
    const uint32_t bin_ids = rb[6], bin_nodes = rb[7];
    C.last_levels = rb[9];

    // ---- wide collapse ----
    // each wide node holds >= 2 binary nodes or is a root; the full capacity when the scene
    // buffers must hold any rebuild (graph-replayed frames)
    const uint32_t wcap = ctx->scene_full_capacity ? C.wide_cap : bin_nodes / 2 + 2 + R;
    if (wcap > C.wide_cap) return ctx->fail(APG_E_CAPACITY, "wide node capacity exceeded");
    if ((rc = ensure_scene_buffers(ctx, wcap, n))) return rc;
    const WideState W = wide_state(S, C, sc, wcap);
    k_init_wide<<<blocks_for(R, 256), 256, 0, s>>>(W, R, nullptr);
    const uint32_t g_wide = std::min<uint32_t>(blocks_for(wcap, 128), 2048u);
    auto enqueue_wide = [&]() { enqueue_wide_level(W, g_wide, s); };
    enqueued = 0;
    for (;;)
    {
        const uint32_t batch =
            enqueued == 0 ? (same && C.last_wide ? C.last_wide : std::max<uint32_t>(2u, (ceil_log2(n) + 2u) / 3u)) : 2u;
        for (uint32_t b = 0; b < batch; ++b) enqueue_wide();
        enqueued += batch;
        APG_HIP_CHECK(ctx, hipMemcpyAsync(rb, W.wl, 5 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        APG_HIP_CHECK(ctx, hipStreamSynchronize(s));
        APG_HIP_CHECK(ctx, hipGetLastError());
        if (rb[4]) return ctx->fail(APG_E_CAPACITY, "wide node capacity exceeded");
        if (rb[0] == rb[1]) break;
        if (enqueued > 4096) return ctx->fail(APG_E_CAPACITY, "wide tree deeper than 4096 levels");
    }
    const uint32_t wide_nodes = rb[1], tri_cursor = rb[2], wlevels = rb[3];
    if (tri_cursor == n) k_tri_gather<<<blocks_for(n, 256), 256, 0, s>>>(W, n, nullptr);
    C.last_wide = wlevels;
    C.last_n = n;
    C.last_r = R;
    if (tri_cursor != n) return ctx->fail(APG_E_STATE, "GPU build lost triangles");
    sc.node_count = wide_nodes;
    sc.tri_count = tri_cursor;
    sc.roots = R;
    sc.wide_depth = wlevels;
    bt.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();

    // ---- scene bounds (union of the roots) ----
    {
        std::vector<uint32_t> rbox(6 * size_t(R));
        APG_HIP_CHECK(ctx, hipMemcpy(rbox.data(), S.nbox, rbox.size() * 4, hipMemcpyDeviceToHost));
        for (int a = 0; a < 3; ++a)
        {
            bt.scene_lo[a] = FLT_MAX;
            bt.scene_hi[a] = -FLT_MAX;
            for (uint32_t r = 0; r < R; ++r)
            {
                bt.scene_lo[a] = std::min(bt.scene_lo[a], float_of(rbox[6 * size_t(r) + a]));
                bt.scene_hi[a] = std::max(bt.scene_hi[a], float_of(rbox[6 * size_t(r) + 3 + a]));
            }
        }
    }
    if (!stats) return APG_OK;

    // ---- binary-tree statistics and leaf keys (host, from the device tree) ----
    std::vector<uint32_t> hbox(6 * size_t(bin_ids)), hfirst(bin_ids), hcount(bin_ids),
        hleft(bin_ids), hflags(bin_ids), hidx(n);
    APG_HIP_CHECK(ctx, hipMemcpy(hbox.data(), S.nbox, hbox.size() * 4, hipMemcpyDeviceToHost));
    APG_HIP_CHECK(ctx, hipMemcpy(hfirst.data(), S.nfirst, bin_ids * 4, hipMemcpyDeviceToHost));
    APG_HIP_CHECK(ctx, hipMemcpy(hcount.data(), S.ncount, bin_ids * 4, hipMemcpyDeviceToHost));
    APG_HIP_CHECK(ctx, hipMemcpy(hleft.data(), S.nleft, bin_ids * 4, hipMemcpyDeviceToHost));
    APG_HIP_CHECK(ctx, hipMemcpy(hflags.data(), S.nflags, bin_ids * 4, hipMemcpyDeviceToHost));
    APG_HIP_CHECK(ctx, hipMemcpy(hidx.data(), S.idx, size_t(n) * 4, hipMemcpyDeviceToHost));
    auto area = [&](uint32_t k) {
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { lo[a] = float_of(hbox[6 * size_t(k) + a]); hi[a] = float_of(hbox[6 * size_t(k) + 3 + a]); }
        const float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
        return double((x * y + y * z + z * x) * 2.0f);
    };
    bt.leaf_keys.assign(n, 0);
    uint32_t sah_nodes = 0, sah_leaves = 0, maxd = 0;
    double sah = 0.0, sah_leaf = 0.0;
    struct Item { uint32_t k, d; };
    for (uint32_t r = 0; r < R; ++r)
    {
        // per tree, normalised by its root's area; summed over the trees
        double inner = 0.0, leaf = 0.0;
        std::vector<Item> stack{ { r, 0 } };
        while (!stack.empty())
        {
            Item it = stack.back();
            stack.pop_back();
            sah_nodes++;
            maxd = std::max(maxd, it.d);
            const double sa = area(it.k);
            if (hflags[it.k] & NF_SAH_LEAF)
            {
                sah_leaves++;
                leaf += sa * double(hcount[it.k]);
                uint32_t mn = 0xFFFFFFFFu;
                for (uint32_t j = 0; j < hcount[it.k]; ++j) mn = std::min(mn, hidx[hfirst[it.k] + j]);
                for (uint32_t j = 0; j < hcount[it.k]; ++j) bt.leaf_keys[hidx[hfirst[it.k] + j]] = mn;
                continue;
            }
            inner += sa;
            stack.push_back({ hleft[it.k] + 1, it.d + 1 });
            stack.push_back({ hleft[it.k], it.d + 1 });
        }
        const double root_sa = area(r);
        sah += (inner + leaf) / root_sa;
        sah_leaf += leaf / root_sa;
    }
    bt.sah_cost = sah;
    bt.sah_cost_leaf = sah_leaf;
    bt.sah_nodes = sah_nodes;
    bt.sah_leaves = sah_leaves;
    bt.max_depth = maxd;
    return APG_OK;
}

// A rebuild for capture into a HIP graph: the geometry on the device rebuilt with no host
// read-back, running the previous eager build's binary and wide level counts plus `extra` of
// each (less `cut` binary levels: a test hook).  *status (device) is the build's status word: 0 a complete tree; otherwise the scene
// keeps the previous tree (binary levels unfinished) or holds an incomplete one (wide levels
// unfinished) and must be rebuilt eagerly.  Needs a previous eager build of the same triangles
// and trees with the scene buffers at full capacity (scene_full_capacity).
int build_scene_gpu_enqueue(Context* ctx, uint32_t leaf_threshold, bool per_shape, uint32_t extra, uint32_t cut,
                            hipStream_t s, uint32_t** status)
{
    const uint32_t n = ctx->total_tris;
    const uint32_t R = per_shape ? uint32_t(ctx->shapes.size()) : 1u;
    BuildCache* Cp = ctx->build_cache;
    if (n == 0 || !Cp || Cp->last_n != n || Cp->last_r != R || !Cp->last_levels || !Cp->last_wide)
        return ctx->fail(APG_E_STATE, "a replayed build needs an eager GPU build of the same scene first");
    const BuildCache& C = *Cp;
    const DeviceScene& sc = ctx->scene;
    if (!sc.nodes || sc.node_cap < C.wide_cap || sc.tri_cap < n)
        return ctx->fail(APG_E_STATE, "the scene buffers are not sized for replayed builds");
    const BuildState S = C.S;
    const uint32_t thr = std::max<uint32_t>(1, leaf_threshold);
    k_prep<<<blocks_for(n, 256), 256, 0, s>>>(ctx->d_xyz, ctx->d_tri, n, S.tri9_src, S.pbox);
    k_fill_u32<<<1, 64, 0, s>>>(C.slack_bits, 1, 0u);
    k_scene_slack<<<256, 256, 0, s>>>(S.pbox, n, C.slack_bits);
    k_slack_final<<<1, 1, 0, s>>>(C.slack_bits, C.slack);
    k_init_slots<<<blocks_for(n, 256), 256, 0, s>>>(n, S.idx);
    k_init_roots<<<blocks_for(R, 256), 256, 0, s>>>(S, C.rfirst, R);
    const uint32_t g_node = node_grid(n), g_chunk = C.chunk_cap;
    const uint32_t levels = C.last_levels + extra > cut ? C.last_levels + extra - cut : 0u;
    for (uint32_t l = 0; l < levels; ++l) enqueue_binary_level(S, thr, g_node, g_chunk, s);
    k_build_gate<<<1, 1, 0, s>>>(S);
    k_finish<<<fin_grid(n), FIN_BLOCK, 0, s>>>(S, thr);
    k_finish_done<<<1, 1, 0, s>>>(S);
    const WideState W = wide_state(S, C, sc, C.wide_cap);
    k_init_wide<<<blocks_for(R, 256), 256, 0, s>>>(W, R, S.lvl + 12);
    const uint32_t g_wide = std::min<uint32_t>(blocks_for(C.wide_cap, 128), 2048u);
    for (uint32_t l = 0; l < C.last_wide + extra; ++l) enqueue_wide_level(W, g_wide, s);
    k_build_status<<<1, 1, 0, s>>>(S.lvl, C.wl, n);
    k_tri_gather<<<blocks_for(n, 256), 256, 0, s>>>(W, n, S.lvl + 13);
    APG_HIP_CHECK(ctx, hipGetLastError());
    *status = S.lvl + 13;
    return APG_OK;
}

// The vertex-form triangles from the vertices the tree was built from, which the watertight
// test reads: gathered after a build when that test is selected, or when it is selected later
// (the buffer is kept across builds while it is large enough).
int gather_vertex_triangles(Context* ctx)
{
    DeviceScene& S = ctx->scene;
    if (S.tri_count == 0) { S.tris_v_valid = true; return APG_OK; }
    if (int rc = ctx->sync_geometry_to_device(nullptr)) return rc;
    if (!S.tris_v || S.tris_v_cap < S.tri_count)
    {
        if (S.tris_v) (void)hipFree(S.tris_v);
        S.tris_v = nullptr;
        S.tris_v_cap = 0;
        APG_HIP_CHECK(ctx, hipMalloc(&S.tris_v, size_t(S.tri_count) * 9 * sizeof(float)));
        S.tris_v_cap = S.tri_count;
    }
    k_gather_vertices<<<(S.tri_count + 255) / 256, 256>>>(S.tri_ids, S.tri_count, ctx->d_tri, ctx->d_xyz, S.tris_v);
    APG_HIP_CHECK(ctx, hipGetLastError());
    APG_HIP_CHECK(ctx, hipDeviceSynchronize());
    S.tris_v_valid = true;
    return APG_OK;
}

} // namespace apg

extern "C" int apg_mesh_transform(apg_context* ctx, uint32_t shape_id, const float* m,
                                  uint32_t first, uint32_t count, void* stream)
{
    using namespace apg;
    if (!ctx || !m) return APG_E_INVALID;
    if (shape_id >= ctx->shapes.size()) return ctx->fail(APG_E_INVALID, "unknown shape id");
    const uint32_t nv = uint32_t(ctx->shapes[shape_id].xyz.size() / 3);
    if (uint64_t(first) + count > nv) return ctx->fail(APG_E_INVALID, "vertex range out of bounds");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (int rc = ctx->sync_geometry_to_device(s)) return rc;
    size_t vbase = 0;
    for (uint32_t k = 0; k < shape_id; ++k) vbase += ctx->shapes[k].xyz.size() / 3;
    if (count)
        k_transform<<<(count + 255) / 256, 256, 0, s>>>(ctx->d_xyz + 3 * (vbase + first), count, m[0], m[1], m[2],
                                                       m[3], m[4], m[5], m[6], m[7], m[8], m[9], m[10], m[11]);
    APG_HIP_CHECK(ctx, hipGetLastError());
    ctx->host_stale = true;
    return APG_OK;
}
