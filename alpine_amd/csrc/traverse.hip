// Persistent wide-BVH traversal for gfx950 (closest hit and any hit).
//
// One ray per lane.  Each wave keeps its lanes busy from a global ray cursor:
// when enough lanes are idle the wave takes a contiguous batch with one
// atomicAdd and hands the indices to its idle lanes in lane order (ballot +
// popcount prefix), so short and long rays share a wave without waiting for
// each other.
//
// Per lane the traversal state is a "node group" (child base, inner-slot mask,
// still-to-visit hit mask in octant traversal order) and a "triangle group"
// (first triangle, 32-bit hit mask), after the compressed wide BVH traversal
// of Ylitie, Karras and Laine (HPG 2017).  Pending node groups go to a
// per-lane stack whose top APG_STACK_LDS entries live in LDS (lane-major, so a
// wave's accesses are bank-conflict free) and whose deeper entries spill to a
// per-thread HBM region.
//
// The box test works on the quantised 64 B nodes (wide_node.h) with one FMA
// per plane; the boxes carry a slack that covers that test's rounding.  The
// triangle test is the reference's Moller-Trumbore (bvh_common.cpp:66-103)
// with the same operation order and IEEE divisions; this file is compiled with
// -ffp-contract=off so it rounds exactly as the reference does.
#include "apg_internal.h"
#include "traverse_core.h"

#include <algorithm>
#include <cfloat>
#include <string>

namespace apg {

// Stack entries per lane: STACK in LDS (8 or 16, chosen by the launch
// configuration) and APG_STACK_MAX - STACK in the per-thread HBM spill area.
// A lane pushes at most one entry per wide level it descends, so the builders
// reject trees deeper than APG_STACK_MAX levels and the push below can only
// overflow under a lowered stack_limit.
constexpr uint32_t APG_STACK_MAX = APG_TRACE_STACK_MAX;
constexpr uint32_t APG_SPILL_ALLOC = 56;  // spill entries allocated per thread (>= 64 - 8)

// Triangle-group queue entries per wave: room for 64 new groups on top of the
// 32 a wave may hold between rounds.
constexpr uint32_t TRACE_QCAP = 96;
// Node visits per lane between the wave's fetch / triangle-round / finish steps.
constexpr uint32_t TRACE_VISITS = 4;
// Rays per taking-part wave: a launch over n rays runs min(grid, ceil(n / 64)) waves.
constexpr uint32_t TRACE_WAVE_RAYS = 64;

// LDS words per wave: the stack's top (2 words per entry and lane), 64 keys of
// 2 words, 64 barycentric pairs, 64 pending counts and the triangle-group queue
// (3 words and an owner byte per entry): 6.5 KiB with an 8-entry stack top, so
// six 256-thread workgroups fit in a CU's 160 KiB.
__host__ __device__ constexpr uint32_t trace_wave_lds_words(uint32_t stack)
{
    return stack * 128u + 128u + 128u + 64u + 3u * TRACE_QCAP + TRACE_QCAP / 4u;
}

__device__ __forceinline__ uint32_t queue_slot(uint32_t head, uint32_t k)
{
    const uint32_t i = head + k;
    return i >= TRACE_QCAP ? i - TRACE_QCAP : i;
}

struct TraceArgs
{
    const ApgWideNode* nodes;
    const float* tris;
    const uint32_t* tri_ids;
    uint32_t tri_count;
    apg_rays rays;
    uint32_t n;
    const uint32_t* n_dev;            // when set, the ray count is read here (queued work)
    const uint32_t* roots;            // when set, ray i starts at tree root roots[i]
    uint32_t root_count;
    uint32_t has_scene;
    apg_hits hits;
    uint8_t* occluded;
    uint32_t* counters;               // [0] ray cursor, [1] stack overflow flag, [2] waves done
    uint64_t* spill;                  // spill_stride entries per persistent thread
    uint32_t spill_stride;
    unsigned long long* fetch_counts; // [0] nodes, [1] triangles
    uint32_t fetch_threshold;
    uint32_t tri_batch;               // queued triangle tests that start a round
    uint32_t stack_max;               // stack entries a lane may hold (<= APG_STACK_MAX)
    uint32_t* flag_out;               // when set, the last wave stores the overflow flag here
};

// Inclusive prefix sum over the 64 lanes of a wave: shift-and-add inside each
// row of 16 lanes, then carry the row totals with the row broadcasts.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x111, 0xf, 0xf, false));  // row_shr:1
    x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x112, 0xf, 0xf, false));  // row_shr:2
    x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x114, 0xf, 0xf, false));  // row_shr:4
    x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x118, 0xf, 0xf, false));  // row_shr:8
    x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x142, 0xa, 0xf, false));  // row_bcast:15
    x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x143, 0xc, 0xf, false));  // row_bcast:31
    return x;
}

__device__ __forceinline__ uint32_t lane_u32(uint32_t v, uint32_t src)
{
    return uint32_t(__builtin_amdgcn_ds_bpermute(int(src << 2), int(v)));
}

__device__ __forceinline__ float lane_f32(float v, uint32_t src)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(int(src << 2), __float_as_int(v)));
}

// WT: the watertight triangle test over the vertex-form triangles instead of the reference's
// Moller-Trumbore over (v0, e0, e1).
template <bool ANY, bool COUNT, uint32_t APG_STACK_LDS, bool WT>
__global__ __launch_bounds__(256, 6) void trace_kernel(TraceArgs A)
{
    constexpr uint32_t APG_STACK_SPILL = APG_STACK_MAX - APG_STACK_LDS;
    static_assert(APG_STACK_SPILL <= APG_SPILL_ALLOC, "spill area smaller than the stack");
    extern __shared__ uint32_t lds[];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    uint32_t* wl = lds + wave * trace_wave_lds_words(APG_STACK_LDS);
    uint32_t* st_base = wl;
    uint32_t* st_bits = st_base + APG_STACK_LDS * 64u;
    // per lane (ray): best hit key so far (|t| bits, then triangle index, then the
    // sign of t), the winning hit's b0, b1, and a pending word: queued triangle
    // tests in the low half, lanes helping with the ray in the high half
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(wl + APG_STACK_LDS * 128u);
    float* rec = reinterpret_cast<float*>(wl + APG_STACK_LDS * 128u + 128u);
    uint32_t* pend = wl + APG_STACK_LDS * 128u + 256u;
    // the wave's ring of triangle groups: base, count mask and pending mask of each
    // group, and the lane that owns it
    uint32_t* queue = wl + APG_STACK_LDS * 128u + 320u;
    uint8_t* q_owner = reinterpret_cast<uint8_t*>(queue + 3u * TRACE_QCAP);
    uint64_t* spill = A.spill + size_t(blockIdx.x * blockDim.x + threadIdx.x) * A.spill_stride;

    const uint32_t nrays = A.n_dev ? *A.n_dev : A.n;
    // Only one wave per TRACE_WAVE_RAYS rays takes part (the first ones, spread over the
    // CUs as the blocks are): the others leave before touching the ray cursor, so a launch
    // over few rays (a late bounce of a render, whose count only the device knows) does not
    // serialise a claim and a finish atomic for every wave of the persistent grid.
    const uint32_t grid_waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t want = (nrays + TRACE_WAVE_RAYS - 1u) / TRACE_WAVE_RAYS;
    const uint32_t total_waves = want == 0u ? 1u : (want < grid_waves ? want : grid_waves);
    const uint32_t gwave = blockIdx.x * (blockDim.x >> 6) + wave;
    if (gwave >= total_waves) return;
    const uint32_t batch = A.tri_batch;
    bool active = false;
    bool exhausted = false;  // wave-uniform: the cursor has passed the last ray
    // Each wave's first pool is a static slice of half its share (whole runs of 64 rays), so
    // the waves of a large launch do not all queue on the cursor at the start; the cursor
    // hands out the rays after the slices.
    const uint32_t slice = (nrays / (2u * total_waves)) & ~63u;
    const uint32_t sliced = slice * total_waves;
    uint32_t pool_next = gwave * slice, pool_end = pool_next + slice, seen = sliced;  // wave-uniform
    uint32_t rid = 0;
    float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0;
    RayBox rb{};
    WtRay wr{};
    float tmax = 0;
    uint32_t g_base = 0, g_bits = 0;
    uint32_t sp = 0;
    uint32_t n_node = 0, n_tri = 0;
    uint32_t q_head = 0, q_count = 0, q_items = 0;  // wave-uniform
    // The lane whose ray this lane traces: itself, or, once the ray cursor is
    // exhausted, another lane whose stack it took a node group from.
    uint32_t owner = lane;
    pend[lane] = 0u;
    keys[lane] = ~0ull;

    for (;;)
    {
        // ---- ray fetch: idle lanes take the next indices of the wave's pool ----
        // A wave claims a run of ray indices from the global cursor with one
        // atomic and hands them out locally: same-address atomics serialise
        // device-wide, so claims are sized to the rays left (guided
        // self-scheduling: a 1/(2 x waves) share, 64..512 rays), which also
        // keeps the last claims small and the waves' ends close together.
        const uint64_t idle = __ballot(!active);
        const uint32_t nidle = __popcll(idle);
        if (nidle >= A.fetch_threshold || nidle == 64u)
        {
            if (pool_next == pool_end && !exhausted)
            {
                const uint32_t left = nrays > seen ? nrays - seen : 0u;
                uint32_t chunk = (left / (2u * total_waves)) & ~63u;
                chunk = chunk < 64u ? 64u : (chunk > 512u ? 512u : chunk);
                uint32_t base = 0;
                if (lane == 0u) base = atomicAdd(&A.counters[0], chunk);
                base = sliced + uint32_t(__builtin_amdgcn_readlane(int(base), 0));
                seen = base + chunk;
                pool_next = base < nrays ? base : nrays;
                pool_end = seen < nrays ? seen : nrays;
                if (seen >= nrays) exhausted = true;
            }
            const uint32_t avail = pool_end - pool_next;
            const uint32_t rank = __popcll(idle & ((1ull << lane) - 1ull));
            if (!active && rank < avail)
            {
                rid = pool_next + rank;
                active = true;
                owner = lane;
                ox = A.rays.ox[rid]; oy = A.rays.oy[rid]; oz = A.rays.oz[rid];
                dx = A.rays.dx[rid]; dy = A.rays.dy[rid]; dz = A.rays.dz[rid];
                tmax = A.rays.tmax ? A.rays.tmax[rid] : FLT_MAX;
                rb = make_raybox(ox, oy, oz, dx, dy, dz);
                if constexpr (WT) wr = make_wtray(ox, oy, oz, dx, dy, dz);
                keys[lane] = ~0ull;
                sp = 0;
                // Root as a group holding one inner child in slot 0.
                g_base = 0;
                if (A.roots)
                {
                    const uint32_t r = A.roots[rid];
                    g_base = r < A.root_count ? r : A.root_count - 1u;
                }
                g_bits = A.has_scene ? root_group_bits(rb.oct) : 0u;
            }
            pool_next += avail < nidle ? avail : nidle;
        }
        const bool drained = exhausted && pool_next == pool_end;  // wave-uniform
        if (__ballot(active) == 0ull)
        {
            if (drained) break;
            continue;
        }

        // ---- next node group from the stack ----
        auto pop = [&](uint32_t& b, uint32_t& bits) {
            sp--;
            const uint32_t ring = (sp % APG_STACK_LDS) * 64u + lane;
            b = st_base[ring];
            bits = st_bits[ring];
            if (sp >= APG_STACK_LDS)
            {
                const uint64_t e = spill[sp - APG_STACK_LDS];
                st_base[ring] = uint32_t(e);
                st_bits[ring] = uint32_t(e >> 32);
            }
        };
        // ---- up to VISITS inner nodes per lane and iteration ----
        // Each visit: one node per lane, its triangle groups queued, the next node group
        // popped.  A further visit runs while the queue has room for another 64 groups and a
        // lane has node work; the fetch, triangle rounds and finish steps below then run once
        // for all of them.
        for (uint32_t visit = 0; visit < TRACE_VISITS; ++visit)
        {
            if (visit && (q_count > TRACE_QCAP - 64u || __ballot(active && (g_bits >> 8) != 0u) == 0ull)) break;
            // ---- one inner node per lane ----
            uint32_t t_base = 0, t_mask = 0, t_cmask = 0;
            if (active && (g_bits >> 8) != 0u)
            {
                uint32_t hits = 0;
                const uint32_t node_idx = group_next_node(g_base, g_bits, rb.oct, hits);
                if (hits)
                {
                    // push the remaining siblings
                    if (sp >= A.stack_max)
                    {
                        // dropped: the host reports APG_E_OVERFLOW for this launch's results
                        A.counters[1] = 1u;
                        hits = 0;
                    }
                    else if (sp >= APG_STACK_LDS)
                    {
                        const uint32_t deep = sp - APG_STACK_LDS;  // < APG_STACK_SPILL
                        const uint32_t ring = (sp % APG_STACK_LDS) * 64u + lane;
                        spill[deep] = (uint64_t(st_bits[ring]) << 32) | st_base[ring];
                    }
                    if (hits)
                    {
                        const uint32_t ring = (sp % APG_STACK_LDS) * 64u + lane;
                        st_base[ring] = g_base;
                        st_bits[ring] = (hits << 8) | (g_bits & 0xffu);
                        sp++;
                    }
                }
                const uint4* np = reinterpret_cast<const uint4*>(A.nodes + node_idx);
                const uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3];
                const uint32_t w[16] = { w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w,
                                         w2.x, w2.y, w2.z, w2.w, w3.x, w3.y, w3.z, w3.w };
                if (COUNT) n_node++;
                node_visit(w, rb, tmax, g_base, g_bits, t_base, t_mask, t_cmask);
            }

            // ---- queue this iteration's triangle groups ----
            {
                const uint32_t cnt = bit_popc(t_mask);
                const uint64_t m = __ballot(cnt != 0u);
                if (m)
                {
                    if (cnt)
                    {
                        const uint32_t pos = __popcll(m & ((1ull << lane) - 1ull));
                        const uint32_t slot = queue_slot(q_head, q_count + pos);
                        queue[3u * slot] = t_base;
                        queue[3u * slot + 1u] = t_cmask;
                        queue[3u * slot + 2u] = t_mask;
                        q_owner[slot] = uint8_t(owner);
                        atomicAdd(&pend[owner], cnt);
                    }
                    q_count += uint32_t(__popcll(m));
                    q_items += uint32_t(__builtin_amdgcn_readlane(int(wave_incl_scan(cnt)), 63));
                }
            }
            if (active && (g_bits >> 8) == 0u && sp != 0u) pop(g_base, g_bits);
        }

        // ---- idle lanes take node groups from other lanes' stacks ----
        // Once no rays are left to fetch, the k-th idle lane takes the top stack
        // entry of the k-th lane with a non-empty stack and traces that subtree
        // for the entry's owner: its triangle groups are queued under the owner,
        // and the owner finishes only after its helpers (counted in the high
        // half of its pending word) are done.
        if (drained)
        {
            const uint64_t idle = __ballot(!active);
            const uint64_t donors = __ballot(active && sp != 0u);
            if (idle != 0ull && donors != 0ull)
            {
                const uint32_t ni = __popcll(idle), nd = __popcll(donors);
                const uint32_t npair = ni < nd ? ni : nd;
                const uint64_t below = (1ull << lane) - 1ull;
                uint32_t xb = 0, xbits = 0, src = lane;
                if (active && sp != 0u && uint32_t(__popcll(donors & below)) < npair) pop(xb, xbits);
                const uint32_t irank = __popcll(idle & below);
                const bool take = !active && irank < npair;
                if (take)
                {
                    const uint32_t lo = __popc(uint32_t(donors));
                    src = irank < lo ? select_bit(uint32_t(donors), irank)
                                     : 32u + select_bit(uint32_t(donors >> 32), irank - lo);
                }
                const uint32_t sb = lane_u32(xb, src), sbits = lane_u32(xbits, src);
                const uint32_t sown = lane_u32(owner, src);
                const float sox = lane_f32(ox, src), soy = lane_f32(oy, src), soz = lane_f32(oz, src);
                const float sdx = lane_f32(dx, src), sdy = lane_f32(dy, src), sdz = lane_f32(dz, src);
                const float stm = lane_f32(tmax, src);
                if (take)
                {
                    active = true;
                    owner = sown;
                    g_base = sb;
                    g_bits = sbits;
                    sp = 0;
                    ox = sox; oy = soy; oz = soz;
                    dx = sdx; dy = sdy; dz = sdz;
                    tmax = stm;
                    rb = make_raybox(ox, oy, oz, dx, dy, dz);
                    if constexpr (WT) wr = make_wtray(ox, oy, oz, dx, dy, dz);
                    atomicAdd(&pend[owner], 1u << 16);
                }
            }
        }

        // ---- triangle rounds: 64 queued tests at a time over all lanes ----
        // A round runs once `batch` tests are queued, when no lane has node work
        // left, or to keep room in the queue.  Item j belongs to the queued group
        // whose inclusive count first exceeds j; every owner keeps the smallest
        // key, i.e. the closest hit (exact-t ties go to the lower triangle index).
        const bool any_work = __ballot(active && (g_bits >> 8) != 0u) != 0ull;
        while (q_items >= batch || (q_items != 0u && !any_work) || q_count > TRACE_QCAP - 64u)
        {
            const uint32_t slot = queue_slot(q_head, lane);
            uint32_t gb = 0, gcm = 0, gm = 0, gown = 0;
            if (lane < q_count)
            {
                gb = queue[3u * slot];
                gcm = queue[3u * slot + 1u];
                gm = queue[3u * slot + 2u];
                gown = q_owner[slot];
            }
            const uint32_t c = bit_popc(gm);
            const uint32_t incl = wave_incl_scan(c);
            const uint32_t total = uint32_t(__builtin_amdgcn_readlane(int(incl), 63));
            const uint32_t take = total < 64u ? total : 64u;
            uint32_t o = 0;
#pragma unroll
            for (uint32_t step = 32u; step; step >>= 1)
                if (lane_u32(incl, o + step - 1u) <= lane) o += step;
            o = o > 63u ? 63u : o;
            const uint32_t om = lane_u32(gm, o);
            const uint32_t oexcl = lane_u32(incl, o) - bit_popc(om);
            const uint32_t ob = lane_u32(gb, o), ocm = lane_u32(gcm, o), own = lane_u32(gown, o);
            const float rox = lane_f32(ox, own), roy = lane_f32(oy, own), roz = lane_f32(oz, own);
            const float rtmax = lane_f32(tmax, own);
            WtRay rw{};
            float rdx = 0.0f, rdy = 0.0f, rdz = 0.0f;
            if constexpr (WT)
            {
                rw.ox = rox; rw.oy = roy; rw.oz = roz;
                rw.sx = lane_f32(wr.sx, own); rw.sy = lane_f32(wr.sy, own); rw.sz = lane_f32(wr.sz, own);
                rw.k = lane_u32(wr.k, own);
            }
            else
            {
                rdx = lane_f32(dx, own); rdy = lane_f32(dy, own); rdz = lane_f32(dz, own);
            }
            bool hit = false;
            float t = 0.0f, b0 = 0.0f, b1 = 0.0f;
            unsigned long long key = 0;
            if (lane < take)
            {
                const uint32_t p = select_bit(om, lane - oexcl);
                uint32_t tri = tri_index(ob, ocm, p);
                tri = tri < A.tri_count ? tri : A.tri_count - 1u;
                if (COUNT) n_tri++;
                if constexpr (WT)
                    hit = tri_hit_watertight(A.tris + size_t(tri) * APG_TRI_FLOATS, rw, rtmax, t, b0, b1);
                else
                    hit = tri_hit(A.tris + size_t(tri) * APG_TRI_FLOATS, rox, roy, roz, rdx, rdy, rdz, rtmax, t,
                                  b0, b1);
                if (hit)
                {
                    const uint32_t tb = __float_as_uint(t);
                    key = (uint64_t(tb & 0x7fffffffu) << 32) | (tri << 1) | (tb >> 31);
                    atomicMin(&keys[own], key);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (hit && keys[own] == key)
            {
                rec[2u * own] = b0;
                rec[2u * own + 1u] = b1;
            }
            // retire what this round consumed: whole groups from the head, and the
            // lowest set bits of the group the round ended in
            const uint32_t excl = incl - c;
            const uint32_t cons = take > excl ? (take - excl < c ? take - excl : c) : 0u;
            if (cons) atomicSub(&pend[gown], cons);
            const bool full = c != 0u && cons == c;
            if (cons && !full)
                queue[3u * slot + 2u] = gm & ~((1u << select_bit(gm, cons)) - 1u);
            const uint32_t popped = uint32_t(__popcll(__ballot(full)));
            q_head = queue_slot(q_head, popped);
            q_count -= popped;
            q_items -= take;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }

        // ---- hits found so far tighten tmax (any hit: stop) ----
        if (active)
        {
            const unsigned long long k = keys[owner];
            if (k != ~0ull)
            {
                if (ANY)
                {
                    g_bits = 0;
                    sp = 0;
                }
                else
                    tmax = __uint_as_float(uint32_t(k >> 32));
            }
        }

        // ---- a helper is done with its subtree ----
        if (active && owner != lane && (g_bits >> 8) == 0u && sp == 0u)
        {
            atomicSub(&pend[owner], 1u << 16);
            active = false;
        }

        // ---- finish: no node work, no queued tests and no helpers left ----
        if (active && owner == lane && (g_bits >> 8) == 0u && sp == 0u && pend[lane] == 0u)
        {
            const unsigned long long k = keys[lane];
            const bool hit = k != ~0ull;
            if (ANY)
                A.occluded[rid] = hit ? 1u : 0u;
            else
            {
                const uint32_t tri = (uint32_t(k) >> 1);
                A.hits.prim[rid] = hit ? A.tri_ids[tri] : 0xFFFFFFFFu;
                A.hits.t[rid] = hit ? __uint_as_float(uint32_t(k >> 32) | (uint32_t(k) << 31)) : FLT_MAX;
                A.hits.b0[rid] = hit ? rec[2u * lane] : 0.0f;
                A.hits.b1[rid] = hit ? rec[2u * lane + 1u] : 0.0f;
            }
            active = false;
        }
    }
    if (COUNT)
    {
        atomicAdd(&A.fetch_counts[0], (unsigned long long)n_node);
        atomicAdd(&A.fetch_counts[1], (unsigned long long)n_tri);
    }
    // The last taking-part wave to finish rewinds the ray cursor for the next launch.
    if (lane == 0u)
    {
        __threadfence();
        if (atomicAdd(&A.counters[2], 1u) == total_waves - 1u)
        {
            A.counters[0] = 0u;
            A.counters[2] = 0u;
            if (A.flag_out)
                A.flag_out[0] = __hip_atomic_load(&A.counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

static int ensure_trace_buffers(Context* ctx, uint32_t threads)
{
    if (!ctx->counters)
    {
        APG_HIP_CHECK(ctx, hipMalloc(&ctx->counters, 64));
        APG_HIP_CHECK(ctx, hipMemset(ctx->counters, 0, 64));
    }
    if (!ctx->overflow_host)
    {
        APG_HIP_CHECK(ctx, hipHostMalloc(&ctx->overflow_host, 64));
        *ctx->overflow_host = 0u;
    }
    if (ctx->spill_threads < threads)
    {
        if (ctx->spill)
        {
            // kernels of earlier calls may still read the old area
            APG_HIP_CHECK(ctx, hipDeviceSynchronize());
            (void)hipFree(ctx->spill);
        }
        ctx->spill = nullptr;
        APG_HIP_CHECK(ctx, hipMalloc(&ctx->spill, size_t(threads) * APG_SPILL_ALLOC * sizeof(uint64_t)));
        ctx->spill_threads = threads;
    }
    return APG_OK;
}

// Size the shared traversal buffers for a launch over up to n rays (callers that
// capture launches into graphs do this first: the capture must not allocate).
int ensure_trace_capacity(Context* ctx, uint32_t n)
{
    const apg_trace_config& cfg = ctx->cfg;
    uint32_t grid = uint32_t(ctx->cu_count) * cfg.blocks_per_cu;
    const uint32_t need = (n + cfg.block_size - 1) / cfg.block_size;
    if (grid > need) grid = need;
    if (grid == 0) grid = 1;
    return ensure_trace_buffers(ctx, grid * cfg.block_size);
}

// ---------------------------------------------------------------------------
// Allocation-free closest hit (apg_intersect_noalloc)
// ---------------------------------------------------------------------------
// The ray cursor / flags and the stack spill live in this file-scope device array instead of
// context buffers, so the entry point never allocates.  Its stack is the LDS top plus
// NOALLOC_SPILL spilled entries; scenes whose wide tree needs more are refused up front.
constexpr uint32_t NOALLOC_THREADS = 256u * 6u * 256u;  // the default persistent grid on 256 CUs
constexpr uint32_t NOALLOC_SPILL = 8u;
constexpr size_t NOALLOC_COUNTER_WORDS = 32u;         // 256 B: 16 counters, then padding
constexpr size_t NOALLOC_WORDS = NOALLOC_COUNTER_WORDS + size_t(NOALLOC_THREADS) * NOALLOC_SPILL;
static __device__ __align__(256) uint64_t scratch[NOALLOC_WORDS];  // 24 MiB + 256 B per GPU

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mutex;
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};  // device address of `scratch` on each GPU

// Look up the device address of `scratch` on `device` (current) once; apg_create calls this.
int noalloc_scratch_init(int device, std::string& err)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
    {
        err = "device index beyond the allocation-free scratch table";
        return APG_E_CAPACITY;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (g_noalloc_ptr[device]) return APG_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
    {
        err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return APG_E_HIP;
    }
    g_noalloc_ptr[device] = p;
    return APG_OK;
}

namespace {
__global__ void k_clear_words(uint32_t* w, uint32_t n)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) w[i] = 0u;
}
} // namespace

// Where a launch keeps its ray cursor / flags and its stack spill, and what that allows.
struct TraceScratch
{
    uint32_t* counters;     // 16 words, [0] and [2] zero between launches
    uint64_t* spill;        // spill_stride entries per persistent thread
    uint32_t spill_stride;
    uint32_t max_threads;   // persistent threads the spill has room for (0: any)
    uint32_t stack_max;     // stack entries per lane
    uint32_t* flag_out;     // optional: receives counters[1] at the end of the launch
};

namespace {

uint32_t trace_grid(const Context* ctx, uint32_t n, uint32_t max_threads)
{
    const apg_trace_config& cfg = ctx->cfg;
    uint32_t grid = uint32_t(ctx->cu_count) * cfg.blocks_per_cu;
    const uint32_t need = (n + cfg.block_size - 1) / cfg.block_size;
    if (grid > need) grid = need;
    if (max_threads && grid > max_threads / cfg.block_size) grid = max_threads / cfg.block_size;
    return grid ? grid : 1u;
}

uint32_t config_stack_max(const apg_trace_config& cfg)
{
    return cfg.stack_limit >= 1u && cfg.stack_limit <= APG_STACK_MAX ? cfg.stack_limit : APG_STACK_MAX;
}

int launch_trace_kernel(Context* ctx, const apg_rays& rays, uint32_t n, const uint32_t* n_dev,
                        const apg_hits* hits, bool any, uint8_t* occ, unsigned long long* counts,
                        hipStream_t stream, const uint32_t* roots, const TraceScratch& scr);

} // namespace

// Launch one traversal over `rays` on `stream` without cross-stream ordering.
// n_dev, when non-null, holds the ray count on the device (n is then the
// capacity used to size the grid).
int launch_trace_raw(Context* ctx, const apg_rays& rays, uint32_t n, const uint32_t* n_dev,
                     const apg_hits* hits, bool any, uint8_t* occ, unsigned long long* counts,
                     hipStream_t stream, const uint32_t* roots)
{
    const uint32_t grid = trace_grid(ctx, n, 0u);
    int rc = ensure_trace_buffers(ctx, grid * ctx->cfg.block_size);
    if (rc) return rc;
    const TraceScratch scr{ ctx->counters, ctx->spill, APG_SPILL_ALLOC, 0u, config_stack_max(ctx->cfg) };
    return launch_trace_kernel(ctx, rays, n, n_dev, hits, any, occ, counts, stream, roots, scr);
}

namespace {

int launch_trace_kernel(Context* ctx, const apg_rays& rays, uint32_t n, const uint32_t* n_dev,
                        const apg_hits* hits, bool any, uint8_t* occ, unsigned long long* counts,
                        hipStream_t stream, const uint32_t* roots, const TraceScratch& scr)
{
    const apg_trace_config& cfg = ctx->cfg;
    const uint32_t block = cfg.block_size;
    const uint32_t grid = trace_grid(ctx, n, scr.max_threads);

    TraceArgs A;
    A.nodes = ctx->scene.nodes;
    A.tris = ctx->scene.tris;
    A.tri_ids = ctx->scene.tri_ids;
    A.tri_count = ctx->scene.tri_count;
    A.rays = rays;
    A.n = n;
    A.n_dev = n_dev;
    A.roots = roots;
    A.root_count = ctx->scene.roots;
    A.has_scene = ctx->scene.node_count > 0 ? 1u : 0u;
    if (hits) A.hits = *hits; else A.hits = apg_hits{ nullptr, nullptr, nullptr, nullptr };
    A.occluded = occ;
    A.counters = scr.counters;
    A.spill = scr.spill;
    A.spill_stride = scr.spill_stride;
    A.fetch_counts = counts;
    A.fetch_threshold = cfg.fetch_threshold;
    A.tri_batch = cfg.tri_batch >= 1u && cfg.tri_batch <= 64u ? cfg.tri_batch : 48u;
    A.stack_max = scr.stack_max;
    A.flag_out = scr.flag_out;

    // the watertight test reads the vertex-form triangles (8-entry LDS stack only)
    const bool wt = cfg.triangle_test == APG_TRI_WATERTIGHT;
    if (wt)
    {
        if (!ctx->scene.tris_v_valid && ctx->scene.tri_count)
            return ctx->fail(APG_E_STATE, "no vertex-form triangles for the watertight test");
        A.tris = ctx->scene.tris_v;
    }
    const uint32_t stack = (cfg.stack_lds == 16 && !wt) ? 16u : 8u;
    const size_t shmem = size_t(block / 64) * trace_wave_lds_words(stack) * sizeof(uint32_t);
#define APG_LAUNCH(S, W)                                                                   \
    if (any)                                                                               \
    {                                                                                      \
        if (counts) trace_kernel<true, true, S, W><<<grid, block, shmem, stream>>>(A);     \
        else trace_kernel<true, false, S, W><<<grid, block, shmem, stream>>>(A);           \
    }                                                                                      \
    else                                                                                   \
    {                                                                                      \
        if (counts) trace_kernel<false, true, S, W><<<grid, block, shmem, stream>>>(A);    \
        else trace_kernel<false, false, S, W><<<grid, block, shmem, stream>>>(A);          \
    }
    if (wt) { APG_LAUNCH(8, true) }
    else if (stack == 16) { APG_LAUNCH(16, false) }
    else { APG_LAUNCH(8, false) }
#undef APG_LAUNCH
    APG_HIP_CHECK(ctx, hipGetLastError());
    return APG_OK;
}

} // namespace

// Launch one traversal of a host batch on the staging's stream with the staging's own counters
// and spill (grown to the batch's persistent grid), so concurrent batches need no ordering
// between them.  The kernel's last wave copies the staging's overflow flag (counters[1]) to
// flag_out, a host-visible word.
int launch_trace_staged(Context* ctx, const apg_rays& rays, uint32_t n, const apg_hits* hits, bool any,
                        uint8_t* occ, uint32_t* flag_out, HostStaging* s)
{
    if (n == 0) return APG_OK;
    if (!ctx->built) return ctx->fail(APG_E_STATE, "apg_scene_build has not been called");
    if (n > 0x7fffffffu) return ctx->fail(APG_E_CAPACITY, "more than 2^31 - 1 rays in one call");
    const uint32_t threads = trace_grid(ctx, n, 0u) * ctx->cfg.block_size;
    if (s->trace_threads < threads)
    {
        if (s->trace_words) APG_HIP_CHECK(ctx, hipFree(s->trace_words));
        s->trace_words = nullptr;
        s->trace_threads = 0;
        const size_t words = NOALLOC_COUNTER_WORDS + size_t(threads) * APG_SPILL_ALLOC;
        APG_HIP_CHECK(ctx, hipMalloc(&s->trace_words, words * sizeof(uint64_t)));
        // counters [0] and [2] return to zero at the end of every launch; [1] is cleared by
        // the caller after it reports an overflow
        APG_HIP_CHECK(ctx, hipMemset(s->trace_words, 0, NOALLOC_COUNTER_WORDS * sizeof(uint64_t)));
        s->trace_threads = threads;
    }
    const TraceScratch scr{ reinterpret_cast<uint32_t*>(s->trace_words), s->trace_words + NOALLOC_COUNTER_WORDS,
                            APG_SPILL_ALLOC, s->trace_threads, config_stack_max(ctx->cfg), flag_out };
    return launch_trace_kernel(ctx, rays, n, nullptr, hits, any, occ, nullptr, s->stream, nullptr, scr);
}

int Context::take_overflow()
{
    if (!counters || !overflow_host) return APG_OK;
    APG_HIP_CHECK(this, hipMemcpy(overflow_host, counters + 1, 4, hipMemcpyDeviceToHost));
    if (*overflow_host == 0u) return APG_OK;
    APG_HIP_CHECK(this, hipMemset(counters + 1, 0, 4));
    *overflow_host = 0u;
    return fail(APG_E_OVERFLOW, "a traversal needed more stack entries than the launch allows "
                                "(stack_limit); its results are not reliable");
}

int launch_trace(Context* ctx, const apg_rays* rays, uint32_t n, const apg_hits* hits, bool any,
                 uint8_t* occ, unsigned long long* counts, hipStream_t stream, const uint32_t* roots = nullptr)
{
    if (n == 0) return APG_OK;
    if (!rays || (!any && (!hits || !hits->prim || !hits->t || !hits->b0 || !hits->b1)) ||
        (any && !occ))
        return ctx->fail(APG_E_INVALID, "null ray or result pointer");
    if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz)
        return ctx->fail(APG_E_INVALID, "null ray component pointer");
    if (!ctx->built) return ctx->fail(APG_E_STATE, "apg_scene_build has not been called");
    if (n == 0) return APG_OK;
    if (n > 0x7fffffffu) return ctx->fail(APG_E_CAPACITY, "more than 2^31 - 1 rays in one call");

    std::lock_guard<std::mutex> lock(ctx->launch_mutex);
    if (!ctx->last_launch)
        APG_HIP_CHECK(ctx, hipEventCreateWithFlags(&ctx->last_launch, hipEventDisableTiming));
    // Kernels from all streams share the ray cursor and the spill area:
    // order this launch after the previous one.
    APG_HIP_CHECK(ctx, hipStreamWaitEvent(stream, ctx->last_launch, 0));
    int rc = launch_trace_raw(ctx, *rays, n, nullptr, hits, any, occ, counts, stream, roots);
    if (rc) return rc;
    APG_HIP_CHECK(ctx, hipEventRecord(ctx->last_launch, stream));
    return APG_OK;
}

} // namespace apg

extern "C" {

int apg_intersect(apg_context* ctx, const apg_rays* rays, uint32_t n, const apg_hits* hits,
                  void* stream)
{
    if (!ctx) return APG_E_INVALID;
    return apg::launch_trace(ctx, rays, n, hits, false, nullptr, nullptr,
                             static_cast<hipStream_t>(stream));
}

int apg_intersect_noalloc(apg_context* ctx, const apg_rays* rays, uint32_t n, const apg_hits* hits,
                          void* stream_)
{
    using namespace apg;
    if (!ctx) return APG_E_INVALID;
    if (n == 0) return APG_OK;
    if (!rays || !hits || !hits->prim || !hits->t || !hits->b0 || !hits->b1)
        return ctx->fail(APG_E_INVALID, "null ray or result pointer");
    if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz)
        return ctx->fail(APG_E_INVALID, "null ray component pointer");
    if (!ctx->built) return ctx->fail(APG_E_STATE, "apg_scene_build has not been called");
    if (n > 0x7fffffffu) return ctx->fail(APG_E_CAPACITY, "more than 2^31 - 1 rays in one call");
    void* base = ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES ? g_noalloc_ptr[ctx->device] : nullptr;
    if (!base) return ctx->fail(APG_E_STATE, "allocation-free scratch not set up for this device");
    const apg_trace_config& cfg = ctx->cfg;
    const bool wt = cfg.triangle_test == APG_TRI_WATERTIGHT;
    if (wt && !ctx->scene.tris_v_valid && ctx->scene.tri_count)
        return ctx->fail(APG_E_STATE, "no vertex-form triangles for the watertight test");
    const uint32_t lds = (cfg.stack_lds == 16 && !wt) ? 16u : 8u;
    const uint32_t stack_max = std::min<uint32_t>(lds + NOALLOC_SPILL, APG_STACK_MAX);
    // a lane holds at most wide_depth - 1 entries: refuse scenes the stack cannot hold
    if (ctx->scene.wide_depth > stack_max + 1u)
        return ctx->fail(APG_E_CAPACITY, "the scene's 8-wide tree has " + std::to_string(ctx->scene.wide_depth) +
                                             " levels; apg_intersect_noalloc holds " +
                                             std::to_string(stack_max + 1u) + " (use apg_intersect)");
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    uint64_t* words = static_cast<uint64_t*>(base);
    const TraceScratch scr{ reinterpret_cast<uint32_t*>(words), words + NOALLOC_COUNTER_WORDS, NOALLOC_SPILL,
                            NOALLOC_THREADS, stack_max };
    // the counters keep the last call's values: start from zero (a kernel, not a memset node,
    // so the call can be captured into a graph)
    k_clear_words<<<1, 64, 0, stream>>>(scr.counters, 16u);
    APG_HIP_CHECK(ctx, hipGetLastError());
    int rc = launch_trace_kernel(ctx, *rays, n, nullptr, hits, false, nullptr, nullptr, stream, nullptr, scr);
    if (rc) return rc;
    // scene builds wait for this launch (not under capture: the replays are the caller's)
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    APG_HIP_CHECK(ctx, hipStreamIsCapturing(stream, &cap));
    if (cap == hipStreamCaptureStatusNone) APG_HIP_CHECK(ctx, hipEventRecord(ctx->noalloc_launch, stream));
    return APG_OK;
}

int apg_occluded(apg_context* ctx, const apg_rays* rays, uint32_t n, uint8_t* occluded,
                 void* stream)
{
    if (!ctx) return APG_E_INVALID;
    return apg::launch_trace(ctx, rays, n, nullptr, true, occluded, nullptr,
                             static_cast<hipStream_t>(stream));
}

int apg_intersect_roots(apg_context* ctx, const apg_rays* rays, const uint32_t* roots, uint32_t n,
                        const apg_hits* hits, void* stream)
{
    if (!ctx) return APG_E_INVALID;
    if (n && !roots) return ctx->fail(APG_E_INVALID, "null roots");
    return apg::launch_trace(ctx, rays, n, hits, false, nullptr, nullptr,
                             static_cast<hipStream_t>(stream), roots);
}

int apg_occluded_roots(apg_context* ctx, const apg_rays* rays, const uint32_t* roots, uint32_t n,
                       uint8_t* occluded, void* stream)
{
    if (!ctx) return APG_E_INVALID;
    if (n && !roots) return ctx->fail(APG_E_INVALID, "null roots");
    return apg::launch_trace(ctx, rays, n, nullptr, true, occluded, nullptr,
                             static_cast<hipStream_t>(stream), roots);
}

int apg_intersect_counted(apg_context* ctx, const apg_rays* rays, uint32_t n,
                          const apg_hits* hits, int any, uint8_t* occluded,
                          uint64_t* node_fetches, uint64_t* tri_fetches, void* stream)
{
    if (!ctx || !node_fetches || !tri_fetches) return APG_E_INVALID;
    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned long long* counts = nullptr;
    APG_HIP_CHECK(ctx, hipMalloc(&counts, 2 * sizeof(unsigned long long)));
    APG_HIP_CHECK(ctx, hipMemsetAsync(counts, 0, 2 * sizeof(unsigned long long), s));
    int rc = apg::launch_trace(ctx, rays, n, hits, any != 0, occluded, counts, s);
    unsigned long long host[2] = { 0, 0 };
    if (rc == APG_OK)
    {
        hipError_t e = hipMemcpyAsync(host, counts, sizeof(host), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = ctx->fail(APG_E_HIP, hipGetErrorString(e));
    }
    (void)hipFree(counts);
    *node_fetches = host[0];
    *tri_fetches = host[1];
    if (rc == APG_OK) rc = ctx->take_overflow();
    return rc;
}

int apg_trace_status(apg_context* ctx)
{
    if (!ctx) return APG_E_INVALID;
    if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
    return ctx->take_overflow();
}

} // extern "C"
