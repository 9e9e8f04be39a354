// Wavefront path tracer for gfx950: the reference's per-pixel loop
// (Alpine::traceRayBatch / estimateDirectIllumination, alpine.cpp:290-432)
// restated as stages over structure-of-arrays queues in HBM.
//
// One pass = one sample for every pixel.  A pixel's sampler state persists
// across passes and calls, so the random numbers each path draws are the ones
// the reference's per-pixel loop would draw, in the same order:
//   generate  jitter (2), camera ray                    -> queue A
//   per bounce
//     extend   closest hit of queue A (traverse.hip, persistent)
//     shade    light-sampler choice (1), light sample (2), BSDF sample for MIS
//              (2), continuation BSDF sample (2): appends up to two shadow rays
//              and the continuation ray (queue B) with wave-aggregated atomics
//     shadow   any hit of the shadow queue
//     resolve  NEE contribution of the visible shadow rays, radiance of
//              finished paths into the per-pixel accumulation
//   A <- B
// Emission enters only through next-event estimation (lights are not
// geometry), as in the reference.  Optional Russian roulette after
// rr_depth (not in the reference) draws one extra number per surviving vertex.
#include "render_internal.h"
#include "guide_core.h"

#include <algorithm>
#include <cstring>

namespace apg {

int guide_ensure(Context* ctx, uint32_t cap);
int guide_train(Context* ctx, const float* feat, const float* q9, const float* dir, const float* weight, uint32_t M,
                uint32_t off, uint32_t stride, bool adam, float* loss, hipStream_t s);
int guide_last_loss(Context* ctx, float* loss);
int guide_infer(Context* ctx, const float* q9, uint32_t M, const uint32_t* M_dev, float* out, hipStream_t s);

struct Queue
{
    float* o[3];
    float* d[3];
    float* tmax;
    uint32_t* path;
    uint32_t* root;   // tree of each ray (per-shape forests)
};

// Tiled frames: tile t (row-major over tiles_x columns) has its own camera
// and names the tree its rays are traced in.
struct Tiles
{
    const DevCamera* cams;
    const uint32_t* roots;
    uint32_t tw, th, tiles_x;
};

struct RenderState
{
    // shading tables
    DevMaterial* mats = nullptr;
    uint32_t nmats = 0;
    uint32_t* tri_mat = nullptr;
    uint32_t tri_mat_n = 0;
    float4* tri_ng = nullptr;
    uint32_t tri_ng_n = 0;
    DevLight* lights = nullptr;
    uint32_t nlights = 0;
    uint32_t ls_kind = LS_UNIFORM;
    bool ls_set = false;
    float* alias_p = nullptr;
    float* alias_q = nullptr;
    uint32_t* alias_a = nullptr;
    DevLightPair* lpairs = nullptr;
    int32_t* lparent = nullptr;
    int32_t* leaf_of = nullptr;
    int32_t root_light = -1;
    DevCamera cam{};
    bool cam_set = false;
    DevCamera* tile_cams = nullptr;
    uint32_t* tile_roots = nullptr;
    uint32_t ntiles = 0, tile_w = 0, tile_h = 0;
    // graph replay of a pass
    hipGraphExec_t gexec = nullptr;
    hipStream_t gstream = nullptr;
    hipEvent_t gev = nullptr;
    std::vector<char> gkey;
    // pixels and paths: P = NPIX x S paths per pass, S samples of each pixel in flight
    uint32_t W = 0, H = 0, P = 0, NPIX = 0, S = 1;
    uint32_t* rng = nullptr;      // per path slot
    float* accum = nullptr;       // per path slot
    double* accum64 = nullptr;    // per path slot
    float* px = nullptr;          // per pixel: slot sums (reads, postprocess)
    double* px64 = nullptr;
    uint32_t total_spp = 0;
    // paths
    float* T = nullptr;
    float* L = nullptr;
    uint32_t* state = nullptr;  // depth | flags
    float* nee = nullptr;       // 10 per path: Tnee xyz, A xyz, B xyz, light pdf
    uint8_t* vis = nullptr;     // 2 per path
    Queue qa{}, qb{}, qs{};
    uint32_t* hprim = nullptr;
    float *ht = nullptr, *hb0 = nullptr, *hb1 = nullptr;
    uint8_t* occ = nullptr;
    uint32_t* counts = nullptr;             // [0] A, [1] B, [2] shadow
    unsigned long long* stats = nullptr;    // [0] closest rays, [1] shadow rays, [2] vertices
    std::vector<void*> pixel_allocs;
    // path guiding
    float* gfeat = nullptr;   // P x 9 network inputs of the shading points of a bounce
    float* gout = nullptr;    // P x 48 proposal parameters
    float* vrec = nullptr;    // P x vrec_depth x VREC vertex records
    uint32_t vrec_depth = 0;
    uint32_t guide_P = 0;     // paths per pass the guide's per-path buffers are sized for
    float* tin = nullptr;     // tcap x 9 training inputs (p01, n, wo)
    float* tdir = nullptr;    // tcap x 3
    float* tw = nullptr;      // tcap
    float train_keep = 1.0f;  // fraction of candidate samples kept in the last pass
    uint32_t tcap = 0;
    uint32_t* tcount = nullptr;
    std::vector<void*> guide_allocs;
    uint32_t last_train_samples = 0;
    uint32_t guided_passes = 0;
    float scene_lo[3] = { 0, 0, 0 }, scene_hi[3] = { 1, 1, 1 };
    // shading attributes of the shapes (pools over all shapes, global triangle order)
    AttrView attr{};
    std::vector<void*> attr_allocs;
    uint64_t attr_version = ~0ull;  // Context::attr_version the pools were built from
    uint32_t attr_tris = 0;         // Context::total_tris the pools were built for
    uint32_t nweights = 0;
    float* weights_dev = nullptr;   // == attr.weights
    uint32_t max_tex_ref = 0;       // largest texture number the materials name
    // textures, environment map
    DevTexture* texs = nullptr;
    float4* texels = nullptr;
    uint32_t ntex = 0;
    float4* env = nullptr;
    uint32_t env_w = 0, env_h = 0;
    // albedo / normal AOVs per path slot (allocated on first use)
    float* aov_albedo = nullptr;
    float* aov_normal = nullptr;
    uint32_t aov_P = 0;
};

namespace {

constexpr uint32_t ST_ENDED = 1u << 31;
constexpr uint32_t ST_NEE = 1u << 30;
constexpr uint32_t ST_DEPTH = 0xFFFFu;
constexpr int VREC = 20;  // p01 3, n 3, wo 3, dir 3, w 3, E 3, has_dir, pad
constexpr uint32_t TRAIN_MAX_STEPS = 1024;  // minibatches per pass

struct LsView
{
    LightSamplerView v;
};

struct ShadeArgs
{
    DevMaterial* mats;
    uint32_t* tri_mat;
    float4* tri_ng;
    LightSamplerView ls;
    bool has_ls;
    uint32_t* rng;
    float* T;
    float* L;
    uint32_t* state;
    float* nee;
    uint8_t* vis;
    Queue in, out, shadow;
    uint32_t* hprim;
    float* ht;
    float* hb0;
    float* hb1;
    uint32_t* counts;
    uint32_t max_depth, rr_depth;
    float bg[3];
    AttrView attr;
    const float4* env;       // environment map on a miss (nullptr: bg)
    uint32_t env_w, env_h;
    float* aov_albedo;       // depth-0 base colour and shading normal sums (nullptr: off)
    float* aov_normal;
    // guiding
    const float* gout;
    float guide_frac;
    float* vrec;
    uint32_t vrec_depth;
    uint32_t bounce;
    float lo[3], inv[3];
};

__device__ __forceinline__ f3 to_unit_box(const ShadeArgs& A, f3 p)
{
    return mk(fminf(fmaxf((p.x - A.lo[0]) * A.inv[0], 0.0f), 1.0f),
              fminf(fmaxf((p.y - A.lo[1]) * A.inv[1], 0.0f), 1.0f),
              fminf(fmaxf((p.z - A.lo[2]) * A.inv[2], 0.0f), 1.0f));
}

// Features of every shading point of the bounce (queue A), for the proposal network.
// Network inputs of every shading point of the bounce (queue A): position
// normalised to the scene bounds, shading normal, outgoing direction.
__global__ void k_guide_features(ShadeArgs A, float* q9)
{
    const uint32_t count = A.counts[0];
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x)
    {
        float* f = q9 + size_t(q) * 9;
        const uint32_t prim = A.hprim[q];
        if (prim == 0xFFFFFFFFu)
        {
            for (int i = 0; i < 9; ++i) f[i] = 0.0f;
            continue;
        }
        const f3 ro = mk(A.in.o[0][q], A.in.o[1][q], A.in.o[2][q]);
        const f3 rd = mk(A.in.d[0][q], A.in.d[1][q], A.in.d[2][q]);
        const float4 ngm = A.tri_ng[prim];
        f3 ns;
        f2 uv;
        shading_attributes(A.attr, prim, A.hb0[q], A.hb1[q], mk(ngm.x, ngm.y, ngm.z), A.mats[A.tri_mat[prim]],
                           ns, uv);
        const f3 p = to_unit_box(A, ro + rd * A.ht[q]);
        f[0] = p.x; f[1] = p.y; f[2] = p.z;
        f[3] = ns.x; f[4] = ns.y; f[5] = ns.z;
        f[6] = -rd.x; f[7] = -rd.y; f[8] = -rd.z;
    }
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Inclusive prefix sum over the wave's lanes.
__device__ __forceinline__ uint32_t wave_scan(uint32_t x)
{
#pragma unroll
    for (uint32_t d = 1; d < 64u; d <<= 1)
    {
        const uint32_t y = __shfl_up(x, d);
        if (lane_id() >= d) x += y;
    }
    return x;
}

// The shade loop's 256 entries of one pass of a block, ordered by a small key (0..SORT_KEYS-1,
// stable within a key): returns the entry (0..255) this thread takes.  With the key the
// branch the shading takes (a miss, a conductor, a Lambert or a coated dielectric), the lanes
// of a wave mostly run the same material code instead of every branch at part occupancy.
// The order changes no result: each path keeps its own state and sample sequence.
constexpr uint32_t SORT_KEYS = 5;
__device__ __forceinline__ uint32_t block_sort_keys(uint32_t key)
{
    __shared__ uint32_t s_kc[SORT_KEYS][4];
    __shared__ uint16_t s_src[256];
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t mine = 0;
    for (uint32_t k = 0; k < SORT_KEYS; ++k)
    {
        const uint64_t m = __ballot(key == k);
        if (key == k) mine = m;
        if (lane == 0u) s_kc[k][w] = uint32_t(__popcll(m));
    }
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t k = 0; k < SORT_KEYS; ++k)
        for (uint32_t v = 0; v < 4u; ++v)
            if (k < key || (k == key && v < w)) off += s_kc[k][v];
    s_src[off + uint32_t(__popcll(mine & below))] = uint16_t(threadIdx.x);
    __syncthreads();
    return s_src[threadIdx.x];
}

// Block-aggregated appends for k_shade: the shadow queue takes both NEE rays of
// a path (a then b), the continuation queue one ray.  Device-scope atomics on
// one address serialise across the chip (~11 ns each), so the whole 256-thread
// workgroup reserves its slots with one atomic per queue.  Every thread of the
// block must call it; `buf` alternates between calls so that a wave still
// reading one call's totals never sees the next call's.
struct ShadeSlots { uint32_t a, b, c; };
__device__ __forceinline__ ShadeSlots block_append3(uint32_t* shadow_count, uint32_t* cont_count, bool ea,
                                                    bool eb, bool ec, uint32_t buf)
{
    __shared__ uint32_t s_cnt[2][3][4];
    __shared__ uint32_t s_base[2][2];
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    const uint64_t ma = __ballot(ea), mb = __ballot(eb), mc = __ballot(ec);
    if (lane == 0u)
    {
        s_cnt[buf][0][w] = uint32_t(__popcll(ma));
        s_cnt[buf][1][w] = uint32_t(__popcll(mb));
        s_cnt[buf][2][w] = uint32_t(__popcll(mc));
    }
    __syncthreads();
    const uint32_t nw = blockDim.x >> 6;
    uint32_t ta = 0, tb = 0, tc = 0, pa = 0, pb = 0, pc = 0;
    for (uint32_t k = 0; k < nw; ++k)
    {
        const uint32_t ca = s_cnt[buf][0][k], cb = s_cnt[buf][1][k], cc = s_cnt[buf][2][k];
        if (k < w) { pa += ca; pb += cb; pc += cc; }
        ta += ca; tb += cb; tc += cc;
    }
    if (threadIdx.x == 0u)
    {
        s_base[buf][0] = ta + tb ? atomicAdd(shadow_count, ta + tb) : 0u;
        s_base[buf][1] = tc ? atomicAdd(cont_count, tc) : 0u;
    }
    __syncthreads();
    const uint32_t sb = s_base[buf][0], cb = s_base[buf][1];
    return { sb + pa + uint32_t(__popcll(ma & below)), sb + ta + pb + uint32_t(__popcll(mb & below)),
             cb + pc + uint32_t(__popcll(mc & below)) };
}

__device__ __forceinline__ void store3(float* a, uint32_t i, f3 v)
{
    a[3 * size_t(i)] = v.x;
    a[3 * size_t(i) + 1] = v.y;
    a[3 * size_t(i) + 2] = v.z;
}
__device__ __forceinline__ f3 load3(const float* a, uint32_t i)
{
    return mk(a[3 * size_t(i)], a[3 * size_t(i) + 1], a[3 * size_t(i) + 2]);
}

// Sampler seeds: slot 0 of pixel i is seeded with i + seed as the reference
// seeds its per-pixel sampler (alpine.cpp:259-265); further slots of the
// pixel get hashed seeds of their own.
__device__ __forceinline__ uint32_t slot_seed(uint32_t pixel, uint32_t slot, uint32_t seed)
{
    uint32_t h = pixel + seed;
    if (slot == 0u) return h;
    h ^= slot * 0x9E3779B9u;
    h ^= h >> 16; h *= 0x7feb352du;
    h ^= h >> 15; h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

__global__ void k_reset_pixels(uint32_t P, uint32_t NPIX, uint32_t seed, uint32_t* rng, float* accum,
                               double* accum64)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    uint32_t s[4];
    sampler_reset(s, slot_seed(p % NPIX, p / NPIX, seed));
    for (int k = 0; k < 4; ++k) rng[4 * size_t(p) + k] = s[k];
    for (int k = 0; k < 3; ++k)
    {
        accum[3 * size_t(p) + k] = 0.0f;
        accum64[3 * size_t(p) + k] = 0.0;
    }
}

// As k_reset_pixels with the seed read from device memory (a frame graph's per-frame seed).
__global__ void k_reset_pixels_dev(uint32_t P, uint32_t NPIX, const uint32_t* seed, uint32_t* rng, float* accum,
                                   double* accum64)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    uint32_t s[4];
    sampler_reset(s, slot_seed(p % NPIX, p / NPIX, *seed));
    for (int k = 0; k < 4; ++k) rng[4 * size_t(p) + k] = s[k];
    for (int k = 0; k < 3; ++k)
    {
        accum[3 * size_t(p) + k] = 0.0f;
        accum64[3 * size_t(p) + k] = 0.0;
    }
}

__global__ void k_zero_words(uint32_t* p, size_t n)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0u;
}

// Per-pixel sums over the slots, in slot order.
template <typename T>
__global__ void k_sum_slots(const T* acc, uint32_t NPIX, uint32_t S, T* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * NPIX) return;
    T v = acc[i];
    for (uint32_t s = 1; s < S; ++s) v += acc[size_t(s) * 3 * NPIX + i];
    out[i] = v;
}

__global__ void k_set_counts(uint32_t* counts, uint32_t a)
{
    if (threadIdx.x == 0)
    {
        counts[0] = a;
        counts[1] = 0;
        counts[2] = 0;
    }
}

// Camera rays with the per-pixel jitter (alpine.cpp:305-312).
__global__ void k_generate(DevCamera cam, Tiles tl, uint32_t W, uint32_t H, uint32_t P, uint32_t* rng, float* T,
                           float* L, uint32_t* state, Queue q, unsigned long long* stats)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    uint32_t s[4];
    for (int k = 0; k < 4; ++k) s[k] = rng[4 * size_t(p) + k];
    const f2 j = get2d(s);
    const uint32_t pix = p % (W * H);
    const uint32_t x = pix % W, y = pix / W;
    f3 o, d;
    uint32_t root = 0;
    if (tl.cams)
    {
        const uint32_t t = (y / tl.th) * tl.tiles_x + x / tl.tw;
        camera_ray(tl.cams[t], (float(x % tl.tw) + j.x) / float(tl.tw), (float(y % tl.th) + j.y) / float(tl.th),
                   o, d);
        root = tl.roots ? tl.roots[t] : 0u;
    }
    else
        camera_ray(cam, (float(x) + j.x) / float(W), (float(y) + j.y) / float(H), o, d);
    for (int k = 0; k < 4; ++k) rng[4 * size_t(p) + k] = s[k];
    store3(T, p, mk(1.0f, 1.0f, 1.0f));
    store3(L, p, mk(0.0f, 0.0f, 0.0f));
    state[p] = 0;
    q.o[0][p] = o.x; q.o[1][p] = o.y; q.o[2][p] = o.z;
    q.d[0][p] = d.x; q.d[1][p] = d.y; q.d[2][p] = d.z;
    q.path[p] = p;
    q.root[p] = root;
    if (p == 0) atomicAdd(&stats[0], (unsigned long long)P);
}

// Plain shading is held to 128 VGPRs (4 waves per SIMD, a few spills) to hide the light-BVH
// descent's load latency; the guided form (vMF mixture sampling and density) to 168 VGPRs, 3
// waves per SIMD with about 100 B of per-lane spills, which measured 2 % faster than its
// unconstrained 210 VGPRs at 2 waves and than 128 VGPRs at 4 with 244 B of spills.
template <bool GUIDED>
__global__ __launch_bounds__(256, GUIDED ? 3 : 4) void k_shade(ShadeArgs A)
{
    const uint32_t count = A.counts[0];
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t buf = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride)
    {
        // this pass's entries ordered by the material branch they take (entries past the end last)
        uint32_t key = SORT_KEYS - 1u;
        if (base + threadIdx.x < count)
        {
            const uint32_t pr = A.hprim[base + threadIdx.x];
            if (pr == 0xFFFFFFFFu) key = 0u;
            else
            {
                const DevMaterial& mm = A.mats[A.tri_mat[pr]];
                key = mm.kind == MAT_CONDUCTOR ? 1u : (mm.nbsdf > 1u ? 3u : 2u);
            }
        }
        const uint32_t q = base + block_sort_keys(key);
        const bool valid = q < count;
        bool emit_a = false, emit_b = false, emit_c = false;
        f3 sa_o{}, sa_d{}, sb_d{}, c_o{}, c_d{};
        float sa_t = 0.0f, sb_t = 0.0f;
        uint32_t p = 0, rt = 0;
        if (valid)
        {
            p = A.in.path[q];
            rt = A.in.root[q];
            uint32_t s[4];
            for (int k = 0; k < 4; ++k) s[k] = A.rng[4 * size_t(p) + k];
            f3 T = load3(A.T, p), L = load3(A.L, p);
            uint32_t st = A.state[p] & ST_DEPTH;
            const uint32_t depth = st;
            const uint32_t prim = A.hprim[q];
            const f3 ro = mk(A.in.o[0][q], A.in.o[1][q], A.in.o[2][q]);
            const f3 rd = mk(A.in.d[0][q], A.in.d[1][q], A.in.d[2][q]);
            A.vis[2 * size_t(p)] = 0;
            A.vis[2 * size_t(p) + 1] = 0;
            bool ended = false;
            if (prim == 0xFFFFFFFFu)
            {
                // background colour or environment map (alpine.cpp:318-324)
                const f3 bg = A.env ? environment_sample(A.env, A.env_w, A.env_h, rd) : mk(A.bg[0], A.bg[1], A.bg[2]);
                L = L + T * bg;
                ended = true;
                if (GUIDED && A.bounce < A.vrec_depth)
                {
                    float* vr = A.vrec + (size_t(p) * A.vrec_depth + A.bounce) * VREC;
                    vr[15] = bg.x; vr[16] = bg.y; vr[17] = bg.z;
                    vr[18] = 0.0f;
                }
            }
            else
            {
                const f3 hit = ro + rd * A.ht[q];
                const float4 ngm = A.tri_ng[prim];
                const f3 ng = mk(ngm.x, ngm.y, ngm.z);
                DevMaterial m = A.mats[A.tri_mat[prim]];
                f3 ns;
                f2 uv;
                shading_attributes(A.attr, prim, A.hb0[q], A.hb1[q], ng, m, ns, uv);
                if (m.tex_bc)
                {
                    const f3 bc = material_base_color(A.attr, m, uv);
                    m.bc[0] = bc.x; m.bc[1] = bc.y; m.bc[2] = bc.z;
                }
                if (A.aov_albedo && depth == 0u)
                {
                    // albedo and normal AOVs of the first hit (alpine.cpp:333-337)
                    store3(A.aov_albedo, p, load3(A.aov_albedo, p) + ld3(m.bc));
                    store3(A.aov_normal, p, load3(A.aov_normal, p) + ns);
                }
                f3 ss, ts;
                get_basis(ns, ss, ts);
                const f3 wo = to_local(-rd, ss, ts, ns);
                // next-event estimation (alpine.cpp:370-432)
                if (A.has_ls)
                {
                    float lsel = 0.0f;
                    const int32_t li = light_sampler_sample(A.ls, get1d(s), hit, ns, lsel);
                    if (li >= 0)
                    {
                        const DevLight& Lt = A.ls.lights[li];
                        const f3 off = ns * kRayOffset;
                        f3 ca = mk(0, 0, 0), cb = mk(0, 0, 0);
                        f3 le, wiw;
                        float dist, pdf;
                        if (light_sample(Lt, get2d(s), hit, le, wiw, dist, pdf))
                        {
                            const f3 wi = to_local(wiw, ss, ts, ns);
                            f3 f;
                            float bpdf;
                            mat_eval(m, wo, wi, f, bpdf);
                            if (bpdf > 0.0f)
                            {
                                const float c = max_s(0.0f, dot(wiw, ns));
                                const float w = !light_is_delta(Lt) ? power_heuristic(pdf, bpdf) : 1.0f;
                                ca = le * w * f * c / pdf;
                                emit_a = true;
                                sa_o = hit + off;
                                sa_d = wiw;
                                sa_t = dist;
                            }
                        }
                        if (!light_is_delta(Lt))
                        {
                            f3 est, wi;
                            float mpdf;
                            if (mat_sample(m, wo, get2d(s), est, wi, mpdf))
                            {
                                const f3 ld = to_world(wi, ss, ts, ns);
                                float d2;
                                const float lp = light_pdf(Lt, hit, ld, d2);
                                if (lp > 0.0f)
                                {
                                    const float w = power_heuristic(mpdf, lp);
                                    cb = ld3(Lt.le) * w * est;
                                    emit_b = true;
                                    sb_d = ld;
                                    sb_t = d2;
                                }
                            }
                        }
                        sa_o = hit + off;
                        float* ne = A.nee + 10 * size_t(p);
                        ne[0] = T.x; ne[1] = T.y; ne[2] = T.z;
                        ne[3] = ca.x; ne[4] = ca.y; ne[5] = ca.z;
                        ne[6] = cb.x; ne[7] = cb.y; ne[8] = cb.z;
                        ne[9] = lsel;
                        st |= ST_NEE;
                    }
                }
                // continuation (alpine.cpp:342-361)
                f3 est, wi, wiw;
                bool ok;
                float* vr = nullptr;
                if (GUIDED)
                {
                    // one-sample MIS of BSDF sampling and the learned proposal (true densities)
                    if (A.bounce < A.vrec_depth) vr = A.vrec + (size_t(p) * A.vrec_depth + A.bounce) * VREC;
                    const VmfMix mix = guide_decode(A.gout + size_t(q) * GUIDE_OUT);
                    const float c = A.guide_frac;
                    if (get1d(s) < c)
                    {
                        const f2 uu = get2d(s);
                        wiw = guide_sample(mix, uu.x, uu.y, get1d(s));
                        wi = to_local(wiw, ss, ts, ns);
                        ok = true;
                    }
                    else
                    {
                        float mpdf;
                        ok = mat_sample(m, wo, get2d(s), est, wi, mpdf);
                        if (ok) wiw = to_world(wi, ss, ts, ns);
                    }
                    if (ok)
                    {
                        f3 f;
                        float dummy;
                        mat_eval(m, wo, wi, f, dummy);
                        const float pm = (1.0f - c) * mat_true_pdf(m, wo, wi) + c * guide_pdf(mix, wiw);
                        ok = wi.z > 0.0f && pm > 0.0f;
                        if (ok) est = f * wi.z / pm;
                    }
                    if (vr)
                    {
                        const f3 p01 = to_unit_box(A, hit);
                        vr[0] = p01.x; vr[1] = p01.y; vr[2] = p01.z;
                        vr[3] = ns.x; vr[4] = ns.y; vr[5] = ns.z;
                        vr[6] = -rd.x; vr[7] = -rd.y; vr[8] = -rd.z;
                        vr[15] = 0.0f; vr[16] = 0.0f; vr[17] = 0.0f;
                        vr[18] = 0.0f;
                    }
                }
                else
                {
                    float mpdf;
                    ok = mat_sample(m, wo, get2d(s), est, wi, mpdf);
                    if (ok) wiw = to_world(wi, ss, ts, ns);
                }
                if (!ok) ended = true;
                else
                {
                    if (!(dot(wiw, ng) > 0.0f)) ended = true;
                    else
                    {
                        T = T * est;
                        if (A.rr_depth && depth + 1 >= A.rr_depth)
                        {
                            const float qs = min_s(maxc(T), 0.95f);
                            if (!(get1d(s) < qs)) ended = true;
                            else T = T / qs;
                        }
                        if (!ended && depth + 1 < A.max_depth)
                        {
                            emit_c = true;
                            c_o = hit + ng * kRayOffset;
                            c_d = wiw;
                            st = (st & ~ST_DEPTH) | (depth + 1);
                            if (GUIDED && vr)
                            {
                                vr[9] = wiw.x; vr[10] = wiw.y; vr[11] = wiw.z;
                                vr[12] = est.x; vr[13] = est.y; vr[14] = est.z;
                                vr[18] = 1.0f;
                            }
                        }
                        else
                            ended = true;
                    }
                }
            }
            for (int k = 0; k < 4; ++k) A.rng[4 * size_t(p) + k] = s[k];
            store3(A.T, p, T);
            store3(A.L, p, L);
            A.state[p] = st | (ended ? ST_ENDED : 0u);
        }
        const ShadeSlots sl = block_append3(&A.counts[2], &A.counts[1], emit_a, emit_b, emit_c, buf);
        buf ^= 1u;
        const uint32_t ia = sl.a, ib = sl.b, ic = sl.c;
        if (emit_a)
        {
            A.shadow.o[0][ia] = sa_o.x; A.shadow.o[1][ia] = sa_o.y; A.shadow.o[2][ia] = sa_o.z;
            A.shadow.d[0][ia] = sa_d.x; A.shadow.d[1][ia] = sa_d.y; A.shadow.d[2][ia] = sa_d.z;
            A.shadow.tmax[ia] = sa_t;
            A.shadow.path[ia] = 2u * p;
            A.shadow.root[ia] = rt;
        }
        if (emit_b)
        {
            A.shadow.o[0][ib] = sa_o.x; A.shadow.o[1][ib] = sa_o.y; A.shadow.o[2][ib] = sa_o.z;
            A.shadow.d[0][ib] = sb_d.x; A.shadow.d[1][ib] = sb_d.y; A.shadow.d[2][ib] = sb_d.z;
            A.shadow.tmax[ib] = sb_t;
            A.shadow.path[ib] = 2u * p + 1u;
            A.shadow.root[ib] = rt;
        }
        if (emit_c)
        {
            A.out.o[0][ic] = c_o.x; A.out.o[1][ic] = c_o.y; A.out.o[2][ic] = c_o.z;
            A.out.d[0][ic] = c_d.x; A.out.d[1][ic] = c_d.y; A.out.d[2][ic] = c_d.z;
            A.out.path[ic] = p;
            A.out.root[ic] = rt;
        }
    }
}

__global__ void k_shadow_vis(const uint32_t* counts, const uint32_t* spath, const uint8_t* occ, uint8_t* vis)
{
    const uint32_t count = counts[2];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        if (!occ[i]) vis[spath[i]] = 1;
}

struct TrainOut
{
    float* vrec;
    uint32_t vrec_depth;
    uint32_t bounce;
    float* tin;
    float* tdir;
    float* tw;
    uint32_t tcap;
    uint32_t* tcount;
    float keep;      // probability of keeping a training sample (spreads tcap over the frame)
    uint32_t seed;   // per pass
};

__device__ __forceinline__ float hash01(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ c * 0xC2B2AE3Du;
    h ^= h >> 16; h *= 0x7feb352du;
    h ^= h >> 15; h *= 0x846ca68bu;
    h ^= h >> 16;
    return float(h >> 8) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float luminance(f3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

__global__ void k_resolve(const uint32_t* counts, const uint32_t* qpath, uint32_t* state, const float* nee,
                          const uint8_t* vis, float* Lbuf, float* accum, double* accum64, int fp64, TrainOut G)
{
    const uint32_t count = counts[0];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride)
    {
        const uint32_t q = base + threadIdx.x;
        const bool valid = q < count;
        const uint32_t p = valid ? qpath[q] : 0u;
        uint32_t st = valid ? state[p] : 0u;
        f3 L = valid ? load3(Lbuf, p) : mk(0.0f, 0.0f, 0.0f);
        if (valid && (st & ST_NEE))
        {
            const float* ne = nee + 10 * size_t(p);
            f3 r = mk(0, 0, 0);
            if (vis[2 * size_t(p)]) r = r + mk(ne[3], ne[4], ne[5]);
            if (vis[2 * size_t(p) + 1]) r = r + mk(ne[6], ne[7], ne[8]);
            const f3 e = r / ne[9];
            L = L + mk(ne[0], ne[1], ne[2]) * e;
            st &= ~ST_NEE;
            store3(Lbuf, p, L);
            if (G.vrec && G.bounce < G.vrec_depth)
            {
                float* vr = G.vrec + (size_t(p) * G.vrec_depth + G.bounce) * VREC;
                vr[15] = e.x; vr[16] = e.y; vr[17] = e.z;
            }
        }
        if (G.vrec)
        {
            // radiance leaving each vertex towards the previous one, back to front; each
            // traced continuation becomes a training sample weighted by F / p = w (x) L_in.
            // Pass 0 counts a path's samples, one atomic reserves the wave's, pass 1 writes.
            const bool ended = valid && (st & ST_ENDED);
            const uint32_t last = G.bounce < G.vrec_depth ? G.bounce : G.vrec_depth - 1;
            uint32_t t = 0;
            for (int pass = 0; pass < 2; ++pass)
            {
                uint32_t n = 0;
                f3 lnext = mk(0.0f, 0.0f, 0.0f);
                for (int j = ended ? int(last) : -1; j >= 0; --j)
                {
                    const float* vr = G.vrec + (size_t(p) * G.vrec_depth + uint32_t(j)) * VREC;
                    const f3 e = mk(vr[15], vr[16], vr[17]);
                    f3 lo = e;
                    if (vr[18] > 0.5f)
                    {
                        const f3 w = mk(vr[12], vr[13], vr[14]);
                        const f3 contrib = w * lnext;
                        lo = e + contrib;
                        const float wt = luminance(contrib);
                        if (wt > 0.0f && (G.keep >= 1.0f || hash01(p, uint32_t(j), G.seed) < G.keep))
                        {
                            if (pass == 1 && t < G.tcap)
                            {
                                for (int k = 0; k < 9; ++k) G.tin[9 * size_t(t) + k] = vr[k];
                                for (int k = 0; k < 3; ++k) G.tdir[3 * size_t(t) + k] = vr[9 + k];
                                G.tw[t] = wt;
                            }
                            ++t;
                            ++n;
                        }
                    }
                    lnext = lo;
                }
                if (pass == 0)
                {
                    const uint32_t incl = wave_scan(n);
                    const uint32_t total = uint32_t(__builtin_amdgcn_readlane(int(incl), 63));
                    uint32_t wbase = 0;
                    if (lane_id() == 0u && total) wbase = atomicAdd(G.tcount, total);
                    wbase = uint32_t(__builtin_amdgcn_readlane(int(wbase), 0));
                    t = wbase + incl - n;
                }
            }
        }
        if (valid && (st & ST_ENDED))
        {
            accum[3 * size_t(p)] += L.x;
            accum[3 * size_t(p) + 1] += L.y;
            accum[3 * size_t(p) + 2] += L.z;
            if (fp64)
            {
                accum64[3 * size_t(p)] += double(L.x);
                accum64[3 * size_t(p) + 1] += double(L.y);
                accum64[3 * size_t(p) + 2] += double(L.z);
            }
        }
        if (valid) state[p] = st;
    }
}

__global__ void k_advance(uint32_t* counts, unsigned long long* stats)
{
    if (threadIdx.x == 0)
    {
        stats[1] += counts[2];
        stats[2] += counts[0];
        stats[0] += counts[1];
        counts[0] = counts[1];
        counts[1] = 0;
        counts[2] = 0;
    }
}

// Alpine::postProcess without denoising (alpine.cpp:446-483): mean colour, optional ACES
// fit (:434-442), clamp, sRGB encode (:461-465), round to bytes.
__global__ void k_postprocess(const float* accum, uint32_t n3, float spp, uint32_t tonemap,
                              uint8_t* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    float c = accum[i] / spp;
    if (tonemap)
        c = (c * (c * 2.51f + 0.03f)) / (c * (c * 2.43f + 0.59f) + 0.14f);
    c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
    c = c <= 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
    out[i] = uint8_t(c * 255.0f + 0.5f);
}

__global__ void k_zero_f32(float* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_tri_ng(const float* tri9, const uint32_t* ids, uint32_t n, float4* ng)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* t = tri9 + 9 * size_t(i);
    // Primitive::updateVertices: ng = normalize(cross(e0, e1)) (bvh_common.cpp:116)
    const f3 g = normalize(cross(mk(t[3], t[4], t[5]), mk(t[6], t[7], t[8])));
    ng[ids[i]] = make_float4(g.x, g.y, g.z, 0.0f);
}

__global__ void k_ls_sample(LightSamplerView S, const float* u, const float* hit, const float* ns, uint32_t n,
                            int32_t* light, float* pdf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float p = 0.0f;
    light[i] = light_sampler_sample(S, u[i], load3(hit, i), load3(ns, i), p);
    pdf[i] = light[i] >= 0 ? p : 0.0f;
}

__global__ void k_ls_pdf(LightSamplerView S, const int32_t* light, const float* hit, const float* ns, uint32_t n,
                         float* pdf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pdf[i] = light_sampler_pdf(S, light[i], load3(hit, i), load3(ns, i));
}

template <typename T>
int dmalloc(Context* ctx, T*& p, size_t count)
{
    p = nullptr;
    if (ctx->fail_allocs_after == 0)
        return ctx->fail(APG_E_HIP, "hipMalloc: out of memory (forced by apg_debug_fail_allocations)");
    if (ctx->fail_allocs_after > 0) ctx->fail_allocs_after--;
    APG_HIP_CHECK(ctx, hipMalloc(&p, (count ? count : 1) * sizeof(T)));
    return APG_OK;
}

template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_pixels(RenderState* R)
{
    for (void* p : R->pixel_allocs) (void)hipFree(p);
    R->pixel_allocs.clear();
    R->P = R->W = R->H = R->NPIX = 0;
    R->S = 1;
}

RenderState* state_of(Context* ctx)
{
    if (!ctx->render) ctx->render = new RenderState;
    return ctx->render;
}

LightSamplerView view_of(const RenderState* R)
{
    LightSamplerView v;
    v.kind = R->ls_kind;
    v.n = R->nlights;
    v.lights = R->lights;
    v.alias_p = R->alias_p;
    v.alias_q = R->alias_q;
    v.alias_a = R->alias_a;
    v.pairs = R->lpairs;
    v.parent_slot = R->lparent;
    v.leaf_of_light = R->leaf_of;
    v.root_light = R->root_light;
    return v;
}

inline uint32_t nblocks(uint32_t n, uint32_t b) { return n ? (n + b - 1) / b : 1; }

} // namespace

void render_state_free(RenderState* R)
{
    if (!R) return;
    // graph replays still queued read every buffer below and the exec itself
    if (R->gstream) (void)hipStreamSynchronize(R->gstream);
    dfree(R->mats);
    dfree(R->tri_mat);
    dfree(R->tri_ng);
    dfree(R->lights);
    dfree(R->alias_p);
    dfree(R->alias_q);
    dfree(R->alias_a);
    dfree(R->lpairs);
    dfree(R->lparent);
    dfree(R->leaf_of);
    dfree(R->tile_cams);
    dfree(R->tile_roots);
    for (void* p : R->attr_allocs) (void)hipFree(p);
    dfree(R->texs);
    dfree(R->texels);
    dfree(R->env);
    dfree(R->aov_albedo);
    dfree(R->aov_normal);
    if (R->gexec) (void)hipGraphExecDestroy(R->gexec);
    if (R->gev) (void)hipEventDestroy(R->gev);
    if (R->gstream) (void)hipStreamDestroy(R->gstream);
    free_pixels(R);
    for (void* p : R->guide_allocs) (void)hipFree(p);
    delete R;
}

int Context::rebuild_shading()
{
    RenderState* R = state_of(this);
    const uint32_t n = scene.tri_count;
    if (R->tri_ng_n != total_tris)
    {
        dfree(R->tri_ng);
        int rc = dmalloc(this, R->tri_ng, total_tris);
        if (rc) return rc;
        R->tri_ng_n = total_tris;
    }
    if (n) k_tri_ng<<<nblocks(n, 256), 256>>>(scene.tris, scene.tri_ids, n, R->tri_ng);
    for (int a = 0; a < 3; ++a)
    {
        const bool ok = scene_lo[a] <= scene_hi[a];
        R->scene_lo[a] = ok ? scene_lo[a] : 0.0f;
        R->scene_hi[a] = ok ? scene_hi[a] : 1.0f;
    }
    APG_HIP_CHECK(this, hipGetLastError());
    APG_HIP_CHECK(this, hipDeviceSynchronize());
    return APG_OK;
}

// Frame graphs (frame.hip): apg_render_reset's clears with the seed read from device memory,
// and the per-triangle normals of a scene build, as kernels on `s` only (capturable).
int render_reset_enqueue(Context* ctx, const uint32_t* seed_dev, hipStream_t s)
{
    RenderState* R = state_of(ctx);
    if (!R->P) return ctx->fail(APG_E_STATE, "apg_render_reset has not been called");
    k_reset_pixels_dev<<<nblocks(R->P, 256), 256, 0, s>>>(R->P, R->NPIX, seed_dev, R->rng, R->accum, R->accum64);
    k_zero_words<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t*>(R->stats), 16);
    if (R->aov_albedo)
    {
        const size_t words = 3 * size_t(R->aov_P);
        k_zero_words<<<256, 256, 0, s>>>(reinterpret_cast<uint32_t*>(R->aov_albedo), words);
        k_zero_words<<<256, 256, 0, s>>>(reinterpret_cast<uint32_t*>(R->aov_normal), words);
    }
    APG_HIP_CHECK(ctx, hipGetLastError());
    return APG_OK;
}

int render_tri_ng_enqueue(Context* ctx, hipStream_t s)
{
    RenderState* R = state_of(ctx);
    const uint32_t n = ctx->scene.tri_count;
    if (R->tri_ng_n != ctx->total_tris) return ctx->fail(APG_E_STATE, "no triangle normals for this scene");
    if (n) k_tri_ng<<<nblocks(n, 256), 256, 0, s>>>(ctx->scene.tris, ctx->scene.tri_ids, n, R->tri_ng);
    APG_HIP_CHECK(ctx, hipGetLastError());
    return APG_OK;
}

uint32_t render_frame_spp(Context* ctx) { return state_of(ctx)->total_spp; }
void render_set_total_spp(Context* ctx, uint32_t spp) { state_of(ctx)->total_spp = spp; }
void render_size(Context* ctx, uint32_t& W, uint32_t& H, uint32_t& S)
{
    RenderState* R = state_of(ctx);
    W = R->W;
    H = R->H;
    S = R->S;
}

namespace {

// Pools of the shapes' shading attributes in global triangle order (rebuilt when a shape's
// attributes or targets change, or shapes are appended); morph weights are refreshed alone.
int ensure_attributes(Context* ctx, RenderState* R)
{
    if (R->attr_version == ctx->attr_version && R->attr_tris == ctx->total_tris)
    {
        if (ctx->weights_dirty && R->nweights)
        {
            std::vector<float> w;
            gather_weights(ctx->shapes, w);
            APG_HIP_CHECK(ctx, hipMemcpy(R->weights_dev, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
        }
        ctx->weights_dirty = false;
        return APG_OK;
    }
    if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
    for (void* p : R->attr_allocs) (void)hipFree(p);
    R->attr_allocs.clear();
    const DevTexture* texs = R->attr.texs;
    const float4* texels = R->attr.texels;
    R->attr = AttrView{};
    R->attr.texs = texs;
    R->attr.texels = texels;
    R->weights_dev = nullptr;
    R->nweights = 0;
    AttrPools P;
    gather_attributes(ctx->shapes, ctx->total_tris, P);
    if (P.any)
    {
        auto up = [&](auto*& dst, const auto& src) -> int {
            using T = typename std::decay_t<decltype(src)>::value_type;
            T* p = nullptr;
            int rc = dmalloc(ctx, p, src.size());
            if (rc) return rc;
            R->attr_allocs.push_back(p);
            if (!src.empty())
                APG_HIP_CHECK(ctx, hipMemcpy(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
            dst = p;
            return APG_OK;
        };
        const DevTriAttr* dta = nullptr;
        const float4 *dn = nullptr, *dt = nullptr, *dtn = nullptr, *dtt = nullptr;
        const float* duv = nullptr;
        const DevMorph* dm = nullptr;
        float* dw = nullptr;
        int rc;
        if ((rc = up(dta, P.tri)) || (rc = up(dn, P.normals)) || (P.any_tangents && (rc = up(dt, P.tangents))) ||
            (rc = up(duv, P.uvs)) || (rc = up(dm, P.morphs)) || (rc = up(dw, P.weights)) ||
            (rc = up(dtn, P.tnormals)) || (rc = up(dtt, P.ttangents)))
            return rc;
        R->attr.tri = dta;
        R->attr.normals = dn;
        R->attr.tangents = dt;
        R->attr.uvs = duv;
        R->attr.morphs = dm;
        R->attr.weights = dw;
        R->attr.tnormals = dtn;
        R->attr.ttangents = dtt;
        R->weights_dev = dw;
        R->nweights = uint32_t(P.weights.size());
    }
    R->attr_version = ctx->attr_version;
    R->attr_tris = ctx->total_tris;
    ctx->weights_dirty = false;
    return APG_OK;
}

} // namespace

// The lights and the sampler over them, shared by apg_lights_set and apg_light_sampler_set.
int install_lights(Context* ctx, const std::vector<DevLight>& dl, uint32_t sampler);

} // namespace apg

using namespace apg;

extern "C" {

int apg_materials_set(apg_context* ctx, const apg_material* mats, uint32_t nmats,
                      const uint32_t* tri_material, uint32_t ntris)
{
    if (!ctx || (nmats && !mats) || (ntris && !tri_material)) return APG_E_INVALID;
    if (nmats == 0) return ctx->fail(APG_E_INVALID, "at least one material is required");
    for (uint32_t i = 0; i < ntris; ++i)
        if (tri_material[i] >= nmats) return ctx->fail(APG_E_INVALID, "triangle material index out of range");
    for (uint32_t i = 0; i < nmats; ++i)
        if (mats[i].kind > APG_MAT_CONDUCTOR) return ctx->fail(APG_E_INVALID, "unknown material kind");
    APG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    RenderState* R = state_of(ctx);
    std::vector<DevMaterial> dm(nmats);
    uint32_t max_tex = 0;
    for (uint32_t i = 0; i < nmats; ++i)
    {
        dm[i] = make_material(mats[i].kind, mats[i].base_color, mats[i].alpha);
        dm[i].tex_bc = mats[i].base_color_tex;
        dm[i].tex_n = mats[i].normal_tex;
        max_tex = std::max(max_tex, std::max(mats[i].base_color_tex, mats[i].normal_tex));
    }
    R->max_tex_ref = max_tex;
    dfree(R->mats);
    dfree(R->tri_mat);
    int rc;
    if ((rc = dmalloc(ctx, R->mats, nmats))) return rc;
    if ((rc = dmalloc(ctx, R->tri_mat, ntris))) return rc;
    APG_HIP_CHECK(ctx, hipMemcpy(R->mats, dm.data(), nmats * sizeof(DevMaterial), hipMemcpyHostToDevice));
    if (ntris)
        APG_HIP_CHECK(ctx, hipMemcpy(R->tri_mat, tri_material, ntris * sizeof(uint32_t), hipMemcpyHostToDevice));
    R->nmats = nmats;
    R->tri_mat_n = ntris;
    return APG_OK;
}

int apg_lights_set(apg_context* ctx, const apg_light* lights, uint32_t n, uint32_t sampler)
{
    if (!ctx || (n && !lights)) return APG_E_INVALID;
    if (sampler > APG_LS_BVH) return ctx->fail(APG_E_INVALID, "unknown light sampler");
    for (uint32_t i = 0; i < n; ++i)
        if (lights[i].kind > APG_LIGHT_POINT_INTENSITY) return ctx->fail(APG_E_INVALID, "unknown light kind");
    APG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    std::vector<DevLight> dl(n);
    for (uint32_t i = 0; i < n; ++i) dl[i] = prepare_light(lights[i]);
    return install_lights(ctx, dl, sampler);
}

int apg_light_sampler_set(apg_context* ctx, const float* power, const float* bounds, uint32_t n,
                          uint32_t sampler)
{
    if (!ctx || (n && (!power || !bounds))) return APG_E_INVALID;
    if (sampler > APG_LS_BVH) return ctx->fail(APG_E_INVALID, "unknown light sampler");
    APG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    // the samplers read a light's power and bounds only (uniform_light_sampler.cpp,
    // power_light_sampler.cpp:6-20, bvh_light_sampler.cpp:167-240)
    std::vector<DevLight> dl(n);
    for (uint32_t i = 0; i < n; ++i)
    {
        DevLight L{};
        L.kind = LIGHT_POINT;
        for (int a = 0; a < 3; ++a)
        {
            L.power[a] = power[3 * size_t(i) + a];
            L.lo[a] = bounds[6 * size_t(i) + a];
            L.hi[a] = bounds[6 * size_t(i) + 3 + a];
        }
        dl[i] = L;
    }
    return install_lights(ctx, dl, sampler);
}

int apg_light_sampler_sample_host(apg_context* ctx, const float* u, const float* hit, const float* ns,
                                  uint32_t n, int32_t* light, float* pdf)
{
    if (!ctx || (n && (!u || !hit || !ns || !light || !pdf))) return APG_E_INVALID;
    RenderState* R = state_of(ctx);
    if (!R->ls_set) return ctx->fail(APG_E_STATE, "no lights set");
    if (n == 0) return APG_OK;
    int rc;
    HostStaging* s = ctx->take_staging(n, rc);
    if (!s) return rc;
    const size_t N = s->cap;
    float* h = s->h_in;
    for (uint32_t i = 0; i < n; ++i)
    {
        h[i] = u[i];
        for (int a = 0; a < 3; ++a)
        {
            h[N + 3 * size_t(i) + a] = hit[3 * size_t(i) + a];
            h[4 * N + 3 * size_t(i) + a] = ns[3 * size_t(i) + a];
        }
    }
    // the kernel reads the queries from and writes the samples to the pinned staging buffers
    float* d = s->dev_in;
    k_ls_sample<<<nblocks(n, 256), 256, 0, s->stream>>>(view_of(R), d, d + N, d + 4 * N, n,
                                                         reinterpret_cast<int32_t*>(s->dev_out), s->dev_out + N);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess)
    {
        (void)hipStreamSynchronize(s->stream);
        ctx->give_staging(s);
        return ctx->fail(APG_E_HIP, std::string("staged light sampling: ") + hipGetErrorString(e));
    }
    std::memcpy(light, s->h_out, n * sizeof(int32_t));
    std::memcpy(pdf, s->h_out + N, n * sizeof(float));
    ctx->give_staging(s);
    return APG_OK;
}

} // extern "C"

namespace apg {

int install_lights(Context* ctx, const std::vector<DevLight>& dl, uint32_t sampler)
{
    RenderState* R = state_of(ctx);
    const uint32_t n = uint32_t(dl.size());
    if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
    dfree(R->lights);
    dfree(R->alias_p);
    dfree(R->alias_q);
    dfree(R->alias_a);
    dfree(R->lpairs);
    dfree(R->lparent);
    dfree(R->leaf_of);
    R->root_light = -1;
    // no light sampler until every buffer below is in place (a failed allocation leaves none)
    R->ls_set = false;
    R->nlights = 0;
    int rc;
    if ((rc = dmalloc(ctx, R->lights, n))) return rc;
    if (n) APG_HIP_CHECK(ctx, hipMemcpy(R->lights, dl.data(), n * sizeof(DevLight), hipMemcpyHostToDevice));
    if (sampler == APG_LS_POWER && n)
    {
        std::vector<float> w(n), p, q;
        std::vector<uint32_t> a;
        for (uint32_t i = 0; i < n; ++i) w[i] = length(ld3(dl[i].power));
        build_alias(w, p, q, a);
        if ((rc = dmalloc(ctx, R->alias_p, n)) || (rc = dmalloc(ctx, R->alias_q, n)) || (rc = dmalloc(ctx, R->alias_a, n)))
            return rc;
        APG_HIP_CHECK(ctx, hipMemcpy(R->alias_p, p.data(), n * 4, hipMemcpyHostToDevice));
        APG_HIP_CHECK(ctx, hipMemcpy(R->alias_q, q.data(), n * 4, hipMemcpyHostToDevice));
        APG_HIP_CHECK(ctx, hipMemcpy(R->alias_a, a.data(), n * 4, hipMemcpyHostToDevice));
    }
    if (sampler == APG_LS_BVH && n)
    {
        std::vector<DevLightNode> nodes;
        std::vector<int32_t> leaf;
        build_light_bvh(dl, nodes, leaf);
        std::vector<DevLightPair> pairs;
        std::vector<int32_t> parent;
        pack_light_bvh(nodes, pairs, parent, R->root_light);
        if ((rc = dmalloc(ctx, R->lpairs, pairs.size())) || (rc = dmalloc(ctx, R->lparent, parent.size())) ||
            (rc = dmalloc(ctx, R->leaf_of, n)))
            return rc;
        APG_HIP_CHECK(ctx, hipMemcpy(R->lpairs, pairs.data(), pairs.size() * sizeof(DevLightPair), hipMemcpyHostToDevice));
        APG_HIP_CHECK(ctx, hipMemcpy(R->lparent, parent.data(), parent.size() * 4, hipMemcpyHostToDevice));
        APG_HIP_CHECK(ctx, hipMemcpy(R->leaf_of, leaf.data(), n * 4, hipMemcpyHostToDevice));
    }
    R->nlights = n;
    R->ls_kind = sampler;
    R->ls_set = true;
    return APG_OK;
}

// The stage kernels' arguments of a pass (no allocation: guide buffers exist already).
void fill_shade_args(RenderState* R, const apg_render_params* prm, ShadeArgs& S)
{
    std::memset(&S, 0, sizeof(S));  // padding included: the graph key compares bytes
    S.mats = R->mats;
    S.tri_mat = R->tri_mat;
    S.tri_ng = R->tri_ng;
    S.ls = view_of(R);
    S.has_ls = R->ls_set;
    S.rng = R->rng;
    S.T = R->T;
    S.L = R->L;
    S.state = R->state;
    S.nee = R->nee;
    S.vis = R->vis;
    S.shadow = R->qs;
    S.hprim = R->hprim;
    S.ht = R->ht;
    S.hb0 = R->hb0;
    S.hb1 = R->hb1;
    S.counts = R->counts;
    S.max_depth = prm->max_depth;
    S.rr_depth = prm->rr_depth;
    for (int k = 0; k < 3; ++k) S.bg[k] = prm->background[k];
    S.attr = R->attr;
    S.env = R->env;
    S.env_w = R->env_w;
    S.env_h = R->env_h;
    S.aov_albedo = prm->aovs ? R->aov_albedo : nullptr;
    S.aov_normal = prm->aovs ? R->aov_normal : nullptr;
    S.gout = R->gout;
    // only training passes record path vertices (the resolve stage reads their radiance
    // targets back); plain and frozen-proposal passes write none
    S.vrec = prm->guide == 1 ? R->vrec : nullptr;
    S.vrec_depth = prm->guide == 1 ? R->vrec_depth : 0u;
    for (int a = 0; a < 3; ++a)
    {
        S.lo[a] = R->scene_lo[a];
        const float ext = R->scene_hi[a] - R->scene_lo[a];
        S.inv[a] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
}

// What a captured frame's passes baked in (frame.hip compares it before each replay): the
// stage arguments, camera, tiles, launch shapes and buffers.
void render_frame_key(Context* ctx, const apg_render_params* prm, std::vector<char>& key)
{
    RenderState* R = state_of(ctx);
    ShadeArgs S;
    fill_shade_args(R, prm, S);
    auto put = [&key](const void* p, size_t n) {
        key.insert(key.end(), static_cast<const char*>(p), static_cast<const char*>(p) + n);
    };
    put(&S, sizeof(S));
    put(&R->cam, sizeof(R->cam));
    put(&ctx->rlaunch, sizeof(ctx->rlaunch));
    const uintptr_t extra[] = { R->P, R->W, R->H, R->S, uintptr_t(R->accum), uintptr_t(R->accum64), uintptr_t(R->stats),
                                uintptr_t(R->tile_cams), uintptr_t(R->tile_roots), R->ntiles, R->tile_w, R->tile_h,
                                uintptr_t(R->gfeat), uintptr_t(R->aov_albedo), R->aov_P, uintptr_t(ctx->spill),
                                ctx->spill_threads, R->tri_ng_n };
    put(extra, sizeof(extra));
}

} // namespace apg

extern "C" {

int apg_light_sampler_sample(apg_context* ctx, const float* u, const float* hit, const float* ns,
                             uint32_t n, int32_t* light, float* pdf, void* stream)
{
    if (!ctx || (n && (!u || !hit || !ns || !light || !pdf))) return APG_E_INVALID;
    RenderState* R = state_of(ctx);
    if (!R->ls_set) return ctx->fail(APG_E_STATE, "apg_lights_set has not been called");
    if (n == 0) return APG_OK;
    k_ls_sample<<<nblocks(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(view_of(R), u, hit, ns, n, light, pdf);
    APG_HIP_CHECK(ctx, hipGetLastError());
    return APG_OK;
}

int apg_light_sampler_pdf(apg_context* ctx, const int32_t* light, const float* hit, const float* ns,
                          uint32_t n, float* pdf, void* stream)
{
    if (!ctx || (n && (!light || !hit || !ns || !pdf))) return APG_E_INVALID;
    RenderState* R = state_of(ctx);
    if (!R->ls_set) return ctx->fail(APG_E_STATE, "apg_lights_set has not been called");
    if (n == 0) return APG_OK;
    k_ls_pdf<<<nblocks(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(view_of(R), light, hit, ns, n, pdf);
    APG_HIP_CHECK(ctx, hipGetLastError());
    return APG_OK;
}

int apg_textures_set(apg_context* ctx, const float* texels, const uint32_t* sizes, uint32_t ntex)
{
    if (!ctx || (ntex && (!texels || !sizes))) return APG_E_INVALID;
    std::vector<DevTexture> dt(ntex);
    size_t total = 0;
    for (uint32_t k = 0; k < ntex; ++k)
    {
        const uint64_t n = uint64_t(sizes[2 * k]) * sizes[2 * k + 1];
        if (n == 0) return ctx->fail(APG_E_INVALID, "empty texture");
        if (total + n > 0xffffffffull) return ctx->fail(APG_E_CAPACITY, "more than 2^32 texels");
        dt[k] = DevTexture{ uint32_t(total), sizes[2 * k], sizes[2 * k + 1], 0u };
        total += n;
    }
    APG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    RenderState* R = state_of(ctx);
    if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
    dfree(R->texs);
    dfree(R->texels);
    R->ntex = 0;
    R->attr.texs = nullptr;
    R->attr.texels = nullptr;
    if (ntex == 0) return APG_OK;
    int rc;
    if ((rc = dmalloc(ctx, R->texs, ntex)) || (rc = dmalloc(ctx, R->texels, total))) return rc;
    APG_HIP_CHECK(ctx, hipMemcpy(R->texs, dt.data(), ntex * sizeof(DevTexture), hipMemcpyHostToDevice));
    APG_HIP_CHECK(ctx, hipMemcpy(R->texels, texels, total * sizeof(float4), hipMemcpyHostToDevice));
    R->attr.texs = R->texs;
    R->attr.texels = R->texels;
    R->ntex = ntex;
    return APG_OK;
}

int apg_render_set_environment(apg_context* ctx, const float* rgb, uint32_t width, uint32_t height)
{
    if (!ctx || (rgb && (width == 0 || height == 0))) return APG_E_INVALID;
    APG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    RenderState* R = state_of(ctx);
    if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
    dfree(R->env);
    R->env_w = R->env_h = 0;
    if (!rgb) return APG_OK;
    const size_t n = size_t(width) * height;
    std::vector<float4> t(n);
    for (size_t i = 0; i < n; ++i) t[i] = make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], 0.0f);
    int rc;
    if ((rc = dmalloc(ctx, R->env, n))) return rc;
    APG_HIP_CHECK(ctx, hipMemcpy(R->env, t.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    R->env_w = width;
    R->env_h = height;
    return APG_OK;
}

int apg_render_set_camera(apg_context* ctx, const apg_camera* cam)
{
    if (!ctx || !cam) return APG_E_INVALID;
    RenderState* R = state_of(ctx);
    camera_setup(*cam, R->cam);
    R->cam_set = true;
    return APG_OK;
}

int apg_render_set_tiles(apg_context* ctx, const apg_camera* cams, const uint32_t* roots, uint32_t ntiles,
                         uint32_t tile_width, uint32_t tile_height)
{
    if (!ctx || (ntiles && (!cams || tile_width == 0 || tile_height == 0))) return APG_E_INVALID;
    RenderState* R = state_of(ctx);
    if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
    dfree(R->tile_cams);
    dfree(R->tile_roots);
    R->ntiles = 0;
    if (ntiles == 0) return APG_OK;
    std::vector<DevCamera> dc(ntiles);
    for (uint32_t t = 0; t < ntiles; ++t) camera_setup(cams[t], dc[t]);
    int rc;
    if ((rc = dmalloc(ctx, R->tile_cams, ntiles))) return rc;
    APG_HIP_CHECK(ctx, hipMemcpy(R->tile_cams, dc.data(), ntiles * sizeof(DevCamera), hipMemcpyHostToDevice));
    if (roots)
    {
        if ((rc = dmalloc(ctx, R->tile_roots, ntiles))) return rc;
        APG_HIP_CHECK(ctx, hipMemcpy(R->tile_roots, roots, ntiles * 4, hipMemcpyHostToDevice));
    }
    R->ntiles = ntiles;
    R->tile_w = tile_width;
    R->tile_h = tile_height;
    return APG_OK;
}

int apg_render_reset_slots(apg_context* ctx, uint32_t width, uint32_t height, uint32_t seed, uint32_t slots)
{
    if (!ctx || width == 0 || height == 0 || slots == 0) return APG_E_INVALID;
    if (uint64_t(width) * height * slots > (1u << 26))
        return ctx->fail(APG_E_CAPACITY, "more than 2^26 paths per pass");
    APG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    RenderState* R = state_of(ctx);
    const uint32_t NPIX = width * height;
    const uint32_t P = NPIX * slots;
    if (P != R->P || NPIX != R->NPIX)
    {
        if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
        free_pixels(R);
        auto get = [&](auto*& ptr, size_t count) -> int {
            int rc = dmalloc(ctx, ptr, count);
            if (rc == APG_OK) R->pixel_allocs.push_back(ptr);
            return rc;
        };
        int rc = 0;
        if ((rc = get(R->rng, 4 * size_t(P))) || (rc = get(R->accum, 3 * size_t(P))) ||
            (rc = get(R->accum64, 3 * size_t(P))) || (rc = get(R->px, 3 * size_t(NPIX))) ||
            (rc = get(R->px64, 3 * size_t(NPIX))) || (rc = get(R->T, 3 * size_t(P))) ||
            (rc = get(R->L, 3 * size_t(P))) || (rc = get(R->state, P)) || (rc = get(R->nee, 10 * size_t(P))) ||
            (rc = get(R->vis, 2 * size_t(P))) || (rc = get(R->hprim, P)) || (rc = get(R->ht, P)) ||
            (rc = get(R->hb0, P)) || (rc = get(R->hb1, P)) || (rc = get(R->occ, 2 * size_t(P))) ||
            (rc = get(R->counts, 16)) || (rc = get(R->stats, 8)))
            return rc;
        for (Queue* q : { &R->qa, &R->qb, &R->qs })
        {
            const size_t cap = q == &R->qs ? 2 * size_t(P) : P;
            for (int a = 0; a < 3; ++a)
                if ((rc = get(q->o[a], cap)) || (rc = get(q->d[a], cap))) return rc;
            if ((rc = get(q->path, cap)) || (rc = get(q->root, cap))) return rc;
            if (q == &R->qs) { if ((rc = get(q->tmax, cap))) return rc; }
            else q->tmax = nullptr;
        }
        R->P = P;
        R->NPIX = NPIX;
    }
    R->W = width;
    R->H = height;
    R->S = slots;
    k_reset_pixels<<<nblocks(P, 256), 256>>>(P, NPIX, seed, R->rng, R->accum, R->accum64);
    APG_HIP_CHECK(ctx, hipMemset(R->stats, 0, 8 * sizeof(unsigned long long)));
    if (R->aov_albedo && R->aov_P != P)
    {
        dfree(R->aov_albedo);
        dfree(R->aov_normal);
        R->aov_P = 0;
    }
    if (R->aov_albedo)
    {
        APG_HIP_CHECK(ctx, hipMemset(R->aov_albedo, 0, 3 * size_t(P) * sizeof(float)));
        APG_HIP_CHECK(ctx, hipMemset(R->aov_normal, 0, 3 * size_t(P) * sizeof(float)));
    }
    APG_HIP_CHECK(ctx, hipDeviceSynchronize());
    R->total_spp = 0;
    return APG_OK;
}

int apg_render_set_launch(apg_context* ctx, const apg_render_launch* launch)
{
    if (!ctx || !launch) return APG_E_INVALID;
    const uint32_t v[3] = { launch->shade_blocks_per_cu, launch->shadow_blocks_per_cu, launch->resolve_blocks_per_cu };
    for (uint32_t x : v)
        if (x > 64u) return ctx->fail(APG_E_INVALID, "render stage grids take 1..64 workgroups per CU (0: default)");
    std::lock_guard<std::mutex> lock(ctx->launch_mutex);
    ctx->rlaunch.shade_blocks_per_cu = v[0] ? v[0] : 8u;
    ctx->rlaunch.shadow_blocks_per_cu = v[1] ? v[1] : 4u;
    ctx->rlaunch.resolve_blocks_per_cu = v[2] ? v[2] : 4u;
    return APG_OK;
}

int apg_render_get_launch(const apg_context* ctx, apg_render_launch* launch)
{
    if (!ctx || !launch) return APG_E_INVALID;
    *launch = ctx->rlaunch;
    return APG_OK;
}

int apg_render_reset(apg_context* ctx, uint32_t width, uint32_t height, uint32_t seed)
{
    return apg_render_reset_slots(ctx, width, height, seed, 1);
}

int apg_render_guide_stats(const apg_context* cctx, uint32_t* samples, float* loss)
{
    if (!cctx) return APG_E_INVALID;
    RenderState* R = state_of(const_cast<apg_context*>(cctx));
    if (samples) *samples = R->last_train_samples;
    if (loss) return guide_last_loss(const_cast<apg_context*>(cctx), loss);
    return APG_OK;
}

static int ensure_guide_buffers(apg_context* ctx, RenderState* R, uint32_t depth, uint32_t want)
{
    const uint32_t P = R->P;
    const uint32_t tcap = std::min<uint32_t>(P * 2, want ? want : (1u << 18));
    if (R->gfeat && R->guide_P == P && R->vrec_depth >= depth && R->tcap == tcap) return APG_OK;
    if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
    for (void* p : R->guide_allocs) (void)hipFree(p);
    R->guide_allocs.clear();
    // Nothing refers to the freed buffers from here on, and the sizes are set only once every
    // allocation has succeeded: a failure part-way leaves the state "not allocated", so the
    // next guided render allocates again instead of reusing freed memory.
    R->gfeat = R->gout = R->vrec = nullptr;
    R->tin = R->tdir = R->tw = nullptr;
    R->tcount = nullptr;
    R->vrec_depth = 0;
    R->tcap = 0;
    R->guide_P = 0;
    auto get = [&](auto*& ptr, size_t count) -> int {
        int rc = dmalloc(ctx, ptr, count);
        if (rc == APG_OK) R->guide_allocs.push_back(ptr);
        return rc;
    };
    int rc;
    if ((rc = get(R->gfeat, size_t(P) * 9)) || (rc = get(R->gout, size_t(P) * GUIDE_OUT)) ||
        (rc = get(R->vrec, size_t(P) * depth * VREC)) || (rc = get(R->tin, size_t(tcap) * 9)) ||
        (rc = get(R->tdir, size_t(tcap) * 3)) || (rc = get(R->tw, tcap)) || (rc = get(R->tcount, 4)))
    {
        R->gfeat = nullptr;  // the check at the top then allocates again
        return rc;
    }
    if ((rc = guide_ensure(ctx, tcap)))
    {
        R->gfeat = nullptr;
        return rc;
    }
    R->vrec_depth = depth;
    R->tcap = tcap;
    R->guide_P = P;
    return APG_OK;
}

int apg_render(apg_context* ctx, const apg_render_params* prm, void* stream_)
{
    if (!ctx || !prm) return APG_E_INVALID;
    RenderState* R = state_of(ctx);
    if (!ctx->built) return ctx->fail(APG_E_STATE, "apg_scene_build has not been called");
    if (!R->P) return ctx->fail(APG_E_STATE, "apg_render_reset has not been called");
    if (!R->cam_set && !R->ntiles) return ctx->fail(APG_E_STATE, "apg_render_set_camera has not been called");
    if (!R->mats || R->tri_mat_n != ctx->total_tris)
        return ctx->fail(APG_E_STATE, "apg_materials_set must cover every triangle");
    if (prm->max_depth == 0 || prm->max_depth > ST_DEPTH) return ctx->fail(APG_E_INVALID, "max_depth out of range");
    if (R->max_tex_ref > R->ntex) return ctx->fail(APG_E_STATE, "a material names a texture that is not set");
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    const uint32_t P = R->P;
    if (int rc = ensure_attributes(ctx, R)) return rc;
    if (prm->aovs && !R->aov_albedo)
    {
        if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
        int rc;
        if ((rc = dmalloc(ctx, R->aov_albedo, 3 * size_t(P))) || (rc = dmalloc(ctx, R->aov_normal, 3 * size_t(P))))
            return rc;
        APG_HIP_CHECK(ctx, hipMemset(R->aov_albedo, 0, 3 * size_t(P) * sizeof(float)));
        APG_HIP_CHECK(ctx, hipMemset(R->aov_normal, 0, 3 * size_t(P) * sizeof(float)));
        R->aov_P = P;
    }

    const uint32_t guide = prm->guide;
    if (guide > 2) return ctx->fail(APG_E_INVALID, "guide must be 0, 1 or 2");
    if (guide)
    {
        int rc = ensure_guide_buffers(ctx, R, prm->max_depth, prm->train_samples);
        if (rc) return rc;
    }
    ShadeArgs S;
    fill_shade_args(R, prm, S);
    const float frac = prm->guide_frac > 0.0f ? std::min(prm->guide_frac, 0.95f) : 0.5f;
    const uint32_t train_steps = prm->train_steps ? prm->train_steps : 4u;
    const apg_render_launch& rl = ctx->rlaunch;
    const uint32_t sgrid = std::min<uint32_t>(nblocks(P, 256), uint32_t(ctx->cu_count) * rl.shade_blocks_per_cu);
    const uint32_t rgrid = std::min<uint32_t>(nblocks(P, 256), uint32_t(ctx->cu_count) * rl.resolve_blocks_per_cu);
    const uint32_t vgrid = std::min<uint32_t>(nblocks(2 * P, 256), uint32_t(ctx->cu_count) * rl.shadow_blocks_per_cu);

    std::lock_guard<std::mutex> lock(ctx->launch_mutex);
    if (!ctx->last_launch) APG_HIP_CHECK(ctx, hipEventCreateWithFlags(&ctx->last_launch, hipEventDisableTiming));
    // a frame graph being captured is ordered after earlier work by its launch
    if (!ctx->capturing) APG_HIP_CHECK(ctx, hipStreamWaitEvent(stream, ctx->last_launch, 0));
    if (prm->spp % R->S) return ctx->fail(APG_E_INVALID, "spp must be a multiple of the slots per pixel");
    Tiles tl;
    std::memset(&tl, 0, sizeof(tl));
    tl.tw = tl.th = tl.tiles_x = 1;
    if (R->ntiles)
    {
        if (R->W % R->tile_w || R->H % R->tile_h || (R->W / R->tile_w) * (R->H / R->tile_h) != R->ntiles)
            return ctx->fail(APG_E_INVALID, "the frame is not a grid of the configured tiles");
        tl.cams = R->tile_cams;
        tl.roots = R->tile_roots;
        tl.tw = R->tile_w;
        tl.th = R->tile_h;
        tl.tiles_x = R->W / R->tile_w;
    }
    const bool forest = ctx->scene.roots > 1;
    // One pass (one sample per path slot) on stream st: generate, then per bounce
    // extend / shade / shadow / resolve.
    auto enqueue_pass = [&](hipStream_t st, const TrainOut& G0) -> int {
        TrainOut G = G0;
        k_set_counts<<<1, 64, 0, st>>>(R->counts, P);
        k_generate<<<nblocks(P, 256), 256, 0, st>>>(R->cam, tl, R->W, R->H, P, R->rng, R->T, R->L, R->state,
                                                      R->qa, R->stats);
        Queue in = R->qa, out = R->qb;
        for (uint32_t b = 0; b < prm->max_depth; ++b)
        {
            apg_rays rays{ in.o[0], in.o[1], in.o[2], in.d[0], in.d[1], in.d[2], nullptr };
            apg_hits hits{ R->hprim, R->ht, R->hb0, R->hb1 };
            int rc = launch_trace_raw(ctx, rays, P, R->counts, &hits, false, nullptr,
                                      prm->count_fetches ? R->stats + 3 : nullptr, st,
                                      forest ? in.root : nullptr);
            if (rc) return rc;
            S.in = in;
            S.out = out;
            S.bounce = b;
            // frozen proposals may stop guiding after guide_depth bounces (no inference there)
            if (guide && (guide == 1 || !prm->guide_depth || b < prm->guide_depth))
            {
                k_guide_features<<<sgrid, 256, 0, st>>>(S, R->gfeat);
                rc = guide_infer(ctx, R->gfeat, P, R->counts, R->gout, st);
                if (rc) return rc;
                k_shade<true><<<sgrid, 256, 0, st>>>(S);
            }
            else
                k_shade<false><<<sgrid, 256, 0, st>>>(S);
            apg_rays srays{ R->qs.o[0], R->qs.o[1], R->qs.o[2], R->qs.d[0], R->qs.d[1], R->qs.d[2], R->qs.tmax };
            rc = launch_trace_raw(ctx, srays, 2 * P, R->counts + 2, nullptr, true, R->occ,
                                  prm->count_fetches ? R->stats + 5 : nullptr, st,
                                  forest ? R->qs.root : nullptr);
            if (rc) return rc;
            k_shadow_vis<<<vgrid, 256, 0, st>>>(R->counts, R->qs.path, R->occ, R->vis);
            G.bounce = b;
            k_resolve<<<rgrid, 256, 0, st>>>(R->counts, in.path, R->state, R->nee, R->vis, R->L, R->accum,
                                             R->accum64, prm->fp64_accum ? 1 : 0, G);
            k_advance<<<1, 64, 0, st>>>(R->counts, R->stats);
            std::swap(in, out);
        }
        return APG_OK;
    };
    const TrainOut no_train{ nullptr, 0, 0, nullptr, nullptr, nullptr, 0, nullptr, 1.0f, 0u };

    if (prm->graph && guide != 1)
    {
        // The pass is captured once into a HIP graph on the library's own stream and
        // replayed; it is captured again when anything it baked in has changed.
        S.guide_frac = guide ? frac : 0.0f;
        int rc = ensure_trace_capacity(ctx, 2 * P);
        if (rc) return rc;
        std::vector<char> key;
        auto put = [&key](const void* p, size_t n) {
            key.insert(key.end(), static_cast<const char*>(p), static_cast<const char*>(p) + n);
        };
        put(&S, sizeof(S));
        put(&tl, sizeof(tl));
        put(&R->cam, sizeof(R->cam));
        put(&ctx->scene, sizeof(ctx->scene));
        put(&ctx->cfg, sizeof(ctx->cfg));
        put(&ctx->rlaunch, sizeof(ctx->rlaunch));
        const uintptr_t extra[] = { P, R->W, R->H, uintptr_t(R->accum), uintptr_t(R->accum64), prm->fp64_accum,
                                    uintptr_t(R->gfeat), uintptr_t(R->gout), guide, uintptr_t(ctx->spill),
                                    prm->guide_depth, prm->count_fetches };
        put(extra, sizeof(extra));
        if (!R->gexec || key != R->gkey)
        {
            if (!R->gstream) APG_HIP_CHECK(ctx, hipStreamCreateWithFlags(&R->gstream, hipStreamNonBlocking));
            if (!R->gev) APG_HIP_CHECK(ctx, hipEventCreateWithFlags(&R->gev, hipEventDisableTiming));
            if (R->gexec)
            {
                // replays of the old exec may still be queued on gstream: destroying it under
                // them frees what the runtime's completion handling still reads
                APG_HIP_CHECK(ctx, hipStreamSynchronize(R->gstream));
                (void)hipGraphExecDestroy(R->gexec);
                R->gexec = nullptr;
            }
            hipGraph_t graph = nullptr;
            APG_HIP_CHECK(ctx, hipStreamBeginCapture(R->gstream, hipStreamCaptureModeRelaxed));
            rc = enqueue_pass(R->gstream, no_train);
            hipError_t e = hipStreamEndCapture(R->gstream, &graph);
            if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
            if (e != hipSuccess) return ctx->fail(APG_E_HIP, hipGetErrorString(e));
            e = hipGraphInstantiate(&R->gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) { R->gexec = nullptr; return ctx->fail(APG_E_HIP, hipGetErrorString(e)); }
            R->gkey = key;
        }
        APG_HIP_CHECK(ctx, hipEventRecord(R->gev, stream));
        APG_HIP_CHECK(ctx, hipStreamWaitEvent(R->gstream, R->gev, 0));
        for (uint32_t sidx = 0; sidx < prm->spp / R->S; ++sidx)
            APG_HIP_CHECK(ctx, hipGraphLaunch(R->gexec, R->gstream));
        APG_HIP_CHECK(ctx, hipEventRecord(R->gev, R->gstream));
        APG_HIP_CHECK(ctx, hipStreamWaitEvent(stream, R->gev, 0));
        if (guide == 2) R->guided_passes += prm->spp / R->S;
    }
    else
    for (uint32_t sidx = 0; sidx < prm->spp / R->S; ++sidx)
    {
        const bool guided = guide != 0 && (guide == 2 || R->guided_passes >= prm->guide_warmup);
        S.guide_frac = guided ? frac : 0.0f;
        TrainOut G = no_train;
        if (guide == 1)
        {
            // keep about tcap of the pass's P x (max_depth - 1) candidate samples, spread
            // uniformly over the frame instead of the first paths to finish
            G = TrainOut{ R->vrec, R->vrec_depth, 0, R->tin, R->tdir, R->tw, R->tcap, R->tcount,
                          R->train_keep, R->guided_passes * 0x632BE5ABu + 17u };
            APG_HIP_CHECK(ctx, hipMemsetAsync(R->tcount, 0, 4, stream));
        }
        int rc = enqueue_pass(stream, G);
        if (rc) return rc;
        if (guide == 1)
        {
            // online training on this pass's radiance samples
            uint32_t tn = 0;
            APG_HIP_CHECK(ctx, hipMemcpyAsync(&tn, R->tcount, 4, hipMemcpyDeviceToHost, stream));
            APG_HIP_CHECK(ctx, hipStreamSynchronize(stream));
            // keep about 95 % of the capacity next pass, spread over the whole frame
            // instead of the paths that happen to finish first
            {
                const float cand = float(tn) / R->train_keep;
                R->train_keep = cand > 0.0f ? std::min(1.0f, 0.95f * float(R->tcap) / cand) : 1.0f;
            }
            tn = std::min(tn, R->tcap);
            R->last_train_samples = tn;
            if (tn)
            {
                // one epoch over the pass's samples in train_steps minibatches: minibatch j
                // takes samples j, j + S, j + 2S, ... so each spreads over the whole frame
                const uint32_t S = std::min(std::min(train_steps, tn), TRAIN_MAX_STEPS);
                for (uint32_t j = 0; j < S; ++j)
                {
                    const uint32_t mb = (tn - j + S - 1) / S;
                    rc = guide_train(ctx, nullptr, R->tin, R->tdir, R->tw, mb, j, S, true, nullptr, stream);
                    if (rc) return rc;
                }
            }
            R->guided_passes++;
        }
    }
    APG_HIP_CHECK(ctx, hipGetLastError());
    // a frame graph being captured records its own completion after each replay
    if (!ctx->capturing) APG_HIP_CHECK(ctx, hipEventRecord(ctx->last_launch, stream));
    R->total_spp += prm->spp;
    return APG_OK;
}

int apg_render_read(apg_context* ctx, float* rgb, double* rgb64, uint32_t* total_spp)
{
    if (!ctx) return APG_E_INVALID;
    RenderState* R = state_of(ctx);
    if (!R->P) return ctx->fail(APG_E_STATE, "apg_render_reset has not been called");
    if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
    if (int rc = ctx->take_overflow()) return rc;
    const uint32_t n3 = 3u * R->NPIX;
    if (rgb)
    {
        k_sum_slots<<<nblocks(n3, 256), 256>>>(R->accum, R->NPIX, R->S, R->px);
        APG_HIP_CHECK(ctx, hipGetLastError());
        APG_HIP_CHECK(ctx, hipMemcpy(rgb, R->px, size_t(n3) * 4, hipMemcpyDeviceToHost));
    }
    if (rgb64)
    {
        k_sum_slots<<<nblocks(n3, 256), 256>>>(R->accum64, R->NPIX, R->S, R->px64);
        APG_HIP_CHECK(ctx, hipGetLastError());
        APG_HIP_CHECK(ctx, hipMemcpy(rgb64, R->px64, size_t(n3) * 8, hipMemcpyDeviceToHost));
    }
    if (total_spp) *total_spp = R->total_spp;
    return APG_OK;
}

int apg_render_read_aovs(apg_context* ctx, float* albedo, float* normal)
{
    if (!ctx) return APG_E_INVALID;
    RenderState* R = state_of(ctx);
    if (!R->P) return ctx->fail(APG_E_STATE, "apg_render_reset has not been called");
    if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
    if (int rc = ctx->take_overflow()) return rc;
    const uint32_t n3 = 3u * R->NPIX;
    for (int k = 0; k < 2; ++k)
    {
        float* out = k == 0 ? albedo : normal;
        if (!out) continue;
        const float* src = k == 0 ? R->aov_albedo : R->aov_normal;
        if (!src)
        {
            std::memset(out, 0, size_t(n3) * sizeof(float));
            continue;
        }
        k_sum_slots<<<nblocks(n3, 256), 256>>>(src, R->NPIX, R->S, R->px);
        APG_HIP_CHECK(ctx, hipGetLastError());
        APG_HIP_CHECK(ctx, hipMemcpy(out, R->px, size_t(n3) * 4, hipMemcpyDeviceToHost));
    }
    return APG_OK;
}

int apg_render_postprocess(apg_context* ctx, int tonemap, uint8_t* frame, void* stream_)
{
    if (!ctx || !frame) return APG_E_INVALID;
    RenderState* R = state_of(ctx);
    if (!R->P) return ctx->fail(APG_E_STATE, "apg_render_reset has not been called");
    if (R->total_spp == 0) return ctx->fail(APG_E_STATE, "no samples accumulated");
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    std::lock_guard<std::mutex> lock(ctx->launch_mutex);
    if (ctx->last_launch) APG_HIP_CHECK(ctx, hipStreamWaitEvent(stream, ctx->last_launch, 0));
    const uint32_t n3 = 3u * R->NPIX;
    k_sum_slots<<<nblocks(n3, 256), 256, 0, stream>>>(R->accum, R->NPIX, R->S, R->px);
    k_postprocess<<<nblocks(n3, 256), 256, 0, stream>>>(R->px, n3, float(R->total_spp),
                                                        tonemap ? 1u : 0u, frame);
    APG_HIP_CHECK(ctx, hipGetLastError());
    return APG_OK;
}

int apg_render_fetch_counts(const apg_context* cctx, uint64_t* closest_nodes, uint64_t* closest_tris,
                            uint64_t* shadow_nodes, uint64_t* shadow_tris)
{
    if (!cctx) return APG_E_INVALID;
    apg_context* ctx = const_cast<apg_context*>(cctx);
    RenderState* R = state_of(ctx);
    unsigned long long h[8] = {};
    if (R->stats)
    {
        if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
        APG_HIP_CHECK(ctx, hipMemcpy(h, R->stats, sizeof(h), hipMemcpyDeviceToHost));
    }
    if (closest_nodes) *closest_nodes = h[3];
    if (closest_tris) *closest_tris = h[4];
    if (shadow_nodes) *shadow_nodes = h[5];
    if (shadow_tris) *shadow_tris = h[6];
    return APG_OK;
}

int apg_render_counters(const apg_context* cctx, uint64_t* closest, uint64_t* shadow, uint64_t* vertices)
{
    if (!cctx) return APG_E_INVALID;
    apg_context* ctx = const_cast<apg_context*>(cctx);
    RenderState* R = state_of(ctx);
    unsigned long long h[8] = {};
    if (R->stats)
    {
        if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
        APG_HIP_CHECK(ctx, hipMemcpy(h, R->stats, sizeof(h), hipMemcpyDeviceToHost));
    }
    if (closest) *closest = h[0];
    if (shadow) *shadow = h[1];
    if (vertices) *vertices = h[2];
    return APG_OK;
}

} // extern "C"
