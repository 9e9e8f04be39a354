// Online-trained path-guiding MLP on gfx950: fused BF16 MFMA forward,
// backward (data and weight gradients) and Adam.  No counterpart in the
// reference; the proposal is mixed with BSDF sampling by one-sample MIS in the
// path tracer (render.hip), so plain path tracing stays the fallback.
//
// Layers 96 -> 64 -> 64 -> 64 -> 64 -> 48 (guide_core.h).  All weights (21.5 K
// BF16 = 43 KB) sit in LDS for the whole kernel; a wave carries 16 samples
// through every layer with v_mfma_f32_16x16x16_bf16 (A: 16 samples x 16
// inputs, B: 16 inputs x 16 outputs, FP32 accumulators), bias and ReLU fused
// into the epilogue.  Weight gradients reduce over samples in the MFMA
// accumulators of each wave and are added to an FP32 gradient vector once per
// workgroup.  Adam keeps FP32 master weights and writes the BF16 copy.
//
// Operand layout of v_mfma_f32_16x16x16_bf16 (wave64): lane l holds
// A[l % 16][4 (l / 16) + i] and B[4 (l / 16) + i][l % 16] for i = 0..3, and the
// result D[4 (l / 16) + i][l % 16].
#include "guide_core.h"
#include "apg_internal.h"

#include <cmath>
#include <vector>

namespace apg {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u;
#if defined(__HIP_DEVICE_COMPILE__)
    u = __float_as_uint(f);
#else
    __builtin_memcpy(&u, &f, 4);
#endif
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return uint16_t(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h)
{
    const uint32_t u = uint32_t(h) << 16;
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(u);
#else
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
#endif
}

struct GuideNet
{
    float* w32 = nullptr;   // master parameters (GUIDE_PARAMS)
    uint16_t* w16 = nullptr; // BF16 weights (first GUIDE_B_OFF[0] entries)
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    float* part = nullptr;     // WG_MAX x GUIDE_PARAMS partial gradients
    float* scratch = nullptr;  // [0] sum of weights, [1] loss sum
    bool ready = false;        // every buffer above is allocated and initialised
    uint32_t step = 0;
    float lr = 1e-3f;
    // training buffers (capacity in samples)
    uint32_t cap = 0;
    uint16_t* acts = nullptr;  // cap x 4 x 64
    uint16_t* feat16 = nullptr; // cap x 96 (fused training path)
    float* out = nullptr;      // cap x 48
    uint16_t* dz = nullptr;    // cap x (48 + 4 x 64)
};

namespace {

constexpr int WAVES = 4;
constexpr int BLOCK = 64 * WAVES;
constexpr int WSZ = 21504;  // BF16 weights

__device__ __forceinline__ s4v load_a_row(const uint16_t* row)
{
    // 4 consecutive BF16 values (8 bytes, 8-byte aligned)
    const uint2 u = *reinterpret_cast<const uint2*>(row);
    s4v r;
    r[0] = short(u.x & 0xFFFFu);
    r[1] = short(u.x >> 16);
    r[2] = short(u.y & 0xFFFFu);
    r[3] = short(u.y >> 16);
    return r;
}

__device__ __forceinline__ s4v gather4(const uint16_t* base, int stride)
{
    s4v r;
    r[0] = short(base[0]);
    r[1] = short(base[stride]);
    r[2] = short(base[2 * stride]);
    r[3] = short(base[3 * stride]);
    return r;
}

__device__ __forceinline__ void load_weights(uint16_t* sw, const uint16_t* w16)
{
    const uint4* src = reinterpret_cast<const uint4*>(w16);
    uint4* dst = reinterpret_cast<uint4*>(sw);
    for (int i = threadIdx.x; i < WSZ / 8; i += BLOCK) dst[i] = src[i];
}

// One dense layer for a 16-sample tile: act_out[16][N] = act_in[16][K] . W^T (+ bias, ReLU).
template <int K, int N, bool RELU, bool LAST>
__device__ __forceinline__ void dense(const uint16_t* act_in, const uint16_t* W, const float* bias,
                                      uint16_t* act_out, float* out_f32, uint32_t row0, uint32_t valid_rows,
                                      uint16_t* acts_g, int acts_off)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt)
    {
        f4v acc = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int ks = 0; ks < K / 16; ++ks)
        {
            const s4v a = load_a_row(act_in + r * K + ks * 16 + 4 * kg);
            const s4v b = load_a_row(W + (nt * 16 + r) * K + ks * 16 + 4 * kg);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
        const int col = nt * 16 + r;
        const float bb = bias[col];
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            const int row = 4 * kg + i;
            float v = acc[i] + bb;
            if (RELU) v = v > 0.0f ? v : 0.0f;
            if (LAST)
            {
                if (uint32_t(row) < valid_rows) out_f32[size_t(row0 + row) * GUIDE_OUT + col] = v;
            }
            else
            {
                const uint16_t h = f2bf(v);
                act_out[row * N + col] = h;
                if (acts_g && uint32_t(row) < valid_rows)
                    acts_g[size_t(row0 + row) * (4 * GUIDE_HID) + acts_off + col] = h;
            }
        }
    }
}

struct FwdArgs
{
    const uint16_t* w16;
    const float* w32;   // biases read from here
    const float* feat;  // M x 96 fp32
    uint32_t M;
    const uint32_t* M_dev;
    float* out;         // M x 48
    uint16_t* acts;     // optional M x 256
};

__global__ __launch_bounds__(BLOCK) void k_guide_fwd(FwdArgs A)
{
    __shared__ __align__(16) uint16_t sw[WSZ];
    __shared__ float sb[GUIDE_PARAMS - 21504];
    __shared__ __align__(16) uint16_t sa[WAVES][2][16 * GUIDE_IN];
    load_weights(sw, A.w16);
    for (int i = threadIdx.x; i < GUIDE_PARAMS - 21504; i += BLOCK) sb[i] = A.w32[21504 + i];
    __syncthreads();
    const uint32_t M = A.M_dev ? *A.M_dev : A.M;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* a0 = sa[wave][0];
    uint16_t* a1 = sa[wave][1];
    for (uint32_t base = blockIdx.x * (16 * WAVES); base < M; base += gridDim.x * 16 * WAVES)
    {
        const uint32_t row0 = base + 16 * wave;
        const uint32_t valid = row0 < M ? min(16u, M - row0) : 0u;
        // features -> BF16 tile
        for (int i = lane; i < 16 * GUIDE_IN; i += 64)
        {
            const int rr = i / GUIDE_IN;
            a0[i] = uint32_t(rr) < valid ? f2bf(A.feat[size_t(row0 + rr) * GUIDE_IN + (i % GUIDE_IN)]) : 0;
        }
        __syncthreads();
        dense<GUIDE_IN, GUIDE_HID, true, false>(a0, sw + guide_w_off(0), sb + 0, a1, nullptr, row0, valid, A.acts, 0);
        __syncthreads();
        dense<GUIDE_HID, GUIDE_HID, true, false>(a1, sw + guide_w_off(1), sb + 64, a0, nullptr, row0, valid, A.acts, 64);
        __syncthreads();
        dense<GUIDE_HID, GUIDE_HID, true, false>(a0, sw + guide_w_off(2), sb + 128, a1, nullptr, row0, valid, A.acts, 128);
        __syncthreads();
        dense<GUIDE_HID, GUIDE_HID, true, false>(a1, sw + guide_w_off(3), sb + 192, a0, nullptr, row0, valid, A.acts, 192);
        __syncthreads();
        dense<GUIDE_HID, GUIDE_OUT, false, true>(a0, sw + guide_w_off(4), sb + 256, nullptr, A.out, row0, valid, nullptr, 0);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Inference with the encoding fused in: queries are (p01, n, wo), 9 floats.
// Each wave carries 16 queries through the network in the transposed form
// D[unit][query] = W[unit][:] . act[query][:], so a layer's MFMA result tile
// (lane (r, kg) holds units 4 kg .. 4 kg + 3 of query r) is exactly the next
// layer's B operand for one 16-wide k-step: activations never leave
// registers.  Weights live in LDS with rows padded by 4 BF16 so the 16 rows a
// wave reads at once fall in different banks.
// ---------------------------------------------------------------------------
constexpr int IPAD = 4;
__host__ __device__ constexpr int infer_row(int l) { return (l == 0 ? GUIDE_IN : GUIDE_HID) + IPAD; }
__host__ __device__ constexpr int infer_off(int l)
{
    return l == 0 ? 0 : GUIDE_HID * infer_row(0) + (l - 1) * GUIDE_HID * infer_row(1);
}
constexpr int INFER_WSZ = infer_off(4) + GUIDE_OUT * infer_row(4);

struct InferArgs
{
    const uint16_t* w16;
    const float* w32;
    const float* q;         // M x 9
    uint32_t M;
    const uint32_t* M_dev;
    float* out;             // M x 48
};

__device__ __forceinline__ uint32_t pack_bf2(float a, float b)
{
    return uint32_t(f2bf(a)) | (uint32_t(f2bf(b)) << 16);
}

__device__ __forceinline__ s4v as_s4v(uint32_t lo, uint32_t hi)
{
    s4v r;
    r[0] = short(lo & 0xFFFFu);
    r[1] = short(lo >> 16);
    r[2] = short(hi & 0xFFFFu);
    r[3] = short(hi >> 16);
    return r;
}

// One layer: act_out tiles (N / 16) from act_in tiles (K / 16).
template <int L, int K, int N, bool RELU>
__device__ __forceinline__ void infer_layer(const uint16_t* sw, const float* sb, const s4v (&in)[K / 16],
                                            f4v (&res)[N / 16], int r, int kg)
{
    const uint16_t* W = sw + infer_off(L);
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt)
    {
        f4v acc = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int ks = 0; ks < K / 16; ++ks)
        {
            const s4v a = load_a_row(W + (nt * 16 + r) * infer_row(L) + ks * 16 + 4 * kg);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, in[ks], acc, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            float v = acc[i] + sb[guide_b_off(L) - 21504 + nt * 16 + 4 * kg + i];
            if (RELU) v = v > 0.0f ? v : 0.0f;
            acc[i] = v;
        }
        res[nt] = acc;
    }
}

template <int T>
__device__ __forceinline__ void to_bf16_tiles(const f4v (&x)[T], s4v (&y)[T])
{
#pragma unroll
    for (int t = 0; t < T; ++t) y[t] = as_s4v(pack_bf2(x[t][0], x[t][1]), pack_bf2(x[t][2], x[t][3]));
}

__global__ __launch_bounds__(BLOCK) void k_guide_infer(InferArgs A)
{
    __shared__ __align__(16) uint16_t sw[INFER_WSZ];
    __shared__ float sb[GUIDE_PARAMS - 21504];
    for (int l = 0; l < GUIDE_LAYERS; ++l)
    {
        const int K = guide_k(l), N = guide_n(l);
        for (int i = threadIdx.x; i < N * K / 4; i += BLOCK)
        {
            const int row = (4 * i) / K, col = (4 * i) % K;
            *reinterpret_cast<uint2*>(sw + infer_off(l) + row * infer_row(l) + col) =
                *reinterpret_cast<const uint2*>(A.w16 + guide_w_off(l) + 4 * i);
        }
    }
    for (int i = threadIdx.x; i < GUIDE_PARAMS - 21504; i += BLOCK) sb[i] = A.w32[21504 + i];
    __syncthreads();
    const uint32_t M = A.M_dev ? *A.M_dev : A.M;
    const int lane = threadIdx.x & 63, r = lane & 15, kg = lane >> 4;
    const uint32_t nwaves = gridDim.x * WAVES;
    for (uint32_t tile = blockIdx.x * WAVES + (threadIdx.x >> 6); tile * 16u < M; tile += nwaves)
    {
        const uint32_t qi = tile * 16u + uint32_t(r);
        float p[3] = { 0, 0, 0 }, n[3] = { 0, 0, 0 }, wo[3] = { 0, 0, 0 };
        if (qi < M)
        {
            const float* qq = A.q + size_t(qi) * 9;
            p[0] = qq[0]; p[1] = qq[1]; p[2] = qq[2];
            n[0] = qq[3]; n[1] = qq[4]; n[2] = qq[5];
            wo[0] = qq[6]; wo[1] = qq[7]; wo[2] = qq[8];
        }
        s4v h0[GUIDE_IN / 16];
#pragma unroll
        for (int ks = 0; ks < GUIDE_IN / 16; ++ks)
        {
            const int f = ks * 16 + 4 * kg;
            h0[ks] = as_s4v(pack_bf2(guide_feature(f, p, n, wo), guide_feature(f + 1, p, n, wo)),
                            pack_bf2(guide_feature(f + 2, p, n, wo), guide_feature(f + 3, p, n, wo)));
        }
        f4v t1[4], t2[4];
        s4v h1[4], h2[4];
        infer_layer<0, GUIDE_IN, GUIDE_HID, true>(sw, sb, h0, t1, r, kg);
        to_bf16_tiles(t1, h1);
        infer_layer<1, GUIDE_HID, GUIDE_HID, true>(sw, sb, h1, t2, r, kg);
        to_bf16_tiles(t2, h2);
        infer_layer<2, GUIDE_HID, GUIDE_HID, true>(sw, sb, h2, t1, r, kg);
        to_bf16_tiles(t1, h1);
        infer_layer<3, GUIDE_HID, GUIDE_HID, true>(sw, sb, h1, t2, r, kg);
        to_bf16_tiles(t2, h2);
        f4v o[GUIDE_OUT / 16];
        infer_layer<4, GUIDE_HID, GUIDE_OUT, false>(sw, sb, h2, o, r, kg);
        if (qi < M)
        {
#pragma unroll
            for (int nt = 0; nt < GUIDE_OUT / 16; ++nt)
                *reinterpret_cast<float4*>(A.out + size_t(qi) * GUIDE_OUT + nt * 16 + 4 * kg) =
                    make_float4(o[nt][0], o[nt][1], o[nt][2], o[nt][3]);
        }
    }
}

// ---------------------------------------------------------------------------
// Training step, fused: encoding, forward, loss head and backward through the
// layers for 16-query tiles, one tile per wave at a time, activations and
// gradients in registers in the transposed tile layout of k_guide_infer.
// Writes what the weight-gradient kernel reads: the BF16 features, the hidden
// activations and the pre-activation gradients of every layer.
// ---------------------------------------------------------------------------
struct TrainArgs
{
    const uint16_t* w16;
    const float* w32;
    const float* q;        // M x 9
    const float* dir;      // M x 3
    const float* weight;   // M
    const float* wsum;     // device: sum of the minibatch's weights
    uint32_t M;
    uint16_t* feat16;      // M x 96
    uint16_t* acts;        // M x 256
    uint16_t* dz;          // M x (48 + 256)
    float* loss_sum;       // optional
};

__global__ __launch_bounds__(BLOCK) void k_guide_train_fused(TrainArgs A)
{
    __shared__ __align__(16) uint16_t sw[INFER_WSZ];
    __shared__ float sb[GUIDE_PARAMS - 21504];
    __shared__ float sh[WAVES][16 * GUIDE_OUT];
    for (int l = 0; l < GUIDE_LAYERS; ++l)
    {
        const int K = guide_k(l), N = guide_n(l);
        for (int i = threadIdx.x; i < N * K / 4; i += BLOCK)
        {
            const int row = (4 * i) / K, col = (4 * i) % K;
            *reinterpret_cast<uint2*>(sw + infer_off(l) + row * infer_row(l) + col) =
                *reinterpret_cast<const uint2*>(A.w16 + guide_w_off(l) + 4 * i);
        }
    }
    for (int i = threadIdx.x; i < GUIDE_PARAMS - 21504; i += BLOCK) sb[i] = A.w32[21504 + i];
    __syncthreads();
    const uint32_t M = A.M;
    const float wsum = *A.wsum;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, kg = lane >> 4;
    float* hw = sh[wave];
    constexpr int DZS = GUIDE_OUT + 4 * GUIDE_HID;
    float loss = 0.0f;
    const uint32_t nwaves = gridDim.x * WAVES;
    for (uint32_t tile = blockIdx.x * WAVES + wave; tile * 16u < M; tile += nwaves)
    {
        const uint32_t qi = tile * 16u + uint32_t(r);
        const bool valid = qi < M;
        float p[3] = { 0, 0, 0 }, n[3] = { 0, 0, 0 }, wo[3] = { 0, 0, 0 };
        if (valid)
        {
            const float* qq = A.q + size_t(qi) * 9;
            p[0] = qq[0]; p[1] = qq[1]; p[2] = qq[2];
            n[0] = qq[3]; n[1] = qq[4]; n[2] = qq[5];
            wo[0] = qq[6]; wo[1] = qq[7]; wo[2] = qq[8];
        }
        // ---- forward ----
        s4v h0[GUIDE_IN / 16];
#pragma unroll
        for (int ks = 0; ks < GUIDE_IN / 16; ++ks)
        {
            const int f = ks * 16 + 4 * kg;
            const uint32_t lo = pack_bf2(guide_feature(f, p, n, wo), guide_feature(f + 1, p, n, wo));
            const uint32_t hi = pack_bf2(guide_feature(f + 2, p, n, wo), guide_feature(f + 3, p, n, wo));
            h0[ks] = as_s4v(lo, hi);
            if (valid) *reinterpret_cast<uint2*>(A.feat16 + size_t(qi) * GUIDE_IN + f) = make_uint2(lo, hi);
        }
        f4v t[4];
        s4v h[4][4];  // hidden layer outputs (post-ReLU, BF16), layer-major
        infer_layer<0, GUIDE_IN, GUIDE_HID, true>(sw, sb, h0, t, r, kg);
        to_bf16_tiles(t, h[0]);
        infer_layer<1, GUIDE_HID, GUIDE_HID, true>(sw, sb, h[0], t, r, kg);
        to_bf16_tiles(t, h[1]);
        infer_layer<2, GUIDE_HID, GUIDE_HID, true>(sw, sb, h[1], t, r, kg);
        to_bf16_tiles(t, h[2]);
        infer_layer<3, GUIDE_HID, GUIDE_HID, true>(sw, sb, h[2], t, r, kg);
        to_bf16_tiles(t, h[3]);
        if (valid)
        {
#pragma unroll
            for (int l = 0; l < 4; ++l)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                {
                    const s4v v = h[l][nt];
                    const uint32_t lo = uint32_t(uint16_t(v[0])) | (uint32_t(uint16_t(v[1])) << 16);
                    const uint32_t hi = uint32_t(uint16_t(v[2])) | (uint32_t(uint16_t(v[3])) << 16);
                    *reinterpret_cast<uint2*>(A.acts + size_t(qi) * (4 * GUIDE_HID) + l * GUIDE_HID + nt * 16 + 4 * kg) =
                        make_uint2(lo, hi);
                }
        }
        f4v o[GUIDE_OUT / 16];
        infer_layer<4, GUIDE_HID, GUIDE_OUT, false>(sw, sb, h[3], o, r, kg);
        // ---- loss head: one lane per query ----
#pragma unroll
        for (int nt = 0; nt < GUIDE_OUT / 16; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i) hw[r * GUIDE_OUT + nt * 16 + 4 * kg + i] = o[nt][i];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (kg == 0)
        {
            float g[GUIDE_OUT];
            if (valid)
            {
                const float scale = wsum > 0.0f ? A.weight[qi] / wsum : 0.0f;
                const f3 d = mk(A.dir[3 * size_t(qi)], A.dir[3 * size_t(qi) + 1], A.dir[3 * size_t(qi) + 2]);
                loss += guide_nll_grad(hw + r * GUIDE_OUT, d, scale, g) * scale;
            }
            else
                for (int k = 0; k < GUIDE_OUT; ++k) g[k] = 0.0f;
            for (int k = 0; k < GUIDE_OUT; ++k) hw[r * GUIDE_OUT + k] = g[k];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // dZ5 in the tile layout, rounded to BF16 as stored
        s4v d5[GUIDE_OUT / 16];
#pragma unroll
        for (int nt = 0; nt < GUIDE_OUT / 16; ++nt)
        {
            const float* gp = hw + r * GUIDE_OUT + nt * 16 + 4 * kg;
            const uint32_t lo = pack_bf2(gp[0], gp[1]), hi = pack_bf2(gp[2], gp[3]);
            d5[nt] = as_s4v(lo, hi);
            if (valid) *reinterpret_cast<uint2*>(A.dz + size_t(qi) * DZS + nt * 16 + 4 * kg) = make_uint2(lo, hi);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // ---- backward through layers 4 .. 1: dA[k][q] = sum_n W[n][k] dZ[n][q] ----
        s4v dcur[4];
        s4v dnext[4];
        auto back = [&](const uint16_t* W, int rowlen, int ksteps, const s4v* din, int hl) {
#pragma unroll
            for (int kt = 0; kt < 4; ++kt)
            {
                f4v acc = { 0.0f, 0.0f, 0.0f, 0.0f };
                for (int ks = 0; ks < ksteps; ++ks)
                {
                    // A[m = input unit k][kdim = output unit n] = W[n][k]
                    const uint16_t* wp = W + (ks * 16 + 4 * kg) * rowlen + kt * 16 + r;
                    s4v a;
                    a[0] = short(wp[0]);
                    a[1] = short(wp[rowlen]);
                    a[2] = short(wp[2 * rowlen]);
                    a[3] = short(wp[3 * rowlen]);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, din[ks], acc, 0, 0, 0);
                }
                // lane holds dA for input units 16 kt + 4 kg + i of query r; ReLU mask of the
                // hidden layer that produced them
                const s4v act = h[hl][kt];
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) v[i] = bf2f(uint16_t(act[i])) > 0.0f ? acc[i] : 0.0f;
                const uint32_t lo = pack_bf2(v[0], v[1]), hi = pack_bf2(v[2], v[3]);
                dnext[kt] = as_s4v(lo, hi);
                if (valid)
                    *reinterpret_cast<uint2*>(A.dz + size_t(qi) * DZS + GUIDE_OUT + hl * GUIDE_HID + kt * 16 + 4 * kg) =
                        make_uint2(lo, hi);
            }
        };
        back(sw + infer_off(4), infer_row(4), GUIDE_OUT / 16, d5, 3);
        for (int k = 0; k < 4; ++k) dcur[k] = dnext[k];
        back(sw + infer_off(3), infer_row(3), GUIDE_HID / 16, dcur, 2);
        for (int k = 0; k < 4; ++k) dcur[k] = dnext[k];
        back(sw + infer_off(2), infer_row(2), GUIDE_HID / 16, dcur, 1);
        for (int k = 0; k < 4; ++k) dcur[k] = dnext[k];
        back(sw + infer_off(1), infer_row(1), GUIDE_HID / 16, dcur, 0);
    }
    if (A.loss_sum)
    {
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off);
        if (lane == 0) atomicAdd(A.loss_sum, loss);
    }
}

// Loss head: dZ5 = d(-w log q)/d out, scaled by 1 / sum(w).
__global__ void k_guide_wsum(const float* weight, uint32_t M, float* scratch)
{
    float s = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) s += weight[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(&scratch[0], s);
}

// wsum: the minibatch's weight sum (device); loss_sum, when set, accumulates the
// weighted mean negative log-likelihood.
__global__ void k_guide_head(const float* out, const float* dir, const float* weight, uint32_t M,
                             const float* wsum_p, float* loss_sum, uint16_t* dz)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float nll = 0.0f;
    if (i < M)
    {
        const float wsum = *wsum_p;
        const float scale = wsum > 0.0f ? weight[i] / wsum : 0.0f;
        float g[GUIDE_OUT];
        nll = guide_nll_grad(out + size_t(i) * GUIDE_OUT, load3c(dir, i), scale, g) * scale;
        uint16_t* d = dz + size_t(i) * (GUIDE_OUT + 4 * GUIDE_HID);
        for (int k = 0; k < GUIDE_OUT; ++k) d[k] = f2bf(g[k]);
    }
    if (!loss_sum) return;
    for (int off = 32; off > 0; off >>= 1) nll += __shfl_down(nll, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss_sum, nll);
}

__global__ void k_guide_zero(float* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// Backward through the layers for 16-sample tiles: dA_{l-1} = dZ_l . W_l, dZ_{l-1} = dA (A > 0).
__global__ __launch_bounds__(BLOCK) void k_guide_bwd_data(const uint16_t* w16, const uint16_t* acts, uint16_t* dz,
                                                          uint32_t M)
{
    __shared__ __align__(16) uint16_t sw[WSZ];
    __shared__ __align__(16) uint16_t sd[WAVES][2][16 * GUIDE_HID];
    load_weights(sw, w16);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, kg = lane >> 4;
    constexpr int DZS = GUIDE_OUT + 4 * GUIDE_HID;
    for (uint32_t base = blockIdx.x * (16 * WAVES); base < M; base += gridDim.x * 16 * WAVES)
    {
        const uint32_t row0 = base + 16 * wave;
        const uint32_t valid = row0 < M ? min(16u, M - row0) : 0u;
        uint16_t* cur = sd[wave][0];
        uint16_t* nxt = sd[wave][1];
        for (int i = lane; i < 16 * GUIDE_OUT; i += 64)
        {
            const int rr = i / GUIDE_OUT;
            cur[i] = uint32_t(rr) < valid ? dz[size_t(row0 + rr) * DZS + (i % GUIDE_OUT)] : 0;
        }
        __syncthreads();
        // layer l (1-based 5..2): in width Kl = 64, out width Nl
#pragma unroll
        for (int l = 4; l >= 1; --l)
        {
            const int N = guide_n(l);          // gradient width of dZ_l
            const uint16_t* W = sw + guide_w_off(l);  // W[n][k], k = 64
#pragma unroll
            for (int nt = 0; nt < GUIDE_HID / 16; ++nt)
            {
                f4v acc = { 0.0f, 0.0f, 0.0f, 0.0f };
                for (int ks = 0; ks < N / 16; ++ks)
                {
                    const s4v a = load_a_row(cur + r * N + ks * 16 + 4 * kg);
                    // B[k = output unit][n = input unit] = W[k][n]
                    const s4v b = gather4(W + (ks * 16 + 4 * kg) * GUIDE_HID + nt * 16 + r, GUIDE_HID);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
                }
                const int col = nt * 16 + r;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                {
                    const int row = 4 * kg + i;
                    float v = 0.0f;
                    if (uint32_t(row) < valid)
                    {
                        const uint16_t act = acts[size_t(row0 + row) * (4 * GUIDE_HID) + (l - 1) * GUIDE_HID + col];
                        v = bf2f(act) > 0.0f ? acc[i] : 0.0f;
                    }
                    const uint16_t h = f2bf(v);
                    nxt[row * GUIDE_HID + col] = h;
                    if (uint32_t(row) < valid)
                        dz[size_t(row0 + row) * DZS + GUIDE_OUT + (l - 1) * GUIDE_HID + col] = h;
                }
            }
            __syncthreads();
            uint16_t* t = cur;
            cur = nxt;
            nxt = t;
        }
    }
}

// dW_l[n][k] = sum_s dZ_l[s][n] A_{l-1}[s][k], db_l[n] = sum_s dZ_l[s][n] over a chunk of
// samples per workgroup (blockIdx.y = layer); added to grad with one atomic per value.
// Weight and bias gradients of one layer over a sample range, written as
// this workgroup's partial sums (no atomics; k_guide_reduce adds the partials
// in a fixed order).  64 samples are staged per round; each wave owns up to
// six 16 x 16 output tiles and runs four MFMA k-steps (16 samples each) per
// tile and round.
constexpr int WG_MAX = 64;  // partial-sum slots
constexpr int WG_STEP = 64;
constexpr int SIN_LD = GUIDE_IN + 8, SDZ_LD = GUIDE_HID + 8;

__device__ __forceinline__ s4v gather4_ld(const uint16_t* base, int ld)
{
    s4v r;
    r[0] = short(base[0]);
    r[1] = short(base[ld]);
    r[2] = short(base[2 * ld]);
    r[3] = short(base[3 * ld]);
    return r;
}

__global__ __launch_bounds__(BLOCK) void k_guide_bwd_weight(const float* feat, const uint16_t* feat16,
                                                            const uint16_t* acts, const uint16_t* dz,
                                                            uint32_t M, uint32_t chunk, float* part)
{
    const int layer = blockIdx.y;
    const int K = guide_k(layer), N = guide_n(layer);
    __shared__ __align__(16) uint16_t s_in[WG_STEP * SIN_LD];
    __shared__ __align__(16) uint16_t s_dz[WG_STEP * SDZ_LD];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, kg = lane >> 4;
    constexpr int DZS = GUIDE_OUT + 4 * GUIDE_HID;
    const int ntile_k = K / 16, ntiles = (N / 16) * ntile_k;
    f4v acc[6];
    for (int t = 0; t < 6; ++t) acc[t] = f4v{ 0.0f, 0.0f, 0.0f, 0.0f };
    float bacc = 0.0f;
    const uint32_t s0 = blockIdx.x * chunk;
    const uint32_t s1 = min(M, s0 + chunk);
    const int dz_off = layer == 4 ? 0 : GUIDE_OUT + layer * GUIDE_HID;
    // rows of the layer's input and of its output gradient, moved as 8-BF16 vectors
    const int in_vec = K / 8, dz_vec = N / 8;
    for (uint32_t sb = s0; sb < s1; sb += WG_STEP)
    {
        __syncthreads();
        for (int i = threadIdx.x; i < WG_STEP * in_vec; i += BLOCK)
        {
            const int rr = i / in_vec, c = (i - rr * in_vec) * 8;
            const uint32_t smp = sb + rr;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (smp < s1)
            {
                if (layer > 0)
                    v = *reinterpret_cast<const uint4*>(acts + size_t(smp) * (4 * GUIDE_HID) + (layer - 1) * GUIDE_HID + c);
                else if (feat16)
                    v = *reinterpret_cast<const uint4*>(feat16 + size_t(smp) * GUIDE_IN + c);
                else
                {
                    const float* f = feat + size_t(smp) * GUIDE_IN + c;
                    v.x = uint32_t(f2bf(f[0])) | (uint32_t(f2bf(f[1])) << 16);
                    v.y = uint32_t(f2bf(f[2])) | (uint32_t(f2bf(f[3])) << 16);
                    v.z = uint32_t(f2bf(f[4])) | (uint32_t(f2bf(f[5])) << 16);
                    v.w = uint32_t(f2bf(f[6])) | (uint32_t(f2bf(f[7])) << 16);
                }
            }
            *reinterpret_cast<uint4*>(&s_in[rr * SIN_LD + c]) = v;
        }
        for (int i = threadIdx.x; i < WG_STEP * dz_vec; i += BLOCK)
        {
            const int rr = i / dz_vec, c = (i - rr * dz_vec) * 8;
            const uint32_t smp = sb + rr;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (smp < s1) v = *reinterpret_cast<const uint4*>(dz + size_t(smp) * DZS + dz_off + c);
            *reinterpret_cast<uint4*>(&s_dz[rr * SDZ_LD + c]) = v;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 6; ++t)
        {
            const int tile = wave + WAVES * t;
            if (tile >= ntiles) break;
            const int tn = tile / ntile_k, tk = tile - tn * ntile_k;
            // A[m = n unit][kdim = sample] = dZ[s][n]; B[kdim = sample][col = k unit] = in[s][k]
#pragma unroll
            for (int ks = 0; ks < WG_STEP / 16; ++ks)
            {
                const s4v a = gather4_ld(&s_dz[(16 * ks + 4 * kg) * SDZ_LD + tn * 16 + r], SDZ_LD);
                const s4v b = gather4_ld(&s_in[(16 * ks + 4 * kg) * SIN_LD + tk * 16 + r], SIN_LD);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
            }
        }
        if (threadIdx.x < N)
            for (int rr = 0; rr < WG_STEP; ++rr) bacc += bf2f(s_dz[rr * SDZ_LD + threadIdx.x]);
    }
    float* out = part + size_t(blockIdx.x) * GUIDE_PARAMS;
    for (int t = 0; t < 6; ++t)
    {
        const int tile = wave + WAVES * t;
        if (tile >= ntiles) break;
        const int tn = tile / ntile_k, tk = tile - tn * ntile_k;
        for (int i = 0; i < 4; ++i)
        {
            const int n = tn * 16 + 4 * kg + i;
            const int k = tk * 16 + r;
            out[guide_w_off(layer) + n * K + k] = acc[t][i];
        }
    }
    if (threadIdx.x < N) out[guide_b_off(layer) + threadIdx.x] = bacc;
}

struct AdamArgs
{
    float* w32;
    uint16_t* w16;
    float* m;
    float* v;
    float lr, b1, b2, eps, c1, c2;
};

// Sum of the partial gradients in slot order; then, when apply is set, the
// Adam step with FP32 master weights and the BF16 copy the kernels read.
__global__ void k_guide_reduce(const float* part, uint32_t G, float* grad, int apply, AdamArgs A)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= uint32_t(GUIDE_PARAMS)) return;
    float g = 0.0f;
    for (uint32_t k = 0; k < G; ++k) g += part[size_t(k) * GUIDE_PARAMS + i];
    grad[i] = g;
    if (!apply) return;
    const float mm = A.b1 * A.m[i] + (1.0f - A.b1) * g;
    const float vv = A.b2 * A.v[i] + (1.0f - A.b2) * g * g;
    A.m[i] = mm;
    A.v[i] = vv;
    const float p = A.w32[i] - A.lr * (mm / A.c1) / (sqrtf(vv / A.c2) + A.eps);
    A.w32[i] = p;
    if (i < uint32_t(WSZ)) A.w16[i] = f2bf(p);
}

__global__ void k_guide_sync16(const float* w32, uint16_t* w16)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < uint32_t(WSZ)) w16[i] = f2bf(w32[i]);
}

// Deterministic initialisation: He-uniform weights, zero biases; the last
// layer starts small so the initial mixture is near uniform.
void init_params(std::vector<float>& p, uint32_t seed)
{
    p.assign(GUIDE_PARAMS, 0.0f);
    uint64_t s = 0x9E3779B97F4A7C15ull ^ seed;
    auto rnd = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return float((z >> 11) * (1.0 / 9007199254740992.0));
    };
    for (int l = 0; l < GUIDE_LAYERS; ++l)
    {
        const float lim = (l == GUIDE_LAYERS - 1 ? 0.1f : 1.0f) * std::sqrt(6.0f / float(guide_k(l)));
        for (int i = guide_w_off(l); i < guide_w_off(l) + guide_k(l) * guide_n(l); ++i) p[i] = (2.0f * rnd() - 1.0f) * lim;
    }
}

inline uint32_t nb(uint32_t n, uint32_t b) { return n ? (n + b - 1) / b : 1; }

} // namespace

GuideNet* guide_of(Context* ctx)
{
    if (!ctx->guide) ctx->guide = new GuideNet;
    return ctx->guide;
}

int guide_ensure(Context* ctx, uint32_t cap)
{
    GuideNet* g = guide_of(ctx);
    if (!g->ready)
    {
        // allocated in full or not at all (a failure part-way frees what it got, so the next
        // call starts again instead of using a half-allocated network)
        auto reset = [g]() {
            for (void** p : { (void**)&g->w32, (void**)&g->w16, (void**)&g->m, (void**)&g->v, (void**)&g->grad,
                              (void**)&g->part, (void**)&g->scratch })
            {
                if (*p) (void)hipFree(*p);
                *p = nullptr;
            }
        };
        hipError_t e = hipSuccess;
        if (e == hipSuccess) e = hipMalloc(&g->w32, GUIDE_PARAMS * 4);
        if (e == hipSuccess) e = hipMalloc(&g->w16, WSZ * 2);
        if (e == hipSuccess) e = hipMalloc(&g->m, GUIDE_PARAMS * 4);
        if (e == hipSuccess) e = hipMalloc(&g->v, GUIDE_PARAMS * 4);
        if (e == hipSuccess) e = hipMalloc(&g->grad, GUIDE_PARAMS * 4);
        if (e == hipSuccess) e = hipMalloc(&g->part, size_t(WG_MAX) * GUIDE_PARAMS * 4);
        if (e == hipSuccess) e = hipMalloc(&g->scratch, 64);
        std::vector<float> p;
        init_params(p, 1);
        if (e == hipSuccess) e = hipMemcpy(g->w32, p.data(), GUIDE_PARAMS * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(g->m, 0, GUIDE_PARAMS * 4);
        if (e == hipSuccess) e = hipMemset(g->v, 0, GUIDE_PARAMS * 4);
        if (e == hipSuccess)
        {
            k_guide_sync16<<<nb(WSZ, 256), 256>>>(g->w32, g->w16);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess)
        {
            reset();
            return ctx->fail(APG_E_HIP, std::string("guide network allocation: ") + hipGetErrorString(e));
        }
        g->ready = true;
    }
    if (cap > g->cap)
    {
        APG_HIP_CHECK(ctx, hipDeviceSynchronize());
        if (g->acts) (void)hipFree(g->acts);
        if (g->out) (void)hipFree(g->out);
        if (g->dz) (void)hipFree(g->dz);
        if (g->feat16) (void)hipFree(g->feat16);
        g->acts = nullptr; g->out = nullptr; g->dz = nullptr; g->feat16 = nullptr;
        g->cap = 0;  // set again once all four are allocated
        APG_HIP_CHECK(ctx, hipMalloc(&g->feat16, size_t(cap) * GUIDE_IN * 2));
        APG_HIP_CHECK(ctx, hipMalloc(&g->acts, size_t(cap) * 4 * GUIDE_HID * 2));
        APG_HIP_CHECK(ctx, hipMalloc(&g->out, size_t(cap) * GUIDE_OUT * 4));
        APG_HIP_CHECK(ctx, hipMalloc(&g->dz, size_t(cap) * (GUIDE_OUT + 4 * GUIDE_HID) * 2));
        g->cap = cap;
    }
    return APG_OK;
}

void guide_free(GuideNet* g)
{
    if (!g) return;
    for (void* p : { (void*)g->w32, (void*)g->w16, (void*)g->m, (void*)g->v, (void*)g->grad, (void*)g->part, (void*)g->scratch,
                     (void*)g->acts, (void*)g->out, (void*)g->dz, (void*)g->feat16 })
        if (p) (void)hipFree(p);
    delete g;
}

int guide_forward(Context* ctx, const float* feat, uint32_t M, const uint32_t* M_dev, float* out,
                  uint16_t* acts, hipStream_t s)
{
    GuideNet* g = guide_of(ctx);
    FwdArgs A{ g->w16, g->w32, feat, M, M_dev, out, acts };
    const uint32_t grid = std::min<uint32_t>(nb(M, 16 * WAVES), uint32_t(ctx->cu_count) * 4);
    k_guide_fwd<<<grid, BLOCK, 0, s>>>(A);
    APG_HIP_CHECK(ctx, hipGetLastError());
    return APG_OK;
}

// Forward + backward into g->grad (zeroed first); optionally the Adam step.
int guide_infer(Context* ctx, const float* q9, uint32_t M, const uint32_t* M_dev, float* out, hipStream_t s)
{
    int rc = guide_ensure(ctx, 1);
    if (rc) return rc;
    GuideNet* g = guide_of(ctx);
    InferArgs A{ g->w16, g->w32, q9, M, M_dev, out };
    const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(nb(M, 16 * WAVES), uint32_t(ctx->cu_count) * 3));
    k_guide_infer<<<grid, BLOCK, 0, s>>>(A);
    APG_HIP_CHECK(ctx, hipGetLastError());
    return APG_OK;
}

int guide_train(Context* ctx, const float* feat, const float* dir, const float* weight, uint32_t M,
                bool adam, float* loss, hipStream_t s, const float* wsum_dev, const float* q9)
{
    int rc = guide_ensure(ctx, M);
    if (rc) return rc;
    GuideNet* g = guide_of(ctx);
    // scratch[0]: weight sum when the caller has none, scratch[1]: loss sum
    if (loss || !wsum_dev) k_guide_zero<<<1, 64, 0, s>>>(g->scratch, 2);
    if (!wsum_dev)
    {
        k_guide_wsum<<<std::min<uint32_t>(nb(M, 256), 1024), 256, 0, s>>>(weight, M, g->scratch);
        wsum_dev = g->scratch;
    }
    const uint16_t* feat16 = nullptr;
    if (q9)
    {
        // queries: encoding, forward, head and backward-data in one kernel
        TrainArgs A{ g->w16, g->w32, q9, dir, weight, wsum_dev, M, g->feat16, g->acts, g->dz,
                     loss ? g->scratch + 1 : nullptr };
        const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(nb(M, 16 * WAVES), uint32_t(ctx->cu_count) * 2));
        k_guide_train_fused<<<grid, BLOCK, 0, s>>>(A);
        feat16 = g->feat16;
    }
    else
    {
        if ((rc = guide_forward(ctx, feat, M, nullptr, g->out, g->acts, s))) return rc;
        k_guide_head<<<nb(M, 256), 256, 0, s>>>(g->out, dir, weight, M, wsum_dev, loss ? g->scratch + 1 : nullptr,
                                                g->dz);
        const uint32_t grid = std::min<uint32_t>(nb(M, 16 * WAVES), uint32_t(ctx->cu_count) * 4);
        k_guide_bwd_data<<<grid, BLOCK, 0, s>>>(g->w16, g->acts, g->dz, M);
    }
    const uint32_t G = std::max<uint32_t>(1u, std::min<uint32_t>(WG_MAX, nb(M, WG_STEP)));
    const uint32_t chunk = nb(nb(M, G), WG_STEP) * WG_STEP;
    k_guide_bwd_weight<<<dim3(G, GUIDE_LAYERS), BLOCK, 0, s>>>(feat, feat16, g->acts, g->dz, M, chunk, g->part);
    AdamArgs A{ g->w32, g->w16, g->m, g->v, g->lr, 0.9f, 0.999f, 1e-8f, 1.0f, 1.0f };
    if (adam)
    {
        g->step++;
        A.c1 = 1.0f - std::pow(A.b1, float(g->step));
        A.c2 = 1.0f - std::pow(A.b2, float(g->step));
    }
    k_guide_reduce<<<nb(GUIDE_PARAMS, 256), 256, 0, s>>>(g->part, G, g->grad, adam ? 1 : 0, A);
    APG_HIP_CHECK(ctx, hipGetLastError());
    if (loss)
    {
        float h[2];
        APG_HIP_CHECK(ctx, hipMemcpyAsync(h, g->scratch, 8, hipMemcpyDeviceToHost, s));
        APG_HIP_CHECK(ctx, hipStreamSynchronize(s));
        *loss = h[1];
    }
    return APG_OK;
}

} // namespace apg

using namespace apg;

extern "C" {

int apg_guide_init(apg_context* ctx, uint32_t seed, float lr)
{
    if (!ctx) return APG_E_INVALID;
    APG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    int rc = guide_ensure(ctx, 1);
    if (rc) return rc;
    GuideNet* g = guide_of(ctx);
    std::vector<float> p;
    init_params(p, seed);
    APG_HIP_CHECK(ctx, hipMemcpy(g->w32, p.data(), GUIDE_PARAMS * 4, hipMemcpyHostToDevice));
    APG_HIP_CHECK(ctx, hipMemset(g->m, 0, GUIDE_PARAMS * 4));
    APG_HIP_CHECK(ctx, hipMemset(g->v, 0, GUIDE_PARAMS * 4));
    k_guide_sync16<<<nb(WSZ, 256), 256>>>(g->w32, g->w16);
    APG_HIP_CHECK(ctx, hipDeviceSynchronize());
    g->step = 0;
    g->lr = lr > 0.0f ? lr : 1e-3f;
    return APG_OK;
}

int apg_guide_infer(apg_context* ctx, const float* q, uint32_t n, float* out, void* stream)
{
    if (!ctx || (n && (!q || !out))) return APG_E_INVALID;
    if (n == 0) return APG_OK;
    return guide_infer(ctx, q, n, nullptr, out, static_cast<hipStream_t>(stream));
}

int apg_guide_params(apg_context* ctx, float* params, int write)
{
    if (!ctx || !params) return APG_E_INVALID;
    int rc = guide_ensure(ctx, 1);
    if (rc) return rc;
    GuideNet* g = guide_of(ctx);
    if (write)
    {
        APG_HIP_CHECK(ctx, hipMemcpy(g->w32, params, GUIDE_PARAMS * 4, hipMemcpyHostToDevice));
        k_guide_sync16<<<nb(WSZ, 256), 256>>>(g->w32, g->w16);
        APG_HIP_CHECK(ctx, hipDeviceSynchronize());
    }
    else
        APG_HIP_CHECK(ctx, hipMemcpy(params, g->w32, GUIDE_PARAMS * 4, hipMemcpyDeviceToHost));
    return APG_OK;
}

int apg_guide_forward(apg_context* ctx, const float* feat, uint32_t n, float* out, void* stream)
{
    if (!ctx || (n && (!feat || !out))) return APG_E_INVALID;
    int rc = guide_ensure(ctx, 1);
    if (rc || n == 0) return rc;
    return guide_forward(ctx, feat, n, nullptr, out, nullptr, static_cast<hipStream_t>(stream));
}

int apg_guide_train_step(apg_context* ctx, const float* feat, const float* dir, const float* weight,
                         uint32_t n, int apply, float* loss, float* grad_host, void* stream)
{
    if (!ctx || (n && (!feat || !dir || !weight))) return APG_E_INVALID;
    if (n == 0) return APG_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = guide_train(ctx, feat, dir, weight, n, apply != 0, loss, s, nullptr, nullptr);
    if (rc) return rc;
    if (grad_host)
    {
        APG_HIP_CHECK(ctx, hipStreamSynchronize(s));
        APG_HIP_CHECK(ctx, hipMemcpy(grad_host, guide_of(ctx)->grad, GUIDE_PARAMS * 4, hipMemcpyDeviceToHost));
    }
    return APG_OK;
}

int apg_guide_train_queries(apg_context* ctx, const float* q, const float* dir, const float* weight,
                            uint32_t n, int apply, float* loss, float* grad_host, void* stream)
{
    if (!ctx || (n && (!q || !dir || !weight))) return APG_E_INVALID;
    if (n == 0) return APG_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = guide_train(ctx, nullptr, dir, weight, n, apply != 0, loss, s, nullptr, q);
    if (rc) return rc;
    if (grad_host)
    {
        APG_HIP_CHECK(ctx, hipStreamSynchronize(s));
        APG_HIP_CHECK(ctx, hipMemcpy(grad_host, guide_of(ctx)->grad, GUIDE_PARAMS * 4, hipMemcpyDeviceToHost));
    }
    return APG_OK;
}

int apg_guide_param_count(void) { return GUIDE_PARAMS; }

} // extern "C"
