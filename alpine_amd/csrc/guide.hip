// Online-trained path-guiding MLP on gfx950.  No counterpart in the reference; the proposal
// is mixed with BSDF sampling by one-sample MIS in the path tracer (render.hip), so plain
// path tracing stays the fallback.
//
// Layers 96 -> 64 -> 64 -> 64 -> 64 -> 48 (guide_core.h): BF16 weights and activations, FP32
// accumulation, every product on v_mfma_f32_16x16x32_bf16.  Operand layout (wave64): lane l
// holds A[l % 16][8 (l / 16) + i] and B[8 (l / 16) + i][l % 16] for i = 0..7, and the result
// D[4 (l / 16) + i][l % 16] for i = 0..3.
//
// A layer runs in the transposed form D[unit][sample] = W[unit][:] . act[:][sample]: weights
// are the A operand (LDS), a wave's activations the B operand (registers).  Lane (r, g) =
// (l % 16, l / 16) of a result tile holds units 4g .. 4g + 3 of sample r, so two consecutive
// result tiles concatenated per lane are a B operand of the next layer over 32 inputs in the
// order
//     position 32p + 8g + i  <->  unit 32p + 4g + i (i < 4), 32p + 16 + 4g + (i - 4) (i >= 4)
// (guide_pair_unit); the next layer's weight columns are stored in that order and
// activations never leave registers.  The 96 inputs are computed in registers as well: lane
// group g evaluates the sin / cos features of frequencies 3g .. 3g + 2 on every axis and six
// of the other 24 (guide_slot_feature), and the layer-0 columns follow that assignment.  The
// bias is the accumulator's initial value and ReLU one packed 16-bit max on the BF16 pair.
//
// The weights reach LDS as images (rows padded by 8 BF16 so the rows one fragment read
// touches fall in distinct banks, then the FP32 biases) that k_guide_pack writes from the
// FP32 master weights after every update; a workgroup copies its image whole.
//
// Training computes a minibatch's gradient in one kernel.  A workgroup takes 128 samples at
// a time and each wave carries 32 of them (two 16-sample tiles sharing every weight fragment)
// through the forward pass, the loss head (a sample's four lane groups take two mixture lobes
// each) and the backward pass.  Layer by layer, the layer's inputs and output gradients for
// the 128 samples are staged in LDS (unit-major, samples contiguous) and the four waves add
// the weight gradient dW[n][k] = sum_s dZ[n][s] A[k][s] into MFMA accumulators, each wave
// owning fixed 16 x 16 tiles over every sample its workgroup sees; the bias gradient is the
// same product with a B operand of ones.  Each workgroup writes its sums once and
// k_guide_reduce adds them in workgroup order (deterministic), divides by the minibatch's
// weight sum and takes the Adam step.
#include "guide_core.h"
#include "apg_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace apg {

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

// Position j (0..63) of a hidden layer's input -> unit of the previous layer's output.
SHD constexpr int guide_pair_unit(int j)
{
    return 32 * (j >> 5) + ((j & 7) < 4 ? 4 * ((j & 31) >> 3) + (j & 7) : 16 + 4 * ((j & 31) >> 3) + (j & 7) - 4);
}
// Input slot s (0..23) of lane group g -> feature index (guide_feature): slots 0..17 are the
// sin / cos pairs of frequency 3g + s / 6 on axis (s % 6) / 2, slots 18..23 six of the
// position, normal, direction and zero-padding features.
SHD constexpr int guide_slot_feature(int g, int s)
{
    return s < 18 ? 3 + 24 * ((s % 6) / 2) + 2 * (3 * g + s / 6) + (s & 1)
                  : (g == 0 ? (s < 21 ? s - 18 : 75 + s - 21)
                            : (g == 1 ? (s < 21 ? 78 + s - 18 : 81 + s - 21) : (g == 2 ? 84 + s - 18 : 90 + s - 18)));
}
// Position j (0..95) of the layer-0 input -> feature index.
SHD constexpr int guide_in_feature(int j) { return guide_slot_feature((j & 31) >> 3, 8 * (j >> 5) + (j & 7)); }
// Position u (0..47) of the output layer in the training layout -> output index: lane group
// g's twelve result values are the outputs of lobes 2g, 2g + 1 and two unused outputs.
SHD constexpr int guide_head_output(int u)
{
    return 4 * (u >> 4) + (u & 3) < 10 ? 10 * ((u & 15) >> 2) + 4 * (u >> 4) + (u & 3)
                                         : 40 + 2 * ((u & 15) >> 2) + 4 * (u >> 4) + (u & 3) - 10;
}

namespace {

constexpr int WAVES = 4;
constexpr int BLOCK = 64 * WAVES;
constexpr int WSZ = GUIDE_B_OFF[0];        // weights
constexpr int NBIAS = GUIDE_PARAMS - WSZ;  // biases
constexpr int LPAD = 8;                    // BF16 padding per LDS row
constexpr int FROW0 = GUIDE_IN + LPAD, FROW = GUIDE_HID + LPAD;
SHD constexpr int fw_row(int l) { return l == 0 ? FROW0 : FROW; }
SHD constexpr int fw_off(int l) { return l == 0 ? 0 : GUIDE_HID * FROW0 + (l - 1) * GUIDE_HID * FROW; }
constexpr int FW_ELEMS = fw_off(4) + GUIDE_OUT * FROW;  // forward weights, rows = output units
SHD constexpr int bw_off(int l) { return (l - 1) * GUIDE_HID * FROW; }
constexpr int BW_ELEMS = 4 * GUIDE_HID * FROW;          // W_l^T (l = 1..4), rows = input units
constexpr int IMG_BIAS = FW_ELEMS * 2;                  // byte offsets in an image
constexpr int IMG_BW = IMG_BIAS + NBIAS * 4;
constexpr int IMG_INFER_BYTES = IMG_BW;
constexpr int IMG_TRAIN_BYTES = IMG_BW + BW_ELEMS * 2;
static_assert(IMG_INFER_BYTES % 16 == 0 && IMG_TRAIN_BYTES % 16 == 0, "images are copied as 16-byte vectors");

// training: 2 tiles per wave, 128 samples per workgroup round, staged unit-major
constexpr int TQ = 2;
constexpr int TS = WAVES * 16 * TQ;
constexpr int SROW = TS + LPAD;                     // staging row (BF16)
constexpr int ST_DZ = GUIDE_IN;                     // first output-gradient row (inputs: 0..95)
constexpr int ST_ELEMS = (GUIDE_IN + GUIDE_HID) * SROW;
constexpr int PART_STRIDE = GUIDE_PARAMS + 4;       // per workgroup: gradient sums, weight sum, loss sum
constexpr int INFER_Q = 4;                          // 16-query tiles per wave (inference)

typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef uint32_t u4v __attribute__((ext_vector_type(4)));
typedef short s2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t cvt2(float a, float b)  // v_cvt_pk_bf16_f32 (nearest even)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f2v){ a, b }, bf2v));
}
__device__ __forceinline__ uint32_t relu2(uint32_t x)  // max(x, 0) of both BF16 halves
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s2v, x), (s2v){ 0, 0 }));
}
// Each 16-bit half of x kept where the same half of h (a ReLU output, >= 0) is non-zero:
// x * min(h, 1) per half.  Written as the two packed instructions (ALU only), since the
// compiler otherwise rewrites the product into per-half compares and selects.
__device__ __forceinline__ uint32_t relu_grad(uint32_t x, uint32_t h)
{
    uint32_t r;
    asm("v_pk_min_u16 %0, %1, %3\n\tv_pk_mul_lo_u16 %0, %2, %0" : "=&v"(r) : "v"(h), "v"(x), "s"(0x00010001u));
    return r;
}
__device__ __forceinline__ f4v mfma(u4v a, u4v b, f4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a), __builtin_bit_cast(bf8v, b), c, 0, 0, 0);
}
__device__ __forceinline__ u4v lds16(const uint16_t* p) { return *reinterpret_cast<const u4v*>(p); }

__device__ __forceinline__ void copy_image(uint8_t* lds, const uint8_t* img, int bytes)
{
    const uint4* src = reinterpret_cast<const uint4*>(img);
    uint4* dst = reinterpret_cast<uint4*>(lds);
    for (int i = threadIdx.x; i < bytes / 16; i += BLOCK) dst[i] = src[i];
}

// The 24 input slots of lane group g for one sample, as three B operands.
template <bool FEAT>
__device__ __forceinline__ void load_inputs(const float* q9, const float* feat, size_t src, bool valid, int g,
                                            u4v (&in)[3])
{
    float s[24];
    if (FEAT)
    {
#pragma unroll
        for (int k = 0; k < 24; ++k) s[k] = valid ? feat[src * GUIDE_IN + guide_slot_feature(g, k)] : 0.0f;
    }
    else
    {
        float p[3] = { 0, 0, 0 }, n[3] = { 0, 0, 0 }, wo[3] = { 0, 0, 0 };
        if (valid)
        {
            const float* qq = q9 + src * 9;
            p[0] = qq[0]; p[1] = qq[1]; p[2] = qq[2];
            n[0] = qq[3]; n[1] = qq[4]; n[2] = qq[5];
            wo[0] = qq[6]; wo[1] = qq[7]; wo[2] = qq[8];
        }
        // sin and cos of 2^k pi p for k = 3g as sin(2 pi frac(2^(k-1) p + c / 4)), as
        // guide_feature; k = 3g + 1, 3g + 2 by the double-angle formulas (each doubles the
        // absolute error, about 1e-7 here, far below the BF16 rounding of the inputs)
        const float sc = ldexpf(1.0f, 3 * g - 1);
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            const float x = p[a] * sc, xc = x + 0.25f;
            float sn = sin_rev(x - floorf(x)), cs = sin_rev(xc - floorf(xc));
#pragma unroll
            for (int j = 0; j < 3; ++j)
            {
                s[6 * j + 2 * a] = sn;
                s[6 * j + 2 * a + 1] = cs;
                const float s2 = 2.0f * sn * cs;
                cs = 1.0f - 2.0f * sn * sn;
                sn = s2;
            }
        }
#pragma unroll
        for (int t = 0; t < 3; ++t)
        {
            s[18 + t] = g == 0 ? p[t] : (g == 1 ? wo[t] : 0.0f);
            s[21 + t] = g == 0 ? n[t] : 0.0f;
        }
    }
#pragma unroll
    for (int kb = 0; kb < 3; ++kb)
        in[kb] = (u4v){ cvt2(s[8 * kb], s[8 * kb + 1]), cvt2(s[8 * kb + 2], s[8 * kb + 3]),
                        cvt2(s[8 * kb + 4], s[8 * kb + 5]), cvt2(s[8 * kb + 6], s[8 * kb + 7]) };
}

// One layer for Q tiles: acc[q][nt] = bias + W (16nt.., :) . in[q].
template <int Q, int KB, int NT>
__device__ __forceinline__ void layer_fwd(const uint16_t* W, int row, const float* bias, const u4v (&in)[Q][KB],
                                          f4v (&acc)[Q][NT], int r, int g)
{
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
    {
        const f4v b = *reinterpret_cast<const f4v*>(bias + 16 * nt + 4 * g);
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[q][nt] = b;
    }
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
        {
            const u4v a = lds16(W + (16 * nt + r) * row + 32 * kb + 8 * g);
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[q][nt] = mfma(a, in[q][kb], acc[q][nt]);
        }
}

// Hidden result tiles -> the next layer's B operands (ReLU, BF16).
template <int Q>
__device__ __forceinline__ void to_hidden(const f4v (&t)[Q][4], u4v (&h)[Q][2])
{
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int p = 0; p < 2; ++p)
            h[q][p] = (u4v){ relu2(cvt2(t[q][2 * p][0], t[q][2 * p][1])), relu2(cvt2(t[q][2 * p][2], t[q][2 * p][3])),
                             relu2(cvt2(t[q][2 * p + 1][0], t[q][2 * p + 1][1])),
                             relu2(cvt2(t[q][2 * p + 1][2], t[q][2 * p + 1][3])) };
}

// ---------------------------------------------------------------------------
// Weight images from the FP32 master parameters.
// ---------------------------------------------------------------------------
__global__ void k_guide_pack(const float* w32, uint8_t* infer, uint8_t* train)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < FW_ELEMS)
    {
        int l = 0;
        while (l < 4 && i >= fw_off(l + 1)) ++l;
        const int loc = i - fw_off(l), u = loc / fw_row(l), j = loc % fw_row(l), K = guide_k(l);
        float vi = 0.0f, vt = 0.0f;
        if (j < K)
        {
            const int k = l == 0 ? guide_in_feature(j) : guide_pair_unit(j);
            vi = w32[guide_w_off(l) + u * K + k];
            vt = w32[guide_w_off(l) + (l == 4 ? guide_head_output(u) : u) * K + k];
        }
        reinterpret_cast<uint16_t*>(infer)[i] = f2bf(vi);
        reinterpret_cast<uint16_t*>(train)[i] = f2bf(vt);
    }
    if (i < NBIAS)
    {
        const int l = i / GUIDE_HID, u = i - GUIDE_HID * l;
        reinterpret_cast<float*>(infer + IMG_BIAS)[i] = w32[WSZ + i];
        reinterpret_cast<float*>(train + IMG_BIAS)[i] = w32[WSZ + GUIDE_HID * l + (l == 4 ? guide_head_output(u) : u)];
    }
    if (i < BW_ELEMS)
    {
        // W_l^T for the backward pass: row = input unit k, column = position of the output unit
        const int l = 1 + i / (GUIDE_HID * FROW), loc = i % (GUIDE_HID * FROW), k = loc / FROW, j = loc % FROW;
        float v = 0.0f;
        if (j < GUIDE_HID)
        {
            const int u = guide_pair_unit(j);
            if (u < guide_n(l)) v = w32[guide_w_off(l) + (l == 4 ? guide_head_output(u) : u) * GUIDE_HID + k];
        }
        reinterpret_cast<uint16_t*>(train + IMG_BW)[i] = f2bf(v);
    }
}

// ---------------------------------------------------------------------------
// Inference: queries (p01, n, wo; 9 floats) or features (96 floats) -> 48 outputs.
// ---------------------------------------------------------------------------
struct InferArgs
{
    const uint8_t* img;
    const float* q;      // M x 9 (queries) ...
    const float* feat;   // ... or M x 96 (features)
    uint32_t M;
    const uint32_t* M_dev;
    float* out;          // M x 48
};

template <int Q, bool FEAT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_guide_infer(InferArgs A)
{
    __shared__ __align__(16) uint8_t lds[IMG_INFER_BYTES];
    copy_image(lds, A.img, IMG_INFER_BYTES);
    __syncthreads();
    const uint32_t M = A.M_dev ? *A.M_dev : A.M;
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const uint32_t step = gridDim.x * WAVES * Q;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform loop bounds
    for (uint32_t t0 = (blockIdx.x * WAVES + wave) * Q; t0 * 16u < M; t0 += step)
    {
        // weight fragments are read from LDS in every iteration rather than kept in registers:
        // the register budget then allows several waves per SIMD, whose conversions and
        // encoding overlap each other's MFMAs
        uint8_t* base = lds;
        asm volatile("" : "+v"(base));
        const uint16_t* sw = reinterpret_cast<const uint16_t*>(base);
        const float* sb = reinterpret_cast<const float*>(base + IMG_BIAS);
        u4v in0[Q][3];
#pragma unroll
        for (int q = 0; q < Q; ++q)
        {
            const uint32_t qi = (t0 + q) * 16u + uint32_t(r);
            load_inputs<FEAT>(A.q, A.feat, qi, qi < M, g, in0[q]);
        }
        f4v t[Q][4];
        u4v h[Q][2];
        layer_fwd<Q, 3, 4>(sw + fw_off(0), FROW0, sb, in0, t, r, g);
        to_hidden(t, h);
        layer_fwd<Q, 2, 4>(sw + fw_off(1), FROW, sb + GUIDE_HID, h, t, r, g);
        to_hidden(t, h);
        layer_fwd<Q, 2, 4>(sw + fw_off(2), FROW, sb + 2 * GUIDE_HID, h, t, r, g);
        to_hidden(t, h);
        layer_fwd<Q, 2, 4>(sw + fw_off(3), FROW, sb + 3 * GUIDE_HID, h, t, r, g);
        to_hidden(t, h);
        f4v o[Q][3];
        layer_fwd<Q, 2, 3>(sw + fw_off(4), FROW, sb + 4 * GUIDE_HID, h, o, r, g);
#pragma unroll
        for (int q = 0; q < Q; ++q)
        {
            const uint32_t qi = (t0 + q) * 16u + uint32_t(r);
            if (qi < M)
#pragma unroll
                for (int nt = 0; nt < 3; ++nt)
                    if (nt < 2 || g < 2)  // outputs 40..47 feed no lobe and are not written
                        *reinterpret_cast<float4*>(A.out + size_t(qi) * GUIDE_OUT + 16 * nt + 4 * g) =
                            make_float4(o[q][nt][0], o[q][nt][1], o[q][nt][2], o[q][nt][3]);
        }
    }
}

// ---------------------------------------------------------------------------
// Training.
// ---------------------------------------------------------------------------
struct TrainArgs
{
    const uint8_t* img;     // training image
    const float* q;         // queries (9 floats per input) ...
    const float* feat;      // ... or features (96)
    const float* dir;       // 3 per input
    const float* weight;    // 1 per input
    uint32_t M;             // minibatch size; sample i is input off + i * stride
    uint32_t off, stride;
    float* part;            // gridDim.x x PART_STRIDE
};

// Hardware transcendentals (v_exp_f32, v_log_f32, v_rcp_f32, v_rsq_f32; about 1 ulp) for the
// training head, whose result only steers the gradient.
__device__ __forceinline__ float hexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ float hlog(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }

// Loss head of one sample on its four lane groups: group g holds outputs 10g .. 10g + 9
// (lobes 2g and 2g + 1) in oo[0..9].  Returns -log q(d) and writes scale * d(-log q)/d out
// into gg in the same layout (gg[10], gg[11] = 0).  The terms are those of guide_nll_grad,
// with the sums over lobes taken across the lane groups, log pi from the log-softmax and
// log kappa = the clamped output.
__device__ __forceinline__ float head_grad(const float (&oo)[12], f3 d, float scale, float (&gg)[12])
{
    float amax = fmaxf(oo[0], oo[5]);
    amax = fmaxf(amax, __shfl_xor(amax, 16));
    amax = fmaxf(amax, __shfl_xor(amax, 32));
    float pi[2], kap[2], lv[2], cth[2], il[2], e2[2];
    bool nz[2];
    f3 mu[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) pi[t] = hexp(oo[5 * t] - amax);
    float sum = pi[0] + pi[1];
    sum += __shfl_xor(sum, 16);
    sum += __shfl_xor(sum, 32);
    const float isum = __builtin_amdgcn_rcpf(sum), lsum = hlog(sum);
    constexpr float kLog2Pi = 1.83787706640934548f, kLogTiny = -69.0775527898213705f;  // log(2 pi), log(1e-30)
#pragma unroll
    for (int t = 0; t < 2; ++t)
    {
        pi[t] *= isum;
        const float lk = clampk(oo[5 * t + 1]);
        kap[t] = hexp(lk);
        const f3 v = mk(oo[5 * t + 2], oo[5 * t + 3], oo[5 * t + 4]);
        const float l2 = dot(v, v);
        il[t] = __builtin_amdgcn_rsqf(l2);
        nz[t] = l2 * il[t] > 1e-6f;  // |v|
        mu[t] = nz[t] ? v * il[t] : mk(0.0f, 0.0f, 1.0f);
        cth[t] = dot(mu[t], d);
        e2[t] = hexp(-2.0f * kap[t]);
        // log pi + log vMF(kappa, c)
        lv[t] = fmaxf(oo[5 * t] - amax - lsum, kLogTiny) + lk - kLog2Pi - hlog(1.0f - e2[t]) + kap[t] * (cth[t] - 1.0f);
    }
    float lmax = fmaxf(lv[0], lv[1]);
    lmax = fmaxf(lmax, __shfl_xor(lmax, 16));
    lmax = fmaxf(lmax, __shfl_xor(lmax, 32));
    float s2 = hexp(lv[0] - lmax) + hexp(lv[1] - lmax);
    s2 += __shfl_xor(s2, 16);
    s2 += __shfl_xor(s2, 32);
    const float logq = lmax + hlog(s2);
#pragma unroll
    for (int t = 0; t < 2; ++t)
    {
        const float r = hexp(lv[t] - logq), k = kap[t];
        gg[5 * t] = scale * (pi[t] - r);
        // d log vMF / d kappa = 1/k - 2 e^{-2k} / (1 - e^{-2k}) + (c - 1); kappa = exp(o)
        const float dlk = __builtin_amdgcn_rcpf(k) - 2.0f * e2[t] * __builtin_amdgcn_rcpf(1.0f - e2[t]) + (cth[t] - 1.0f);
        const float ok = oo[5 * t + 1];
        gg[5 * t + 1] = ok > GUIDE_LOGK_MIN && ok < GUIDE_LOGK_MAX ? scale * (-r * dlk * k) : 0.0f;
        // d/d mu = -r k d, through mu = v / |v|
        const f3 dmu = d * (-r * k);
        const f3 dv = (dmu - mu[t] * dot(mu[t], dmu)) * (nz[t] ? scale * il[t] : 0.0f);
        gg[5 * t + 2] = dv.x;
        gg[5 * t + 3] = dv.y;
        gg[5 * t + 4] = dv.z;
    }
    gg[10] = gg[11] = 0.0f;
    return -logq;
}

// Staging: one dword per unit holds the unit's values of the lane's two samples (tiles 0, 1);
// samples 32 w + 2 r + q of the workgroup round sit at column 32 w + 2 r + q.
__device__ __forceinline__ uint32_t pair_word(uint32_t a, uint32_t b, int half)
{
    return half ? ((a >> 16) | (b & 0xFFFF0000u)) : ((a & 0xFFFFu) | (b << 16));
}

// Result-tile operands x[q][p] (units 16 tile + 4g + i, tiles < NT) to rows row0 + unit.
template <int NT>
__device__ __forceinline__ void stage_tiles(uint32_t* st32, int row0, const u4v (&x)[TQ][2], int g, int col_dw)
{
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int e = 0; e < 8; ++e)
        {
            const int tile = 2 * p + (e >> 2);
            if (tile < NT)
                st32[(row0 + 16 * tile + 4 * g + (e & 3)) * (SROW / 2) + col_dw] =
                    pair_word(x[0][p][e >> 1], x[1][p][e >> 1], e & 1);
        }
}

// Layer-0 inputs (positions 32 kb + 8g + e) to rows 0..95.
__device__ __forceinline__ void stage_inputs(uint32_t* st32, const u4v (&x)[TQ][3], int g, int col_dw)
{
#pragma unroll
    for (int kb = 0; kb < 3; ++kb)
#pragma unroll
        for (int e = 0; e < 8; ++e)
            st32[(32 * kb + 8 * g + e) * (SROW / 2) + col_dw] = pair_word(x[0][kb][e >> 1], x[1][kb][e >> 1], e & 1);
}

constexpr u4v kOnes = { 0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u };  // BF16 1.0

// Weight gradient of a layer whose n-tile w this wave owns (layers 0..3): KT input tiles.
template <int KT>
__device__ __forceinline__ void dw_rows(const uint16_t* st, f4v (&acc)[KT], f4v& accb, int w, int r, int g)
{
#pragma unroll
    for (int jb = 0; jb < TS / 32; ++jb)
    {
        const int col = 32 * jb + 8 * g;
        const u4v a = lds16(st + (ST_DZ + 16 * w + r) * SROW + col);
#pragma unroll
        for (int tk = 0; tk < KT; ++tk) acc[tk] = mfma(a, lds16(st + (16 * tk + r) * SROW + col), acc[tk]);
        accb = mfma(a, kOnes, accb);
    }
}

// Output layer: this wave owns input tile w and the three output tiles (bias: output tile w).
__device__ __forceinline__ void dw_out(const uint16_t* st, f4v (&acc)[3], f4v& accb, int w, int r, int g)
{
#pragma unroll
    for (int jb = 0; jb < TS / 32; ++jb)
    {
        const int col = 32 * jb + 8 * g;
        const u4v b = lds16(st + (16 * w + r) * SROW + col);
#pragma unroll
        for (int tn = 0; tn < 3; ++tn)
        {
            const u4v a = lds16(st + (ST_DZ + 16 * tn + r) * SROW + col);
            acc[tn] = mfma(a, b, acc[tn]);
            if (tn == w) accb = mfma(a, kOnes, accb);
        }
    }
}

// Backward through layer L: dZ_{L-1} = (W_L^T dZ_L) masked by the layer's input h > 0.
template <int L>
__device__ __forceinline__ void back(const uint16_t* swb, const u4v (&dz)[TQ][2], const u4v (&h)[TQ][2],
                                     u4v (&out)[TQ][2], int r, int g)
{
    f4v da[TQ][4];
#pragma unroll
    for (int q = 0; q < TQ; ++q)
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) da[q][kt] = (f4v){ 0.0f, 0.0f, 0.0f, 0.0f };
    const uint16_t* W = swb + bw_off(L);
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
        {
            const u4v a = lds16(W + (16 * kt + r) * FROW + 32 * p + 8 * g);
#pragma unroll
            for (int q = 0; q < TQ; ++q) da[q][kt] = mfma(a, dz[q][p], da[q][kt]);
        }
#pragma unroll
    for (int q = 0; q < TQ; ++q)
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                const int kt = 2 * p + (e >> 1), i0 = 2 * (e & 1);
                out[q][p][e] = relu_grad(cvt2(da[q][kt][i0], da[q][kt][i0 + 1]), h[q][p][e]);
            }
}

template <int L, int KT>
__device__ __forceinline__ void put_rows(float* P, const f4v (&acc)[KT], const f4v& accb, int w, int r, int g)
{
    constexpr int K = GUIDE_K[L];
#pragma unroll
    for (int tk = 0; tk < KT; ++tk)
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            const int kp = 16 * tk + r;
            P[guide_w_off(L) + (16 * w + 4 * g + i) * K + (L == 0 ? guide_in_feature(kp) : kp)] = acc[tk][i];
        }
    if (r == 0)
#pragma unroll
        for (int i = 0; i < 4; ++i) P[guide_b_off(L) + 16 * w + 4 * g + i] = accb[i];
}

template <bool FEAT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_guide_train(TrainArgs A)
{
    __shared__ __align__(16) uint8_t lds[IMG_TRAIN_BYTES];
    __shared__ __align__(16) uint16_t st[ST_ELEMS];
    __shared__ float red[2 * WAVES];
    copy_image(lds, A.img, IMG_TRAIN_BYTES);
    __syncthreads();
    const uint16_t* sw = reinterpret_cast<const uint16_t*>(lds);
    const float* sb = reinterpret_cast<const float*>(lds + IMG_BIAS);
    const uint16_t* swb = reinterpret_cast<const uint16_t*>(lds + IMG_BW);
    uint32_t* st32 = reinterpret_cast<uint32_t*>(st);
    // the wave index in a scalar register: branches on it (dw_out) must be scalar branches,
    // since an MFMA under a partial EXEC mask still runs on every lane
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int col_dw = 16 * w + r;
    const f4v z = { 0.0f, 0.0f, 0.0f, 0.0f };
    f4v aw0[6], aw1[4], aw2[4], aw3[4], aw4[3], ab[5];
#pragma unroll
    for (int k = 0; k < 6; ++k) aw0[k] = z;
#pragma unroll
    for (int k = 0; k < 4; ++k) aw1[k] = aw2[k] = aw3[k] = z;
#pragma unroll
    for (int k = 0; k < 3; ++k) aw4[k] = z;
#pragma unroll
    for (int k = 0; k < 5; ++k) ab[k] = z;
    float wsum = 0.0f, lsum = 0.0f;
    const uint32_t rounds = (A.M + TS - 1) / TS;
    for (uint32_t c = blockIdx.x; c < rounds; c += gridDim.x)
    {
        // ---- forward, this wave's 32 samples ----
        u4v in0[TQ][3];
        f3 dq[TQ];
        float wq[TQ];
        bool valid[TQ];
#pragma unroll
        for (int q = 0; q < TQ; ++q)
        {
            const uint32_t i = c * TS + 32u * w + 2u * r + q;
            valid[q] = i < A.M;
            const size_t src = size_t(A.off) + size_t(i) * A.stride;
            load_inputs<FEAT>(A.q, A.feat, src, valid[q], g, in0[q]);
            dq[q] = valid[q] ? load3c(A.dir, uint32_t(src)) : mk(0.0f, 0.0f, 0.0f);
            wq[q] = valid[q] ? A.weight[src] : 0.0f;
        }
        f4v t[TQ][4];
        u4v h0[TQ][2], h1[TQ][2], h2[TQ][2], h3[TQ][2];
        layer_fwd<TQ, 3, 4>(sw + fw_off(0), FROW0, sb, in0, t, r, g);
        to_hidden(t, h0);
        layer_fwd<TQ, 2, 4>(sw + fw_off(1), FROW, sb + GUIDE_HID, h0, t, r, g);
        to_hidden(t, h1);
        layer_fwd<TQ, 2, 4>(sw + fw_off(2), FROW, sb + 2 * GUIDE_HID, h1, t, r, g);
        to_hidden(t, h2);
        layer_fwd<TQ, 2, 4>(sw + fw_off(3), FROW, sb + 3 * GUIDE_HID, h2, t, r, g);
        to_hidden(t, h3);
        f4v o[TQ][3];
        layer_fwd<TQ, 2, 3>(sw + fw_off(4), FROW, sb + 4 * GUIDE_HID, h3, o, r, g);
        // ---- loss head: dZ of the output layer (weight w_i; 1 / sum w comes in the reduction) ----
        u4v dz[TQ][2], dn[TQ][2];
#pragma unroll
        for (int q = 0; q < TQ; ++q)
        {
            float oo[12], gg[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) oo[k] = o[q][k >> 2][k & 3];
            const float nll = head_grad(oo, dq[q], wq[q], gg);
            if (g == 0 && valid[q])
            {
                wsum += wq[q];
                lsum += wq[q] * nll;
            }
            dz[q][0] = (u4v){ cvt2(gg[0], gg[1]), cvt2(gg[2], gg[3]), cvt2(gg[4], gg[5]), cvt2(gg[6], gg[7]) };
            dz[q][1] = (u4v){ cvt2(gg[8], gg[9]), cvt2(gg[10], gg[11]), 0u, 0u };
        }
        // ---- layer 4 ----
        __syncthreads();  // the previous round's layer-0 gradient has read the staging buffer
        stage_tiles<4>(st32, 0, h3, g, col_dw);
        stage_tiles<3>(st32, ST_DZ, dz, g, col_dw);
        __syncthreads();
        dw_out(st, aw4, ab[4], w, r, g);
        back<4>(swb, dz, h3, dn, r, g);
        // ---- layers 3, 2, 1 ----
        __syncthreads();
        stage_tiles<4>(st32, 0, h2, g, col_dw);
        stage_tiles<4>(st32, ST_DZ, dn, g, col_dw);
        __syncthreads();
        dw_rows<4>(st, aw3, ab[3], w, r, g);
        back<3>(swb, dn, h2, dz, r, g);
        __syncthreads();
        stage_tiles<4>(st32, 0, h1, g, col_dw);
        stage_tiles<4>(st32, ST_DZ, dz, g, col_dw);
        __syncthreads();
        dw_rows<4>(st, aw2, ab[2], w, r, g);
        back<2>(swb, dz, h1, dn, r, g);
        __syncthreads();
        stage_tiles<4>(st32, 0, h0, g, col_dw);
        stage_tiles<4>(st32, ST_DZ, dn, g, col_dw);
        __syncthreads();
        dw_rows<4>(st, aw1, ab[1], w, r, g);
        back<1>(swb, dn, h0, dz, r, g);
        // ---- layer 0 ----
        __syncthreads();
        stage_inputs(st32, in0, g, col_dw);
        stage_tiles<4>(st32, ST_DZ, dz, g, col_dw);
        __syncthreads();
        dw_rows<6>(st, aw0, ab[0], w, r, g);
    }
    // ---- this workgroup's sums ----
    float* P = A.part + size_t(blockIdx.x) * PART_STRIDE;
    put_rows<0, 6>(P, aw0, ab[0], w, r, g);
    put_rows<1, 4>(P, aw1, ab[1], w, r, g);
    put_rows<2, 4>(P, aw2, ab[2], w, r, g);
    put_rows<3, 4>(P, aw3, ab[3], w, r, g);
#pragma unroll
    for (int tn = 0; tn < 3; ++tn)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            P[guide_w_off(4) + guide_head_output(16 * tn + 4 * g + i) * GUIDE_HID + 16 * w + r] = aw4[tn][i];
    if (w < 3 && r == 0)
#pragma unroll
        for (int i = 0; i < 4; ++i) P[guide_b_off(4) + guide_head_output(16 * w + 4 * g + i)] = ab[4][i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        wsum += __shfl_xor(wsum, off);
        lsum += __shfl_xor(lsum, off);
    }
    if (lane == 0)
    {
        red[w] = wsum;
        red[WAVES + w] = lsum;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        float a = 0.0f, b = 0.0f;
        for (int k = 0; k < WAVES; ++k)
        {
            a += red[k];
            b += red[WAVES + k];
        }
        P[GUIDE_PARAMS] = a;
        P[GUIDE_PARAMS + 1] = b;
    }
}

struct AdamArgs
{
    float* w32;
    float* m;
    float* v;
    float lr, b1, b2, eps, c1, c2;
};

// Sum of the workgroup sums, divided by the minibatch's weight sum; then, when apply is set,
// the Adam step on the FP32 master weights.  A block takes RED_P consecutive entries; its
// threads add every RED_K-th workgroup's sums and the RED_K partial results are combined in
// a fixed order, so the gradient does not depend on timing.  Entry GUIDE_PARAMS is the
// weight sum, GUIDE_PARAMS + 1 the weighted loss: scratch[0] = weight sum, scratch[1] =
// weighted mean negative log-likelihood.
constexpr int RED_P = 64, RED_K = 4;
__global__ __launch_bounds__(RED_P * RED_K) void k_guide_reduce(const float* part, uint32_t G, float* grad, float* scratch,
                                                              int apply, AdamArgs A)
{
    __shared__ float sp[RED_K][RED_P];
    __shared__ float sw[RED_P * RED_K];
    const int t = threadIdx.x, p = t % RED_P, k0 = t / RED_P;
    // the weight sum, added in the same order by every block
    float a = 0.0f;
    for (uint32_t k = t; k < G; k += RED_P * RED_K) a += part[size_t(k) * PART_STRIDE + GUIDE_PARAMS];
    sw[t] = a;
    const uint32_t i = blockIdx.x * RED_P + p;
    float s = 0.0f;
    if (i < uint32_t(GUIDE_PARAMS + 2))
    {
#pragma unroll 8
        for (uint32_t k = k0; k < G; k += RED_K) s += part[size_t(k) * PART_STRIDE + i];
    }
    sp[k0][p] = s;
    __syncthreads();
    for (int h = RED_P * RED_K / 2; h > 0; h >>= 1)
    {
        if (t < h) sw[t] += sw[t + h];
        __syncthreads();
    }
    if (k0 != 0 || i >= uint32_t(GUIDE_PARAMS + 2)) return;
    const float ws = sw[0];
    const float inv = ws > 0.0f ? 1.0f / ws : 0.0f;
    float tot = sp[0][p];
#pragma unroll
    for (int k = 1; k < RED_K; ++k) tot += sp[k][p];
    if (i >= uint32_t(GUIDE_PARAMS))
    {
        if (i == uint32_t(GUIDE_PARAMS)) scratch[0] = ws;
        else scratch[1] = tot * inv;
        return;
    }
    const float gv = tot * inv;
    grad[i] = gv;
    if (!apply) return;
    const float mm = A.b1 * A.m[i] + (1.0f - A.b1) * gv;
    const float vv = A.b2 * A.v[i] + (1.0f - A.b2) * gv * gv;
    A.m[i] = mm;
    A.v[i] = vv;
    A.w32[i] = A.w32[i] - A.lr * (mm / A.c1) / (sqrtf(vv / A.c2) + A.eps);
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

struct GuideNet
{
    float* w32 = nullptr;        // master parameters (GUIDE_PARAMS)
    uint8_t* img_infer = nullptr; // LDS images (k_guide_pack)
    uint8_t* img_train = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    float* part = nullptr;       // G_max x PART_STRIDE workgroup sums
    float* scratch = nullptr;    // [0] weight sum, [1] loss of the last step
    uint32_t G_max = 0;
    uint32_t infer_blocks[2] = { 0, 0 };  // resident inference workgroups per CU (queries, features)
    bool ready = false;          // every buffer above is allocated and initialised
    uint32_t step = 0;
    float lr = 1e-3f;
};

GuideNet* guide_of(Context* ctx)
{
    if (!ctx->guide) ctx->guide = new GuideNet;
    return ctx->guide;
}

static void guide_pack(GuideNet* g, hipStream_t s)
{
    k_guide_pack<<<nb(FW_ELEMS, 256), 256, 0, s>>>(g->w32, g->img_infer, g->img_train);
}

// `cap` (training samples per step) is accepted for the callers' sizing; the network keeps no
// per-sample buffers.
int guide_ensure(Context* ctx, uint32_t cap)
{
    (void)cap;
    GuideNet* g = guide_of(ctx);
    if (g->ready) return APG_OK;
    // allocated in full or not at all (a failure part-way frees what it got, so the next call
    // starts again instead of using a half-allocated network)
    auto reset = [g]() {
        for (void** p : { (void**)&g->w32, (void**)&g->img_infer, (void**)&g->img_train, (void**)&g->m, (void**)&g->v,
                          (void**)&g->grad, (void**)&g->part, (void**)&g->scratch })
        {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
    };
    const uint32_t G = std::max(1, ctx->cu_count);
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMalloc(&g->w32, GUIDE_PARAMS * 4);
    if (e == hipSuccess) e = hipMalloc(&g->img_infer, IMG_INFER_BYTES);
    if (e == hipSuccess) e = hipMalloc(&g->img_train, IMG_TRAIN_BYTES);
    if (e == hipSuccess) e = hipMalloc(&g->m, GUIDE_PARAMS * 4);
    if (e == hipSuccess) e = hipMalloc(&g->v, GUIDE_PARAMS * 4);
    if (e == hipSuccess) e = hipMalloc(&g->grad, GUIDE_PARAMS * 4);
    if (e == hipSuccess) e = hipMalloc(&g->part, size_t(G) * PART_STRIDE * 4);
    if (e == hipSuccess) e = hipMalloc(&g->scratch, 64);
    std::vector<float> p;
    init_params(p, 1);
    if (e == hipSuccess) e = hipMemcpy(g->w32, p.data(), GUIDE_PARAMS * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(g->m, 0, GUIDE_PARAMS * 4);
    if (e == hipSuccess) e = hipMemset(g->v, 0, GUIDE_PARAMS * 4);
    if (e == hipSuccess) e = hipMemset(g->scratch, 0, 64);
    if (e == hipSuccess)
    {
        guide_pack(g, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess)
    {
        reset();
        return ctx->fail(APG_E_HIP, std::string("guide network allocation: ") + hipGetErrorString(e));
    }
    g->G_max = G;
    g->ready = true;
    return APG_OK;
}

void guide_free(GuideNet* g)
{
    if (!g) return;
    for (void* p : { (void*)g->w32, (void*)g->img_infer, (void*)g->img_train, (void*)g->m, (void*)g->v, (void*)g->grad,
                     (void*)g->part, (void*)g->scratch })
        if (p) (void)hipFree(p);
    delete g;
}

// Forward pass over M (or *M_dev) queries (q9) or feature vectors (feat).
static int guide_run(Context* ctx, const float* q9, const float* feat, uint32_t M, const uint32_t* M_dev, float* out,
                     hipStream_t s)
{
    int rc = guide_ensure(ctx, 1);
    if (rc) return rc;
    GuideNet* g = guide_of(ctx);
    InferArgs A{ g->img_infer, q9, feat, M, M_dev, out };
    // a persistent grid of the workgroups that fit at once (the weights are loaded once per
    // workgroup); with a device count the grid covers the capacity M
    const void* fn = feat ? reinterpret_cast<const void*>(&k_guide_infer<INFER_Q, true>)
                          : reinterpret_cast<const void*>(&k_guide_infer<INFER_Q, false>);
    uint32_t& per_cu = g->infer_blocks[feat ? 1 : 0];
    if (!per_cu)
    {
        int nblk = 0;
        APG_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, fn, BLOCK, 0));
        per_cu = uint32_t(std::max(1, nblk));
    }
    const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(nb(M, 16 * WAVES * INFER_Q), uint32_t(ctx->cu_count) * per_cu));
    if (feat)
        k_guide_infer<INFER_Q, true><<<grid, BLOCK, 0, s>>>(A);
    else
        k_guide_infer<INFER_Q, false><<<grid, BLOCK, 0, s>>>(A);
    APG_HIP_CHECK(ctx, hipGetLastError());
    return APG_OK;
}

int guide_forward(Context* ctx, const float* feat, uint32_t M, const uint32_t* M_dev, float* out, hipStream_t s)
{
    return guide_run(ctx, nullptr, feat, M, M_dev, out, s);
}

int guide_infer(Context* ctx, const float* q9, uint32_t M, const uint32_t* M_dev, float* out, hipStream_t s)
{
    return guide_run(ctx, q9, nullptr, M, M_dev, out, s);
}

// One training step on M samples, sample i being input off + i * stride of q9 (queries) or feat
// (features), dir and weight.  Enqueued on s; when loss is set, waits for the step and stores
// its loss there.  The loss also stays in device memory for guide_last_loss.
int guide_train(Context* ctx, const float* feat, const float* q9, const float* dir, const float* weight, uint32_t M,
                uint32_t off, uint32_t stride, bool adam, float* loss, hipStream_t s)
{
    int rc = guide_ensure(ctx, M);
    if (rc) return rc;
    GuideNet* g = guide_of(ctx);
    const uint32_t G = std::max<uint32_t>(1u, std::min<uint32_t>(g->G_max, nb(M, TS)));
    TrainArgs T{ g->img_train, q9, feat, dir, weight, M, off, stride, g->part };
    if (feat)
        k_guide_train<true><<<G, BLOCK, 0, s>>>(T);
    else
        k_guide_train<false><<<G, BLOCK, 0, s>>>(T);
    AdamArgs A{ g->w32, g->m, g->v, g->lr, 0.9f, 0.999f, 1e-8f, 1.0f, 1.0f };
    if (adam)
    {
        g->step++;
        A.c1 = 1.0f - std::pow(A.b1, float(g->step));
        A.c2 = 1.0f - std::pow(A.b2, float(g->step));
    }
    k_guide_reduce<<<nb(GUIDE_PARAMS + 2, RED_P), RED_P * RED_K, 0, s>>>(g->part, G, g->grad, g->scratch, adam ? 1 : 0, A);
    if (adam) guide_pack(g, s);
    APG_HIP_CHECK(ctx, hipGetLastError());
    if (loss)
    {
        APG_HIP_CHECK(ctx, hipMemcpyAsync(loss, g->scratch + 1, 4, hipMemcpyDeviceToHost, s));
        APG_HIP_CHECK(ctx, hipStreamSynchronize(s));
    }
    return APG_OK;
}

// Loss of the last training step (waits for the device).
int guide_last_loss(Context* ctx, float* loss)
{
    GuideNet* g = ctx->guide;
    if (!g || !g->ready)
    {
        *loss = 0.0f;
        return APG_OK;
    }
    APG_HIP_CHECK(ctx, hipDeviceSynchronize());
    APG_HIP_CHECK(ctx, hipMemcpy(loss, g->scratch + 1, 4, hipMemcpyDeviceToHost));
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
    APG_HIP_CHECK(ctx, hipDeviceSynchronize());
    APG_HIP_CHECK(ctx, hipMemcpy(g->w32, p.data(), GUIDE_PARAMS * 4, hipMemcpyHostToDevice));
    APG_HIP_CHECK(ctx, hipMemset(g->m, 0, GUIDE_PARAMS * 4));
    APG_HIP_CHECK(ctx, hipMemset(g->v, 0, GUIDE_PARAMS * 4));
    guide_pack(g, nullptr);
    APG_HIP_CHECK(ctx, hipGetLastError());
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
    APG_HIP_CHECK(ctx, hipDeviceSynchronize());
    if (write)
    {
        APG_HIP_CHECK(ctx, hipMemcpy(g->w32, params, GUIDE_PARAMS * 4, hipMemcpyHostToDevice));
        guide_pack(g, nullptr);
        APG_HIP_CHECK(ctx, hipGetLastError());
        APG_HIP_CHECK(ctx, hipDeviceSynchronize());
    }
    else
        APG_HIP_CHECK(ctx, hipMemcpy(params, g->w32, GUIDE_PARAMS * 4, hipMemcpyDeviceToHost));
    return APG_OK;
}

int apg_guide_forward(apg_context* ctx, const float* feat, uint32_t n, float* out, void* stream)
{
    if (!ctx || (n && (!feat || !out))) return APG_E_INVALID;
    if (n == 0) return APG_OK;
    return guide_forward(ctx, feat, n, nullptr, out, static_cast<hipStream_t>(stream));
}

static int train_entry(apg_context* ctx, const float* feat, const float* q, const float* dir, const float* weight,
                       uint32_t n, int apply, float* loss, float* grad_host, void* stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = guide_train(ctx, feat, q, dir, weight, n, 0, 1, apply != 0, loss, s);
    if (rc) return rc;
    if (grad_host)
    {
        APG_HIP_CHECK(ctx, hipStreamSynchronize(s));
        APG_HIP_CHECK(ctx, hipMemcpy(grad_host, guide_of(ctx)->grad, GUIDE_PARAMS * 4, hipMemcpyDeviceToHost));
    }
    return APG_OK;
}

int apg_guide_train_step(apg_context* ctx, const float* feat, const float* dir, const float* weight,
                         uint32_t n, int apply, float* loss, float* grad_host, void* stream)
{
    if (!ctx || (n && (!feat || !dir || !weight))) return APG_E_INVALID;
    if (n == 0) return APG_OK;
    return train_entry(ctx, feat, nullptr, dir, weight, n, apply, loss, grad_host, stream);
}

int apg_guide_train_queries(apg_context* ctx, const float* q, const float* dir, const float* weight,
                            uint32_t n, int apply, float* loss, float* grad_host, void* stream)
{
    if (!ctx || (n && (!q || !dir || !weight))) return APG_E_INVALID;
    if (n == 0) return APG_OK;
    return train_entry(ctx, nullptr, q, dir, weight, n, apply, loss, grad_host, stream);
}

int apg_guide_param_count(void) { return GUIDE_PARAMS; }

} // extern "C"
