// C ABI: host-memory, synchronous variants of the traversal entry points for
// per-ray callers.
//
// The reference calls Accelerator::intersect(const Ray&) and intersectAny
// (accelerator.h:33-35) one ray at a time from up to 256 tile threads
// (alpine.cpp:276-285).  These entry points take rays in alpine's Ray layout
// (org xyz, dir xyz; 24 B, ray.h:6-10) in host memory, stage them through a
// pooled stream with pinned buffers (one per concurrently calling thread) and
// return when the results are in host memory.  The kernels are the batched
// ones; launches from all threads are ordered on the GPU by launch_trace.
#include "apg_internal.h"

#include <cfloat>
#include <cstring>

namespace apg {

int launch_trace(Context* ctx, const apg_rays* rays, uint32_t n, const apg_hits* hits, bool any,
                 uint8_t* occ, unsigned long long* counts, hipStream_t stream, const uint32_t* roots);

HostStaging* Context::take_staging(size_t n, int& rc)
{
    rc = APG_OK;
    HostStaging* s = nullptr;
    {
        std::lock_guard<std::mutex> lock(staging_mutex);
        if (!staging_free.empty())
        {
            s = staging_free.back();
            staging_free.pop_back();
        }
        else
        {
            s = new HostStaging;
            staging_all.push_back(s);
        }
    }
    auto fail = [&](hipError_t e, const char* what) {
        give_staging(s);
        rc = this->fail(APG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
        return static_cast<HostStaging*>(nullptr);
    };
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return fail(e, "hipSetDevice");
    if (!s->stream && (e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(e, "hipStreamCreateWithFlags");
    if (s->cap < n)
    {
        size_t cap = s->cap ? s->cap : 64;
        while (cap < n) cap *= 2;
        if (s->d_in) (void)hipFree(s->d_in);
        if (s->d_out) (void)hipFree(s->d_out);
        if (s->h_in) (void)hipHostFree(s->h_in);
        if (s->h_out) (void)hipHostFree(s->h_out);
        s->d_in = s->d_out = s->h_in = s->h_out = nullptr;
        s->cap = 0;
        if ((e = hipMalloc(&s->d_in, cap * 8 * sizeof(float))) != hipSuccess) return fail(e, "hipMalloc");
        if ((e = hipMalloc(&s->d_out, cap * 4 * sizeof(float))) != hipSuccess) return fail(e, "hipMalloc");
        if ((e = hipHostMalloc(&s->h_in, cap * 8 * sizeof(float))) != hipSuccess) return fail(e, "hipHostMalloc");
        // 4 result words per ray, then the overflow flag of the launch
        if ((e = hipHostMalloc(&s->h_out, (cap * 4 + 1) * sizeof(float))) != hipSuccess) return fail(e, "hipHostMalloc");
        s->cap = cap;
    }
    return s;
}

void Context::give_staging(HostStaging* s)
{
    std::lock_guard<std::mutex> lock(staging_mutex);
    staging_free.push_back(s);
}

namespace {

// Rays in alpine's layout (6 floats each) -> SoA in the staging buffer, traced on its stream.
int trace_host(Context* ctx, const float* rays6, const float* tmax, uint32_t n, bool any, uint32_t* prim,
               float* t, float* b0, float* b1, uint8_t* occ)
{
    if (n == 0) return APG_OK;
    if (!rays6 || (!any && (!prim || !t || !b0 || !b1)) || (any && !occ))
        return ctx->fail(APG_E_INVALID, "null ray or result pointer");
    int rc;
    HostStaging* s = ctx->take_staging(n, rc);
    if (!s) return rc;
    const size_t N = s->cap;
    float* h = s->h_in;
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 6; ++k) h[k * N + i] = rays6[6 * size_t(i) + k];
    for (uint32_t i = 0; i < n; ++i) h[6 * N + i] = tmax ? tmax[i] : FLT_MAX;
    float* d = s->d_in;
    apg_rays rays{ d, d + N, d + 2 * N, d + 3 * N, d + 4 * N, d + 5 * N, d + 6 * N };
    apg_hits hits{ reinterpret_cast<uint32_t*>(s->d_out), s->d_out + N, s->d_out + 2 * N, s->d_out + 3 * N };
    uint8_t* docc = reinterpret_cast<uint8_t*>(s->d_out);
    hipError_t e = hipSuccess;
    for (int k = 0; k < 7 && e == hipSuccess; ++k)
        e = hipMemcpyAsync(d + k * N, h + k * N, n * sizeof(float), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
    {
        rc = launch_trace(ctx, &rays, n, any ? nullptr : &hits, any, any ? docc : nullptr, nullptr, s->stream,
                          nullptr);
        if (rc == APG_OK)
        {
            if (any)
                e = hipMemcpyAsync(s->h_out, docc, n, hipMemcpyDeviceToHost, s->stream);
            else
                for (int k = 0; k < 4 && e == hipSuccess; ++k)
                    e = hipMemcpyAsync(s->h_out + k * N, s->d_out + k * N, n * sizeof(float),
                                       hipMemcpyDeviceToHost, s->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(s->h_out + 4 * N, ctx->counters + 1, 4, hipMemcpyDeviceToHost, s->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        }
    }
    if (e != hipSuccess)
    {
        (void)hipStreamSynchronize(s->stream);
        ctx->give_staging(s);
        return ctx->fail(APG_E_HIP, std::string("staged trace: ") + hipGetErrorString(e));
    }
    // the overflow flag read behind the results: set, it is reported (and cleared) here
    uint32_t overflow = 0;
    std::memcpy(&overflow, s->h_out + 4 * N, 4);
    if (rc == APG_OK && overflow) rc = ctx->take_overflow();
    if (rc == APG_OK)
    {
        if (any)
            std::memcpy(occ, s->h_out, n);
        else
        {
            std::memcpy(prim, s->h_out, n * sizeof(float));
            std::memcpy(t, s->h_out + N, n * sizeof(float));
            std::memcpy(b0, s->h_out + 2 * N, n * sizeof(float));
            std::memcpy(b1, s->h_out + 3 * N, n * sizeof(float));
        }
    }
    ctx->give_staging(s);
    return rc;
}

} // namespace
} // namespace apg

using namespace apg;

extern "C" {

int apg_intersect_host(apg_context* ctx, const float* rays6, uint32_t n, uint32_t* prim, float* t,
                       float* b0, float* b1)
{
    if (!ctx) return APG_E_INVALID;
    return trace_host(ctx, rays6, nullptr, n, false, prim, t, b0, b1, nullptr);
}

int apg_occluded_host(apg_context* ctx, const float* rays6, const float* tmax, uint32_t n,
                      uint8_t* occluded)
{
    if (!ctx) return APG_E_INVALID;
    return trace_host(ctx, rays6, tmax, n, true, nullptr, nullptr, nullptr, nullptr, occluded);
}

} // extern "C"
