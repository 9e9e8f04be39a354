// C ABI: host-memory, synchronous variants of the traversal entry points for
// per-ray callers.
//
// The reference calls Accelerator::intersect(const Ray&) and intersectAny
// (accelerator.h:33-35) one ray at a time from up to 256 tile threads
// (alpine.cpp:276-285).  These entry points take rays in alpine's Ray layout
// (org xyz, dir xyz; 24 B, ray.h:6-10) in host memory and return when the
// results are in host memory.  Concurrent calls are coalesced: a call appends
// its rays to the context's pending list (one for closest hit, one for any hit)
// and, while fewer than HOST_BATCHES_IN_FLIGHT batches of its kind are running,
// becomes the leader of a batch holding every pending call.  The leader packs
// the batch into a pooled pinned buffer, launches the batched kernel once on a
// pooled stream (the kernel reads the rays from and writes the hits to pinned
// host memory directly), copies each call's results into that call's own
// output buffers and wakes the callers; the others sleep on a futex word of
// their kind that every finished batch increments.  Calls that arrive while a
// batch runs form the next one, so batches grow to the number of calling
// threads without a timing window, and a lone caller is launched at once.
#include "apg_internal.h"
#include "../../include/apg_debug.h"

#include <linux/futex.h>
#include <sys/syscall.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

namespace apg {

int launch_trace_staged(Context* ctx, const apg_rays& rays, uint32_t n, const apg_hits* hits, bool any,
                        uint8_t* occ, uint32_t* flag_out, HostStaging* s);

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
        if (s->h_in) (void)hipHostFree(s->h_in);
        if (s->h_out) (void)hipHostFree(s->h_out);
        s->h_in = s->h_out = s->dev_in = s->dev_out = nullptr;
        s->cap = 0;
        if ((e = hipHostMalloc(&s->h_in, cap * 7 * sizeof(float))) != hipSuccess) return fail(e, "hipHostMalloc");
        // 4 result words per ray, then the overflow flag of the launch
        if ((e = hipHostMalloc(&s->h_out, (cap * 4 + 1) * sizeof(float))) != hipSuccess) return fail(e, "hipHostMalloc");
        if ((e = hipHostGetDevicePointer(reinterpret_cast<void**>(&s->dev_in), s->h_in, 0)) != hipSuccess ||
            (e = hipHostGetDevicePointer(reinterpret_cast<void**>(&s->dev_out), s->h_out, 0)) != hipSuccess)
            return fail(e, "hipHostGetDevicePointer");
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

constexpr uint32_t HOST_BATCHES_IN_FLIGHT = 8;  // per kind; each on its own stream and trace scratch
constexpr uint32_t HOST_BATCH_MAX = 1u << 22;    // rays a leader takes into one batch

// One caller's rays and result buffers (host memory, owned by the blocked caller).
struct HostRequest
{
    const float* rays6;
    const float* tmax;
    uint32_t n;
    uint32_t* prim;
    float* t;
    float* b0;
    float* b1;
    uint8_t* occ;
    int rc = APG_OK;
    // REQ_LEAD: the caller is handed the next batch; REQ_DONE: results and rc are in place.
    // The caller sleeps on this word.
    std::atomic<uint32_t> state{ 0 };
    // two other calls of the same batch whose callers this caller wakes once it sees REQ_DONE
    // (the batch's wake-ups fan out as a binary tree over the callers' threads)
    HostRequest* kids[2] = { nullptr, nullptr };
};

constexpr uint32_t REQ_LEAD = 1u;
constexpr uint32_t REQ_DONE = 2u;

// Trace one batch of requests of one kind (all with rc set on return).  The kernel reads the
// rays from and writes the hits and its overflow flag to the pinned staging buffers directly
// (SoA with the batch's size as stride), so a batch costs one launch and one synchronisation.
void run_batch(Context* ctx, const std::vector<HostRequest*>& batch, bool any)
{
    size_t total = 0;
    for (const HostRequest* r : batch) total += r->n;
    auto finish = [&](int rc) {
        for (HostRequest* r : batch) r->rc = rc;
    };
    int rc;
    HostStaging* s = ctx->take_staging(total, rc);
    if (!s) return finish(rc);
    const size_t n = total;
    float* h = s->h_in;
    size_t o = 0;
    for (const HostRequest* r : batch)
    {
        for (uint32_t i = 0; i < r->n; ++i)
            for (int k = 0; k < 6; ++k) h[k * n + o + i] = r->rays6[6 * size_t(i) + k];
        for (uint32_t i = 0; i < r->n; ++i) h[6 * n + o + i] = r->tmax ? r->tmax[i] : FLT_MAX;
        o += r->n;
    }
    float* d = s->dev_in;
    float* dout = s->dev_out;
    apg_rays rays{ d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, d + 6 * n };
    apg_hits hits{ reinterpret_cast<uint32_t*>(dout), dout + n, dout + 2 * n, dout + 3 * n };
    uint8_t* docc = reinterpret_cast<uint8_t*>(dout);
    float* hflag = s->h_out + 4 * s->cap;
    hipError_t e = hipSuccess;
    rc = launch_trace_staged(ctx, rays, uint32_t(n), any ? nullptr : &hits, any, any ? docc : nullptr,
                             reinterpret_cast<uint32_t*>(s->dev_out + 4 * s->cap), s);
    uint32_t* dflag = s->trace_words ? reinterpret_cast<uint32_t*>(s->trace_words) + 1 : nullptr;
    if (rc == APG_OK) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess)
    {
        (void)hipStreamSynchronize(s->stream);
        ctx->give_staging(s);
        return finish(ctx->fail(APG_E_HIP, std::string("staged trace: ") + hipGetErrorString(e)));
    }
    // the staging's overflow flag, read behind the results: set, it is reported (and cleared)
    // here for every call of the batch
    uint32_t overflow = 0;
    if (rc == APG_OK) std::memcpy(&overflow, hflag, 4);
    if (rc == APG_OK && overflow)
    {
        e = hipMemsetAsync(dflag, 0, 4, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        rc = e != hipSuccess ? ctx->fail(APG_E_HIP, std::string("staged trace: ") + hipGetErrorString(e))
                             : ctx->fail(APG_E_OVERFLOW, "a traversal needed more stack entries than the launch "
                                                         "allows (stack_limit); its results are not reliable");
    }
    if (rc == APG_OK)
    {
        const uint8_t* hocc = reinterpret_cast<const uint8_t*>(s->h_out);
        o = 0;
        for (HostRequest* r : batch)
        {
            if (any)
                std::memcpy(r->occ, hocc + o, r->n);
            else
            {
                std::memcpy(r->prim, s->h_out + o, r->n * sizeof(float));
                std::memcpy(r->t, s->h_out + n + o, r->n * sizeof(float));
                std::memcpy(r->b0, s->h_out + 2 * n + o, r->n * sizeof(float));
                std::memcpy(r->b1, s->h_out + 3 * n + o, r->n * sizeof(float));
            }
            o += r->n;
        }
    }
    ctx->give_staging(s);
    finish(rc);
}

void futex_wait(std::atomic<uint32_t>* word, uint32_t seen)
{
    (void)syscall(SYS_futex, reinterpret_cast<uint32_t*>(word), FUTEX_WAIT_PRIVATE, seen, nullptr, nullptr, 0);
}

// Wake the thread sleeping on a request's word.  The request may already have been seen done
// and its stack slot reused: the wake is then a spurious one that the waiter there re-checks.
void futex_wake_one(HostRequest* r)
{
    (void)syscall(SYS_futex, reinterpret_cast<uint32_t*>(&r->state), FUTEX_WAKE_PRIVATE, 1, nullptr, nullptr, 0);
}

// Lead batches of pending calls of one kind until none is left to take for this slot, then
// give the slot to the oldest pending call's thread (or free it).  Called with a leader slot
// held.
void lead_batches(Context* ctx, HostCoalescer& C, int kind)
{
    std::vector<HostRequest*> batch;
    {
        std::lock_guard<std::mutex> lock(C.mutex);
        auto& q = C.pending[kind];
        size_t rays = 0, take = 0;
        while (take < q.size() && (take == 0 || rays + static_cast<HostRequest*>(q[take])->n <= HOST_BATCH_MAX))
        {
            rays += static_cast<HostRequest*>(q[take])->n;
            batch.push_back(static_cast<HostRequest*>(q[take]));
            ++take;
        }
        q.erase(q.begin(), q.begin() + long(take));
    }
    if (!batch.empty())
    {
        run_batch(ctx, batch, kind == 1);
        const size_t m = batch.size();
        for (size_t i = 0; i < m; ++i)
        {
            batch[i]->kids[0] = 2 * i + 1 < m ? batch[2 * i + 1] : nullptr;
            batch[i]->kids[1] = 2 * i + 2 < m ? batch[2 * i + 2] : nullptr;
        }
        for (HostRequest* r : batch) r->state.fetch_or(REQ_DONE, std::memory_order_release);
        futex_wake_one(batch[0]);
    }
    HostRequest* next = nullptr;
    {
        std::lock_guard<std::mutex> lock(C.mutex);
        auto& q = C.pending[kind];
        if (q.empty())
            C.leaders[kind]--;
        else
        {
            // the slot passes to a caller still waiting to be taken (it takes itself along)
            next = static_cast<HostRequest*>(q.front());
            next->state.fetch_or(REQ_LEAD, std::memory_order_release);
        }
    }
    if (next) futex_wake_one(next);
}

int trace_host(Context* ctx, const float* rays6, const float* tmax, uint32_t n, bool any, uint32_t* prim,
               float* t, float* b0, float* b1, uint8_t* occ)
{
    if (n == 0) return APG_OK;
    if (!rays6 || (!any && (!prim || !t || !b0 || !b1)) || (any && !occ))
        return ctx->fail(APG_E_INVALID, "null ray or result pointer");
    HostRequest req{ rays6, tmax, n, prim, t, b0, b1, occ };
    HostCoalescer& C = ctx->host_batches;
    const int kind = any ? 1 : 0;
    bool lead;
    {
        std::lock_guard<std::mutex> lock(C.mutex);
        C.pending[kind].push_back(&req);
        lead = C.leaders[kind] < HOST_BATCHES_IN_FLIGHT;
        if (lead) C.leaders[kind]++;
    }
    if (lead) lead_batches(ctx, C, kind);
    for (;;)
    {
        uint32_t st = req.state.load(std::memory_order_acquire);
        // a handed-over slot is used before anything else, so that no slot is lost
        if (st & REQ_LEAD)
        {
            req.state.fetch_and(~REQ_LEAD, std::memory_order_acq_rel);
            lead_batches(ctx, C, kind);
            continue;
        }
        if (st & REQ_DONE) break;
        futex_wait(&req.state, st);
    }
    for (HostRequest* k : req.kids)
        if (k) futex_wake_one(k);
    return req.rc;
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

extern "C" {

// Per-ray callers measured (include/apg_debug.h): nthreads threads, each over a contiguous share
// of the rays, calling apg_intersect_host (any = 0) or apg_occluded_host (any = 1) with
// rays_per_call rays per call, as alpine's tile tasks call Accelerator::intersect per ray.
int apg_debug_host_rate(apg_context* ctx, const float* rays6, uint32_t n, uint32_t nthreads,
                        uint32_t rays_per_call, int any, uint32_t* prim, float* t, uint8_t* occluded,
                        double* seconds)
{
    if (!ctx || (n && !rays6) || !nthreads || !rays_per_call || (any ? !occluded : (!prim || !t)))
        return APG_E_INVALID;
    std::vector<float> b(any ? 0 : 2 * size_t(n));
    std::vector<int> rcs(nthreads, APG_OK);
    std::atomic<uint32_t> ready{ 0 };
    std::atomic<bool> go{ false };
    auto work = [&](uint32_t k) {
        const uint32_t chunk = (n + nthreads - 1) / nthreads;
        const uint32_t lo = std::min(n, k * chunk), hi = std::min(n, lo + chunk);
        ready++;
        while (!go.load()) std::this_thread::yield();
        for (uint32_t i = lo; i < hi && rcs[k] == APG_OK; i += rays_per_call)
        {
            const uint32_t m = std::min(rays_per_call, hi - i);
            rcs[k] = any ? apg_occluded_host(ctx, rays6 + 6 * size_t(i), nullptr, m, occluded + i)
                         : apg_intersect_host(ctx, rays6 + 6 * size_t(i), m, prim + i, t + i, b.data() + i,
                                              b.data() + n + i);
        }
    };
    std::vector<std::thread> th;
    th.reserve(nthreads);
    for (uint32_t k = 0; k < nthreads; ++k) th.emplace_back(work, k);
    while (ready.load() < nthreads) std::this_thread::yield();
    const auto t0 = std::chrono::steady_clock::now();
    go = true;
    for (auto& x : th) x.join();
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    for (int rc : rcs)
        if (rc != APG_OK) return rc;
    return APG_OK;
}

} // extern "C"
