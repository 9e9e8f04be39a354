// C ABI: shading attributes and morph targets of meshes.
//
//  - apg_mesh_set_shading keeps the Mesh::Data normals, tangents, uvs and index
//    buffers (shapes/mesh.h:16-33) that the render path interpolates at a hit as
//    Mesh::getIntersectionAttributes does (shapes/mesh.cpp:57-154; device code in
//    shade_core.h shading_attributes).
//  - apg_mesh_set_targets / apg_mesh_morph are Mesh::Data::Target and
//    Mesh::update (mesh.cpp:29-55): the blend runs on the GPU into the library's
//    device vertex copy, so a per-frame morph needs no host round trip; the next
//    apg_scene_build rebuilds from it.
#include "apg_internal.h"

#include <algorithm>
#include <cstring>

namespace apg {
namespace {

struct MorphWeights
{
    float w[APG_MAX_MORPH_TARGETS];
};

// Mesh::update (mesh.cpp:43-52): vertex i = base_i, then += dpos_j,i * w_j in target order.
__global__ void k_morph(float* xyz, const float* base, const float* dpos, MorphWeights W, uint32_t nv,
                        uint32_t ntargets)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    float x = base[3 * size_t(i)], y = base[3 * size_t(i) + 1], z = base[3 * size_t(i) + 2];
    for (uint32_t j = 0; j < ntargets; ++j)
    {
        const float* d = dpos + (size_t(j) * nv + i) * 3;
        const float w = W.w[j];
        x = x + d[0] * w;
        y = y + d[1] * w;
        z = z + d[2] * w;
    }
    xyz[3 * size_t(i)] = x;
    xyz[3 * size_t(i) + 1] = y;
    xyz[3 * size_t(i) + 2] = z;
}

void free_morph_device(Shape& sh)
{
    if (sh.d_base) (void)hipFree(sh.d_base);
    if (sh.d_tpos) (void)hipFree(sh.d_tpos);
    sh.d_base = sh.d_tpos = nullptr;
}

bool indices_below(const uint32_t* idx, size_t n, uint32_t bound)
{
    for (size_t i = 0; i < n; ++i)
        if (idx[i] >= bound) return false;
    return true;
}

} // namespace
} // namespace apg

using namespace apg;

extern "C" {

int apg_mesh_set_shading(apg_context* ctx, uint32_t shape_id, const apg_mesh_shading* s)
{
    if (!ctx || !s) return APG_E_INVALID;
    if (shape_id >= ctx->shapes.size()) return ctx->fail(APG_E_INVALID, "unknown shape id");
    Shape& sh = ctx->shapes[shape_id];
    const size_t nt = sh.idx.size() / 3;
    const uint32_t nv = uint32_t(sh.xyz.size() / 3);
    if (s->tangents && !s->normals) return ctx->fail(APG_E_INVALID, "tangents need normals");
    if (s->normals)
    {
        if (s->nn == 0) return ctx->fail(APG_E_INVALID, "normals given with nn = 0");
        if (s->normal_idx ? !indices_below(s->normal_idx, 3 * nt, s->nn) : nv > s->nn)
            return ctx->fail(APG_E_INVALID, "normal index out of range");
    }
    if (s->uvs)
    {
        if (s->nuv == 0) return ctx->fail(APG_E_INVALID, "uvs given with nuv = 0");
        if (s->uv_idx ? !indices_below(s->uv_idx, 3 * nt, s->nuv) : nv > s->nuv)
            return ctx->fail(APG_E_INVALID, "uv index out of range");
    }
    if (sh.ntargets && s->normals && s->nn * size_t(3) * sh.ntargets != sh.tnrm.size())
        return ctx->fail(APG_E_INVALID, "normal count differs from the morph targets' (set the targets again)");
    sh.normals.clear();
    sh.nidx.clear();
    sh.tangents.clear();
    sh.uvs.clear();
    sh.uvidx.clear();
    if (s->normals)
    {
        sh.normals.assign(s->normals, s->normals + size_t(s->nn) * 3);
        if (s->normal_idx) sh.nidx.assign(s->normal_idx, s->normal_idx + 3 * nt);
        if (s->tangents) sh.tangents.assign(s->tangents, s->tangents + size_t(s->nn) * 4);
    }
    if (s->uvs)
    {
        sh.uvs.assign(s->uvs, s->uvs + size_t(s->nuv) * 2);
        if (s->uv_idx) sh.uvidx.assign(s->uv_idx, s->uv_idx + 3 * nt);
    }
    ctx->attr_version++;
    return APG_OK;
}

int apg_mesh_set_targets(apg_context* ctx, uint32_t shape_id, uint32_t ntargets, const float* dpos,
                         const float* dnormals, const float* dtangents)
{
    if (!ctx) return APG_E_INVALID;
    if (shape_id >= ctx->shapes.size()) return ctx->fail(APG_E_INVALID, "unknown shape id");
    if (ntargets > APG_MAX_MORPH_TARGETS)
        return ctx->fail(APG_E_CAPACITY, "more than APG_MAX_MORPH_TARGETS morph targets");
    if (ntargets && !dpos) return ctx->fail(APG_E_INVALID, "null position deltas");
    if (int rc = ctx->sync_geometry_to_host()) return rc;
    if (ctx->last_launch) APG_HIP_CHECK(ctx, hipEventSynchronize(ctx->last_launch));
    Shape& sh = ctx->shapes[shape_id];
    free_morph_device(sh);
    const size_t nv = sh.xyz.size() / 3, nn = sh.normals.size() / 3;
    sh.ntargets = ntargets;
    sh.tpos.assign(dpos, dpos + (ntargets ? ntargets * nv * 3 : 0));
    sh.tnrm.assign(ntargets * nn * 3, 0.0f);
    sh.ttan.assign(ntargets * nn * 3, 0.0f);
    if (dnormals && nn) std::memcpy(sh.tnrm.data(), dnormals, sh.tnrm.size() * sizeof(float));
    if (dtangents && nn) std::memcpy(sh.ttan.data(), dtangents, sh.ttan.size() * sizeof(float));
    sh.weights.assign(ntargets, 0.0f);
    if (ntargets)
    {
        // the current vertices are Mesh::Data::vertices, the base of every blend
        APG_HIP_CHECK(ctx, hipMalloc(&sh.d_base, std::max<size_t>(1, nv) * 3 * sizeof(float)));
        APG_HIP_CHECK(ctx, hipMalloc(&sh.d_tpos, std::max<size_t>(1, sh.tpos.size()) * sizeof(float)));
        if (nv)
        {
            APG_HIP_CHECK(ctx, hipMemcpy(sh.d_base, sh.xyz.data(), nv * 3 * sizeof(float), hipMemcpyHostToDevice));
            APG_HIP_CHECK(ctx, hipMemcpy(sh.d_tpos, sh.tpos.data(), sh.tpos.size() * sizeof(float),
                                         hipMemcpyHostToDevice));
        }
    }
    ctx->attr_version++;
    return APG_OK;
}

int apg_mesh_morph(apg_context* ctx, uint32_t shape_id, const float* weights, void* stream)
{
    if (!ctx || !weights) return APG_E_INVALID;
    if (shape_id >= ctx->shapes.size()) return ctx->fail(APG_E_INVALID, "unknown shape id");
    Shape& sh = ctx->shapes[shape_id];
    if (sh.ntargets == 0) return ctx->fail(APG_E_STATE, "the shape has no morph targets");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (int rc = ctx->sync_geometry_to_device(s)) return rc;
    MorphWeights W;
    std::memset(&W, 0, sizeof(W));
    for (uint32_t j = 0; j < sh.ntargets; ++j) W.w[j] = weights[j];
    size_t vbase = 0;
    for (uint32_t k = 0; k < shape_id; ++k) vbase += ctx->shapes[k].xyz.size() / 3;
    const uint32_t nv = uint32_t(sh.xyz.size() / 3);
    if (nv)
        k_morph<<<(nv + 255) / 256, 256, 0, s>>>(ctx->d_xyz + 3 * vbase, sh.d_base, sh.d_tpos, W, nv, sh.ntargets);
    APG_HIP_CHECK(ctx, hipGetLastError());
    sh.weights.assign(weights, weights + sh.ntargets);
    ctx->host_stale = true;
    ctx->weights_dirty = true;
    return APG_OK;
}

} // extern "C"
