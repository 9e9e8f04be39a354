// Render-path state shared by render_host.cpp and render.hip.
#pragma once

#include "apg_internal.h"
#include "shade_core.h"

#include <vector>

namespace apg {

DevLight prepare_light(const apg_light& s);
void build_light_bvh(const std::vector<DevLight>& lights, std::vector<DevLightNode>& nodes,
                     std::vector<int32_t>& leaf_of);
void pack_light_bvh(const std::vector<DevLightNode>& nodes, std::vector<DevLightPair>& pairs,
                    std::vector<int32_t>& parent_slot, int32_t& root_light);
void build_alias(const std::vector<float>& w, std::vector<float>& p, std::vector<float>& q,
                 std::vector<uint32_t>& alias);
void camera_setup(const apg_camera& c, DevCamera& out);

// The shapes' shading attributes gathered into pools in global triangle order (the layout
// AttrView reads); `any` is false when no shape has normals or uvs.
struct AttrPools
{
    bool any = false, any_tangents = false;
    std::vector<DevTriAttr> tri;
    std::vector<float4> normals, tangents, tnormals, ttangents;
    std::vector<float> uvs, weights;
    std::vector<DevMorph> morphs;
};
void gather_attributes(const std::vector<Shape>& shapes, uint32_t total_tris, AttrPools& out);
// The morph weights of the shapes whose normals are morphed, in pool order.
void gather_weights(const std::vector<Shape>& shapes, std::vector<float>& out);

} // namespace apg
