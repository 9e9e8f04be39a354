// Render-path state shared by render_host.cpp and render.hip.
#pragma once

#include "apg_internal.h"
#include "shade_core.h"

#include <vector>

namespace apg {

DevLight prepare_light(const apg_light& s);
void build_light_bvh(const std::vector<DevLight>& lights, std::vector<DevLightNode>& nodes,
                     std::vector<int32_t>& leaf_of);
void build_alias(const std::vector<float>& w, std::vector<float>& p, std::vector<float>& q,
                 std::vector<uint32_t>& alias);
void camera_setup(const apg_camera& c, DevCamera& out);

} // namespace apg
