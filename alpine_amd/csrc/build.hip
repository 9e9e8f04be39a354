// GPU binned-SAH builder (placeholder until the device builder lands).
#include "apg_internal.h"

namespace apg {
int build_scene_gpu(Context* ctx, const std::vector<TriPrep>&, uint32_t, BinaryTree&,
                    std::vector<ApgWideNode>&, std::vector<float>&, std::vector<uint32_t>&)
{
    return ctx->fail(APG_E_STATE, "GPU builder not available in this build");
}
}
