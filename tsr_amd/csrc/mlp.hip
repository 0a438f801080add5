// Learned node culling (no reference counterpart).
#include "handles.h"

using namespace tsr;

struct tsr_culler {
    tsr_context *ctx = nullptr;
};

namespace tsr {
void tsrCullerMaskImpl(tsr_context *, const tsr_culler *, const tsr_tree *, const tsr_scene *, const tsr_params &,
                       DBuf<uint8_t> &) {
    throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "learned culling is not available in this build");
}
}  // namespace tsr

extern "C" {
tsr_status tsr_culler_create(tsr_context *, const float *, float, tsr_culler **out) {
    if (out) *out = nullptr;
    return TSR_ERR_INVALID_ARGUMENT;
}
tsr_status tsr_culler_set_weights(tsr_context *, tsr_culler *, const float *, float) { return TSR_ERR_INVALID_ARGUMENT; }
void tsr_culler_destroy(tsr_culler *c) { delete c; }
tsr_status tsr_node_features(tsr_context *, const tsr_tree *, const tsr_scene *, const tsr_field *, float *) {
    return TSR_ERR_INVALID_ARGUMENT;
}
}
