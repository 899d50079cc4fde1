// backward.hip — placeholder entry points filled in by the adjoint kernels.
#include "gsr_internal.h"

using namespace gsr;

extern "C" {

int gsr_backward(gsr_context* ctx, const float*, uint32_t, const gsr_camera*, const float*, float*, float*) {
    Context* c = reinterpret_cast<Context*>(ctx);
    return fail(c, GSR_ERR_STATE, "gsr_backward: not available in this build");
}

int gsr_render_batch(gsr_context* ctx, const float*, uint32_t, const gsr_camera*, uint32_t, const gsr_options*, int,
                     float*, uint64_t*) {
    Context* c = reinterpret_cast<Context*>(ctx);
    return fail(c, GSR_ERR_STATE, "gsr_render_batch: not available in this build");
}

}  // extern "C"
