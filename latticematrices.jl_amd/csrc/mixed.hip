// mixed.hip -- batched / mixed-precision D^dagger D CG (placeholder until the
// mixed solver lands).
#include "lmx_internal.h"

namespace lmx {

size_t cg_batched_workspace_impl(lmx_wilson*, int, const lmx_cg_params*) { return 0; }

int cg_batched_impl(lmx_wilson*, double2*, const double2*, int, void*, size_t, const lmx_cg_params*,
                    lmx_cg_result*) {
    set_error("batched CG not built yet");
    return LMX_ERR_UNSUPPORTED;
}

}  // namespace lmx

extern "C" int lmx_autotune(lmx_wilson*, int, int, const char*, lmx_tune_result*) {
    lmx::set_error("autotuner not built yet");
    return LMX_ERR_UNSUPPORTED;
}
