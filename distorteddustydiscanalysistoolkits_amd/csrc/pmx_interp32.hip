// pmx_interp32.hip — FP32 compensated interpolation path (placeholder until implemented).
#include <hip/hip_runtime.h>

#include "pmx_internal.h"

namespace pmx {

int prepare_f32(pmx_ctx* c) { return set_error(c, PMX_ERR_ARGUMENT, "PMX_PREC_F32C is not available in this build"); }
int interp_points_f32c(pmx_ctx* c, int, const CatChunk&, const double*, const double*, const double*, int64_t, float*,
                       int64_t) {
    return set_error(c, PMX_ERR_ARGUMENT, "PMX_PREC_F32C is not available in this build");
}
int interp_lines_f32c(pmx_ctx* c, const int*, const int*, int, const double*, const double*, int64_t, float*, int64_t) {
    return set_error(c, PMX_ERR_ARGUMENT, "PMX_PREC_F32C is not available in this build");
}

}  // namespace pmx
