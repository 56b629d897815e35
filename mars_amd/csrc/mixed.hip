// Mixed-precision candidate scoring (BF16-split / FP32 MFMA bulk, FP64 refinement).
#include "internal.h"
#include "kernels.h"

namespace mh {

void mixed_orthonormalize(mars_hip_ctx *ctx, const XSrc &xs, int P)
{
    (void)ctx;
    (void)xs;
    (void)P;
    throw Error(MARS_HIP_EINVAL, "mixed-precision form not built yet");
}

} // namespace mh
