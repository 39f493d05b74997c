// Step kernels of the bf16s storage form (templates in oal_step.h).
#include "oal_step.h"

namespace oal {

hipError_t launch_step_bf16s(int stream_mode, bool wale, const StepArgs& a, hipStream_t stream, hipStream_t side,
                        hipEvent_t bulk0, hipEvent_t bulk1) {
    return launch_sp<StBF16S>(stream_mode, wale, a, stream, side, bulk0, bulk1);
}

}  // namespace oal
