// MTTKRP launcher instantiated for bf16_t factors (float values, float arithmetic).
#include "mttkrp_kernels.hpp"

namespace alto {

int launch_mttkrp_bf16(const alto_tensor* t, int mode, int R, const bf16_t* const* U, const float* val, float* out,
                          int seg, const MttkrpTune& k, hipStream_t s)
{
    return launch_mttkrp<bf16_t, float, float, float>(t, mode, R, U, val, out, seg, k, s);
}

}  // namespace alto
