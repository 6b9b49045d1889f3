// MTTKRP launcher instantiated for float factors (float values, float arithmetic).
#include "mttkrp_kernels.hpp"

namespace alto {

int launch_mttkrp_f32(const alto_tensor* t, int mode, int R, const float* const* U, const float* val, float* out,
                         int seg, const MttkrpTune& k, hipStream_t s)
{
    return launch_mttkrp<float, float, float, float>(t, mode, R, U, val, out, seg, k, s);
}

}  // namespace alto
