// MTTKRP launcher instantiated for bf16_t factors (float values, float arithmetic).
#include "mttkrp_kernels.hpp"

namespace alto {

int launch_mttkrp_bf16(const alto_tensor* t, int mode, int R, const bf16_t* const* U, const float* val, double* out,
                          int seg, const MttkrpTune& k, hipStream_t s)
{
    return launch_mttkrp<bf16_t, float, float, double>(t, mode, R, U, val, out, seg, k, s);
}

int launch_mttkrp_batch_bf16(int n, alto_tensor* const* ts, const int* modes, int R, const bf16_t* const* const* U,
                            const float* const* vals, double* const* outs, int seg, const MttkrpTune& k, hipStream_t s)
{
    return launch_mttkrp_batch<bf16_t, float, float, double>(n, ts, modes, R, U, vals, outs, seg, k, s);
}

}  // namespace alto
