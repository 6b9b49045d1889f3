// MTTKRP launcher instantiated for double factors (double values, double arithmetic).
#include "mttkrp_kernels.hpp"

namespace alto {

int launch_mttkrp_f64(const alto_tensor* t, int mode, int R, const double* const* U, const double* val, double* out,
                         int seg, const MttkrpTune& k, hipStream_t s)
{
    return launch_mttkrp<double, double, double, double>(t, mode, R, U, val, out, seg, k, s);
}

int launch_mttkrp_batch_f64(int n, alto_tensor* const* ts, const int* modes, int R, const double* const* const* U,
                            const double* const* vals, double* const* outs, int seg, const MttkrpTune& k, hipStream_t s)
{
    return launch_mttkrp_batch<double, double, double, double>(n, ts, modes, R, U, vals, outs, seg, k, s);
}

}  // namespace alto
