// MTTKRP launcher instantiated for double factors (double values, double arithmetic).
#include "mttkrp_kernels.hpp"

namespace alto {

int launch_mttkrp_f64(const alto_tensor* t, int mode, int R, const double* const* U, const double* val, double* out,
                         int seg, const MttkrpTune& k, hipStream_t s)
{
    return launch_mttkrp<double, double, double, double>(t, mode, R, U, val, out, seg, k, s);
}

}  // namespace alto
