// Long-fiber MTTKRP launches for double factors (double values and
// arithmetic).
#include "mttkrp_kernels.hpp"

namespace alto {

int fiber_part(const alto_tensor* t, int mode, int R, const double* const* U,
               const StreamArgs<double, double, double>& a, int G, int E, int seg, hipStream_t s)
{
    if (E != 2) {
        set_error("fiber_part: %d columns per lane for FP64", E);
        return ALTO_ERR_UNSUPPORTED;
    }
    return launch_fiber_part<double, double, double, double, 2>(t, mode, R, U, a, G, seg, s);
}

}  // namespace alto
