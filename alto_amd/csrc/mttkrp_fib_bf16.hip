// Long-fiber MTTKRP launches for bf16_t factors (float values, float
// arithmetic, FP64 output where the factors are not double).
#include "mttkrp_kernels.hpp"

namespace alto {

int fiber_part(const alto_tensor* t, int mode, int R, const bf16_t* const* U,
               const StreamArgs<bf16_t, float, double>& a, int G, int E, int seg, hipStream_t s)
{
    if (E == 2) return launch_fiber_part<bf16_t, float, float, double, 2>(t, mode, R, U, a, G, seg, s);
    if (E == 4) return launch_fiber_part<bf16_t, float, float, double, 4>(t, mode, R, U, a, G, seg, s);
    set_error("fiber_part: %d columns per lane", E);
    return ALTO_ERR_UNSUPPORTED;
}

}  // namespace alto
