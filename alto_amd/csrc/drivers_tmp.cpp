// placeholder entry points, replaced by the CP-ALS / CP-APR drivers
#include "alto_hip.h"
extern "C" {
int alto_cpd_als(alto_tensor_t*, uint64_t, int, double, int, double* const*, double*, double*, int*, alto_cpd_timers_t*, int, alto_stream_t) { return ALTO_ERR_UNSUPPORTED; }
int alto_cpd_fit(alto_tensor_t*, uint64_t, double* const*, const double*, double*) { return ALTO_ERR_UNSUPPORTED; }
int alto_cp_apr(alto_tensor_t*, uint64_t, int, int, double, double, double, double, int, double* const*, double*, alto_apr_result_t*, int, alto_stream_t) { return ALTO_ERR_UNSUPPORTED; }
}
