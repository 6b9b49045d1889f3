// Host-side Kruskal model helpers of the C ABI (the reference runs these on
// the host too): KruskalModelRandomInit (src/kruskal_model.cpp:109-125),
// glibc srand/rand, U[n][i] = rand() / RAND_MAX for i < I_n * rank, lambda = 1.
#include <stdlib.h>

#include "alto_hip.h"

extern "C" int alto_kruskal_random_init(int nmodes, const uint64_t* dims, uint64_t rank, unsigned seed, double* const* U,
                                        double* lambda)
{
    if (nmodes < 1 || !dims || !U) return ALTO_ERR_ARG;
    if (lambda)
        for (uint64_t r = 0; r < rank; ++r) lambda[r] = 1.0;
    srand(seed);
    for (int n = 0; n < nmodes; ++n)
        for (uint64_t i = 0; i < dims[n] * rank; ++i) U[n][i] = (double)rand() / RAND_MAX;
    return ALTO_OK;
}
