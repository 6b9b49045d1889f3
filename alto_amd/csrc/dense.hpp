// Host wrappers of the dense kernels (dense.hip) used by the CP drivers.
#pragma once

#include "alto_internal.hpp"
#include "kernels_common.hpp"

namespace alto {

int gram_blocks(uint64_t I);
template <typename T>
int gram_device(const T* U, uint64_t I, int R, double* G, double* part, hipStream_t s, bool f64);
int chol_device(const double* grams, int nmode, int mode, int R, double* L, double* P, double* wsA, double* wsQ,
                int* status, hipStream_t s);
int solve_blocks(uint64_t I, int R);
template <typename MT, typename UT>
int solve_device(const MT* M, uint64_t I, int R, const double* L, const double* P, const int* status, UT* U,
                 double* psum, double* pmax, int nb, hipStream_t s, double* pin = nullptr);
int lambda_device(const double* psum, const double* pmax, int nb, int R, int which, double* lambda, hipStream_t s);
template <typename UT, typename ST>
int scale_device(UT* U, uint64_t I, int R, const double* lambda, ST* S, hipStream_t s);
// U /= lambda (per column) with the storage copy S (null when the types
// agree), then the Gram of the stored values, in one pass for R <= 64;
// zero_bytes of `zero` are cleared in the same launches
template <typename UT, typename ST>
int gram_scale_device(UT* U, uint64_t I, int R, const double* lambda, ST* S, double* G, double* part, hipStream_t s,
                      bool f64, void* zero = nullptr, uint64_t zero_bytes = 0);
// column sums of partials and the clearing of a buffer, one launch
int sum_cols_device(const double* part, int nparts, int n, double* out, void* zero, uint64_t zero_bytes,
                    hipStream_t s);
template <typename UT, typename MT>
int inner_partial_device(const UT* U, const MT* Mm, uint64_t I, int R, double* part, hipStream_t s);
int fit_device(const double* grams, int nmode, int R, const double* lambda, const double* ipart, double normX,
               double* out, hipStream_t s);
int fit_scalar_device(const double* grams, int nmode, int R, const double* lambda, const double* pin, int nb,
                      double normX, double* out, hipStream_t s);
int random_factor_device(double* U, uint64_t n, uint64_t seed, int mode, hipStream_t s);
int normsq_device(const double* v, uint64_t n, double* part, double* out, hipStream_t s);
// alto_update_gram_noalloc (partials in a per-device scratch array)
int gram_noalloc(const void* U, uint64_t dim, int R, int precision, double* gram, hipStream_t s);

}  // namespace alto
