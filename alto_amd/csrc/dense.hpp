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
// zero_bytes of `zero` are cleared in the same launches.  With Mi (this
// mode's MTTKRP), also the fit's inner product from the stored values:
// R <= 64: per block sum of lambda_c S[i,c] Mi[i,c] into pin[gram_blocks(I)];
// R > 64: per-column partials (64 x R, unweighted) into pin
template <typename UT, typename ST>
int gram_scale_device(UT* U, uint64_t I, int R, const double* lambda, ST* S, double* G, double* part, hipStream_t s,
                      bool f64, void* zero = nullptr, uint64_t zero_bytes = 0, const double* Mi = nullptr,
                      double* pin = nullptr);
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

// One CP-ALS model's mode update after its MTTKRP (M = M_mode): V^-1 from
// the other modes' Grams, U = M V^-1 with the column statistics (and the
// fit's inner-product partials when pin is set), lambda (which: 0 = 2-norm,
// 1 = max-norm), U /= lambda with the storage copy S and the new Gram G,
// and the clearing of zero_bytes of `zero` -- the launches of chol_device,
// solve_device, lambda_device and gram_scale_device for R <= 64
struct DenseModeJob {
    const double* grams;
    int nmode, mode;
    double *P, *wsA, *wsQ;
    int* status;
    const double* M;
    uint64_t I;
    void* U;
    double *psum, *pmax, *pin;
    int nb;  // row-solve blocks (the psum / pmax partials)
    int which;
    double* lambda;
    void* S;
    double *G, *gpart;
    void* zero;
    uint64_t zero_bytes;
    // storage that rounds: this mode's MTTKRP and the per-block fit inner
    // products taken from the stored values (null: none)
    const double* Mgram;
    double* pin_gram;
};
constexpr int DENSE_BATCH_MAX = 16;
// the mode updates of n <= DENSE_BATCH_MAX models of one rank (<= 64) and
// factor / storage types in 5 launches (one per kernel, all models each)
template <typename UT, typename ST>
int dense_mode_batch(int n, const DenseModeJob* jobs, int R, bool f64, hipStream_t s);
// the fits of n <= DENSE_BATCH_MAX models (fit_scalar_device each), one
// launch; each fit is also copied to gather[0]
struct DenseFitJob {
    const double *grams, *lambda, *pin;
    double* out;
    double* gather;
    double normX;
    int nmode, nb;
};
int fit_batch(int n, const DenseFitJob* jobs, int R, hipStream_t s);

}  // namespace alto
