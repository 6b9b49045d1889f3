// CP-ALS of the reference ALTO library (include/cpd.hpp) over libalto_hip:
// the ALTO driver cpd_alto and its fit, and the coordinate (COO) forms cpd,
// cpd_fit, mttkrp and mttkrp_par.  Models and tensors are given in host
// memory as there; the arithmetic runs on the GPU.
#ifndef ALTO_DROPIN_CPD_HPP_
#define ALTO_DROPIN_CPD_HPP_

#include <math.h>
#include <stdio.h>

#include <vector>

#include "alto.hpp"
#include "common.hpp"
#include "device.hpp"
#include "kruskal_model.hpp"
#include "sptensor.hpp"

// cpd_alto (include/cpd.hpp:43-186): max_iters CP-ALS sweeps on the device,
// stopping after the first sweep whose fit changed by less than epsilon
// (the first sweep always runs on); prints the reference's "it:" lines and
// timer table; M (host) is read as the initial model and holds the result.
template <typename LIT>
void cpd_alto(AltoTensor<LIT>* AT, KruskalModel* M, int max_iters, double epsilon)
{
    int iters = 0;
    alto_cpd_timers_t tm;
    alto_dropin::check(alto_cpd_als(AT->dev, M->rank, max_iters, epsilon, ALTO_PREC_F64, M->U, M->lambda, nullptr,
                                    &iters, &tm, 1, nullptr),
                       "cpd_alto");
}

// cpd_fit_alto (include/cpd.hpp:189-271): 1 - ||X - M|| / ||X|| from the
// model's lambda and last factor, the Grams (lower triangles read), the
// last mode's MTTKRP U_mttkrp and normAT = ||X||^2, all in host memory
template <typename LIT>
double cpd_fit_alto(AltoTensor<LIT>* AT, KruskalModel* M, FType** grams, FType* U_mttkrp, FType normAT)
{
    double fit = 0;
    alto_dropin::check(alto_cpd_fit_parts(AT->nmode, (const uint64_t*)AT->dims, M->rank, grams, M->lambda,
                                          M->U[AT->nmode - 1], U_mttkrp, normAT, ALTO_PTR_HOST, &fit),
                       "cpd_fit_alto");
    return fit;
}

// cpd_fit (include/cpd.hpp:374-458): cpd_fit_alto with ||X||^2 taken from
// the COO tensor's values
inline double cpd_fit(SparseTensor* X, KruskalModel* M, FType** grams, FType* U_mttkrp)
{
    double normX = 0.0;
    for (IType i = 0; i < X->nnz; ++i) normX += X->vals[i] * X->vals[i];
    double fit = 0;
    alto_dropin::check(alto_cpd_fit_parts(X->nmodes, (const uint64_t*)X->dims, M->rank, grams, M->lambda,
                                          M->U[X->nmodes - 1], U_mttkrp, normX, ALTO_PTR_HOST, &fit),
                       "cpd_fit");
    return fit;
}

// mttkrp (include/cpd.hpp:509-543): M->U[mode] += the mode-`mode` MTTKRP of
// the COO tensor X with M's other factors
inline void mttkrp(SparseTensor* X, KruskalModel* M, IType mode)
{
    alto_dropin::check(alto_coo_mttkrp(X->nmodes, (const uint64_t*)X->dims, X->nnz,
                                       (const uint64_t* const*)X->cidx, X->vals, (int)mode, M->rank, M->U,
                                       M->U[mode], 1, ALTO_PTR_HOST, nullptr),
                       "mttkrp");
}

// mttkrp_par (include/cpd.hpp:460-507): the same; the row locks are not
// needed (the device adds atomically)
inline void mttkrp_par(SparseTensor* X, KruskalModel* M, IType mode, void* /*writelocks*/)
{
    mttkrp(X, M, mode);
}

// cpd (include/cpd.hpp:274-372): CP-ALS straight from the COO tensor, every
// step on the device: per mode the COO MTTKRP, the solve against the
// Hadamard product of the other Grams (Cholesky, least-squares fallback),
// the column normalisation (2-norm in the first sweep, max-norm after) and
// the new Gram; then the fit.  Same stop rule and "it:" lines as there.
inline void cpd(SparseTensor* X, KruskalModel* M, int max_iters, double epsilon)
{
    using namespace alto_dropin;
    fprintf(stdout, "Running CP-ALS with %d max iterations and %.2e epsilon\n", max_iters, epsilon);
    const int N = X->nmodes;
    const IType R = M->rank;
    const uint64_t* dims = (const uint64_t*)X->dims;
    std::vector<DeviceBuffer*> ci(N);
    std::vector<const uint64_t*> cip(N);
    for (int m = 0; m < N; ++m) {
        ci[m] = new DeviceBuffer(X->cidx[m], X->nnz * sizeof(IType));
        cip[m] = ci[m]->as<uint64_t>();
    }
    DeviceBuffer vals(X->vals, X->nnz * sizeof(FType));
    DeviceFactors U(N, X->dims, R, M->U);
    DeviceBuffer lambda(M->lambda, R * sizeof(FType)), grams(N * R * R * sizeof(FType));
    DeviceBuffer last(X->dims[N - 1] * R * sizeof(FType));
    std::vector<const double*> gp(N);
    for (int m = 0; m < N; ++m) gp[m] = grams.as<double>() + (size_t)m * R * R;
    double normX = 0.0;
    for (IType i = 0; i < X->nnz; ++i) normX += X->vals[i] * X->vals[i];
    for (int m = 0; m < N; ++m)
        check(alto_update_gram(U.p[m], dims[m], R, ALTO_PREC_F64, (double*)gp[m], nullptr), "update_gram");
    double fit = 0.0, prev_fit = 0.0;
    for (int i = 0; i < max_iters; ++i) {
        for (int j = 0; j < N; ++j) {
            check(alto_coo_mttkrp(N, dims, X->nnz, cip.data(), vals.as<double>(), j, R,
                                  (const double* const*)U.p.data(), (double*)U.p[j], 0, ALTO_PTR_DEVICE, nullptr),
                  "mttkrp_par");
            if (j == N - 1) check(alto_memcpy_d2d(last.p, U.p[j], dims[j] * R * sizeof(FType), nullptr), "memcpy");
            int fallback = 0;
            check(alto_pseudo_inverse(grams.as<double>(), N, j, R, (double*)U.p[j], dims[j], &fallback, nullptr),
                  "pseudo_inverse");
            check(alto_kruskal_normalize(N, dims, R, (double* const*)U.p.data(), lambda.as<double>(), j,
                                         i == 0 ? ALTO_KRUSKAL_NORM2 : ALTO_KRUSKAL_NORMMAX, ALTO_PTR_DEVICE,
                                         nullptr),
                  "KruskalModelNorm");
            check(alto_update_gram(U.p[j], dims[j], R, ALTO_PREC_F64, (double*)gp[j], nullptr), "update_gram");
        }
        check(alto_cpd_fit_parts(N, dims, R, gp.data(), lambda.as<double>(), (const double*)U.p[N - 1],
                                 last.as<double>(), normX, ALTO_PTR_DEVICE, &fit),
              "cpd_fit");
        printf("it: %d\t fit: %g\t fit-delta: %g\n", i, fit, fabs(prev_fit - fit));
        if (i > 0 && fabs(prev_fit - fit) < epsilon) break;
        prev_fit = fit;
    }
    check(alto_device_synchronize(), "synchronize");
    for (int m = 0; m < N; ++m)
        check(alto_memcpy_d2h(M->U[m], U.p[m], X->dims[m] * R * sizeof(FType)), "alto_memcpy_d2h");
    check(alto_memcpy_d2h(M->lambda, lambda.p, R * sizeof(FType)), "alto_memcpy_d2h");
    for (DeviceBuffer* b : ci) delete b;
}

#endif  // ALTO_DROPIN_CPD_HPP_
