// CP-APR (Poisson CP by multiplicative updates) of the reference ALTO
// library (include/apr.hpp, src/apr.cpp) over libalto_hip: the ALTO driver
// cp_apr_alto and the COO drivers cp_apr_mu_alto / cp_apr_mu_alto_otf.
// Models are given in host memory as there; the iterations run on the GPU.
#ifndef ALTO_DROPIN_APR_HPP_
#define ALTO_DROPIN_APR_HPP_

#include "alto.hpp"
#include "common.hpp"
#include "device.hpp"
#include "kruskal_model.hpp"
#include "sptensor.hpp"

// include/apr.hpp:10
#ifndef MIN_APR_REUSE
#define MIN_APR_REUSE 4
#endif

// cp_apr_alto (include/apr.hpp:984-1376): up to max_iters outer iterations
// of max_inner_iters multiplicative updates per mode; stops when the KKT
// violation changed by at most epsilon^2, or never when dobench is set;
// prints the reference's progress and timer lines; M (host) is read as the
// initial model and holds the result.
template <typename LIT>
static inline void cp_apr_alto(AltoTensor<LIT>* AT, KruskalModel* M, int max_iters, int max_inner_iters,
                               FType epsilon, FType eps_div_zero, FType kappa_tol, FType kappa, bool dobench)
{
    alto_apr_result_t res;
    alto_dropin::check(alto_cp_apr(AT->dev, M->rank, max_iters, max_inner_iters, epsilon, eps_div_zero, kappa_tol,
                                   kappa, dobench ? 1 : 0, M->U, M->lambda, &res, 1, nullptr),
                       "cp_apr_alto");
}

namespace alto_dropin {
// the COO drivers: the tensor is moved to the device once, then run with the
// COO stop rule (an outer iteration in which every mode's first inner step
// met epsilon ends the run, and counts)
inline void cp_apr_coo(SparseTensor* X, KruskalModel* M, int max_iters, int max_inner_iters, FType epsilon,
                       FType eps_div_zero, FType kappa_tol, FType kappa, const char* what)
{
    alto_tensor_t* t = nullptr;
    check(alto_tensor_create(X->nmodes, (const uint64_t*)X->dims, X->nnz, (const uint64_t* const*)X->cidx, X->vals,
                             ALTO_PTR_HOST, ALTO_VAL_F64, 0, 1, nullptr, &t),
          what);
    alto_apr_result_t res;
    const int st = alto_cp_apr_ex(t, M->rank, max_iters, max_inner_iters, epsilon, eps_div_zero, kappa_tol, kappa,
                                  ALTO_APR_STOP_COO, M->U, M->lambda, &res, 1, nullptr);
    alto_tensor_destroy(t);
    check(st, what);
}
}  // namespace alto_dropin

// cp_apr_mu_alto (src/apr.cpp:239-489): CP-APR on a COO tensor
inline void cp_apr_mu_alto(SparseTensor* X, KruskalModel* M, int max_iters, int max_inner_iters, FType epsilon,
                           FType eps_div_zero, FType kappa_tol, FType kappa)
{
    alto_dropin::cp_apr_coo(X, M, max_iters, max_inner_iters, epsilon, eps_div_zero, kappa_tol, kappa,
                            "cp_apr_mu_alto");
}

// cp_apr_mu_alto_otf (src/apr.cpp:544-779): the on-the-fly-Pi variant of the
// same iteration; on the device Pi is never stored, so both are one call
inline void cp_apr_mu_alto_otf(SparseTensor* X, KruskalModel* M, int max_iters, int max_inner_iters, FType epsilon,
                               FType eps_div_zero, FType kappa_tol, FType kappa)
{
    alto_dropin::cp_apr_coo(X, M, max_iters, max_inner_iters, epsilon, eps_div_zero, kappa_tol, kappa,
                            "cp_apr_mu_alto_otf");
}

#endif  // ALTO_DROPIN_APR_HPP_
