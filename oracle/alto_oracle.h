/*
 * CPU oracle for the ALTO hot path: a plain-C (+OpenMP) restatement of the
 * reference algorithm, used ONLY as a test checker and as the timed CPU
 * baseline.  Nothing in the product library links or calls this code.
 *
 * Every function cites the reference file:line it restates.  Keys of width
 * W words (W=1: 64-bit LIT, W=2: 128-bit LIT) are stored little-endian as
 * W consecutive uint64_t words (word 0 = bits 0..63).
 */
#ifndef ALTO_ORACLE_H_
#define ALTO_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_MODES 5
#define OR_MIN_FIBER_REUSE 4

/* ---- bit operations (include/bitops.hpp) ---- */
uint64_t or_pdep64(uint64_t x, uint64_t mask);
uint64_t or_pext64(uint64_t x, uint64_t mask);
void     or_pdep128(uint64_t x, const uint64_t mask[2], uint64_t out[2]);
uint64_t or_pext128(const uint64_t x[2], const uint64_t mask[2]);

/* ---- mask setup (include/alto.hpp:669-770) ---- */
int or_setup_masks(int nmode, const uint64_t *dims, int W,
                   uint64_t *masks /* nmode*W */, uint64_t *alto_mask /* W */);

/* ---- ALTO tensor ---- */
typedef struct or_alto or_alto;

/* create_alto (include/alto.hpp:108-294).  cidx is nmode x nnz, row major.
 * val_int != 0 reproduces ValType = unsigned long long (values truncated). */
or_alto *or_alto_create(int nmode, const uint64_t *dims, uint64_t nnz,
                        const uint64_t *cidx, const double *vals, int val_int,
                        int W, int nprtn);
/* Wrap already-linearized, sorted ALTO arrays (e.g. exported from the GPU):
 * masks, idx (nnz*W, CR bits included), vals, prtn_ptr, intervals, cr masks.
 * oidx may be NULL (then it is rebuilt for low-reuse modes). */
or_alto *or_alto_wrap(int nmode, const uint64_t *dims, uint64_t nnz, int W, int nprtn,
                      const uint64_t *masks, const uint64_t *idx, const double *vals,
                      const uint64_t *prtn_ptr, const uint64_t *intervals,
                      const uint64_t *cr_masks, const uint64_t *oidx);
void or_alto_destroy(or_alto *at);

/* accessors: copy state out into caller buffers */
void or_alto_get_masks(const or_alto *at, uint64_t *masks, uint64_t *alto_mask,
                       uint64_t *cr_masks, uint64_t *alto_cr_mask);
void or_alto_get_idx(const or_alto *at, uint64_t *idx);
void or_alto_get_vals(const or_alto *at, double *vals);
void or_alto_get_prtn(const or_alto *at, uint64_t *prtn_ptr, uint64_t *intervals);
int  or_alto_is_oidx(const or_alto *at);
int  or_alto_get_oidx(const or_alto *at, int mode, uint64_t *oidx); /* 0 if none */
void or_alto_delinearize(const or_alto *at, uint64_t *coords /* nmode x nnz */);

/* ---- MTTKRP ---- */
enum { OR_MTTKRP_DISPATCH = 0, OR_MTTKRP_SEQ = 1, OR_MTTKRP_ATOMIC = 2,
       OR_MTTKRP_ATOMIC_CR = 3, OR_MTTKRP_DA_MEM_PULL = 4, OR_MTTKRP_OP_SCHED = 5 };
/* Adds the mode-n MTTKRP into factors[mode] (caller zeroes it), as the
 * reference does.  Returns the variant actually run. */
int or_mttkrp_alto(or_alto *at, int mode, double **factors, uint64_t rank, int variant);
/* COO sequential MTTKRP (include/cpd.hpp:509-543), adds into out. */
void or_mttkrp_coo(int nmode, uint64_t nnz, const uint64_t *cidx, const double *vals,
                   double **factors, uint64_t rank, int mode, double *out);

/* ---- Kruskal model ops (src/kruskal_model.cpp) ---- */
void or_random_init(int nmode, const uint64_t *dims, uint64_t rank, unsigned seed,
                    double **U, double *lambda);
void or_gram(const double *U, uint64_t dim, uint64_t rank, double *G);
void or_normalize_all_l1(int nmode, const uint64_t *dims, uint64_t rank, double **U, double *lambda);
void or_normalize_mode_l1(uint64_t dim, uint64_t rank, double *U, double *lambda);
void or_mat_norm(uint64_t dim, uint64_t rank, double *U, double *lambda, int which /*0=2,1=max*/);
void or_redistribute_lambda(uint64_t dim, uint64_t rank, double *U, double *lambda);

/* pseudo_inverse (include/cpd.hpp:546-646): U[mode] <- U[mode] * V^-1 with
 * V the Hadamard product of the other modes' Grams.  Returns 0 when the
 * Cholesky path was taken, 1 when the minimum-norm fallback was used. */
int or_pseudo_inverse(int nmode, uint64_t rank, double **grams, double *U,
                      uint64_t dim, int mode);

double or_cpd_fit(int nmode, uint64_t rank, const uint64_t *dims, double **U, const double *lambda,
                  double **grams, const double *U_mttkrp, double normX);

/* cpd_alto (include/cpd.hpp:43-186).  U/lambda are updated in place.
 * fit_trace receives the fit of each executed iteration.  Returns the
 * number of iterations run. */
int or_cpd_als(or_alto *at, uint64_t rank, int max_iters, double epsilon,
               double **U, double *lambda, double *fit_trace, int variant);

/* ---- CP-APR ---- */
/* Phi for mode n (include/apr.hpp:383-501 semantics: Phi fully overwritten). */
void or_phi(or_alto *at, int mode, double **U, uint64_t rank, double eps_div_zero, double *Phi);
double or_loglikelihood(or_alto *at, double **U, double *lambda, uint64_t rank, double eps_div_zero);

typedef struct {
    double loglik, ls_fit, kkt;
    int outer_iters, total_inner;
    int inner_per_mode[OR_MAX_MODES];
} or_apr_result;

int or_cp_apr(or_alto *at, uint64_t rank, int max_iters, int max_inner, double epsilon,
              double eps_div_zero, double kappa_tol, double kappa, int dobench,
              double **U, double *lambda, or_apr_result *res);

/* Observer of the CP-APR driver's steps (tests only).  U / lambda are the
 * live model (U[mode] is B); phi is Phi of `mode` (at PRE_FIX / POST_FIX:
 * the Phi the kappa fix reads, from the previous outer iteration); kkt is
 * the mode's KKT value (PHI, MU, NORM) or the outer maximum (OUTER). */
enum { OR_APR_EV_START = 0, OR_APR_EV_PRE_FIX = 1, OR_APR_EV_POST_FIX = 2, OR_APR_EV_PHI = 3,
       OR_APR_EV_MU = 4, OR_APR_EV_NORM = 5, OR_APR_EV_OUTER = 6 };
typedef void (*or_apr_hook)(void *ctx, int event, int iter, int mode, int inner, double **U,
                            const double *lambda, const double *phi, double kkt);
int or_cp_apr_hooked(or_alto *at, uint64_t rank, int max_iters, int max_inner, double epsilon,
                     double eps_div_zero, double kappa_tol, double kappa, int dobench,
                     double **U, double *lambda, or_apr_result *res, or_apr_hook hook, void *ctx);

/* norm of the tensor values, sum of squares */
double or_normsq(const or_alto *at);

int or_omp_threads(void);
void or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
