// Test-only harness: instantiates the reference's own header-only templates
// (include/alto.hpp, include/apr.hpp of IntelLabs/ALTO, compiled from the
// reference tree by oracle/Makefile) behind an extern "C" surface so the
// tests can compare the oracle and the product against the real reference.
// Nothing here is shipped or called by the product.
#include <stdint.h>
#include <typeinfo>
#include <vector>

#include "alto.hpp"
#include "apr.hpp"

#if REF_LIT_BITS == 128
typedef unsigned __int128 LIT;
#define W 2
#else
typedef unsigned long long LIT;
#define W 1
#endif

static void put_lit(uint64_t *dst, LIT v)
{
    dst[0] = (uint64_t)v;
#if W == 2
    dst[1] = (uint64_t)(v >> 64);
#endif
}

extern "C" {

int ref_lit_words(void) { return W; }

void *ref_create(int nmodes, const uint64_t *dims, uint64_t nnz, const uint64_t *cidx,
                 const double *vals, int nprtn)
{
    SparseTensor X;
    X.nmodes = nmodes;
    X.nnz = nnz;
    X.dims = (IType *)dims;
    std::vector<IType *> rows(nmodes);
    for (int n = 0; n < nmodes; ++n) rows[n] = (IType *)(cidx + (uint64_t)n * nnz);
    X.cidx = rows.data();
    X.vals = (FType *)vals;
    AltoTensor<LIT> *at = nullptr;
    create_alto(&X, &at, nprtn);
    return at;
}

void ref_destroy(void *h) { destroy_alto((AltoTensor<LIT> *)h); }

void ref_get(void *h, uint64_t *masks, uint64_t *idx, double *vals, uint64_t *prtn_ptr,
             uint64_t *intervals, uint64_t *cr_masks, int *is_oidx, uint64_t *oidx)
{
    AltoTensor<LIT> *at = (AltoTensor<LIT> *)h;
    int N = at->nmode;
    for (int n = 0; n < N; ++n) {
        put_lit(masks + n * W, at->mode_masks[n]);
        put_lit(cr_masks + n * W, at->cr_masks[n]);
    }
    for (uint64_t i = 0; i < at->nnz; ++i) {
        put_lit(idx + i * W, at->idx[i]);
        vals[i] = (double)at->vals[i];
    }
    for (int p = 0; p <= at->nprtn; ++p) prtn_ptr[p] = at->prtn_ptr[p];
    for (int k = 0; k < at->nprtn * N; ++k) {
        intervals[2 * k] = at->prtn_intervals[k].start;
        intervals[2 * k + 1] = at->prtn_intervals[k].stop;
    }
    *is_oidx = at->is_oidx ? 1 : 0;
    if (at->is_oidx && oidx)
        for (int n = 0; n < N; ++n)
            if (at->oidx[n])
                for (uint64_t i = 0; i < at->nnz; ++i) oidx[(uint64_t)n * at->nnz + i] = at->oidx[n][i];
}

// mttkrp_alto_par with the reference's own dispatch (include/alto.hpp:617-623);
// adds into factors[mode].
void ref_mttkrp_par(void *h, int mode, double **factors, uint64_t rank)
{
    AltoTensor<LIT> *at = (AltoTensor<LIT> *)h;
    FType **ofibs = nullptr;
    create_da_mem(mode, rank, at, &ofibs);
    mttkrp_alto_par(mode, factors, rank, at, nullptr, ofibs);
    destroy_da_mem(at, ofibs, rank, mode);
}

// sequential mttkrp_alto (include/alto.hpp:457-543)
void ref_mttkrp_seq(void *h, int mode, double **factors, uint64_t rank)
{
    mttkrp_alto(mode, factors, rank, (AltoTensor<LIT> *)h);
}

// CalculatePhi_alto_par (include/apr.hpp:295-303) with on-the-fly products
void ref_phi_par(void *h, int mode, double **U, uint64_t rank, double eps, double *Phi)
{
    AltoTensor<LIT> *at = (AltoTensor<LIT> *)h;
    KruskalModel M;
    M.mode = at->nmode;
    M.dims = at->dims;
    M.rank = rank;
    M.U = U;
    M.lambda = nullptr;
    FType **ofibs = nullptr;
    create_da_mem(-1, rank, at, &ofibs);
    CalculatePhi_alto_par(Phi, at, &M, mode, eps, ofibs, rank, false);
    destroy_da_mem(at, ofibs, rank, -1);
}

}  // extern "C"
