// Test-only harness: instantiates the reference's own header-only templates
// (include/alto.hpp, include/apr.hpp of IntelLabs/ALTO, compiled from the
// reference tree by oracle/Makefile) behind an extern "C" surface so the
// tests can compare the oracle and the product against the real reference.
// Nothing here is shipped or called by the product.
#include <stdint.h>
#include <string.h>
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

static LIT get_lit(const uint64_t *src)
{
#if W == 2
    return (LIT)src[0] | ((LIT)src[1] << 64);
#else
    return (LIT)src[0];
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

// The reference's AltoTensor filled from an ALTO state produced elsewhere
// (the GPU engine's export, bit-identical to create_alto's), so that the
// reference's mttkrp_alto_par can be timed on a large tensor without its
// CPU sort; freed by ref_destroy (destroy_alto).
void *ref_wrap(int nmodes, const uint64_t *dims, uint64_t nnz, const uint64_t *masks, const uint64_t *alto_mask,
               const uint64_t *idx, const double *vals, int nprtn, const uint64_t *prtn_ptr,
               const uint64_t *intervals, const uint64_t *cr_masks, const uint64_t *alto_cr_mask)
{
    AltoTensor<LIT> *at = (AltoTensor<LIT> *)AlignedMalloc(sizeof(AltoTensor<LIT>));
    memset((void *)at, 0, sizeof(AltoTensor<LIT>));
    at->nmode = nmodes;
    at->nprtn = nprtn;
    at->nnz = nnz;
    at->dims = (IType *)AlignedMalloc(nmodes * sizeof(IType));
    at->mode_masks = (LIT *)AlignedMalloc(nmodes * sizeof(LIT));
    at->cr_masks = (LIT *)AlignedMalloc(nmodes * sizeof(LIT));
    for (int n = 0; n < nmodes; ++n) {
        at->dims[n] = dims[n];
        at->mode_masks[n] = get_lit(masks + n * W);
        at->cr_masks[n] = get_lit(cr_masks + n * W);
    }
    at->alto_mask = get_lit(alto_mask);
    at->alto_cr_mask = get_lit(alto_cr_mask);
    at->idx = (LIT *)AlignedMalloc(nnz * sizeof(LIT));
    at->vals = (ValType *)AlignedMalloc(nnz * sizeof(ValType));
#pragma omp parallel for
    for (uint64_t i = 0; i < nnz; ++i) {
        at->idx[i] = get_lit(idx + i * W);
        at->vals[i] = (ValType)vals[i];
    }
    at->prtn_ptr = (IType *)AlignedMalloc((nprtn + 1) * sizeof(IType));
    for (int p = 0; p <= nprtn; ++p) at->prtn_ptr[p] = prtn_ptr[p];
    at->prtn_intervals = (Interval *)AlignedMalloc((size_t)nprtn * nmodes * sizeof(Interval));
    for (int k = 0; k < nprtn * nmodes; ++k) {
        at->prtn_intervals[k].start = intervals[2 * k];
        at->prtn_intervals[k].stop = intervals[2 * k + 1];
    }
    at->is_oidx = false;
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
