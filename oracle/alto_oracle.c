/*
 * ORACLE — test infrastructure only.  A plain C11 + OpenMP restatement of
 * the reference ALTO hot path (IntelLabs/ALTO), used by tests/ as the parity
 * checker and by bench.py's cpu_baseline leg as the timed CPU baseline.
 * The product library (alto_amd/) never links, loads or calls it.
 *
 * Each function names the reference file:line it restates.  Deviations from
 * the reference, all deliberate and listed here:
 *   - sort_alto uses a stable LSD radix sort (include/alto.hpp:772-803 uses
 *     std::sort); identical output for unique keys, ties keep input order.
 *   - oidx ties are ordered by ascending position (the reference's std::sort
 *     leaves tie order unspecified, include/alto.hpp:273-277).
 *   - the CR ownership scan uses per-row min/max partition ids instead of the
 *     reference's CAS loop (include/alto.hpp:975-995); same shared/owned result.
 *   - the Cholesky-failure fallback of pseudo_inverse solves the minimum-norm
 *     least-squares problem on the symmetric V via a Jacobi eigen-decomposition
 *     (the reference calls LAPACK dgelsy after copying a stale upper triangle,
 *     include/cpd.hpp:612-633).
 *   - CP-APR's outer convergence test is skipped at iteration 0 (the reference
 *     reads kktViolations[-1] there, include/apr.hpp:1261).
 *   - Phi is always fully overwritten; the reference's op_sched Phi leaves
 *     rows without nonzeros stale (include/apr.hpp:790-830).
 */
#include "alto_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#if defined(__BMI2__)
#include <immintrin.h>
#endif

/* ------------------------------------------------------------------------ */
/* bit operations: include/bitops.hpp:18-42 (soft pdep/pext), :81-106 (u128) */

uint64_t or_pdep64(uint64_t x, uint64_t y)
{
    uint64_t res = 0;
    for (uint64_t b = 1; y != 0; b += b) {
        if (x & b) res |= y & (~y + 1);
        y &= (y - 1);
    }
    return res;
}

uint64_t or_pext64(uint64_t x, uint64_t y)
{
    uint64_t res = 0;
    for (uint64_t b = 1; y != 0; b += b) {
        if (x & y & (~y + 1)) res |= b;
        y &= (y - 1);
    }
    return res;
}

/* split at bit 64: include/bitops.hpp:81-100 */
void or_pdep128(uint64_t x, const uint64_t m[2], uint64_t out[2])
{
    int shift = __builtin_popcountll(m[0]);
    uint64_t xhi = shift >= 64 ? 0 : (x >> shift);
    out[1] = or_pdep64(xhi, m[1]);
    out[0] = or_pdep64(x, m[0]);
}

uint64_t or_pext128(const uint64_t x[2], const uint64_t m[2])
{
    int shift = __builtin_popcountll(m[0]);
    uint64_t hi = or_pext64(x[1], m[1]);
    hi = shift >= 64 ? 0 : (hi << shift);
    return hi | or_pext64(x[0], m[0]);
}

/* hot-loop variants; BMI2 where the compiler targets it, identical results */
static inline uint64_t pext_fast64(uint64_t x, uint64_t m)
{
#if defined(__BMI2__)
    return _pext_u64(x, m);
#else
    return or_pext64(x, m);
#endif
}

static inline uint64_t pdep_fast64(uint64_t x, uint64_t m)
{
#if defined(__BMI2__)
    return _pdep_u64(x, m);
#else
    return or_pdep64(x, m);
#endif
}

static inline uint64_t clz64(uint64_t x) { return x ? (uint64_t)__builtin_clzll(x) : 64; }

/* ------------------------------------------------------------------------ */
/* masks: setup_packed_alto(at, LSB_FIRST, SHORT_FIRST), include/alto.hpp:669-770 */

int or_setup_masks(int nmode, const uint64_t *dims, int W, uint64_t *masks, uint64_t *alto_mask)
{
    int mode[OR_MAX_MODES], bits[OR_MAX_MODES];
    int total = 0, maxb = 0;
    if (nmode < 1 || nmode > OR_MAX_MODES || (W != 1 && W != 2)) return -1;
    for (int n = 0; n < nmode; ++n) {
        mode[n] = n;
        bits[n] = (int)(64 - clz64(dims[n] - 1));
        total += bits[n];
        if (bits[n] > maxb) maxb = bits[n];
    }
    if (total > 64 * W) return -2;
    /* stable ascending sort by bit count (libstdc++ insertion sort for n<16) */
    for (int i = 1; i < nmode; ++i) {
        int m = mode[i], b = bits[i], j = i - 1;
        while (j >= 0 && bits[j] > b) { bits[j + 1] = bits[j]; mode[j + 1] = mode[j]; --j; }
        bits[j + 1] = b; mode[j + 1] = m;
    }
    memset(masks, 0, sizeof(uint64_t) * nmode * W);
    int shift = 0;
    for (int level = 0; level < maxb; ++level) {
        for (int k = 0; k < nmode; ++k) {
            if (level < bits[k]) {
                masks[mode[k] * W + (shift >> 6)] |= (uint64_t)1 << (shift & 63);
                ++shift;
            }
        }
    }
    for (int w = 0; w < W; ++w) {
        alto_mask[w] = 0;
        for (int n = 0; n < nmode; ++n) alto_mask[w] |= masks[n * W + w];
    }
    return total;
}

/* ------------------------------------------------------------------------ */
/* tensor state: AltoTensor<LIT>, include/alto.hpp:27-54 */

struct or_alto {
    int nmode, nprtn, W, alto_bits;
    uint64_t nnz;
    uint64_t dims[OR_MAX_MODES];
    uint64_t masks[OR_MAX_MODES * 2];
    uint64_t alto_mask[2];
    uint64_t cr_masks[OR_MAX_MODES * 2];
    uint64_t alto_cr_mask[2];
    uint64_t *idx;
    double *vals;
    uint64_t *prtn_ptr;
    uint64_t *intervals; /* (p*nmode+n)*2 + {0:start,1:stop} */
    int is_oidx;
    uint64_t *oidx[OR_MAX_MODES];
    /* da_mem (create_da_mem(-1), include/alto.hpp:324-370) */
    double **ofibs;
    uint64_t ofibs_rank;
};

static inline uint64_t key_pext(const or_alto *at, uint64_t i, int n)
{
    if (at->W == 1) return pext_fast64(at->idx[i], at->masks[n]);
    const uint64_t *x = &at->idx[2 * i];
    const uint64_t *m = &at->masks[2 * n];
    int shift = __builtin_popcountll(m[0]);
    uint64_t hi = pext_fast64(x[1], m[1]);
    hi = shift >= 64 ? 0 : (hi << shift);
    return hi | pext_fast64(x[0], m[0]);
}

static inline int key_has(const or_alto *at, uint64_t i, const uint64_t *m)
{
    if (at->W == 1) return (at->idx[i] & m[0]) != 0;
    return ((at->idx[2 * i] & m[0]) | (at->idx[2 * i + 1] & m[1])) != 0;
}

/* stable LSD radix sort of W-word keys over the low `bits` bits, carrying perm */
static void radix_sort_keys(uint64_t *keys, uint64_t *perm, uint64_t n, int W, int bits)
{
    uint64_t *k2 = (uint64_t *)malloc(sizeof(uint64_t) * n * W);
    uint64_t *p2 = (uint64_t *)malloc(sizeof(uint64_t) * n);
    const int DB = 16, NB = 1 << 16;
    uint64_t *cnt = (uint64_t *)malloc(sizeof(uint64_t) * NB);
    for (int shift = 0; shift < bits; shift += DB) {
        int w = shift >> 6, s = shift & 63;
        memset(cnt, 0, sizeof(uint64_t) * NB);
        for (uint64_t i = 0; i < n; ++i) cnt[(keys[i * W + w] >> s) & (NB - 1)]++;
        uint64_t sum = 0;
        for (int d = 0; d < NB; ++d) { uint64_t c = cnt[d]; cnt[d] = sum; sum += c; }
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t d = (keys[i * W + w] >> s) & (NB - 1);
            uint64_t o = cnt[d]++;
            for (int q = 0; q < W; ++q) k2[o * W + q] = keys[i * W + q];
            p2[o] = perm[i];
        }
        memcpy(keys, k2, sizeof(uint64_t) * n * W);
        memcpy(perm, p2, sizeof(uint64_t) * n);
    }
    free(cnt); free(k2); free(p2);
}

/* oidx for one mode: positions sorted by row index (include/alto.hpp:246-285),
 * stable counting sort */
static void build_oidx(or_alto *at, int n)
{
    uint64_t nnz = at->nnz, dim = at->dims[n];
    uint64_t *cnt = (uint64_t *)calloc(dim + 1, sizeof(uint64_t));
    uint64_t *o = (uint64_t *)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));
    for (uint64_t i = 0; i < nnz; ++i) cnt[key_pext(at, i, n) + 1]++;
    for (uint64_t r = 0; r < dim; ++r) cnt[r + 1] += cnt[r];
    for (uint64_t i = 0; i < nnz; ++i) o[cnt[key_pext(at, i, n)]++] = i;
    free(cnt);
    at->oidx[n] = o;
}

static void setup_oidx(or_alto *at)
{
    at->is_oidx = 0;
    for (int n = 0; n < at->nmode; ++n)
        if (at->nnz / at->dims[n] <= OR_MIN_FIBER_REUSE) { at->is_oidx = 1; break; }
    if (!at->is_oidx) return;
    for (int n = 0; n < at->nmode; ++n)
        if (at->nnz / at->dims[n] <= OR_MIN_FIBER_REUSE) build_oidx(at, n);
}

/* prtn_alto: pointers, intervals, CR bits (include/alto.hpp:805-1028) */
static void prtn(or_alto *at)
{
    int P = at->nprtn, N = at->nmode, W = at->W;
    uint64_t nnz = at->nnz;
    uint64_t per = (nnz + P - 1) / P;
    at->prtn_ptr[0] = 0;
    for (int p = 0; p < P; ++p) {
        uint64_t e = (uint64_t)(p + 1) * per;
        at->prtn_ptr[p + 1] = e > nnz ? nnz : e;
    }
    #pragma omp parallel for schedule(static, 1)
    for (int p = 0; p < P; ++p) {
        uint64_t st[OR_MAX_MODES], sp[OR_MAX_MODES];
        for (int n = 0; n < N; ++n) { st[n] = at->dims[n]; sp[n] = 0; }
        uint64_t s = p == 0 ? 0 : at->prtn_ptr[p];
        if (s > at->prtn_ptr[p + 1]) s = at->prtn_ptr[p + 1];
        for (uint64_t i = s; i < at->prtn_ptr[p + 1]; ++i)
            for (int n = 0; n < N; ++n) {
                uint64_t r = key_pext(at, i, n);
                if (r < st[n]) st[n] = r;
                if (r > sp[n]) sp[n] = r;
            }
        for (int n = 0; n < N; ++n) {
            at->intervals[(p * N + n) * 2] = st[n];
            at->intervals[(p * N + n) * 2 + 1] = sp[n];
        }
    }
    /* CR masks: include/alto.hpp:934-967 */
    int free_bits = 64 * W - at->alto_bits;
    memset(at->cr_masks, 0, sizeof(at->cr_masks));
    at->alto_cr_mask[0] = at->alto_cr_mask[1] = 0;
    for (int n = 0; n < N; ++n) {
        if (nnz / at->dims[n] <= OR_MIN_FIBER_REUSE && free_bits) {
            int pos = 64 * W - free_bits;
            at->cr_masks[n * W + (pos >> 6)] = (uint64_t)1 << (pos & 63);
            at->alto_cr_mask[pos >> 6] |= (uint64_t)1 << (pos & 63);
            free_bits--;
        }
    }
    if (!(at->alto_cr_mask[0] | at->alto_cr_mask[1])) return;
    /* ownership: a row touched by more than one partition is shared
     * (include/alto.hpp:970-995), then set the CR bit (:997-1020) */
    for (int n = 0; n < N; ++n) {
        const uint64_t *cm = &at->cr_masks[n * W];
        if (!(cm[0] | (W == 2 ? cm[1] : 0))) continue;
        uint64_t dim = at->dims[n];
        int *lo = (int *)malloc(sizeof(int) * dim);
        int *hi = (int *)malloc(sizeof(int) * dim);
        for (uint64_t r = 0; r < dim; ++r) { lo[r] = P; hi[r] = -1; }
        for (int p = 0; p < P; ++p)
            for (uint64_t i = at->prtn_ptr[p]; i < at->prtn_ptr[p + 1]; ++i) {
                uint64_t r = key_pext(at, i, n);
                if (p < lo[r]) lo[r] = p;
                if (p > hi[r]) hi[r] = p;
            }
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < nnz; ++i) {
            uint64_t r = key_pext(at, i, n);
            if (lo[r] != hi[r])
                for (int w = 0; w < W; ++w) at->idx[i * W + w] |= cm[w];
        }
        free(lo); free(hi);
    }
}

static or_alto *alloc_alto(int nmode, const uint64_t *dims, uint64_t nnz, int W, int nprtn)
{
    or_alto *at = (or_alto *)calloc(1, sizeof(or_alto));
    at->nmode = nmode; at->nprtn = nprtn; at->W = W; at->nnz = nnz;
    for (int n = 0; n < nmode; ++n) at->dims[n] = dims[n];
    at->idx = (uint64_t *)malloc(sizeof(uint64_t) * (nnz ? nnz : 1) * W);
    at->vals = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
    at->prtn_ptr = (uint64_t *)malloc(sizeof(uint64_t) * (nprtn + 1));
    at->intervals = (uint64_t *)malloc(sizeof(uint64_t) * nprtn * nmode * 2);
    return at;
}

or_alto *or_alto_create(int nmode, const uint64_t *dims, uint64_t nnz, const uint64_t *cidx,
                        const double *vals, int val_int, int W, int nprtn)
{
    if (nprtn < 1) return NULL;
    or_alto *at = alloc_alto(nmode, dims, nnz, W, nprtn);
    at->alto_bits = or_setup_masks(nmode, dims, W, at->masks, at->alto_mask);
    if (at->alto_bits < 0) { or_alto_destroy(at); return NULL; }
    /* linearize: include/alto.hpp:184-204 */
    uint64_t *perm = (uint64_t *)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nnz; ++i) {
        uint64_t k[2] = {0, 0};
        for (int n = 0; n < nmode; ++n) {
            uint64_t c = cidx[(uint64_t)n * nnz + i];
            if (W == 1) k[0] |= pdep_fast64(c, at->masks[n]);
            else {
                uint64_t o[2];
                or_pdep128(c, &at->masks[2 * n], o);
                k[0] |= o[0]; k[1] |= o[1];
            }
        }
        for (int w = 0; w < W; ++w) at->idx[i * W + w] = k[w];
        perm[i] = i;
    }
    /* sort: include/alto.hpp:772-803 */
    radix_sort_keys(at->idx, perm, nnz, W, at->alto_bits);
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nnz; ++i) {
        double v = vals[perm[i]];
        at->vals[i] = val_int ? (double)(unsigned long long)v : v;
    }
    free(perm);
    setup_oidx(at);
    prtn(at);
    return at;
}

or_alto *or_alto_wrap(int nmode, const uint64_t *dims, uint64_t nnz, int W, int nprtn,
                      const uint64_t *masks, const uint64_t *idx, const double *vals,
                      const uint64_t *prtn_ptr, const uint64_t *intervals,
                      const uint64_t *cr_masks, const uint64_t *oidx)
{
    or_alto *at = alloc_alto(nmode, dims, nnz, W, nprtn);
    at->alto_mask[0] = at->alto_mask[1] = 0;
    at->alto_cr_mask[0] = at->alto_cr_mask[1] = 0;
    memset(at->cr_masks, 0, sizeof(at->cr_masks));
    for (int n = 0; n < nmode; ++n)
        for (int w = 0; w < W; ++w) {
            at->masks[n * W + w] = masks[n * W + w];
            at->alto_mask[w] |= masks[n * W + w];
            at->cr_masks[n * W + w] = cr_masks ? cr_masks[n * W + w] : 0;
            at->alto_cr_mask[w] |= at->cr_masks[n * W + w];
        }
    at->alto_bits = __builtin_popcountll(at->alto_mask[0]) + (W == 2 ? __builtin_popcountll(at->alto_mask[1]) : 0);
    memcpy(at->idx, idx, sizeof(uint64_t) * nnz * W);
    memcpy(at->vals, vals, sizeof(double) * nnz);
    memcpy(at->prtn_ptr, prtn_ptr, sizeof(uint64_t) * (nprtn + 1));
    memcpy(at->intervals, intervals, sizeof(uint64_t) * nprtn * nmode * 2);
    at->is_oidx = 0;
    for (int n = 0; n < nmode; ++n)
        if (nnz / dims[n] <= OR_MIN_FIBER_REUSE) at->is_oidx = 1;
    if (at->is_oidx)
        for (int n = 0; n < nmode; ++n)
            if (nnz / dims[n] <= OR_MIN_FIBER_REUSE) {
                if (oidx) {
                    at->oidx[n] = (uint64_t *)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));
                    memcpy(at->oidx[n], oidx + (uint64_t)n * nnz, sizeof(uint64_t) * nnz);
                } else build_oidx(at, n);
            }
    return at;
}

void or_alto_destroy(or_alto *at)
{
    if (!at) return;
    free(at->idx); free(at->vals); free(at->prtn_ptr); free(at->intervals);
    for (int n = 0; n < OR_MAX_MODES; ++n) free(at->oidx[n]);
    if (at->ofibs) {
        for (int p = 0; p < at->nprtn; ++p) free(at->ofibs[p]);
        free(at->ofibs);
    }
    free(at);
}

void or_alto_get_masks(const or_alto *at, uint64_t *masks, uint64_t *alto_mask,
                       uint64_t *cr_masks, uint64_t *alto_cr_mask)
{
    int W = at->W;
    if (masks) memcpy(masks, at->masks, sizeof(uint64_t) * at->nmode * W);
    if (alto_mask) memcpy(alto_mask, at->alto_mask, sizeof(uint64_t) * W);
    if (cr_masks) memcpy(cr_masks, at->cr_masks, sizeof(uint64_t) * at->nmode * W);
    if (alto_cr_mask) memcpy(alto_cr_mask, at->alto_cr_mask, sizeof(uint64_t) * W);
}

void or_alto_get_idx(const or_alto *at, uint64_t *idx) { memcpy(idx, at->idx, sizeof(uint64_t) * at->nnz * at->W); }
void or_alto_get_vals(const or_alto *at, double *vals) { memcpy(vals, at->vals, sizeof(double) * at->nnz); }

void or_alto_get_prtn(const or_alto *at, uint64_t *prtn_ptr, uint64_t *intervals)
{
    memcpy(prtn_ptr, at->prtn_ptr, sizeof(uint64_t) * (at->nprtn + 1));
    memcpy(intervals, at->intervals, sizeof(uint64_t) * at->nprtn * at->nmode * 2);
}

int or_alto_is_oidx(const or_alto *at) { return at->is_oidx; }

int or_alto_get_oidx(const or_alto *at, int mode, uint64_t *oidx)
{
    if (!at->oidx[mode]) return 0;
    memcpy(oidx, at->oidx[mode], sizeof(uint64_t) * at->nnz);
    return 1;
}

void or_alto_delinearize(const or_alto *at, uint64_t *coords)
{
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < at->nnz; ++i)
        for (int n = 0; n < at->nmode; ++n) coords[(uint64_t)n * at->nnz + i] = key_pext(at, i, n);
}

double or_normsq(const or_alto *at)
{
    double s = 0.0;
    #pragma omp parallel for reduction(+ : s) schedule(static)
    for (uint64_t i = 0; i < at->nnz; ++i) s += at->vals[i] * at->vals[i];
    return s;
}

int or_omp_threads(void) { return omp_get_max_threads(); }
void or_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

/* ------------------------------------------------------------------------ */
/* MTTKRP variants (include/alto.hpp) */

/* row <- val * prod_{m != mode} U_m[row_m, :] */
static inline void krp_row(const or_alto *at, uint64_t i, int mode, double **U, uint64_t R, double *row)
{
    double v = at->vals[i];
    for (uint64_t r = 0; r < R; ++r) row[r] = v;
    for (int m = 0; m < at->nmode; ++m) {
        if (m == mode) continue;
        const double *u = U[m] + key_pext(at, i, m) * R;
        for (uint64_t r = 0; r < R; ++r) row[r] *= u[r];
    }
}

/* mttkrp_alto, include/alto.hpp:457-543 */
static void mttkrp_seq(or_alto *at, int mode, double **U, uint64_t R)
{
    double *row = (double *)malloc(sizeof(double) * R);
    double *out = U[mode];
    for (uint64_t i = 0; i < at->nnz; ++i) {
        krp_row(at, i, mode, U, R, row);
        double *o = out + key_pext(at, i, mode) * R;
        for (uint64_t r = 0; r < R; ++r) o[r] += row[r];
    }
    free(row);
}

/* mttkrp_alto_atomic / _atomic_cr, include/alto.hpp:1030-1158 */
static void mttkrp_atomic(or_alto *at, int mode, double **U, uint64_t R, int use_cr)
{
    const uint64_t *cm = &at->cr_masks[mode * at->W];
    double *out = U[mode];
    #pragma omp parallel for schedule(static, 1)
    for (int p = 0; p < at->nprtn; ++p) {
        double *row = (double *)malloc(sizeof(double) * R);
        for (uint64_t i = at->prtn_ptr[p]; i < at->prtn_ptr[p + 1]; ++i) {
            krp_row(at, i, mode, U, R, row);
            double *o = out + key_pext(at, i, mode) * R;
            if (!use_cr || key_has(at, i, cm)) {
                for (uint64_t r = 0; r < R; ++r) {
                    #pragma omp atomic update
                    o[r] += row[r];
                }
            } else {
                for (uint64_t r = 0; r < R; ++r) o[r] += row[r];
            }
        }
        free(row);
    }
}

/* create_da_mem(-1, ...), include/alto.hpp:324-370 */
static void ensure_da_mem(or_alto *at, uint64_t R)
{
    if (at->ofibs && at->ofibs_rank == R) return;
    if (at->ofibs) {
        for (int p = 0; p < at->nprtn; ++p) free(at->ofibs[p]);
        free(at->ofibs);
    }
    at->ofibs = (double **)calloc(at->nprtn, sizeof(double *));
    at->ofibs_rank = R;
    for (int p = 0; p < at->nprtn; ++p) {
        uint64_t nf = 0;
        /* sized over every mode (the reference sizes only modes with reuse
         * above MIN_FIBER_REUSE); a superset, so a forced da_mem_pull on a
         * low-reuse mode stays in bounds */
        for (int n = 0; n < at->nmode; ++n) {
            uint64_t s = at->intervals[(p * at->nmode + n) * 2], e = at->intervals[(p * at->nmode + n) * 2 + 1];
            uint64_t f = e >= s ? e - s + 1 : 0; /* empty partition: no fibers */
            if (f > nf) nf = f;
        }
        at->ofibs[p] = nf ? (double *)malloc(sizeof(double) * nf * R) : NULL;
    }
}

/* mttkrp_alto_da_mem_pull, include/alto.hpp:1227-1318 */
static void mttkrp_da_mem_pull(or_alto *at, int mode, double **U, uint64_t R)
{
    int N = at->nmode, P = at->nprtn;
    double *out = U[mode];
    ensure_da_mem(at, R);
    #pragma omp parallel
    {
        double *row = (double *)malloc(sizeof(double) * R);
        #pragma omp for schedule(static, 1)
        for (int p = 0; p < P; ++p) {
            uint64_t s = at->intervals[(p * N + mode) * 2], e = at->intervals[(p * N + mode) * 2 + 1];
            if (e < s) continue;
            double *o = at->ofibs[p];
            memset(o, 0, sizeof(double) * (e - s + 1) * R);
            for (uint64_t i = at->prtn_ptr[p]; i < at->prtn_ptr[p + 1]; ++i) {
                krp_row(at, i, mode, U, R, row);
                double *t = o + (key_pext(at, i, mode) - s) * R;
                for (uint64_t r = 0; r < R; ++r) t[r] += row[r];
            }
        }
        free(row);
        #pragma omp for schedule(static)
        for (uint64_t i = 0; i < at->dims[mode]; ++i) {
            for (int p = 0; p < P; ++p) {
                uint64_t s = at->intervals[(p * N + mode) * 2], e = at->intervals[(p * N + mode) * 2 + 1];
                if (i >= s && i <= e) {
                    const double *t = at->ofibs[p] + (i - s) * R;
                    for (uint64_t r = 0; r < R; ++r) out[i * R + r] += t[r];
                }
            }
        }
    }
}

/* mttkrp_alto_op_sched, include/alto.hpp:1320-1415 */
static void mttkrp_op_sched(or_alto *at, int mode, double **U, uint64_t R)
{
    const uint64_t *oi = at->oidx[mode];
    double *out = U[mode];
    #pragma omp parallel for schedule(static, 1)
    for (int p = 0; p < at->nprtn; ++p) {
        uint64_t s = at->prtn_ptr[p], e = at->prtn_ptr[p + 1];
        if (s >= e) continue;
        double *row = (double *)malloc(sizeof(double) * R);
        double *acc = (double *)calloc(R, sizeof(double));
        uint64_t last = key_pext(at, oi[s], mode);
        int boundary = 1;
        for (uint64_t i = s; i < e; ++i) {
            krp_row(at, oi[i], mode, U, R, row);
            uint64_t rid = key_pext(at, oi[i], mode);
            if (rid != last) {
                if (boundary) {
                    for (uint64_t r = 0; r < R; ++r) {
                        #pragma omp atomic update
                        out[last * R + r] += acc[r];
                    }
                    boundary = 0;
                } else {
                    for (uint64_t r = 0; r < R; ++r) out[last * R + r] += acc[r];
                }
                memset(acc, 0, sizeof(double) * R);
            }
            last = rid;
            for (uint64_t r = 0; r < R; ++r) acc[r] += row[r];
        }
        for (uint64_t r = 0; r < R; ++r) {
            #pragma omp atomic update
            out[last * R + r] += acc[r];
        }
        free(row); free(acc);
    }
}

/* dispatch, include/alto.hpp:545-623 */
int or_mttkrp_alto(or_alto *at, int mode, double **U, uint64_t R, int variant)
{
    if (variant == OR_MTTKRP_DISPATCH) {
        uint64_t reuse = at->nnz / at->dims[mode];
        if (reuse <= OR_MIN_FIBER_REUSE) {
            const uint64_t *cm = &at->cr_masks[mode * at->W];
            if (at->is_oidx && at->oidx[mode]) variant = OR_MTTKRP_OP_SCHED;
            else if (cm[0] | (at->W == 2 ? cm[1] : 0)) variant = OR_MTTKRP_ATOMIC_CR;
            else variant = OR_MTTKRP_ATOMIC;
        } else variant = OR_MTTKRP_DA_MEM_PULL;
    }
    /* atomic_cr is only sound when the mode owns a CR bit (the reference's
     * dispatch only picks it then); without one every row may be shared */
    if (variant == OR_MTTKRP_ATOMIC_CR) {
        const uint64_t *cm = &at->cr_masks[mode * at->W];
        if (!(cm[0] | (at->W == 2 ? cm[1] : 0))) variant = OR_MTTKRP_ATOMIC;
    }
    switch (variant) {
    case OR_MTTKRP_SEQ: mttkrp_seq(at, mode, U, R); break;
    case OR_MTTKRP_ATOMIC: mttkrp_atomic(at, mode, U, R, 0); break;
    case OR_MTTKRP_ATOMIC_CR: mttkrp_atomic(at, mode, U, R, 1); break;
    case OR_MTTKRP_DA_MEM_PULL: mttkrp_da_mem_pull(at, mode, U, R); break;
    case OR_MTTKRP_OP_SCHED:
        if (!at->oidx[mode]) build_oidx(at, mode);
        mttkrp_op_sched(at, mode, U, R);
        break;
    default: return -1;
    }
    return variant;
}

/* sequential COO MTTKRP, include/cpd.hpp:509-543 */
void or_mttkrp_coo(int nmode, uint64_t nnz, const uint64_t *cidx, const double *vals,
                   double **U, uint64_t R, int mode, double *out)
{
    double *row = (double *)malloc(sizeof(double) * R);
    for (uint64_t i = 0; i < nnz; ++i) {
        for (uint64_t r = 0; r < R; ++r) row[r] = vals[i];
        for (int m = 0; m < nmode; ++m) {
            if (m == mode) continue;
            const double *u = U[m] + cidx[(uint64_t)m * nnz + i] * R;
            for (uint64_t r = 0; r < R; ++r) row[r] *= u[r];
        }
        double *o = out + cidx[(uint64_t)mode * nnz + i] * R;
        for (uint64_t r = 0; r < R; ++r) o[r] += row[r];
    }
    free(row);
}

/* ------------------------------------------------------------------------ */
/* Kruskal model (src/kruskal_model.cpp) */

/* KruskalModelRandomInit, src/kruskal_model.cpp:109-125 (glibc rand) */
void or_random_init(int nmode, const uint64_t *dims, uint64_t R, unsigned seed, double **U, double *lambda)
{
    for (uint64_t r = 0; r < R; ++r) lambda[r] = 1.0;
    srand(seed);
    for (int n = 0; n < nmode; ++n)
        for (uint64_t i = 0; i < dims[n] * R; ++i) U[n][i] = (double)rand() / RAND_MAX;
}

/* update_gram: SYRK(RowMajor, Lower, Trans) = U^T U, src/kruskal_model.cpp:11-24.
 * Both triangles are written here (the reference defines only the lower). */
void or_gram(const double *U, uint64_t dim, uint64_t R, double *G)
{
    memset(G, 0, sizeof(double) * R * R);
    #pragma omp parallel
    {
        double *g = (double *)calloc(R * R, sizeof(double));
        #pragma omp for schedule(static)
        for (uint64_t i = 0; i < dim; ++i) {
            const double *u = U + i * R;
            for (uint64_t a = 0; a < R; ++a)
                for (uint64_t b = 0; b <= a; ++b) g[a * R + b] += u[a] * u[b];
        }
        #pragma omp critical
        for (uint64_t k = 0; k < R * R; ++k) G[k] += g[k];
        free(g);
    }
    for (uint64_t a = 0; a < R; ++a)
        for (uint64_t b = a + 1; b < R; ++b) G[a * R + b] = G[b * R + a];
}

/* KruskalModelNormalizeMode, src/kruskal_model.cpp:178-222 (L1 norms into lambda) */
void or_normalize_mode_l1(uint64_t dim, uint64_t R, double *U, double *lambda)
{
    double *cn = (double *)calloc(R, sizeof(double));
    for (uint64_t i = 0; i < dim; ++i)
        for (uint64_t r = 0; r < R; ++r) cn[r] += fabs(U[i * R + r]);
    for (uint64_t i = 0; i < dim; ++i)
        for (uint64_t r = 0; r < R; ++r) U[i * R + r] /= cn[r];
    for (uint64_t r = 0; r < R; ++r) lambda[r] *= cn[r];
    free(cn);
}

/* KruskalModelNormalize, src/kruskal_model.cpp:128-176 */
void or_normalize_all_l1(int nmode, const uint64_t *dims, uint64_t R, double **U, double *lambda)
{
    for (int n = 0; n < nmode; ++n) or_normalize_mode_l1(dims[n], R, U[n], lambda);
}

/* Mat2Norm / MatMaxNorm via KruskalModelNorm, src/kruskal_model.cpp:224-341.
 * Max norm: lambda_r = max(1, max_i U[i,r]) (signed max). */
void or_mat_norm(uint64_t dim, uint64_t R, double *U, double *lambda, int which)
{
    for (uint64_t r = 0; r < R; ++r) lambda[r] = 0.0;
    if (which == 0) {
        for (uint64_t i = 0; i < dim; ++i)
            for (uint64_t r = 0; r < R; ++r) lambda[r] += U[i * R + r] * U[i * R + r];
        for (uint64_t r = 0; r < R; ++r) lambda[r] = sqrt(lambda[r]);
    } else {
        for (uint64_t i = 0; i < dim; ++i)
            for (uint64_t r = 0; r < R; ++r)
                if (U[i * R + r] > lambda[r]) lambda[r] = U[i * R + r];
        for (uint64_t r = 0; r < R; ++r) if (lambda[r] < 1.0) lambda[r] = 1.0;
    }
    for (uint64_t i = 0; i < dim; ++i)
        for (uint64_t r = 0; r < R; ++r) U[i * R + r] /= lambda[r];
}

/* RedistributeLambda, src/kruskal_model.cpp:355-378 */
void or_redistribute_lambda(uint64_t dim, uint64_t R, double *U, double *lambda)
{
    for (uint64_t i = 0; i < dim; ++i)
        for (uint64_t r = 0; r < R; ++r) U[i * R + r] *= lambda[r];
    for (uint64_t r = 0; r < R; ++r) lambda[r] = 1.0;
}

/* ------------------------------------------------------------------------ */
/* pseudo_inverse, include/cpd.hpp:546-646 */

/* symmetric Jacobi eigen-decomposition (cyclic sweeps) of an RxR matrix */
static void jacobi_eig(double *A, double *Q, int R)
{
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < R; ++j) Q[i * R + j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < R; ++i)
            for (int j = 0; j < R; ++j) {
                tot += A[i * R + j] * A[i * R + j];
                if (i != j) off += A[i * R + j] * A[i * R + j];
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < R - 1; ++p)
            for (int q = p + 1; q < R; ++q) {
                double apq = A[p * R + q];
                if (fabs(apq) < 1e-300) continue;
                double app = A[p * R + p], aqq = A[q * R + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < R; ++k) {
                    double akp = A[k * R + p], akq = A[k * R + q];
                    A[k * R + p] = c * akp - s * akq;
                    A[k * R + q] = s * akp + c * akq;
                }
                for (int k = 0; k < R; ++k) {
                    double apk = A[p * R + k], aqk = A[q * R + k];
                    A[p * R + k] = c * apk - s * aqk;
                    A[q * R + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < R; ++k) {
                    double qkp = Q[k * R + p], qkq = Q[k * R + q];
                    Q[k * R + p] = c * qkp - s * qkq;
                    Q[k * R + q] = s * qkp + c * qkq;
                }
            }
    }
}

int or_pseudo_inverse(int nmode, uint64_t R, double **grams, double *U, uint64_t dim, int mode)
{
    double *V = (double *)malloc(sizeof(double) * R * R);
    int first = 1;
    for (int m = 0; m < nmode; ++m) {
        if (m == mode) continue;
        for (uint64_t k = 0; k < R * R; ++k) V[k] = first ? grams[m][k] : V[k] * grams[m][k];
        first = 0;
    }
    if (first) for (uint64_t k = 0; k < R * R; ++k) V[k] = 1.0;
    /* POTRF('L'): unblocked Cholesky of the lower triangle */
    double *L = (double *)calloc(R * R, sizeof(double));
    int ok = 1;
    for (uint64_t j = 0; j < R && ok; ++j) {
        double d = V[j * R + j];
        for (uint64_t k = 0; k < j; ++k) d -= L[j * R + k] * L[j * R + k];
        if (!(d > 0.0)) { ok = 0; break; }
        d = sqrt(d);
        L[j * R + j] = d;
        for (uint64_t i = j + 1; i < R; ++i) {
            double s = V[i * R + j];
            for (uint64_t k = 0; k < j; ++k) s -= L[i * R + k] * L[j * R + k];
            L[i * R + j] = s / d;
        }
    }
    int used_fallback = 0;
    if (ok) {
        /* POTRS: each row u of U solves (L L^T) x = u^T */
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < dim; ++i) {
            double *u = U + i * R;
            for (uint64_t a = 0; a < R; ++a) {
                double s = u[a];
                for (uint64_t k = 0; k < a; ++k) s -= L[a * R + k] * u[k];
                u[a] = s / L[a * R + a];
            }
            for (uint64_t a = R; a-- > 0;) {
                double s = u[a];
                for (uint64_t k = a + 1; k < R; ++k) s -= L[k * R + a] * u[k];
                u[a] = s / L[a * R + a];
            }
        }
    } else {
        /* minimum-norm least squares: x = V^+ u with V symmetrized from its lower triangle */
        used_fallback = 1;
        double *A = (double *)malloc(sizeof(double) * R * R);
        double *Q = (double *)malloc(sizeof(double) * R * R);
        double *P = (double *)calloc(R * R, sizeof(double));
        for (uint64_t i = 0; i < R; ++i)
            for (uint64_t j = 0; j < R; ++j) A[i * R + j] = i >= j ? V[i * R + j] : V[j * R + i];
        jacobi_eig(A, Q, (int)R);
        double emax = 0.0;
        for (uint64_t i = 0; i < R; ++i) if (fabs(A[i * R + i]) > emax) emax = fabs(A[i * R + i]);
        double tol = emax * (double)R * DBL_EPSILON;
        for (uint64_t k = 0; k < R; ++k) {
            double e = A[k * R + k];
            if (fabs(e) <= tol) continue;
            for (uint64_t i = 0; i < R; ++i)
                for (uint64_t j = 0; j < R; ++j) P[i * R + j] += Q[i * R + k] * Q[j * R + k] / e;
        }
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < dim; ++i) {
            double tmp[512];
            double *u = U + i * R;
            for (uint64_t a = 0; a < R; ++a) {
                double s = 0.0;
                for (uint64_t k = 0; k < R; ++k) s += P[a * R + k] * u[k];
                tmp[a] = s;
            }
            memcpy(u, tmp, sizeof(double) * R);
        }
        free(A); free(Q); free(P);
    }
    free(L); free(V);
    return used_fallback;
}

/* cpd_fit_alto, include/cpd.hpp:189-271 */
double or_cpd_fit(int nmode, uint64_t R, const uint64_t *dims, double **U, const double *lambda,
                  double **grams, const double *Um, double normX)
{
    const double *UN = U[nmode - 1];
    uint64_t dN = dims[nmode - 1];
    double *acc = (double *)calloc(R, sizeof(double));
    for (uint64_t i = 0; i < dN; ++i)
        for (uint64_t r = 0; r < R; ++r) acc[r] += UN[i * R + r] * Um[i * R + r];
    double inner = 0.0;
    for (uint64_t r = 0; r < R; ++r) inner += acc[r] * lambda[r];
    double normU = 0.0;
    for (uint64_t i = 0; i < R; ++i)
        for (uint64_t j = 0; j <= i; ++j) {
            double t = lambda[i] * lambda[j];
            for (int n = 0; n < nmode; ++n) t *= grams[n][i * R + j];
            normU += i == j ? t : 2.0 * t;
        }
    normU = fabs(normU);
    double res = normX + normU - 2.0 * inner;
    if (res > 0.0) res = sqrt(res);
    free(acc);
    return 1.0 - res / sqrt(normX);
}

/* cpd_alto, include/cpd.hpp:43-186 */
int or_cpd_als(or_alto *at, uint64_t R, int max_iters, double eps, double **U, double *lambda,
               double *fit_trace, int variant)
{
    int N = at->nmode;
    double normX = or_normsq(at);
    double *grams[OR_MAX_MODES];
    for (int n = 0; n < N; ++n) {
        grams[n] = (double *)malloc(sizeof(double) * R * R);
        or_gram(U[n], at->dims[n], R, grams[n]);
    }
    double *scratch = (double *)malloc(sizeof(double) * at->dims[N - 1] * R);
    double fit = 0.0, prev = 0.0;
    int iters = max_iters;
    for (int it = 0; it < max_iters; ++it) {
        for (int j = 0; j < N; ++j) {
            memset(U[j], 0, sizeof(double) * at->dims[j] * R);
            or_mttkrp_alto(at, j, U, R, variant);
            if (j == N - 1) memcpy(scratch, U[j], sizeof(double) * at->dims[j] * R);
            or_pseudo_inverse(N, R, grams, U[j], at->dims[j], j);
            or_mat_norm(at->dims[j], R, U[j], lambda, it == 0 ? 0 : 1);
            or_gram(U[j], at->dims[j], R, grams[j]);
        }
        fit = or_cpd_fit(N, R, at->dims, U, lambda, grams, scratch, normX);
        if (fit_trace) fit_trace[it] = fit;
        if (it > 0 && fabs(prev - fit) < eps) { iters = it + 1; break; }
        prev = fit;
    }
    for (int n = 0; n < N; ++n) free(grams[n]);
    free(scratch);
    return iters;
}

/* ------------------------------------------------------------------------ */
/* CP-APR (include/apr.hpp) */

/* CalculatePhi (da_mem_pull semantics, include/apr.hpp:383-501):
 * Phi[i,:] = sum_{x: row_n(x)=i} (v_x / max(<B[i,:], pi_x>, eps)) * pi_x */
void or_phi(or_alto *at, int mode, double **U, uint64_t R, double eps, double *Phi)
{
    int N = at->nmode, P = at->nprtn;
    const double *B = U[mode];
    ensure_da_mem(at, R);
    uint64_t dim = at->dims[mode];
    int pull = at->nnz / dim > OR_MIN_FIBER_REUSE;
    if (!pull) memset(Phi, 0, sizeof(double) * dim * R);
    #pragma omp parallel
    {
        double *row = (double *)malloc(sizeof(double) * R);
        #pragma omp for schedule(static, 1)
        for (int p = 0; p < P; ++p) {
            uint64_t s = at->intervals[(p * N + mode) * 2], e = at->intervals[(p * N + mode) * 2 + 1];
            if (pull) {
                if (e < s) continue;
                memset(at->ofibs[p], 0, sizeof(double) * (e - s + 1) * R);
            }
            for (uint64_t i = at->prtn_ptr[p]; i < at->prtn_ptr[p + 1]; ++i) {
                for (uint64_t r = 0; r < R; ++r) row[r] = 1.0;
                for (int m = 0; m < N; ++m) {
                    if (m == mode) continue;
                    const double *u = U[m] + key_pext(at, i, m) * R;
                    for (uint64_t r = 0; r < R; ++r) row[r] *= u[r];
                }
                uint64_t rid = key_pext(at, i, mode);
                double v = 0.0;
                for (uint64_t r = 0; r < R; ++r) v += B[rid * R + r] * row[r];
                v = at->vals[i] / fmax(v, eps);
                if (pull) {
                    double *t = at->ofibs[p] + (rid - s) * R;
                    for (uint64_t r = 0; r < R; ++r) t[r] += v * row[r];
                } else {
                    for (uint64_t r = 0; r < R; ++r) {
                        #pragma omp atomic update
                        Phi[rid * R + r] += v * row[r];
                    }
                }
            }
        }
        free(row);
        if (pull) {
            #pragma omp for schedule(static)
            for (uint64_t i = 0; i < dim; ++i) {
                for (uint64_t r = 0; r < R; ++r) Phi[i * R + r] = 0.0;
                for (int p = 0; p < P; ++p) {
                    uint64_t s = at->intervals[(p * N + mode) * 2], e = at->intervals[(p * N + mode) * 2 + 1];
                    if (i >= s && i <= e) {
                        const double *t = at->ofibs[p] + (i - s) * R;
                        for (uint64_t r = 0; r < R; ++r) Phi[i * R + r] += t[r];
                    }
                }
            }
        }
    }
}

/* tt_logLikelihood<LIT>, include/apr.hpp:850-982 */
double or_loglikelihood(or_alto *at, double **U, double *lambda, uint64_t R, double eps)
{
    int N = at->nmode;
    or_redistribute_lambda(at->dims[0], R, U[0], lambda);
    double logsum = 0.0;
    #pragma omp parallel
    {
        double *s = (double *)malloc(sizeof(double) * R);
        #pragma omp for schedule(static) reduction(+ : logsum)
        for (uint64_t i = 0; i < at->nnz; ++i) {
            for (uint64_t r = 0; r < R; ++r) s[r] = 1.0;
            for (int n = 0; n < N; ++n) {
                const double *u = U[n] + key_pext(at, i, n) * R;
                for (uint64_t r = 0; r < R; ++r) s[r] *= u[r];
            }
            double t = 0.0;
            for (uint64_t r = 0; r < R; ++r) t += s[r];
            if (t < eps) t = eps;
            logsum += at->vals[i] * log(t);
        }
        free(s);
    }
    double fsum = 0.0;
    for (uint64_t k = 0; k < at->dims[0] * R; ++k) fsum += U[0][k];
    for (uint64_t r = 0; r < R; ++r) {
        double t = 0.0;
        for (uint64_t i = 0; i < at->dims[0]; ++i) t += fabs(U[0][i * R + r]);
        for (uint64_t i = 0; i < at->dims[0]; ++i) U[0][i * R + r] /= t;
        lambda[r] *= t;
    }
    return logsum - fsum;
}

/* Planted mistakes, compiled only into the mutant libraries the oracle's own
 * tests build (OR_MUTANT = 1..4) to show that those tests catch them. */
#ifndef OR_MUTANT
#define OR_MUTANT 0
#endif
#if OR_MUTANT == 1
#define OR_KKT_PICK fmax /* instead of fmin */
#else
#define OR_KKT_PICK fmin
#endif

/* cp_apr_alto, include/apr.hpp:984-1376, with an optional observer called at
 * the driver's steps (test infrastructure: the observer only reads). */
int or_cp_apr_hooked(or_alto *at, uint64_t R, int max_iters, int max_inner, double eps, double eps_div_zero,
                     double kappa_tol, double kappa, int dobench, double **U, double *lambda, or_apr_result *res,
                     or_apr_hook hook, void *ctx)
{
    int N = at->nmode;
    double *Phi[OR_MAX_MODES];
    for (int n = 0; n < N; ++n) Phi[n] = (double *)calloc(at->dims[n] * R, sizeof(double));
    double kkt_mode[OR_MAX_MODES] = {0};
    double *kkt = (double *)calloc(max_iters + 1, sizeof(double));
    int *n_inner = (int *)calloc(max_iters + 1, sizeof(int));
    memset(res, 0, sizeof(*res));
    /* KruskalModelNormalize, include/apr.hpp:1105 */
    or_normalize_all_l1(N, at->dims, R, U, lambda);
    if (hook) hook(ctx, OR_APR_EV_START, 0, -1, -1, U, lambda, NULL, 0.0);
    int iter;
    for (iter = 0; iter < max_iters; ++iter) {
        int converged = 0;
        for (int n = 0; n < N; ++n) {
            uint64_t dim = at->dims[n];
            double *B = U[n], *Ph = Phi[n];
            /* kappa fix, include/apr.hpp:1176-1189 */
            if (iter > 0) {
                if (hook) hook(ctx, OR_APR_EV_PRE_FIX, iter, n, -1, U, lambda, Ph, 0.0);
                for (uint64_t k = 0; k < dim * R; ++k)
#if OR_MUTANT == 2
                    if (B[k] < kappa_tol && Ph[k] < 1.0) B[k] += kappa;
#else
                    if (B[k] < kappa_tol && Ph[k] > 1.0) B[k] += kappa;
#endif
                if (hook) hook(ctx, OR_APR_EV_POST_FIX, iter, n, -1, U, lambda, Ph, 0.0);
            }
            or_redistribute_lambda(dim, R, B, lambda); /* :1192 */
            for (int inner = 0; inner < max_inner; ++inner) {
                n_inner[iter]++;
                res->inner_per_mode[n]++;
                or_phi(at, n, U, R, eps_div_zero, Ph); /* :1205 */
                /* KKT violation, :1207-1215 */
                double kv = 0.0;
                for (uint64_t k = 0; k < dim * R; ++k) {
                    double mc = fabs(OR_KKT_PICK(B[k], 1.0 - Ph[k]));
                    kv = fmax(kv, mc);
                }
                kkt_mode[n] = kv;
                if (hook) hook(ctx, OR_APR_EV_PHI, iter, n, inner, U, lambda, Ph, kv);
                if (kv < eps && !dobench) break; /* :1217-1221 */
                for (uint64_t k = 0; k < dim * R; ++k) B[k] *= Ph[k]; /* multiplicative update, :1225-1229 */
                if (hook) hook(ctx, OR_APR_EV_MU, iter, n, inner, U, lambda, Ph, kv);
            }
            /* KruskalModelNormalizeMode, :1243 */
#if OR_MUTANT == 3
            for (uint64_t r = 0; r < R; ++r) { /* 2-norms into lambda: the wrong normalisation */
                double c = 0.0;
                for (uint64_t i = 0; i < dim; ++i) c += B[i * R + r] * B[i * R + r];
                c = sqrt(c);
                for (uint64_t i = 0; i < dim; ++i) B[i * R + r] /= c;
                lambda[r] *= c;
            }
#else
            or_normalize_mode_l1(dim, R, B, lambda);
#endif
            if (hook) hook(ctx, OR_APR_EV_NORM, iter, n, -1, U, lambda, Ph, kkt_mode[n]);
        }
        /* :1248-1252 and the outer stop :1261 (kkt[-1] at iter 0 is read as "no stop") */
        kkt[iter] = kkt_mode[0];
        for (int n = 1; n < N; ++n) kkt[iter] = fmax(kkt[iter], kkt_mode[n]);
        if (hook) hook(ctx, OR_APR_EV_OUTER, iter, -1, -1, U, lambda, NULL, kkt[iter]);
#if OR_MUTANT == 4
        if (iter > 0 && fabs(kkt[iter] - kkt[iter - 1]) > eps * eps && !dobench) converged = 1;
#else
        if (iter > 0 && fabs(kkt[iter] - kkt[iter - 1]) <= eps * eps && !dobench) converged = 1;
#endif
        if (converged) break; /* as the reference: the converged iteration is not counted */
    }
    int total = 0;
    for (int i = 0; i < iter; ++i) total += n_inner[i];
    res->outer_iters = iter;
    res->total_inner = total;
    res->kkt = iter > 0 ? kkt[iter - 1] : 0.0;
    res->loglik = or_loglikelihood(at, U, lambda, R, eps_div_zero);
    /* least-squares fit of the final model, include/apr.hpp:1305-1330 */
    double *grams[OR_MAX_MODES];
    for (int n = 0; n < N; ++n) {
        grams[n] = (double *)malloc(sizeof(double) * R * R);
        or_gram(U[n], at->dims[n], R, grams[n]);
    }
    double *tm = (double *)calloc(at->dims[N - 1] * R, sizeof(double));
    double *saved = U[N - 1];
    U[N - 1] = tm;
    mttkrp_atomic(at, N - 1, U, R, 0);
    U[N - 1] = saved;
    res->ls_fit = or_cpd_fit(N, R, at->dims, U, lambda, grams, tm, or_normsq(at));
    for (int n = 0; n < N; ++n) { free(grams[n]); free(Phi[n]); }
    free(tm); free(kkt); free(n_inner);
    return iter;
}

int or_cp_apr(or_alto *at, uint64_t R, int max_iters, int max_inner, double eps, double eps_div_zero,
              double kappa_tol, double kappa, int dobench, double **U, double *lambda, or_apr_result *res)
{
    return or_cp_apr_hooked(at, R, max_iters, max_inner, eps, eps_div_zero, kappa_tol, kappa, dobench, U, lambda,
                            res, NULL, NULL);
}
