// Drop-in C++ API of the reference ALTO library (IntelLabs/ALTO) over
// libalto_hip.  The names, parameter order, data layouts and in-place
// semantics follow the reference headers; the work runs on the GPU.
//
//   reference                                   here
//   SparseTensor      (include/sptensor.hpp:8)  same fields
//   KruskalModel      (include/kruskal_model.hpp:8)  same fields
//   AltoTensor<LIT>   (include/alto.hpp:27)     same scalar fields + device handle
//   create_alto       (include/alto.hpp:108)    device linearize / sort / partition
//   mttkrp_alto[_par] (include/alto.hpp:457,617) device MTTKRP, adds into factors[mode]
//   mttkrp_alto_batched (none)                  many tensors' MTTKRPs in one launch
//   cpd_alto          (include/cpd.hpp:43)      device CP-ALS, updates M in place
//   cp_apr_alto       (include/apr.hpp:984)     device CP-APR, updates M in place
//
// Errors: the reference asserts / exits; here a failed call prints
// alto_last_error() and aborts, so a misuse stops the program as there.
#ifndef ALTO_DROPIN_ALTO_HPP_
#define ALTO_DROPIN_ALTO_HPP_

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <memory>
#include <vector>

#include "../alto_hip.h"

#ifndef MAX_NUM_MODES
#define MAX_NUM_MODES 5
#endif

typedef unsigned long long IType;  // include/common.hpp:38-42
typedef double FType;              // include/common.hpp:45-58
#ifdef VAL_TYPE_INT64
typedef unsigned long long ValType;
#else
typedef FType ValType;
#endif

typedef enum Model_ { ALS, APR } Model;

struct SparseTensor {  // include/sptensor.hpp:8-26
    int nmodes;
    IType* dims;
    IType nnz;
    IType** cidx;
    FType* vals;
};

struct KruskalModel {  // include/kruskal_model.hpp:8-15
    int mode;
    IType* dims;
    IType rank;
    FType** U;
    FType* lambda;
};

typedef struct Interval_ {
    IType start;
    IType stop;
} Interval;

template <typename LIT>
struct AltoTensor {  // include/alto.hpp:27-54 (device-resident nonzeros)
    int nmode;
    int nprtn;
    IType nnz;
    IType* dims;
    LIT alto_mask;
    LIT* mode_masks;
    IType* prtn_ptr;
    Interval* prtn_intervals;
    LIT alto_cr_mask;
    LIT* cr_masks;
    bool is_oidx;
    alto_tensor_t* dev;  // the device tensor (keys, values, per-mode streams)
};

namespace alto_dropin {

inline void check(int status, const char* what)
{
    if (status != ALTO_OK) {
        fprintf(stderr, "%s failed: %s\n", what, alto_last_error());
        abort();
    }
}

template <typename LIT> inline int lit_bits() { return (int)(8 * sizeof(LIT)); }

template <typename LIT> inline LIT lit_from(const uint64_t* w, int words)
{
    LIT v = (LIT)w[0];
    if constexpr (sizeof(LIT) == 16)
        if (words == 2) v |= ((LIT)w[1]) << 64;
    return v;
}

// device copies of host factor matrices for one call
struct DeviceFactors {
    std::vector<void*> p;
    DeviceFactors(int nmodes, const IType* dims, IType rank, FType** U)
    {
        p.resize(nmodes);
        for (int m = 0; m < nmodes; ++m) {
            check(alto_device_malloc(&p[m], dims[m] * rank * sizeof(FType)), "alto_device_malloc");
            check(alto_memcpy_h2d(p[m], U[m], dims[m] * rank * sizeof(FType)), "alto_memcpy_h2d");
        }
    }
    ~DeviceFactors()
    {
        for (void* q : p) alto_device_free(q);
    }
};

}  // namespace alto_dropin

// ---------------------------------------------------------------- ALTO format
template <typename LIT>
static inline void create_alto(SparseTensor* spt, AltoTensor<LIT>** at, int nprtn)
{
    using namespace alto_dropin;
    alto_tensor_t* t = nullptr;
#ifdef VAL_TYPE_INT64
    const int vk = ALTO_VAL_U64;
#else
    const int vk = ALTO_VAL_F64;
#endif
    check(alto_tensor_create(spt->nmodes, (const uint64_t*)spt->dims, spt->nnz, (const uint64_t* const*)spt->cidx, spt->vals,
                             ALTO_PTR_HOST, vk, lit_bits<LIT>(), nprtn, nullptr, &t),
          "create_alto");
    alto_tensor_info_t inf;
    check(alto_tensor_info(t, &inf), "alto_tensor_info");
    AltoTensor<LIT>* a = new AltoTensor<LIT>();
    a->nmode = inf.nmodes;
    a->nprtn = inf.nprtn;
    a->nnz = inf.nnz;
    a->dims = new IType[inf.nmodes];
    a->mode_masks = new LIT[inf.nmodes];
    a->cr_masks = new LIT[inf.nmodes];
    for (int m = 0; m < inf.nmodes; ++m) {
        a->dims[m] = inf.dims[m];
        a->mode_masks[m] = lit_from<LIT>(&inf.mode_masks[m * inf.lit_words], inf.lit_words);
        a->cr_masks[m] = lit_from<LIT>(&inf.cr_masks[m * inf.lit_words], inf.lit_words);
    }
    a->alto_mask = lit_from<LIT>(inf.alto_mask, inf.lit_words);
    a->alto_cr_mask = lit_from<LIT>(inf.alto_cr_mask, inf.lit_words);
    a->is_oidx = inf.is_oidx != 0;
    a->prtn_ptr = new IType[nprtn + 1];
    std::vector<uint64_t> iv((size_t)nprtn * inf.nmodes * 2);
    check(alto_tensor_export(t, nullptr, nullptr, (uint64_t*)a->prtn_ptr, iv.data()), "alto_tensor_export");
    a->prtn_intervals = new Interval[(size_t)nprtn * inf.nmodes];
    for (size_t k = 0; k < (size_t)nprtn * inf.nmodes; ++k) a->prtn_intervals[k] = {iv[2 * k], iv[2 * k + 1]};
    a->dev = t;
    *at = a;
}

template <typename LIT>
static inline void destroy_alto(AltoTensor<LIT>* at)
{
    alto_tensor_destroy(at->dev);
    delete[] at->dims;
    delete[] at->mode_masks;
    delete[] at->cr_masks;
    delete[] at->prtn_ptr;
    delete[] at->prtn_intervals;
    delete at;
}

// The device MTTKRP needs no per-partition fiber copies: create_da_mem /
// destroy_da_mem keep their signatures (include/alto.hpp:324-398) and do nothing.
template <typename LIT>
static inline void create_da_mem(int, IType, AltoTensor<LIT>*, FType*** ofibs)
{
    *ofibs = nullptr;
}
template <typename LIT>
static inline void destroy_da_mem(AltoTensor<LIT>*, FType**, IType, int)
{
}

// mttkrp_alto_par (include/alto.hpp:617-623): factors[target_mode] += MTTKRP
template <typename LIT>
static inline void mttkrp_alto_par(int target_mode, FType** factors, IType rank, AltoTensor<LIT>* at, void* /*wlocks*/,
                                   FType** /*ofibs*/)
{
    using namespace alto_dropin;
    DeviceFactors d(at->nmode, at->dims, rank, factors);
    check(alto_mttkrp(at->dev, target_mode, rank, (const void* const*)d.p.data(), d.p[target_mode], ALTO_PREC_F64, 1,
                      nullptr),
          "mttkrp_alto_par");
    check(alto_device_synchronize(), "synchronize");
    check(alto_memcpy_d2h(factors[target_mode], d.p[target_mode], at->dims[target_mode] * rank * sizeof(FType)),
          "alto_memcpy_d2h");
}

// mttkrp_alto (include/alto.hpp:457-543): same result
template <typename LIT>
void mttkrp_alto(int target_mode, FType** factors, IType rank, AltoTensor<LIT>* at)
{
    mttkrp_alto_par(target_mode, factors, rank, at, nullptr, nullptr);
}

// Many MTTKRPs in one device launch (no reference counterpart; there each
// tensor takes its own mttkrp_alto_par call): for every job i,
// factors[i][target_modes[i]] += the MTTKRP of ats[i] along target_modes[i]
// with factors[i]; one rank for all jobs.
template <typename LIT>
static inline void mttkrp_alto_batched(int njobs, const int* target_modes, FType*** factors, IType rank,
                                       AltoTensor<LIT>** ats)
{
    using namespace alto_dropin;
    std::vector<std::unique_ptr<DeviceFactors>> d;
    std::vector<alto_tensor_t*> ts(njobs);
    std::vector<const void* const*> fp(njobs);
    std::vector<void*> outs(njobs);
    for (int i = 0; i < njobs; ++i) {
        d.emplace_back(new DeviceFactors(ats[i]->nmode, ats[i]->dims, rank, factors[i]));
        ts[i] = ats[i]->dev;
        fp[i] = (const void* const*)d[i]->p.data();
        outs[i] = d[i]->p[target_modes[i]];
    }
    check(alto_mttkrp_batched(njobs, ts.data(), target_modes, rank, fp.data(), outs.data(), ALTO_PREC_F64, 1,
                              nullptr),
          "mttkrp_alto_batched");
    check(alto_device_synchronize(), "synchronize");
    for (int i = 0; i < njobs; ++i) {
        const int m = target_modes[i];
        check(alto_memcpy_d2h(factors[i][m], d[i]->p[m], ats[i]->dims[m] * rank * sizeof(FType)), "alto_memcpy_d2h");
    }
}

// ---------------------------------------------------------------- Kruskal model
static inline void CreateKruskalModel(int mode, IType* dims, IType rank, KruskalModel** M_)
{
    KruskalModel* M = new KruskalModel();
    M->mode = mode;
    M->rank = rank;
    M->dims = new IType[mode];
    memcpy(M->dims, dims, sizeof(IType) * mode);
    M->U = new FType*[mode];
    for (int n = 0; n < mode; ++n) M->U[n] = new FType[dims[n] * rank];
    M->lambda = new FType[rank];
    *M_ = M;
}

static inline void KruskalModelRandomInit(KruskalModel* M, unsigned int seed)
{
    alto_dropin::check(alto_kruskal_random_init(M->mode, (const uint64_t*)M->dims, M->rank, seed, M->U, M->lambda),
                       "KruskalModelRandomInit");
}

static inline void DestroyKruskalModel(KruskalModel* M)
{
    for (int n = 0; n < M->mode; ++n) delete[] M->U[n];
    delete[] M->U;
    delete[] M->dims;
    delete[] M->lambda;
    delete M;
}

// update_gram (src/kruskal_model.cpp:11-24): both triangles are written
static inline void update_gram(FType* gram, KruskalModel* M, int mode)
{
    using namespace alto_dropin;
    void *u = nullptr, *g = nullptr;
    const IType R = M->rank;
    check(alto_device_malloc(&u, M->dims[mode] * R * sizeof(FType)), "alto_device_malloc");
    check(alto_device_malloc(&g, R * R * sizeof(FType)), "alto_device_malloc");
    check(alto_memcpy_h2d(u, M->U[mode], M->dims[mode] * R * sizeof(FType)), "h2d");
    check(alto_update_gram(u, M->dims[mode], R, ALTO_PREC_F64, (double*)g, nullptr), "update_gram");
    check(alto_memcpy_d2h(gram, g, R * R * sizeof(FType)), "d2h");
    alto_device_free(u);
    alto_device_free(g);
}

// ---------------------------------------------------------------- CP-ALS / CP-APR
template <typename LIT>
void cpd_alto(AltoTensor<LIT>* AT, KruskalModel* M, int max_iters, double epsilon)
{
    int iters = 0;
    alto_cpd_timers_t tm;
    alto_dropin::check(alto_cpd_als(AT->dev, M->rank, max_iters, epsilon, ALTO_PREC_F64, M->U, M->lambda, nullptr,
                                    &iters, &tm, 1, nullptr),
                       "cpd_alto");
}

// cpd_fit_alto (include/cpd.hpp:189-271): the fit of model M; the Grams and
// the last-mode MTTKRP are recomputed on the device from M.
template <typename LIT>
double cpd_fit_alto(AltoTensor<LIT>* AT, KruskalModel* M, FType** /*grams*/, FType* /*U_mttkrp*/, FType /*normAT*/)
{
    double fit = 0;
    alto_dropin::check(alto_cpd_fit(AT->dev, M->rank, M->U, M->lambda, &fit), "cpd_fit_alto");
    return fit;
}

template <typename LIT>
static inline void cp_apr_alto(AltoTensor<LIT>* AT, KruskalModel* M, int max_iters, int max_inner_iters,
                               FType epsilon, FType eps_div_zero, FType kappa_tol, FType kappa, bool dobench)
{
    alto_apr_result_t res;
    alto_dropin::check(alto_cp_apr(AT->dev, M->rank, max_iters, max_inner_iters, epsilon, eps_div_zero, kappa_tol,
                                   kappa, dobench ? 1 : 0, M->U, M->lambda, &res, 1, nullptr),
                       "cp_apr_alto");
}

#endif  // ALTO_DROPIN_ALTO_HPP_
