// ALTO tensors of the reference library (include/alto.hpp) over libalto_hip.
// The names, parameter order, data layouts and in-place semantics follow the
// reference; linearization, sorting, partitioning and the MTTKRP run on the
// GPU, where the tensor's nonzeros live (AltoTensor::dev).
//
//   reference                                   here
//   AltoTensor<LIT>   (include/alto.hpp:27-54)  same fields + device handle
//   create_alto       (include/alto.hpp:108)    device linearize / sort / partition
//   destroy_alto      (include/alto.hpp:296)    frees host fields and device tensor
//   create_da_mem / destroy_da_mem (:324-398)   no per-partition buffers needed
//   mttkrp_alto[_par] (include/alto.hpp:457,617) device MTTKRP, adds into factors[mode]
//   mttkrp_alto_batched (none)                  many tensors' MTTKRPs in one launch
//
// The host copies idx / vals / oidx that the reference's structure carries
// are filled by create_alto from the device tensor (the reference builds
// them on the host); define ALTO_DROPIN_DEVICE_ONLY to leave them NULL and
// skip the copies, for programs that only hand the tensor to the device.
// Including this header also declares the model, COO tensor and common
// helpers (kruskal_model.hpp, sptensor.hpp, common.hpp).
#ifndef ALTO_DROPIN_ALTO_HPP_
#define ALTO_DROPIN_ALTO_HPP_

#include <assert.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <memory>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "kruskal_model.hpp"
#include "sptensor.hpp"

// include/alto.hpp:90: modes with at most this many nonzeros per row get an
// output-ordered index (oidx)
#ifndef MIN_FIBER_REUSE
#define MIN_FIBER_REUSE 4
#endif

// include/alto.hpp:22-25
typedef struct Interval_ {
    IType start;
    IType stop;
} Interval;

template <typename LIT>
struct AltoTensor {  // include/alto.hpp:27-54
    int nmode;                  // number of modes
    int nprtn;                  // number of partitions
    IType nnz;                  // number of nonzeros
    IType* dims;                // dimensions
    LIT alto_mask;
    LIT* mode_masks;            // gather/scatter masks (nmode)
    LIT* idx;                   // ALTO keys, CR bits included (host copy)
    ValType* vals;              // nonzeros in key order (host copy)
    IType* prtn_ptr;            // nprtn + 1 partition bounds
    Interval* prtn_intervals;   // per-partition index intervals (nprtn * nmode)
    LIT alto_cr_mask;
    LIT* cr_masks;              // conflict resolution masks (nmode)
    FType* precomp_row;         // unused: CP-APR keeps its buffers on the device
    bool is_oidx;               // is oidx data available?
    IType** oidx;               // positions by row index, for modes that need it
    alto_tensor_t* dev;         // the device tensor (keys, values, per-mode streams)
};

namespace alto_dropin {

template <typename LIT> inline int lit_bits() { return (int)(8 * sizeof(LIT)); }

template <typename LIT> inline LIT lit_from(const uint64_t* w, int words)
{
    LIT v = (LIT)w[0];
    if constexpr (sizeof(LIT) == 16)
        if (words == 2) v |= ((LIT)w[1]) << 64;
    return v;
}

template <typename T> inline T* amalloc(size_t n)
{
    T* p = (T*)AlignedMalloc(n * sizeof(T));
    if (!p) {
        fprintf(stderr, "out of host memory (%zu bytes)\n", n * sizeof(T));
        abort();
    }
    return p;
}

}  // namespace alto_dropin

// ---------------------------------------------------------------- ALTO format
// create_alto (include/alto.hpp:108-294): the tensor is built on the device
// from spt (which the caller may destroy afterwards); nprtn sets the
// partition bounds and intervals the reference's OpenMP threads would use.
template <typename LIT>
static inline void create_alto(SparseTensor* spt, AltoTensor<LIT>** at, int nprtn)
{
    using namespace alto_dropin;
    static_assert(sizeof(LIT) == 8 || sizeof(LIT) == 16, "LIT must be a 64- or 128-bit unsigned integer");
    alto_tensor_t* t = nullptr;
#if defined(VAL_TYPE_INT64) || defined(VAL_TYPE_INT32)
    const int vk = ALTO_VAL_U64;
#else
    const int vk = ALTO_VAL_F64;
#endif
    check(alto_tensor_create(spt->nmodes, (const uint64_t*)spt->dims, spt->nnz, (const uint64_t* const*)spt->cidx,
                             spt->vals, ALTO_PTR_HOST, vk, lit_bits<LIT>(), nprtn, nullptr, &t),
          "create_alto");
    alto_tensor_info_t inf;
    check(alto_tensor_info(t, &inf), "alto_tensor_info");
    AltoTensor<LIT>* a = amalloc<AltoTensor<LIT>>(1);
    memset((void*)a, 0, sizeof(*a));
    a->nmode = inf.nmodes;
    a->nprtn = inf.nprtn;
    a->nnz = inf.nnz;
    a->dims = amalloc<IType>(inf.nmodes);
    a->mode_masks = amalloc<LIT>(inf.nmodes);
    a->cr_masks = amalloc<LIT>(inf.nmodes);
    for (int m = 0; m < inf.nmodes; ++m) {
        a->dims[m] = inf.dims[m];
        a->mode_masks[m] = lit_from<LIT>(&inf.mode_masks[m * inf.lit_words], inf.lit_words);
        a->cr_masks[m] = lit_from<LIT>(&inf.cr_masks[m * inf.lit_words], inf.lit_words);
    }
    a->alto_mask = lit_from<LIT>(inf.alto_mask, inf.lit_words);
    a->alto_cr_mask = lit_from<LIT>(inf.alto_cr_mask, inf.lit_words);
    a->prtn_ptr = amalloc<IType>(nprtn + 1);
    std::vector<uint64_t> iv((size_t)nprtn * inf.nmodes * 2);
    a->prtn_intervals = amalloc<Interval>((size_t)nprtn * inf.nmodes);
#ifndef ALTO_DROPIN_DEVICE_ONLY
    std::vector<uint64_t> keys((size_t)inf.nnz * inf.lit_words);
    std::vector<double> v(inf.nnz);
    check(alto_tensor_export(t, keys.data(), v.data(), (uint64_t*)a->prtn_ptr, iv.data()), "alto_tensor_export");
    a->idx = amalloc<LIT>(inf.nnz);
    a->vals = amalloc<ValType>(inf.nnz);
    for (IType i = 0; i < inf.nnz; ++i) {
        a->idx[i] = lit_from<LIT>(&keys[(size_t)i * inf.lit_words], inf.lit_words);
        a->vals[i] = (ValType)v[i];
    }
    a->is_oidx = inf.is_oidx != 0;
    if (a->is_oidx) {
        a->oidx = amalloc<IType*>(inf.nmodes);
        for (int m = 0; m < inf.nmodes; ++m) {
            a->oidx[m] = nullptr;
            if (inf.nnz / inf.dims[m] <= MIN_FIBER_REUSE) {
                a->oidx[m] = amalloc<IType>(inf.nnz);
                check(alto_tensor_export_oidx(t, m, (uint64_t*)a->oidx[m]), "alto_tensor_export_oidx");
            }
        }
    }
#else
    check(alto_tensor_export(t, nullptr, nullptr, (uint64_t*)a->prtn_ptr, iv.data()), "alto_tensor_export");
#endif
    for (size_t k = 0; k < (size_t)nprtn * inf.nmodes; ++k) a->prtn_intervals[k] = {iv[2 * k], iv[2 * k + 1]};
    a->dev = t;
    *at = a;
}

// destroy_alto (include/alto.hpp:296-322)
template <typename LIT>
static inline void destroy_alto(AltoTensor<LIT>* at)
{
    alto_tensor_destroy(at->dev);
    AlignedFree(at->dims);
    AlignedFree(at->mode_masks);
    AlignedFree(at->cr_masks);
    AlignedFree(at->idx);
    AlignedFree(at->vals);
    AlignedFree(at->prtn_ptr);
    AlignedFree(at->prtn_intervals);
    if (at->is_oidx && at->oidx) {
        for (int n = 0; n < at->nmode; ++n) AlignedFree(at->oidx[n]);
        AlignedFree(at->oidx);
    }
    AlignedFree(at);
}

// The device MTTKRP needs no per-partition fiber copies: create_da_mem /
// destroy_da_mem keep their signatures (include/alto.hpp:324-398) and do
// nothing.
template <typename LIT>
static inline void create_da_mem(int, IType, AltoTensor<LIT>*, FType*** ofibs)
{
    *ofibs = nullptr;
}
template <typename LIT>
static inline void destroy_da_mem(AltoTensor<LIT>*, FType**, IType, int)
{
}

// mttkrp_alto_par (include/alto.hpp:617-623): factors[target_mode] += the
// mode-target_mode MTTKRP of at with factors (host, dims[n] x rank)
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

// mttkrp_alto (include/alto.hpp:457-543): the same result
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

#endif  // ALTO_DROPIN_ALTO_HPP_
