// Synthetic Poisson tensors of the reference (include/poisson_generator.hpp,
// src/poisson_generator.cpp:336-376, called by MakeSparseTensor,
// src/main.cpp:653-675), restated step by step so that a run reproduces the
// reference's tensor and model for the same thread count T.  Host code, as
// there: it makes test data, not decompositions.
//
//  - T uniform streams, MT19937-64 seeded 1234 + t (src/rng.cpp is that
//    published generator; std::mt19937_64 is the same algorithm and seeding),
//    created by CreatePoissonGenerator and kept until
//    DestroyPoissonGenerator; uniform draws ((x >> 12) + 0.5) 2^-52 scaled to
//    [low, high) (include/rng.hpp:20, include/rng_stream.hpp:21-28); a batch
//    of n draws is split into T chunks of ceil(n / T), chunk t from stream t
//    (poisson_generator.cpp:20-34).
//  - Model: every factor entry 0.1, then per column ceil(0.2 I_n) entries
//    picked by a partial Fisher-Yates shuffle on glibc rand() get uniform
//    values in [1, 10 R); lambda uniform in [0, 1); columns L1-normalised
//    into lambda, lambda scaled to sum 1 (:144-192).
//  - Edges: counts per component from lambda's cumulative bins, then every
//    index of a component's samples from its factor column's cumulative
//    bins, in batches of 100000 draws (:195-260, :40-84).
//  - Coordinates: least-significant-mode-first counting sort, equal
//    coordinates merged into counts (:263-333).
// The column sums are taken in row order, as one thread does; with T > 1
// the reference's OpenMP reductions may round differently.
#ifndef ALTO_DROPIN_POISSON_GENERATOR_HPP_
#define ALTO_DROPIN_POISSON_GENERATOR_HPP_

#include <assert.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <random>
#include <thread>
#include <utility>
#include <vector>

#include "common.hpp"
#include "kruskal_model.hpp"
#include "sptensor.hpp"

namespace alto_gen {

constexpr uint64_t kBatch = 100000;
constexpr uint64_t kSeed = 1234;

// the reference's stream count: the OpenMP thread count (OMP_NUM_THREADS,
// else the hardware threads)
inline int default_threads()
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    int t = (int)std::thread::hardware_concurrency();
    if (const char* e = getenv("OMP_NUM_THREADS")) t = atoi(e);
    return t < 1 ? 1 : t;
#endif
}

struct Streams {
    std::vector<std::mt19937_64> s;
    explicit Streams(int threads)
    {
        for (int t = 0; t < threads; ++t) s.emplace_back(kSeed + (uint64_t)t);
    }
    void uniform(uint64_t n, double low, double high, double* a)
    {
        const uint64_t T = s.size(), chunk = (n + T - 1) / T;
        for (uint64_t t = 0; t < T; ++t) {
            const uint64_t b = t * chunk, e = b + chunk < n ? b + chunk : n;
            for (uint64_t i = b; i < e; ++i) {
                const double u = ((double)(s[t]() >> 12) + 0.5) * (1.0 / 4503599627370496.0);
                a[i] = u * (high - low) + low;
            }
        }
    }
};

// largest j with bins[j] < rv (bins ascending, bins[0] = 0 < rv); a draw equal
// to a bin edge takes the bin below it (poisson_generator.cpp:37-56)
inline uint64_t binning(double rv, uint64_t nbins, const double* bins)
{
    int64_t lo = 0, hi = (int64_t)nbins - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) / 2;
        if (rv > bins[mid]) {
            lo = mid + 1;
        } else if (rv == bins[mid]) {
            if (mid == (int64_t)nbins - 1) mid -= 1;
            return (uint64_t)mid;
        } else {
            hi = mid - 1;
        }
    }
    return (uint64_t)hi;
}

// X with 0-based coordinates and integer counts, drawing from rng; the
// model the edges were drawn from goes into *model when given (a model of
// nmodes modes, these dims and this rank); prints the reference's progress
// lines
inline SparseTensor* poisson_tensor(int nmodes, const IType* dims, IType num_edges, IType rank, Streams& rng,
                                    KruskalModel* model = nullptr)
{
    IType maxd = 0;
    for (int n = 0; n < nmodes; ++n) maxd = dims[n] > maxd ? dims[n] : maxd;
    printf("Generating a sparse tensor: ");
    for (int n = 0; n < nmodes - 1; n++) printf("%llux", (unsigned long long)dims[n]);
    printf("%llu, rank = %llu ...\n", (unsigned long long)dims[nmodes - 1], (unsigned long long)rank);

    // the model
    std::vector<std::vector<double>> U(nmodes);
    std::vector<double> lambda(rank), vals(maxd + 1);
    std::vector<IType> perm(maxd + 1);
    for (int n = 0; n < nmodes; ++n) {
        const IType d = dims[n];
        U[n].assign(d * rank, 0.1);
        const IType nbig = (IType)ceil(0.2 * d);
        for (IType r = 0; r < rank; ++r) {
            for (IType i = 0; i < d; ++i) perm[i] = i;
            for (IType i = 0; i < nbig; ++i) {
                const IType j = (IType)rand() % (d - i) + i;
                const IType t = perm[i];
                perm[i] = perm[j];
                perm[j] = t;
            }
            rng.uniform(nbig, 1.0, 10.0 * (double)rank, vals.data());
            for (IType i = 0; i < nbig; ++i) U[n][perm[i] * rank + r] = vals[i];
        }
    }
    rng.uniform(rank, 0.0, 1.0, lambda.data());
    for (int n = 0; n < nmodes; ++n) {  // KruskalModelNormalize (src/kruskal_model.cpp:128-176)
        std::vector<double> cn(rank, 0.0);
        for (IType i = 0; i < dims[n]; ++i)
            for (IType j = 0; j < rank; ++j) cn[j] += fabs(U[n][i * rank + j]);
        for (IType i = 0; i < dims[n]; ++i)
            for (IType j = 0; j < rank; ++j) U[n][i * rank + j] /= cn[j];
        for (IType j = 0; j < rank; ++j) lambda[j] *= cn[j];
    }
    double tot = 0.0;
    for (IType j = 0; j < rank; ++j) tot += lambda[j];
    tot = 1.0 / tot;
    for (IType j = 0; j < rank; ++j) lambda[j] *= tot;
    if (model) {
        for (int n = 0; n < nmodes; ++n) memcpy(model->U[n], U[n].data(), dims[n] * rank * sizeof(FType));
        memcpy(model->lambda, lambda.data(), rank * sizeof(FType));
    }

    // the edges: counts per component, then each component's indices
    std::vector<double> bins(maxd + 1), work(kBatch);
    bins[0] = 0.0;
    for (IType i = 0; i < rank; ++i) {
        bins[i + 1] = lambda[i] + bins[i];
        bins[i + 1] = bins[i + 1] < 1.0 ? bins[i + 1] : 1.0;
    }
    std::vector<IType> rcount(rank + 1, 0);  // [rank]: draws above the last bin (dropped)
    for (IType b = 0; b < num_edges; b += kBatch) {
        const IType len = b + kBatch < num_edges ? kBatch : num_edges - b;
        rng.uniform(len, 0.0, 1.0, work.data());
        for (IType i = 0; i < len; ++i) rcount[binning(work[i], rank + 1, bins.data())]++;
    }
    std::vector<IType> eidx((size_t)nmodes * num_edges, 0);
    IType count = 0;
    for (IType r = 0; r < rank; ++r) {
        const IType nc = rcount[r];
        for (int n = 0; n < nmodes; ++n) {
            bins[0] = 0.0;
            for (IType i = 0; i < dims[n]; ++i) {
                bins[i + 1] = U[n][i * rank + r] + bins[i];
                bins[i + 1] = bins[i + 1] < 1.0 ? bins[i + 1] : 1.0;
            }
            for (IType b = 0; b < nc; b += kBatch) {
                const IType len = b + kBatch < nc ? kBatch : nc - b;
                rng.uniform(len, 0.0, 1.0, work.data());
                for (IType i = 0; i < len; ++i)
                    eidx[(count + b + i) * nmodes + n] = binning(work[i], dims[n] + 1, bins.data());
            }
        }
        count += nc;
    }

    // coordinates: counting sort by mode 0, 1, ..., then merge equal ones
    const IType ne = count;
    std::vector<IType> tmp((size_t)nmodes * ne), off(maxd + 2);
    IType *src = eidx.data(), *dst = tmp.data();
    for (int n = 0; n < nmodes; ++n) {
        std::fill(off.begin(), off.end(), 0);
        for (IType i = 0; i < ne; ++i) off[src[i * nmodes + n] + 1] += 1;
        for (IType i = 0; i < dims[n]; ++i) off[i + 1] += off[i];
        for (IType i = 0; i < ne; ++i) {
            const IType pos = off[src[i * nmodes + n]]++;
            for (int m = 0; m < nmodes; ++m) dst[pos * nmodes + m] = src[i * nmodes + m];
        }
        std::swap(src, dst);
    }
    std::vector<IType> keep;
    std::vector<double> cnt;
    for (IType i = 0; i < ne; ++i) {
        bool dup = i > 0;
        for (int m = 0; m < nmodes && dup; ++m) dup = src[i * nmodes + m] == src[(i - 1) * nmodes + m];
        if (dup) {
            cnt.back() += 1.0;
        } else {
            keep.push_back(i);
            cnt.push_back(1.0);
        }
    }
    SparseTensor* X = alto_io::alloc_tensor(nmodes, dims, keep.size());
    for (int n = 0; n < nmodes; ++n)
        for (IType k = 0; k < X->nnz; ++k) X->cidx[n][k] = src[keep[k] * nmodes + n];
    for (IType k = 0; k < X->nnz; ++k) X->vals[k] = cnt[k];
    printf("    Nonzeros = %llu\n", (unsigned long long)X->nnz);
    return X;
}

// the same with T fresh streams
inline SparseTensor* poisson_tensor(int nmodes, const IType* dims, IType num_edges, IType rank, int threads)
{
    Streams rng(threads);
    return poisson_tensor(nmodes, dims, num_edges, rank, rng);
}

// the streams CreatePoissonGenerator makes (a global there too,
// src/poisson_generator.cpp:17-18)
inline Streams*& global_streams()
{
    static Streams* s = nullptr;
    return s;
}

}  // namespace alto_gen

// include/poisson_generator.hpp:9-13
typedef struct PoissonGenerator_ {
    int mode;
    IType* dims;
} PoissonGenerator;

// src/poisson_generator.cpp:102-130: one MT19937-64 stream per OpenMP thread
inline void CreatePoissonGenerator(int mode, IType* dims, PoissonGenerator** pg_)
{
    assert(mode > 0 && mode <= MAX_NUM_MODES);
    PoissonGenerator* pg = (PoissonGenerator*)AlignedMalloc(sizeof(PoissonGenerator));
    assert(pg);
    pg->mode = mode;
    pg->dims = (IType*)AlignedMalloc(mode * sizeof(IType));
    memcpy(pg->dims, dims, mode * sizeof(IType));
    delete alto_gen::global_streams();
    alto_gen::global_streams() = new alto_gen::Streams(alto_gen::default_threads());
    *pg_ = pg;
}

// src/poisson_generator.cpp:132-145
inline void DestroyPoissonGenerator(PoissonGenerator* pg)
{
    delete alto_gen::global_streams();
    alto_gen::global_streams() = nullptr;
    AlignedFree(pg->dims);
    AlignedFree(pg);
}

// src/poisson_generator.cpp:336-376: a random rank-`rank` model *M_ and
// num_edges Poisson draws from it, merged into the count tensor *X_
inline void PoissonGeneratorRun(PoissonGenerator* pg, IType num_edges, IType rank, KruskalModel** M_,
                                SparseTensor** X_)
{
    for (int n = 0; n < pg->mode; n++) assert(pg->dims[n] >= rank);
    assert(num_edges > 0);
    assert(alto_gen::global_streams());
    KruskalModel* M;
    CreateKruskalModel(pg->mode, pg->dims, rank, &M);
    *X_ = alto_gen::poisson_tensor(pg->mode, pg->dims, num_edges, rank, *alto_gen::global_streams(), M);
    *M_ = M;
}

#endif  // ALTO_DROPIN_POISSON_GENERATOR_HPP_
