// Kruskal (CP) models of the reference ALTO library (include/kruskal_model.hpp,
// src/kruskal_model.cpp): nmodes row-major factor matrices U[n] (dims[n] x
// rank) and the weights lambda.  The model lives in host memory as there;
// the normalisations and the Gram run on the device through libalto_hip.
#ifndef ALTO_DROPIN_KRUSKAL_MODEL_HPP_
#define ALTO_DROPIN_KRUSKAL_MODEL_HPP_

#include <assert.h>
#include <stdio.h>
#include <string.h>

#include "common.hpp"
#include "device.hpp"

// include/kruskal_model.hpp:8-15
typedef struct KruskalModel_ {
    int mode;        // number of modes
    IType* dims;     // length of each mode
    IType rank;      // number of components
    FType** U;       // factor matrices, dims[n] x rank, row-major
    FType* lambda;   // rank weights
} KruskalModel;

// include/kruskal_model.hpp:17-21
typedef enum { MAT_NORM_2, MAT_NORM_MAX } mat_norm_type;

// src/kruskal_model.cpp:80-106
inline void CreateKruskalModel(int mode, IType* dims, IType rank, KruskalModel** M_)
{
    assert(mode >= 1 && mode <= MAX_NUM_MODES);
    assert(rank >= 1);
    KruskalModel* M = (KruskalModel*)AlignedMalloc(sizeof(KruskalModel));
    assert(M);
    M->mode = mode;
    M->rank = rank;
    M->dims = (IType*)AlignedMalloc(mode * sizeof(IType));
    memcpy(M->dims, dims, sizeof(IType) * mode);
    M->U = (FType**)AlignedMalloc(mode * sizeof(FType*));
    for (int n = 0; n < mode; n++) {
        assert(dims[n] >= 1);
        M->U[n] = (FType*)AlignedMalloc(dims[n] * rank * sizeof(FType));
        assert(M->U[n]);
    }
    M->lambda = (FType*)AlignedMalloc(rank * sizeof(FType));
    assert(M->lambda);
    *M_ = M;
}

// src/kruskal_model.cpp:343-352
inline void DestroyKruskalModel(KruskalModel* M)
{
    if (!M) return;
    for (int n = 0; n < M->mode; n++) AlignedFree(M->U[n]);
    AlignedFree(M->U);
    AlignedFree(M->dims);
    AlignedFree(M->lambda);
    AlignedFree(M);
}

// src/kruskal_model.cpp:109-125: glibc srand(seed), then every factor
// entry rand() / RAND_MAX in row-major order mode by mode; lambda = 1
inline void KruskalModelRandomInit(KruskalModel* M, unsigned int seed)
{
    alto_dropin::check(alto_kruskal_random_init(M->mode, (const uint64_t*)M->dims, M->rank, seed, M->U, M->lambda),
                       "KruskalModelRandomInit");
}

namespace alto_dropin {
inline void kruskal_op(KruskalModel* M, int mode, int kind, const char* what)
{
    check(alto_kruskal_normalize(M->mode, (const uint64_t*)M->dims, M->rank, M->U, M->lambda, mode, kind,
                                 ALTO_PTR_HOST, nullptr),
          what);
}
}  // namespace alto_dropin

// src/kruskal_model.cpp:128-176: every column divided by its L1 norm, the
// norms multiplied into lambda, mode by mode
inline void KruskalModelNormalize(KruskalModel* M)
{
    alto_dropin::kruskal_op(M, -1, ALTO_KRUSKAL_L1, "KruskalModelNormalize");
}

// src/kruskal_model.cpp:178-222: the same for mode n only
inline void KruskalModelNormalizeMode(KruskalModel* M, int n)
{
    alto_dropin::kruskal_op(M, n, ALTO_KRUSKAL_L1, "KruskalModelNormalizeMode");
}

// src/kruskal_model.cpp:307-341: lambda = the 2-norms of mode's columns
// (MAT_NORM_2) or max(1, the largest entry) (MAT_NORM_MAX), the columns
// divided by lambda.  The scratchpad (per-thread partial norms there) is
// not used: the reduction runs on the device.
inline void KruskalModelNorm(KruskalModel* M, IType mode, mat_norm_type which, FType** scratchpad)
{
    (void)scratchpad;
    alto_dropin::kruskal_op(M, (int)mode, which == MAT_NORM_2 ? ALTO_KRUSKAL_NORM2 : ALTO_KRUSKAL_NORMMAX,
                            "KruskalModelNorm");
}

// src/kruskal_model.cpp:355-378: U[n] columns multiplied by lambda, lambda = 1
inline void RedistributeLambda(KruskalModel* M, int n)
{
    alto_dropin::kruskal_op(M, n, ALTO_KRUSKAL_REDISTRIBUTE, "RedistributeLambda");
}

// src/kruskal_model.cpp:380-383
inline double KruskalTensorFit()
{
    return 0.0;
}

// src/kruskal_model.cpp:11-24: gram (rank x rank) = U[mode]^T U[mode]; the
// reference's SYRK fills the lower triangle, here both triangles are written
inline void update_gram(FType* gram, KruskalModel* M, int mode)
{
    using namespace alto_dropin;
    const IType R = M->rank;
    DeviceBuffer u(M->U[mode], M->dims[mode] * R * sizeof(FType)), g(R * R * sizeof(FType));
    check(alto_update_gram(u.p, M->dims[mode], R, ALTO_PREC_F64, g.as<double>(), nullptr), "update_gram");
    check(alto_memcpy_d2h(gram, g.p, R * R * sizeof(FType)), "alto_memcpy_d2h");
}

// src/kruskal_model.cpp:27-58: "<path>.<n>.out" per factor ("matrix", "2",
// "dims rank", then the entries column by column) and "<path>.lambda.out"
inline void ExportKruskalModel(KruskalModel* M, char* file_path)
{
    char fn[1100];
    for (int n = 0; n < M->mode; n++) {
        snprintf(fn, sizeof fn, "%s.%d.out", file_path, n);
        FILE* fp = fopen(fn, "w");
        assert(fp);
        fprintf(fp, "matrix\n2\n%llu %llu\n", M->dims[n], M->rank);
        for (IType i = 0; i < M->rank; i++)
            for (IType j = 0; j < M->dims[n]; j++) fprintf(fp, "%.20lf\n", M->U[n][j * M->rank + i]);
        fclose(fp);
    }
    snprintf(fn, sizeof fn, "%s.lambda.out", file_path);
    FILE* fp = fopen(fn, "w");
    assert(fp);
    fprintf(fp, "vector\n1\n%llu\n", M->rank);
    for (IType i = 0; i < M->rank; i++) fprintf(fp, "%.20lf\n", M->lambda[i]);
    fclose(fp);
}

// src/kruskal_model.cpp:61-78
inline void PrintKruskalModel(KruskalModel* M)
{
    for (int n = 0; n < M->mode; n++) {
        printf("mode %d:\n", n);
        for (IType j = 0; j < M->dims[n]; j++) {
            for (IType i = 0; i < M->rank; i++) printf("%g ", M->U[n][j * M->rank + i]);
            printf("\n");
        }
    }
    printf("lambda:");
    for (IType r = 0; r < M->rank; r++) printf("%g ", M->lambda[r]);
    printf("\n");
}

#endif  // ALTO_DROPIN_KRUSKAL_MODEL_HPP_
