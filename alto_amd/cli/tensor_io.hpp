// Tensor file I/O of the cpd64 / cpd128 command line: the reference's .tns
// text and binary formats (src/sptensor.cpp), header-only so that it can be
// used without the GPU library.
#pragma once

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "alto/alto.hpp"

// .tns text reader (src/sptensor.cpp:75-161): dims are the largest index
// (+ zero_based), '#' lines and empty lines are skipped, values via strtod.
static inline SparseTensor* read_tns(const char* path, int zero_based)
{
    FILE* fp = fopen(path, "r");
    if (!fp) {
        fprintf(stderr, "cannot open %s\n", path);
        exit(1);
    }
    char* line = NULL;
    size_t len = 0;
    ssize_t nread;
    int nmodes = -1;
    std::vector<std::vector<IType>> idx;
    std::vector<FType> vals;
    while ((nread = getline(&line, &len, fp)) != -1) {
        if (nread <= 1 || line[0] == '#') continue;
        if (nmodes < 0) {
            int tok = 0;
            char* copy = strdup(line);
            for (char* p = strtok(copy, " \t\n"); p; p = strtok(NULL, " \t\n")) ++tok;
            free(copy);
            nmodes = tok - 1;
            if (nmodes < 1 || nmodes > MAX_NUM_MODES) {
                fprintf(stderr, "unsupported number of modes %d\n", nmodes);
                exit(1);
            }
            idx.resize(nmodes);
        }
        char* p = line;
        for (int m = 0; m < nmodes; ++m) idx[m].push_back(strtoull(p, &p, 10) - (zero_based ? 0 : 1));
        vals.push_back(strtod(p, &p));
    }
    free(line);
    fclose(fp);
    SparseTensor* X = new SparseTensor();
    X->nmodes = nmodes;
    X->nnz = vals.size();
    X->dims = new IType[nmodes];
    X->cidx = new IType*[nmodes];
    for (int m = 0; m < nmodes; ++m) {
        IType mx = 0;
        X->cidx[m] = new IType[X->nnz];
        for (IType i = 0; i < X->nnz; ++i) {
            X->cidx[m][i] = idx[m][i];
            if (idx[m][i] + 1 > mx) mx = idx[m][i] + 1;
        }
        X->dims[m] = mx;
    }
    X->vals = new FType[X->nnz];
    memcpy(X->vals, vals.data(), X->nnz * sizeof(FType));
    return X;
}

// binary reader (src/sptensor.cpp:211-258): nmodes (8 bytes, low 32 bits
// meaningful), dims, nnz, the index arrays, the values
static inline SparseTensor* read_bin(const char* path)
{
    FILE* fp = fopen(path, "rb");
    if (!fp) {
        fprintf(stderr, "cannot open %s\n", path);
        exit(1);
    }
    IType nm = 0, nnz = 0;
    if (fread(&nm, 8, 1, fp) != 1) exit(1);
    nm &= 0xffffffffull;  // the writer stores a 4-byte int and its padding
    if (nm < 1 || nm > MAX_NUM_MODES) {
        fprintf(stderr, "%s: unsupported number of modes %llu\n", path, (unsigned long long)nm);
        exit(1);
    }
    SparseTensor* X = new SparseTensor();
    X->nmodes = (int)nm;
    X->dims = new IType[nm];
    if (fread(X->dims, 8, nm, fp) != nm || fread(&nnz, 8, 1, fp) != 1) exit(1);
    if (nnz < 1) {
        fprintf(stderr, "%s: no nonzeros\n", path);
        exit(1);
    }
    X->nnz = nnz;
    X->cidx = new IType*[nm];
    for (IType m = 0; m < nm; ++m) {
        X->cidx[m] = new IType[nnz];
        if (fread(X->cidx[m], 8, nnz, fp) != nnz) exit(1);
    }
    X->vals = new FType[nnz];
    if (fread(X->vals, 8, nnz, fp) != nnz) exit(1);
    fclose(fp);
    return X;
}

// ExportSparseTensor (src/sptensor.cpp:23-72): binary = the layout read_bin
// reads; text = 1-based indices and "%g" values (lossy, as the reference)
static inline void write_tensor(const char* path, SparseTensor* X, bool binary)
{
    FILE* fp = fopen(path, binary ? "wb" : "w");
    if (!fp) return;
    if (binary) {
        IType nm = (IType)X->nmodes;
        fwrite(&nm, 8, 1, fp);
        fwrite(X->dims, 8, X->nmodes, fp);
        fwrite(&X->nnz, 8, 1, fp);
        for (int m = 0; m < X->nmodes; ++m) fwrite(X->cidx[m], 8, X->nnz, fp);
        fwrite(X->vals, 8, X->nnz, fp);
    } else {
        for (IType i = 0; i < X->nnz; ++i) {
            for (int m = 0; m < X->nmodes; ++m) fprintf(fp, "%llu ", X->cidx[m][i] + 1);
            fprintf(fp, "%g\n", X->vals[i]);
        }
    }
    fclose(fp);
}

// the file name ExportSparseTensor uses when given no path
// (src/sptensor.cpp:35-40): "<nmodes>D_<nnz>.tns" / ".bin" in the working
// directory
static inline std::string export_name(const SparseTensor* X, bool binary)
{
    char fn[64];
    snprintf(fn, sizeof fn, "%lluD_%llu.%s", (unsigned long long)X->nmodes, (unsigned long long)X->nnz,
             binary ? "bin" : "tns");
    return fn;
}
