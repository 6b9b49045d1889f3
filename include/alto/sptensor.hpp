// Coordinate (COO) sparse tensors of the reference ALTO library
// (include/sptensor.hpp, src/sptensor.cpp): the SparseTensor structure, its
// .tns text and binary file formats and its create / destroy calls.
// Header-only host code; every block is taken from AlignedMalloc and
// released with AlignedFree, as there, so structures built by either side
// can be destroyed by the other.
#ifndef ALTO_DROPIN_SPTENSOR_HPP_
#define ALTO_DROPIN_SPTENSOR_HPP_

#include <assert.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "common.hpp"

// include/sptensor.hpp:8-26
struct sptensor_struct {
    int nmodes;     // number of modes
    IType* dims;    // length of each mode
    IType nnz;      // number of nonzeros
    IType** cidx;   // nmodes arrays of nnz zero-based coordinates
    FType* vals;    // nnz values
};
typedef struct sptensor_struct SparseTensor;

// include/sptensor.hpp:28-32
typedef enum FileFormat_ { TEXT_FORMAT = 0, BINARY_FORMAT = 1 } FileFormat;

namespace alto_io {

[[noreturn]] inline void fail(const char* what, const char* path)
{
    fprintf(stderr, "%s: %s\n", path ? path : "(null)", what);
    exit(1);
}

// an nmodes x nnz tensor with uninitialised coordinates and values
inline SparseTensor* alloc_tensor(int nmodes, const IType* dims, IType nnz)
{
    SparseTensor* X = (SparseTensor*)AlignedMalloc(sizeof(SparseTensor));
    if (!X) fail("out of host memory", "SparseTensor");
    X->nmodes = nmodes;
    X->nnz = nnz;
    X->dims = (IType*)AlignedMalloc(sizeof(IType) * nmodes);
    X->cidx = (IType**)AlignedMalloc(sizeof(IType*) * nmodes);
    X->vals = (FType*)AlignedMalloc(sizeof(FType) * nnz);
    if (!X->dims || !X->cidx || !X->vals) fail("out of host memory", "SparseTensor");
    for (int m = 0; m < nmodes; ++m) {
        X->dims[m] = dims ? dims[m] : 0;
        X->cidx[m] = (IType*)AlignedMalloc(sizeof(IType) * nnz);
        if (!X->cidx[m]) fail("out of host memory", "SparseTensor");
    }
    return X;
}

// .tns text reader (src/sptensor.cpp:75-161, 164-197): one nonzero per line,
// nmodes indices then the value; '#' lines and empty lines are skipped;
// the mode count comes from the first data line; dims are the largest
// index (one-based files) or the largest index + 1 (zero-based files)
inline SparseTensor* read_tns(const char* path, int zero_based)
{
    FILE* fp = fopen(path, "r");
    if (!fp) fail("cannot open", path);
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
            if (nmodes < 1 || nmodes > MAX_NUM_MODES) fail("unsupported number of modes", path);
            idx.resize(nmodes);
        }
        char* p = line;
        for (int m = 0; m < nmodes; ++m) idx[m].push_back(strtoull(p, &p, 10) - (zero_based ? 0 : 1));
        vals.push_back(strtod(p, &p));
    }
    free(line);
    fclose(fp);
    if (nmodes < 0) fail("no nonzeros", path);
    SparseTensor* X = alloc_tensor(nmodes, nullptr, vals.size());
    for (int m = 0; m < nmodes; ++m) {
        IType mx = 0;
        for (IType i = 0; i < X->nnz; ++i) {
            X->cidx[m][i] = idx[m][i];
            if (idx[m][i] + 1 > mx) mx = idx[m][i] + 1;
        }
        X->dims[m] = mx;
    }
    memcpy(X->vals, vals.data(), X->nnz * sizeof(FType));
    return X;
}

// binary reader (src/sptensor.cpp:211-258): nmodes (8 bytes, of which the
// low 32 bits are meaningful: the writer stores a 4-byte int and its
// padding), dims, nnz, the index arrays mode by mode, the values
inline SparseTensor* read_bin(const char* path)
{
    FILE* fp = fopen(path, "rb");
    if (!fp) fail("cannot open", path);
    IType nm = 0, nnz = 0;
    if (fread(&nm, 8, 1, fp) != 1) fail("truncated header", path);
    nm &= 0xffffffffull;
    if (nm < 1 || nm > MAX_NUM_MODES) fail("unsupported number of modes", path);
    IType dims[MAX_NUM_MODES];
    if (fread(dims, 8, nm, fp) != nm || fread(&nnz, 8, 1, fp) != 1) fail("truncated header", path);
    if (nnz < 1) fail("no nonzeros", path);
    SparseTensor* X = alloc_tensor((int)nm, dims, nnz);
    for (IType m = 0; m < nm; ++m)
        if (fread(X->cidx[m], 8, nnz, fp) != nnz) fail("truncated indices", path);
    if (fread(X->vals, 8, nnz, fp) != nnz) fail("truncated values", path);
    fclose(fp);
    return X;
}

// ExportSparseTensor's writers (src/sptensor.cpp:23-72): binary = the layout
// read_bin reads; text = one-based indices and "%g" values (lossy, as there)
inline void write_tensor(const char* path, const SparseTensor* X, bool binary)
{
    FILE* fp = fopen(path, binary ? "wb" : "w");
    if (!fp) fail("cannot create", path);
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
// (src/sptensor.cpp:35-40): "<nmodes>D_<nnz>.tns" / ".bin"
inline std::string export_name(const SparseTensor* X, bool binary)
{
    char fn[64];
    snprintf(fn, sizeof fn, "%lluD_%llu.%s", (unsigned long long)X->nmodes, (unsigned long long)X->nnz,
             binary ? "bin" : "tns");
    return fn;
}

}  // namespace alto_io

// src/sptensor.cpp:23-72; a NULL path writes "<nmodes>D_<nnz>.tns" / ".bin"
// in the working directory
inline void ExportSparseTensor(const char* file_path, FileFormat f, SparseTensor* X)
{
    assert(f == TEXT_FORMAT || f == BINARY_FORMAT);
    const bool bin = f == BINARY_FORMAT;
    if (file_path)
        alto_io::write_tensor(file_path, X, bin);
    else
        alto_io::write_tensor(alto_io::export_name(X, bin).c_str(), X, bin);
}

// src/sptensor.cpp:200-264
inline void ImportSparseTensor(const char* file_path, FileFormat f, SparseTensor** X_, IType zero_based)
{
    assert(f == TEXT_FORMAT || f == BINARY_FORMAT);
    *X_ = f == BINARY_FORMAT ? alto_io::read_bin(file_path) : alto_io::read_tns(file_path, zero_based ? 1 : 0);
}

// src/sptensor.cpp:12-21
inline void DestroySparseTensor(SparseTensor* X)
{
    if (!X) return;
    for (int m = 0; m < X->nmodes; ++m) AlignedFree(X->cidx[m]);
    AlignedFree(X->cidx);
    AlignedFree(X->dims);
    AlignedFree(X->vals);
    AlignedFree(X);
}

// src/sptensor.cpp:266-311: cidx is row-major nnz x nmodes
// (cidx[i * nmodes + n]); coordinates are checked against dims
inline void CreateSparseTensor(IType nmodes, IType* dims, IType nnz, IType* cidx, FType* vals, SparseTensor** X_)
{
    assert(nmodes > 0 && nmodes <= MAX_NUM_MODES);
    assert(nnz > 0);
    for (IType n = 0; n < nmodes; n++) assert(dims[n] > 0);
    SparseTensor* X = alto_io::alloc_tensor((int)nmodes, dims, nnz);
    for (IType i = 0; i < nnz; i++)
        for (IType n = 0; n < nmodes; n++) {
            assert(cidx[i * nmodes + n] < dims[n]);
            X->cidx[n][i] = cidx[i * nmodes + n];
        }
    memcpy(X->vals, vals, sizeof(FType) * nnz);
    *X_ = X;
}

#endif  // ALTO_DROPIN_SPTENSOR_HPP_
