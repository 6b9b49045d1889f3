// Dense R x R and I x R kernels of CP-ALS / CP-APR on gfx950.
//
//  - Gram matrices U^T U on the matrix cores: v_mfma_f64_16x16x4f64 for FP64
//    factors, v_mfma_f32_16x16x4f32 for FP32 / BF16 factors; per-block
//    partial Grams reduced in a fixed order (deterministic).
//    Reference: update_gram = SYRK (src/kruskal_model.cpp:11-24).
//  - V = Hadamard product of the other modes' Grams and V^-1 by Gauss-Jordan
//    elimination in one workgroup (FP64; its pivots fail exactly when a
//    Cholesky factorization would); then the minimum-norm pseudo-inverse
//    from a Jacobi eigen-decomposition.
//    Reference: pseudo_inverse (include/cpd.hpp:546-646).
//  - Row solves U_n[i,:] = M[i,:] V^-1 as a GEMM with the explicit inverse
//    (or the pseudo-inverse), MFMA FP64 (FP32 for FP32 / BF16 models), fused
//    with the per-block column norms.
//  - Column normalisation (2-norm / max-norm), the fit, ||X||^2.
//    Reference: KruskalModelNorm (src/kruskal_model.cpp:224-341),
//    cpd_fit_alto (include/cpd.hpp:189-271).
#include <math.h>

#include <mutex>
#include <type_traits>

#include "alto_internal.hpp"
#include "dense.hpp"
#include "kernels_common.hpp"

namespace alto {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T> __device__ inline double as_f64(T v) { return (double)to_acc<double>(v); }

// ------------------------------------------------------------------ Gram
// Each wave owns 16x16 output tiles (t = wave, wave+4, ...).  One MFMA step
// consumes 4 rows: lane l supplies U[i0 + l/16][a0 + l%16] as the A operand
// (A = U^T block, 16 x 4) and U[i0 + l/16][b0 + l%16] as the B operand.
// The 16x16 result is spread over the 4 result registers of each lane (see
// the store below for the per-instruction row mapping).
constexpr int GRAM_ROWS = 256;  // rows per block per sweep
constexpr int GRAM_MAXT = 16;   // tiles per wave (R <= 128)

template <typename T, bool F64>
__global__ void __launch_bounds__(256) k_gram_partial(const T* __restrict__ U, uint64_t I, int R, double* __restrict__ part)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int T16 = (R + 15) / 16;
    const int ntiles = T16 * T16;
    const int r16 = lane & 15, kq = lane >> 4;
    using acc_t = typename std::conditional<F64, d4_t, f4_t>::type;
    acc_t acc[GRAM_MAXT];
#pragma unroll
    for (int t = 0; t < GRAM_MAXT; ++t) acc[t] = acc_t{0, 0, 0, 0};
    for (uint64_t base = (uint64_t)blockIdx.x * GRAM_ROWS; base < I; base += (uint64_t)gridDim.x * GRAM_ROWS) {
        const uint64_t end = base + GRAM_ROWS < I ? base + GRAM_ROWS : I;
        for (uint64_t i0 = base; i0 < end; i0 += 4) {
            const uint64_t i = i0 + kq;
            const bool rv = i < end;
#pragma unroll
            for (int tt = 0; tt < GRAM_MAXT; ++tt) {
                const int t = wave + 4 * tt;
                if (t < ntiles) {
                    const int a = (t / T16) * 16 + r16, b = (t % T16) * 16 + r16;
                    const double av = rv && a < R ? as_f64(U[i * R + a]) : 0.0;
                    const double bv = rv && b < R ? as_f64(U[i * R + b]) : 0.0;
                    if constexpr (F64)
                        acc[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[tt], 0, 0, 0);
                    else
                        acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)av, (float)bv, acc[tt], 0, 0, 0);
                }
            }
        }
    }
    double* P = part + (size_t)blockIdx.x * R * R;
#pragma unroll
    for (int tt = 0; tt < GRAM_MAXT; ++tt) {
        const int t = wave + 4 * tt;
        if (t < ntiles) {
            const int b = (t % T16) * 16 + r16;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                // v_mfma_f64_16x16x4f64: lane l, register v holds row l/16 + 4v;
                // v_mfma_f32_16x16x4f32: row 4*(l/16) + v (both column l%16)
                const int a = (t / T16) * 16 + (F64 ? kq + 4 * v : 4 * kq + v);
                if (a < R && b < R) P[a * R + b] = (double)acc[tt][v];
            }
        }
    }
}

// Ranks up to 64 (T16 <= 4 column blocks of 16): each wave owns 4-row steps
// of the block's rows, loads each 16-column block of a step once (lane l:
// U[i0 + l/16][16 cb + l%16]) and multiplies every block pair ta <= tb into
// its accumulators (the Gram is symmetric); the 4 waves' tiles are summed in
// LDS and the block writes both triangles of its partial.
//
// SCALE: the block first divides its rows of U by lambda (per column) in
// place and writes the storage copy S (when the storage type differs), and
// the Gram is taken of the stored values — the column normalisation and the
// Gram update of a mode in one pass over U.
// INNER (with SCALE): the block also sums lambda_c S[i,c] Mi[i,c] over its
// rows into pin[bid] -- the fit's inner product <X, model> taken from the
// stored values (for storage that rounds, so that the fit is that of the
// model the MTTKRPs and Grams see); skipped when Mi is null.
template <typename T, typename ST, bool F64, int T16, bool SCALE, bool INNER>
__device__ __forceinline__ void gram_rows_body(T* __restrict__ U, uint64_t I, int R, uint64_t rows_per_block,
                                               double* __restrict__ part, const double* __restrict__ lambda,
                                               ST* __restrict__ S, const double* __restrict__ Mi,
                                               double* __restrict__ pin, uint32_t bid)
{
    constexpr int NP = T16 * (T16 + 1) / 2;
    using acc_t = typename std::conditional<F64, d4_t, f4_t>::type;
    extern __shared__ double sG[];  // R x R
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    acc_t acc[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[q] = acc_t{0, 0, 0, 0};
    const uint64_t base = (uint64_t)bid * rows_per_block;
    const uint64_t end = base + rows_per_block < I ? base + rows_per_block : I;
    // FP32 / BF16 factors are scaled by the reciprocal of lambda (within
    // their rounding); FP64 factors are divided, as the reference divides
    constexpr bool RECIP = !std::is_same<T, double>::value;
    double lc[T16], lr[T16];
    if constexpr (SCALE) {
#pragma unroll
        for (int cb = 0; cb < T16; ++cb) {
            lc[cb] = cb * 16 + r16 < R ? lambda[cb * 16 + r16] : 1.0;
            if constexpr (RECIP) lr[cb] = 1.0 / lc[cb];
        }
    }
    double inacc = 0;
    // two 16-row steps per pass: both steps' loads are issued before the products
    for (uint64_t i00 = base + 4 * wave; i00 < end; i00 += 32) {
        double raw[2][T16];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const uint64_t i = i00 + 16 * u + kq;
#pragma unroll
            for (int cb = 0; cb < T16; ++cb) {
                const int c = cb * 16 + r16;
                raw[u][cb] = (i < end && c < R) ? as_f64(U[i * R + c]) : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const uint64_t i = i00 + 16 * u + kq;
            double x[T16];
#pragma unroll
            for (int cb = 0; cb < T16; ++cb) {
                const int c = cb * 16 + r16;
                if constexpr (SCALE) {
                    x[cb] = 0.0;
                    if (i < end && c < R) {
                        const double v = RECIP ? raw[u][cb] * lr[cb] : raw[u][cb] / lc[cb];
                        U[i * R + c] = (T)v;
                        if constexpr (std::is_same<T, ST>::value) {
                            x[cb] = (double)(T)v;
                        } else {
                            const ST sv = store_as<ST>(v);
                            S[i * R + c] = sv;
                            x[cb] = as_f64(sv);
                        }
                        if constexpr (INNER)
                            if (Mi) inacc += x[cb] * lc[cb] * Mi[i * R + c];
                    }
                } else {
                    x[cb] = raw[u][cb];
                }
            }
            int q = 0;
#pragma unroll
            for (int ta = 0; ta < T16; ++ta)
#pragma unroll
                for (int tb = ta; tb < T16; ++tb, ++q) {
                    if constexpr (F64)
                        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[ta], x[tb], acc[q], 0, 0, 0);
                    else
                        acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)x[ta], (float)x[tb], acc[q], 0, 0, 0);
                }
        }
    }
    __shared__ double s_in[4];
    if constexpr (INNER) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) inacc += __shfl_xor(inacc, o);
        if (lane == 0) s_in[wave] = inacc;
    }
    for (int k = threadIdx.x; k < R * R; k += blockDim.x) sG[k] = 0.0;
    __syncthreads();
    if constexpr (INNER)
        if (Mi && threadIdx.x == 0) pin[bid] = (s_in[0] + s_in[1]) + (s_in[2] + s_in[3]);
    for (int w = 0; w < 4; ++w) {
        if (wave == w) {
            int q = 0;
#pragma unroll
            for (int ta = 0; ta < T16; ++ta)
#pragma unroll
                for (int tb = ta; tb < T16; ++tb, ++q) {
                    const int b = tb * 16 + r16;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        // v_mfma_f64_16x16x4f64: lane l, register v holds row l/16 + 4v;
                        // v_mfma_f32_16x16x4f32: row 4*(l/16) + v (both column l%16)
                        const int a = ta * 16 + (F64 ? kq + 4 * v : 4 * kq + v);
                        if (a < R && b < R) sG[a * R + b] += (double)acc[q][v];
                    }
                }
        }
        __syncthreads();
    }
    double* P = part + (size_t)bid * R * R;
    for (int k = threadIdx.x; k < R * R; k += blockDim.x) {
        const int a = k / R, b = k % R;
        // the upper block triangle holds the sums; mirror the lower one
        P[k] = (a / 16 <= b / 16) ? sG[k] : sG[b * R + a];
    }
}

template <typename T, typename ST, bool F64, int T16, bool SCALE, bool INNER>
__global__ void __launch_bounds__(256) k_gram_rows(T* __restrict__ U, uint64_t I, int R, uint64_t rows_per_block,
                                                   double* __restrict__ part, const double* __restrict__ lambda,
                                                   ST* __restrict__ S, const double* __restrict__ Mi,
                                                   double* __restrict__ pin)
{
    gram_rows_body<T, ST, F64, T16, SCALE, INNER>(U, I, R, rows_per_block, part, lambda, S, Mi, pin, blockIdx.x);
}

__global__ void k_sum_partials(const double* __restrict__ part, int nparts, int n, double* __restrict__ out)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        double s = 0;
        for (int p = 0; p < nparts; ++p) s += part[(size_t)p * n + k];
        out[k] = s;
    }
}

// out[k] = sum_p part[p n + k]: a block covers 32 consecutive k, its 8
// wave halves stride over the partials, and the 8 sums are added in a fixed
// order (deterministic).  Blocks from nsum on clear `zero` (nz bytes, a
// multiple of 4): a buffer the caller needs zeroed next, for free in this
// launch.
__device__ __forceinline__ void sum_cols_body(const double* __restrict__ part, int nparts, int n,
                                              double* __restrict__ out, int nsum, void* __restrict__ zero, uint64_t nz,
                                              uint32_t bid, uint32_t nblk)
{
    if ((int)bid >= nsum) {
        const uint64_t t = (uint64_t)(bid - nsum) * blockDim.x + threadIdx.x;
        const uint64_t stride = (uint64_t)(nblk - nsum) * blockDim.x;
        uint4* z16 = reinterpret_cast<uint4*>(zero);
        const uint64_t n16 = nz / 16;
        for (uint64_t i = t; i < n16; i += stride) z16[i] = make_uint4(0, 0, 0, 0);
        uint32_t* z4 = reinterpret_cast<uint32_t*>(zero);
        for (uint64_t i = n16 * 4 + t; i < nz / 4; i += stride) z4[i] = 0;
        return;
    }
    __shared__ double s[8][33];
    const int kk = threadIdx.x & 31, pg = threadIdx.x >> 5;
    const int k = bid * 32 + kk;
    double a = 0;
    if (k < n)
        for (int p = pg; p < nparts; p += 8) a += part[(size_t)p * n + k];
    s[pg][kk] = a;
    __syncthreads();
    if (threadIdx.x < 32 && k < n) {
        double v = 0;
#pragma unroll
        for (int g = 0; g < 8; ++g) v += s[g][kk];
        out[k] = v;
    }
}

__global__ void __launch_bounds__(256) k_sum_cols(const double* __restrict__ part, int nparts, int n,
                                                  double* __restrict__ out, int nsum, void* __restrict__ zero,
                                                  uint64_t nz)
{
    sum_cols_body(part, nparts, n, out, nsum, zero, nz, blockIdx.x, gridDim.x);
}

// the column sums of n-entry partials (n > 0) and the clearing of `zero`
// (zero_bytes, a multiple of 4) in one launch
int sum_cols_device(const double* part, int nparts, int n, double* out, void* zero, uint64_t zero_bytes,
                    hipStream_t s)
{
    const int nsum = (n + 31) / 32;
    uint64_t nzb = (zero_bytes / 16 + 255) / 256;
    if (nzb > 1024) nzb = 1024;
    if (zero_bytes && nzb == 0) nzb = 1;
    if (nsum + nzb == 0) return ALTO_OK;
    hipLaunchKernelGGL(k_sum_cols, dim3((unsigned)(nsum + nzb)), dim3(256), 0, s, part, nparts, n, out, nsum, zero,
                       zero_bytes);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

int gram_blocks(uint64_t I)
{
    uint64_t b = (I + GRAM_ROWS - 1) / GRAM_ROWS;
    if (b > 512) b = 512;
    return b < 1 ? 1 : (int)b;
}

template <typename T>
int gram_device(const T* U, uint64_t I, int R, double* G, double* part, hipStream_t s, bool f64)
{
    if (R > 128) {
        set_error("gram: rank above 128");
        return ALTO_ERR_UNSUPPORTED;
    }
    int nb = gram_blocks(I);
    if (R <= 64) {
        const uint64_t rpb = ((I + nb - 1) / nb + 15) / 16 * 16;
        const size_t lds = (size_t)R * R * 8;
    T* Uw = const_cast<T*>(U);  // read only without SCALE
#define GR_(TT)                                                                                                  \
    do {                                                                                                         \
        if (f64)                                                                                                 \
            hipLaunchKernelGGL((k_gram_rows<T, T, true, TT, false, false>), dim3(nb), dim3(256), lds, s, Uw, I, R, \
                               rpb, part, nullptr, nullptr, nullptr, nullptr);                                   \
        else                                                                                                     \
            hipLaunchKernelGGL((k_gram_rows<T, T, false, TT, false, false>), dim3(nb), dim3(256), lds, s, Uw, I, \
                               R, rpb, part, nullptr, nullptr, nullptr, nullptr);                                \
    } while (0)
        switch ((R + 15) / 16) {
        case 1: GR_(1); break;
        case 2: GR_(2); break;
        case 3: GR_(3); break;
        default: GR_(4); break;
        }
#undef GR_
    } else if (f64) {
        hipLaunchKernelGGL((k_gram_partial<T, true>), dim3(nb), dim3(256), 0, s, U, I, R, part);
    } else {
        hipLaunchKernelGGL((k_gram_partial<T, false>), dim3(nb), dim3(256), 0, s, U, I, R, part);
    }
    ALTO_TRY(sum_cols_device(part, nb, R * R, G, nullptr, 0, s));
    return ALTO_OK;
}
template <typename UT, typename ST>
int gram_scale_device(UT* U, uint64_t I, int R, const double* lambda, ST* S, double* G, double* part, hipStream_t s,
                      bool f64, void* zero, uint64_t zero_bytes, const double* Mi, double* pin)
{
    constexpr bool same = std::is_same<UT, ST>::value;
    if (R > 64) {
        ALTO_TRY((scale_device<UT, ST>(U, I, R, lambda, same ? nullptr : S, s)));
        ALTO_TRY(gram_device<ST>(same ? reinterpret_cast<const ST*>(U) : S, I, R, G, part, s, f64));
        if (Mi) ALTO_TRY((inner_partial_device<ST, double>(same ? reinterpret_cast<const ST*>(U) : S, Mi, I, R, pin, s)));
        if (zero_bytes) ALTO_TRY(sum_cols_device(nullptr, 0, 0, nullptr, zero, zero_bytes, s));
        return ALTO_OK;
    }
    const int nb = gram_blocks(I);
    const uint64_t rpb = ((I + nb - 1) / nb + 15) / 16 * 16;
    const size_t lds = (size_t)R * R * 8;
#define GSI_(TT, F, IN)                                                                                        \
    hipLaunchKernelGGL((k_gram_rows<UT, ST, F, TT, true, IN>), dim3(nb), dim3(256), lds, s, U, I, R, rpb, part, \
                       lambda, S, Mi, pin)
#define GS_(TT)                           \
    do {                                  \
        if (Mi) {                         \
            if (f64) GSI_(TT, true, true); \
            else GSI_(TT, false, true);    \
        } else if (f64) {                 \
            GSI_(TT, true, false);        \
        } else {                          \
            GSI_(TT, false, false);       \
        }                                 \
    } while (0)
    switch ((R + 15) / 16) {
    case 1: GS_(1); break;
    case 2: GS_(2); break;
    case 3: GS_(3); break;
    default: GS_(4); break;
    }
#undef GS_
#undef GSI_
    return sum_cols_device(part, nb, R * R, G, zero, zero_bytes, s);
}
template int gram_scale_device<double, double>(double*, uint64_t, int, const double*, double*, double*, double*,
                                               hipStream_t, bool, void*, uint64_t, const double*, double*);
template int gram_scale_device<float, float>(float*, uint64_t, int, const double*, float*, double*, double*,
                                             hipStream_t, bool, void*, uint64_t, const double*, double*);
template int gram_scale_device<float, bf16_t>(float*, uint64_t, int, const double*, bf16_t*, double*, double*,
                                              hipStream_t, bool, void*, uint64_t, const double*, double*);

template int gram_device<double>(const double*, uint64_t, int, double*, double*, hipStream_t, bool);
template int gram_device<float>(const float*, uint64_t, int, double*, double*, hipStream_t, bool);
template int gram_device<bf16_t>(const bf16_t*, uint64_t, int, double*, double*, hipStream_t, bool);

// ------------------------------------------------------------------ Gram without allocation
// Per-block partial Grams of alto_update_gram_noalloc live in this array:
// one copy per GPU, held until the process ends.  Host code never uses the
// array's name as an address: the device address comes from
// hipGetSymbolAddress, resolved once per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = 16u << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

static int scratch_address(double** out)
{
    static std::mutex mu;
    static double* addr[64] = {};
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    if (!addr[dev & 63]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        addr[dev & 63] = static_cast<double*>(p);
    }
    *out = addr[dev & 63];
    return ALTO_OK;
}

// update_gram into a device R x R FP64 matrix with the block partials in the
// per-device scratch array; asynchronous on `s`.  The blocks are the ones
// alto_update_gram uses, so the two agree bit for bit.
int gram_noalloc(const void* U, uint64_t dim, int R, int precision, double* gram, hipStream_t s)
{
    if (!U || !gram || R < 1 || R > 128 || dim < 1) {
        set_error("alto_update_gram_noalloc: invalid argument (rank 1..128, dim >= 1, non-null pointers)");
        return ALTO_ERR_ARG;
    }
    if (precision != ALTO_PREC_F64 && precision != ALTO_PREC_F32 && precision != ALTO_PREC_BF16) {
        set_error("alto_update_gram_noalloc: unknown precision %d", precision);
        return ALTO_ERR_ARG;
    }
    const size_t need = (size_t)gram_blocks(dim) * (size_t)R * (size_t)R * sizeof(double);
    if (need > NOALLOC_SCRATCH_BYTES) {
        set_error("alto_update_gram_noalloc: %zu bytes of partial Grams (dim %llu, rank %d) exceed the %zu-byte "
                  "scratch; use alto_update_gram",
                  need, (unsigned long long)dim, R, NOALLOC_SCRATCH_BYTES);
        return ALTO_ERR_UNSUPPORTED;
    }
    double* part = nullptr;
    ALTO_TRY(scratch_address(&part));
    if (precision == ALTO_PREC_F64) return gram_device<double>((const double*)U, dim, R, gram, part, s, true);
    if (precision == ALTO_PREC_F32) return gram_device<float>((const float*)U, dim, R, gram, part, s, false);
    return gram_device<bf16_t>((const bf16_t*)U, dim, R, gram, part, s, false);
}

// ------------------------------------------------------------------ V and its Cholesky factor
// V^-1 for symmetric positive definite V = prod_{m != mode} G_m by in-place
// Gauss-Jordan elimination without pivoting (one workgroup, R <= 128): the
// k-th pivot is the k-th Schur complement diagonal, i.e. the square of the
// Cholesky diagonal, so a pivot <= 0 is exactly the Cholesky failure of the
// reference's dpotrf test; then status = 1 and the same workgroup computes
// the pseudo-inverse (pinv_body).  status 0: W = V^-1.
//
// Thread t keeps the entries (r0 + rstep n, jc) of the matrix in registers,
// rstep = BT / R, r0 = t / R, jc = t % R (BT threads: 1024 from R = 17 on,
// so that a thread holds at most 16 entries and the waves of all four SIMDs
// share each step).  Step k reads column k (the elimination factors, pivot
// included) and row k from LDS buffer k & 1 and updates every register; the
// owners of column and row k + 1 write their new values into the other
// buffer as they compute them, so one barrier per step orders everything
// (buffer (k + 1) & 1 is written after the barrier of step k, which every
// reader of step k - 1, the buffer's previous use, has passed).
__device__ __attribute__((noinline)) void pinv_body(const double* __restrict__ grams, int nmode, int mode, int R,
                                                    double* __restrict__ A, double* __restrict__ Q,
                                                    double* __restrict__ Pout);

template <int BT, int NPT>
__device__ __forceinline__ void spd_inverse_body(const double* __restrict__ grams, int nmode, int mode, int R,
                                                 double* __restrict__ W, int* __restrict__ status,
                                                 double* __restrict__ wsA, double* __restrict__ wsQ)
{
    __shared__ double s_col[2][128], s_row[2][128];
    const int tid = threadIdx.x;
    const int rstep = BT / R, r0 = tid / R, jc = tid % R;
    const bool active = tid < rstep * R;
    // rows of this thread are clamped into the matrix so that every load and
    // LDS read below is unconditional (entries of rows >= R are never stored)
    auto rc = [&](int n) { return min(r0 + rstep * n, R - 1); };
    double a[NPT];
#pragma unroll
    for (int n = 0; n < NPT; ++n) a[n] = 1.0;
#pragma unroll
    for (int m = 0; m < ALTO_MAXN; ++m) {
        if (m < nmode && m != mode) {
            const double* g = grams + (size_t)m * R * R + (active ? jc : 0);
#pragma unroll
            for (int n = 0; n < NPT; ++n) a[n] *= g[rc(n) * R];
        }
    }
    // column 0 and row 0 for the first step; every later step's column and
    // row are written by their owners as they update them
    if (active) {
#pragma unroll
        for (int n = 0; n < NPT; ++n) {
            const int i = r0 + rstep * n;
            if (i < R && jc == 0) s_col[0][i] = a[n];
        }
        if (r0 == 0) s_row[0][jc] = a[0];
    }
    int ok = 1;
    for (int k = 0; k < R; ++k) {
        const int b = k & 1, nb = b ^ 1;
        __syncthreads();
        const double piv = s_col[b][k];
        if (!(piv > 0.0)) {  // uniform: every thread read the same pivot
            ok = 0;
            break;
        }
        // row k scaled by 1 / pivot, its diagonal slot holding 1 / pivot;
        // the other rows lose s_col[i] times it (column k becomes -s_col / piv)
        const double rk = (jc == k ? 1.0 : s_row[b][jc]) / piv;
        constexpr int CH = NPT < 16 ? NPT : 16;  // LDS reads in flight per batch
#pragma unroll
        for (int n0 = 0; n0 < NPT; n0 += CH) {
            double cv[CH];
#pragma unroll
            for (int n = 0; n < CH; ++n) cv[n] = s_col[b][rc(n0 + n)];
#pragma unroll
            for (int n = 0; n < CH; ++n) {
                const int i = r0 + rstep * (n0 + n);
                const double v = i == k ? rk : (jc == k ? 0.0 : a[n0 + n]) - cv[n] * rk;
                a[n0 + n] = v;
                if (active && i < R) {
                    if (jc == k + 1) s_col[nb][i] = v;
                    if (i == k + 1) s_row[nb][jc] = v;
                }
            }
        }
    }
    if (ok) {
        if (active) {
#pragma unroll
            for (int n = 0; n < NPT; ++n)
                if (r0 + rstep * n < R) W[(r0 + rstep * n) * R + jc] = a[n];
        }
        if (tid == 0) *status = 0;
        return;
    }
    if (tid == 0) *status = 1;
    __syncthreads();
    pinv_body(grams, nmode, mode, R, wsA, wsQ, W);
}

template <int BT, int NPT>
__global__ void __launch_bounds__(BT) k_spd_inverse(const double* __restrict__ grams, int nmode, int mode, int R,
                                                    double* __restrict__ W, int* __restrict__ status,
                                                    double* __restrict__ wsA, double* __restrict__ wsQ)
{
    spd_inverse_body<BT, NPT>(grams, nmode, mode, R, W, status, wsA, wsQ);
}

// Minimum-norm pseudo-inverse V^+ = Q diag(1/e) Q^T over |e| > R*eps*max|e|
// (parallel cyclic Jacobi; runs only after a failed Cholesky).  A and Q live in the
// workspace (global memory), shared through work-group barriers.
__device__ __attribute__((noinline)) void pinv_body(const double* __restrict__ grams, int nmode, int mode, int R,
                                                    double* __restrict__ A, double* __restrict__ Q,
                                                    double* __restrict__ Pout)
{
    const int tid = threadIdx.x;
    __shared__ double s_red[1024];
    __shared__ double s_c[64], s_s[64];
    __shared__ int s_p[64], s_q[64];
    __shared__ int s_stop;
    for (int k = tid; k < R * R; k += blockDim.x) {
        const int i = k / R, c = k % R;
        const int lo = i >= c ? k : c * R + i;  // symmetric from the lower triangle
        double v = 1.0;
        for (int m = 0; m < nmode; ++m)
            if (m != mode) v *= grams[(size_t)m * R * R + lo];
        A[k] = v;
        Q[k] = i == c ? 1.0 : 0.0;
    }
    __syncthreads();
    // a non-finite V has no pseudo-inverse: return NaN at once
    __shared__ int s_bad;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    for (int k = tid; k < R * R; k += blockDim.x)
        if (!isfinite(A[k])) s_bad = 1;
    __syncthreads();
    if (s_bad) {
        for (int k = tid; k < R * R; k += blockDim.x) Pout[k] = NAN;
        return;
    }
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, tot = 0;
        for (int k = tid; k < R * R; k += blockDim.x) {
            const double v = A[k] * A[k];
            tot += v;
            if (k / R != k % R) off += v;
        }
        s_red[tid] = off;
        __syncthreads();
        for (int o = (int)blockDim.x / 2; o > 0; o >>= 1) {
            if (tid < o) s_red[tid] += s_red[tid + o];
            __syncthreads();
        }
        const double offs = s_red[0];
        __syncthreads();
        s_red[tid] = tot;
        __syncthreads();
        for (int o = (int)blockDim.x / 2; o > 0; o >>= 1) {
            if (tid < o) s_red[tid] += s_red[tid + o];
            __syncthreads();
        }
        if (tid == 0) s_stop = (offs <= 1e-30 * s_red[0]) || offs == 0.0;
        __syncthreads();
        if (s_stop) break;
        // one sweep in Rp - 1 steps of Rp / 2 disjoint rotations (round-robin
        // pairing: index Rp - 1 stays, the others circulate, so every pair
        // meets once per sweep; with R odd, index R is a dummy).  A step
        // computes all its rotations from the same A, then applies
        // A <- A J, Q <- Q J and A <- J^T A, each in parallel.
        const int Rp = R + (R & 1);
        for (int t = 0; t < Rp - 1; ++t) {
            for (int k = tid; k < Rp / 2; k += blockDim.x) {
                const int u = k == 0 ? Rp - 1 : (t + k) % (Rp - 1);
                const int v = (t - k + Rp - 1) % (Rp - 1);
                const int p = min(u, v), q = max(u, v);
                double c = 1.0, sn = 0.0;
                if (q < R) {
                    const double apq = A[p * R + q];
                    if (fabs(apq) >= 1e-300) {
                        const double th = (A[q * R + q] - A[p * R + p]) / (2.0 * apq);
                        const double tt = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                        c = 1.0 / sqrt(tt * tt + 1.0);
                        sn = tt * c;
                    }
                }
                s_p[k] = p;
                s_q[k] = q;
                s_c[k] = c;
                s_s[k] = sn;
            }
            __syncthreads();
            for (int w = tid; w < (Rp / 2) * R; w += blockDim.x) {  // columns p, q of A and Q
                const int k = w / R, r = w % R;
                const double c = s_c[k], sn = s_s[k];
                if (sn != 0.0) {
                    const int p = s_p[k], q = s_q[k];
                    const double arp = A[r * R + p], arq = A[r * R + q];
                    A[r * R + p] = c * arp - sn * arq;
                    A[r * R + q] = sn * arp + c * arq;
                    const double qrp = Q[r * R + p], qrq = Q[r * R + q];
                    Q[r * R + p] = c * qrp - sn * qrq;
                    Q[r * R + q] = sn * qrp + c * qrq;
                }
            }
            __syncthreads();
            for (int w = tid; w < (Rp / 2) * R; w += blockDim.x) {  // rows p, q of A
                const int k = w / R, r = w % R;
                const double c = s_c[k], sn = s_s[k];
                if (sn != 0.0) {
                    const int p = s_p[k], q = s_q[k];
                    const double apr = A[p * R + r], aqr = A[q * R + r];
                    A[p * R + r] = c * apr - sn * aqr;
                    A[q * R + r] = sn * apr + c * aqr;
                }
            }
            __syncthreads();
        }
    }
    double emax = 0;
    for (int k = 0; k < R; ++k) emax = fmax(emax, fabs(A[k * R + k]));
    const double tol = emax * R * 2.220446049250313e-16;
    for (int k = tid; k < R * R; k += blockDim.x) {
        const int i = k / R, j = k % R;
        double v = 0;
        for (int e = 0; e < R; ++e) {
            const double ev = A[e * R + e];
            if (fabs(ev) > tol) v += Q[i * R + e] * Q[j * R + e] / ev;
        }
        Pout[k] = v;
    }
}

// ------------------------------------------------------------------ row solves
// U[i,:] = M[i,:] W with W = V^-1 (or V^+), as a GEMM on the matrix cores
// (v_mfma_f64_16x16x4f64; v_mfma_f32_16x16x4f32 for FP32 / BF16 models):
// each wave takes 16 rows at a time; per k-step of
// 4 it loads one M element per lane (A operand: lane l holds M[row0 + l%16]
// [k0 + l/16]) and multiplies it into every 16-column tile of the output
// (B operand from the LDS copy of W: lane l holds W[k0 + l/16][c0 + l%16]).
// Fused with the per-block column sums of squares / maxima of the stored
// values (for lambda) and, when pin is given, the per-block sum of
// U[i,c] M[i,c] over the block's rows (the fit's inner product: with
// U = lambda * U_normalised it equals sum_c lambda_c <U_normalised[:,c], M[:,c]>).
// The 16 x R block of M is loaded into registers before the products, so
// its loads are in flight together.  T16: 16-column tiles (R <= 16 T16).
template <typename MT, typename UT, int T16, bool LOWP>
__device__ __forceinline__ void apply_body(const MT* __restrict__ M, uint64_t I, int R, const double* __restrict__ Wg,
                                           UT* __restrict__ U, double* __restrict__ psum, double* __restrict__ pmax,
                                           double* __restrict__ pin, uint32_t bid, uint32_t nblk)
{
    // LOWP (FP32 / BF16 models with ALTO_TUNE_DENSE_F64 off): the products
    // on FP32 MFMA with M and the FP64 inverse rounded to FP32; otherwise
    // FP64 MFMA, the result rounded to U's type when it is stored
    constexpr bool F32 = LOWP;
    using mma_t = typename std::conditional<F32, float, double>::type;
    using acc_t = typename std::conditional<F32, f4_t, d4_t>::type;
    constexpr int KS = 4 * T16;  // k-steps of 4 covering 16 T16 >= R columns
    extern __shared__ double sWd[];  // W, row stride WS
    mma_t* sW = reinterpret_cast<mma_t*>(sWd);
    const int WS = R + 2;
    for (int k = threadIdx.x; k < R * R; k += blockDim.x) sW[(k / R) * WS + k % R] = (mma_t)Wg[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    double ss[T16], mx[T16];
#pragma unroll
    for (int t = 0; t < T16; ++t) {
        ss[t] = 0;
        mx[t] = -INFINITY;
    }
    double in = 0;
    for (uint64_t row0 = ((uint64_t)bid * 4 + wave) * 16; row0 < I; row0 += (uint64_t)nblk * 64) {
        const uint64_t ra = row0 + r16;
        mma_t av[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int k = ks * 4 + kq;
            av[ks] = (ra < I && k < R) ? (mma_t)M[ra * R + k] : (mma_t)0;
        }
        acc_t acc[T16];
#pragma unroll
        for (int t = 0; t < T16; ++t) acc[t] = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            if (ks * 4 < R) {  // uniform
                const int k = ks * 4 + kq;
#pragma unroll
                for (int t = 0; t < T16; ++t) {
                    const int c = t * 16 + r16;
                    const mma_t b = (k < R && c < R) ? sW[k * WS + c] : (mma_t)0;
                    if constexpr (F32)
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[ks], b, acc[t], 0, 0, 0);
                    else
                        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ks], b, acc[t], 0, 0, 0);
                }
            }
        }
        // register v of lane l holds C[l/16 + 4v][l%16] (FP64) or
        // C[4 (l/16) + v][l%16] (FP32)
#pragma unroll
        for (int t = 0; t < T16; ++t) {
            const int c = t * 16 + r16;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint64_t r = row0 + (F32 ? 4 * kq + v : kq + 4 * v);
                if (r < I && c < R) {
                    const UT x = (UT)acc[t][v];
                    U[r * R + c] = x;
                    const double xd = (double)x;
                    ss[t] += xd * xd;
                    mx[t] = fmax(mx[t], xd);
                    if (pin) in += xd * (double)M[r * R + c];
                }
            }
        }
    }
    // column statistics: lanes with equal l%16 hold the same column
    __shared__ double s_ss[4 * 16 * T16], s_mx[4 * 16 * T16];
#pragma unroll
    for (int t = 0; t < T16; ++t) {
        double a = ss[t], b = mx[t];
        a += __shfl_xor(a, 16);
        b = fmax(b, __shfl_xor(b, 16));
        a += __shfl_xor(a, 32);
        b = fmax(b, __shfl_xor(b, 32));
        if (kq == 0) {
            s_ss[(wave * T16 + t) * 16 + r16] = a;
            s_mx[(wave * T16 + t) * 16 + r16] = b;
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < R; c += blockDim.x) {
        const int t = c / 16, l = c % 16;
        double a = 0, b = -INFINITY;
        for (int w = 0; w < 4; ++w) {
            a += s_ss[(w * T16 + t) * 16 + l];
            b = fmax(b, s_mx[(w * T16 + t) * 16 + l]);
        }
        psum[(size_t)bid * R + c] = a;
        pmax[(size_t)bid * R + c] = b;
    }
    if (pin) {
        __shared__ double s_in[4];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) in += __shfl_xor(in, o);
        if (lane == 0) s_in[wave] = in;
        __syncthreads();
        if (threadIdx.x == 0) pin[bid] = (s_in[0] + s_in[1]) + (s_in[2] + s_in[3]);
    }
}

template <typename MT, typename UT, int T16, bool LOWP>
__global__ void __launch_bounds__(256) k_apply(const MT* __restrict__ M, uint64_t I, int R, const double* __restrict__ Wg,
                                               UT* __restrict__ U, double* __restrict__ psum, double* __restrict__ pmax,
                                               double* __restrict__ pin)
{
    apply_body<MT, UT, T16, LOWP>(M, I, R, Wg, U, psum, pmax, pin, blockIdx.x, gridDim.x);
}

// lambda: iteration 0 -> 2-norm of each column; later -> max(1, max_i U[i,r]).
// One block per column, tree reduction over the per-block partials.
__device__ __forceinline__ void lambda_body(const double* __restrict__ psum, const double* __restrict__ pmax, int nb,
                                            int R, int which, double* __restrict__ lambda, int c)
{
    __shared__ double sa[256], sb[256];
    double a = 0, b = -INFINITY;
    for (int p = threadIdx.x; p < nb; p += blockDim.x) {
        a += psum[(size_t)p * R + c];
        b = fmax(b, pmax[(size_t)p * R + c]);
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) {
            sa[threadIdx.x] += sa[threadIdx.x + o];
            sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) lambda[c] = which == 0 ? sqrt(sa[0]) : fmax(1.0, fmax(sb[0], 0.0));
}

__global__ void __launch_bounds__(256) k_lambda(const double* __restrict__ psum, const double* __restrict__ pmax, int nb,
                                                int R, int which, double* __restrict__ lambda)
{
    lambda_body(psum, pmax, nb, R, which, lambda, blockIdx.x);
}

// U /= lambda (per column), with the storage copy used by MTTKRP
template <typename UT, typename ST>
__global__ void k_scale_cols(UT* __restrict__ U, uint64_t n, int R, const double* __restrict__ lambda, ST* __restrict__ S)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const double v = (double)U[k] / lambda[k % R];
        U[k] = (UT)v;
        if (S) S[k] = store_as<ST>(v);
    }
}

// fit inner-product partials: sum_i U[i,r] * M[i,r]; thread t covers
// column t % R of rows t / R + k * (256 / R) (coalesced)
template <typename UT, typename MT>
__global__ void k_inner_partial(const UT* __restrict__ U, const MT* __restrict__ Mm, uint64_t I, int R,
                                double* __restrict__ part)
{
    __shared__ double s[256];
    const int RP = 256 / R;
    const int t = threadIdx.x;
    double a = 0;
    if (t < RP * R) {
        const int c = t % R, ri = t / R;
        for (uint64_t i = (uint64_t)blockIdx.x * RP + ri; i < I; i += (uint64_t)gridDim.x * RP)
            a += as_f64(U[i * R + c]) * (double)Mm[i * R + c];
    }
    s[t] = a;
    __syncthreads();
    for (int c = t; c < R; c += blockDim.x) {
        double v = 0;
        for (int ri = 0; ri < RP; ++ri) v += s[ri * R + c];
        part[(size_t)blockIdx.x * R + c] = v;
    }
}

// fit = 1 - sqrt(res)/sqrt(normX), res = normX + normU - 2 inner (sqrt only
// when res > 0, as in the reference)
// inner from per-column partials (nb x R, weighted by lambda) or, with
// scalar set, from nb per-block totals already carrying lambda
__device__ __forceinline__ void fit_body(const double* __restrict__ grams, int nmode, int R,
                                         const double* __restrict__ lambda, const double* __restrict__ ipart, int nb,
                                         int scalar, double normX, double* __restrict__ out)
{
    __shared__ double s_in[256], s_nu[256];
    double inner = 0, nu = 0;
    if (scalar) {
        for (int p = threadIdx.x; p < nb; p += blockDim.x) inner += ipart[p];
    } else {
        for (int c = threadIdx.x; c < R; c += blockDim.x) {
            double a = 0;
            for (int p = 0; p < nb; ++p) a += ipart[(size_t)p * R + c];
            inner += a * lambda[c];
        }
    }
    for (int k = threadIdx.x; k < R * R; k += blockDim.x) {
        const int i = k / R, j = k % R;
        if (j > i) continue;
        double t = lambda[i] * lambda[j];
        for (int m = 0; m < nmode; ++m) t *= grams[(size_t)m * R * R + k];
        nu += i == j ? t : 2.0 * t;
    }
    s_in[threadIdx.x] = inner;
    s_nu[threadIdx.x] = nu;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) {
            s_in[threadIdx.x] += s_in[threadIdx.x + o];
            s_nu[threadIdx.x] += s_nu[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double normU = fabs(s_nu[0]);
        double res = normX + normU - 2.0 * s_in[0];
        if (res > 0.0) res = sqrt(res);
        out[0] = 1.0 - res / sqrt(normX);
        out[1] = s_in[0];
        out[2] = normU;
    }
}

__global__ void __launch_bounds__(256) k_fit(const double* __restrict__ grams, int nmode, int R,
                                             const double* __restrict__ lambda, const double* __restrict__ ipart,
                                             int nb, int scalar, double normX, double* __restrict__ out)
{
    fit_body(grams, nmode, R, lambda, ipart, nb, scalar, normX, out);
}

__global__ void k_sumsq_partial(const double* __restrict__ v, uint64_t n, double* __restrict__ part)
{
    __shared__ double s[256];
    double a = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a += v[i] * v[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Counter-based uniform (0, 1] factor entries: SplitMix64 of (seed, mode,
// index).  Used for models seeded on the device (served jobs).
__global__ void k_random_factor(double* __restrict__ U, uint64_t n, uint64_t seed, int mode)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed * 0x9E3779B97F4A7C15ull + ((uint64_t)mode << 56) + i + 1;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        U[i] = ((double)(z >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    }
}

int random_factor_device(double* U, uint64_t n, uint64_t seed, int mode, hipStream_t s)
{
    uint64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL(k_random_factor, dim3((unsigned)(b ? b : 1)), dim3(256), 0, s, U, n, seed, mode);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

// ------------------------------------------------------------------ host wrappers
int solve_blocks(uint64_t I, int R)
{
    (void)R;
    uint64_t b = (I + 63) / 64;  // 64 rows per workgroup pass
    if (b > 2048) b = 2048;
    return b < 1 ? 1 : (int)b;
}

int chol_device(const double* grams, int nmode, int mode, int R, double* L, double* P, double* wsA, double* wsQ,
                int* status, hipStream_t s)
{
    if (R > 128) {
        set_error("pseudo-inverse: rank above 128");
        return ALTO_ERR_UNSUPPORTED;
    }
    (void)L;
    // registers per thread: ceil(R / (BT / R)) matrix entries
#define SPD_(BT, NPT) \
    hipLaunchKernelGGL((k_spd_inverse<BT, NPT>), dim3(1), dim3(BT), 0, s, grams, nmode, mode, R, P, status, wsA, wsQ)
    if (R <= 16) SPD_(256, 1);
    else if (R <= 32) SPD_(1024, 1);
    else if (R <= 64) SPD_(1024, 4);
    else SPD_(1024, 16);
#undef SPD_
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

template <typename MT, typename UT>
int solve_device(const MT* M, uint64_t I, int R, const double* L, const double* P, const int* status, UT* U,
                 double* psum, double* pmax, int nb, hipStream_t s, double* pin)
{
    (void)L;
    (void)status;
    const size_t lds = (size_t)R * (R + 2) * 8;
    const bool lowp = std::is_same<UT, float>::value && !dense_f64_enabled();
#define AP_(TT)                                                                                                     \
    do {                                                                                                            \
        if (lowp)                                                                                                   \
            hipLaunchKernelGGL((k_apply<MT, UT, TT, true>), dim3(nb), dim3(256), lds, s, M, I, R, P, U, psum, pmax, \
                               pin);                                                                                \
        else                                                                                                        \
            hipLaunchKernelGGL((k_apply<MT, UT, TT, false>), dim3(nb), dim3(256), lds, s, M, I, R, P, U, psum,     \
                               pmax, pin);                                                                          \
    } while (0)
    switch ((R + 15) / 16) {
    case 1: AP_(1); break;
    case 2: AP_(2); break;
    case 3: AP_(3); break;
    case 4: AP_(4); break;
    case 5: AP_(5); break;
    case 6: AP_(6); break;
    case 7: AP_(7); break;
    default: AP_(8); break;
    }
#undef AP_
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}
template int solve_device<double, double>(const double*, uint64_t, int, const double*, const double*, const int*,
                                          double*, double*, double*, int, hipStream_t, double*);
template int solve_device<double, float>(const double*, uint64_t, int, const double*, const double*, const int*, float*,
                                         double*, double*, int, hipStream_t, double*);

int lambda_device(const double* psum, const double* pmax, int nb, int R, int which, double* lambda, hipStream_t s)
{
    hipLaunchKernelGGL(k_lambda, dim3(R), dim3(256), 0, s, psum, pmax, nb, R, which, lambda);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

template <typename UT, typename ST>
int scale_device(UT* U, uint64_t I, int R, const double* lambda, ST* S, hipStream_t s)
{
    uint64_t n = I * (uint64_t)R;
    uint64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL((k_scale_cols<UT, ST>), dim3((unsigned)(b ? b : 1)), dim3(256), 0, s, U, n, R, lambda, S);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}
template int scale_device<double, double>(double*, uint64_t, int, const double*, double*, hipStream_t);
template int scale_device<float, float>(float*, uint64_t, int, const double*, float*, hipStream_t);
template int scale_device<float, bf16_t>(float*, uint64_t, int, const double*, bf16_t*, hipStream_t);
template int scale_device<double, float>(double*, uint64_t, int, const double*, float*, hipStream_t);

constexpr int INNER_BLOCKS = 64;

template <typename UT, typename MT>
int inner_partial_device(const UT* U, const MT* Mm, uint64_t I, int R, double* part, hipStream_t s)
{
    hipLaunchKernelGGL((k_inner_partial<UT, MT>), dim3(INNER_BLOCKS), dim3(256), 0, s, U, Mm, I, R, part);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}
template int inner_partial_device<double, double>(const double*, const double*, uint64_t, int, double*, hipStream_t);
template int inner_partial_device<float, float>(const float*, const float*, uint64_t, int, double*, hipStream_t);
template int inner_partial_device<bf16_t, double>(const bf16_t*, const double*, uint64_t, int, double*, hipStream_t);
template int inner_partial_device<float, double>(const float*, const double*, uint64_t, int, double*, hipStream_t);

int fit_device(const double* grams, int nmode, int R, const double* lambda, const double* ipart, double normX,
               double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_fit, dim3(1), dim3(256), 0, s, grams, nmode, R, lambda, ipart, INNER_BLOCKS, 0, normX, out);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

int fit_scalar_device(const double* grams, int nmode, int R, const double* lambda, const double* pin, int nb,
                      double normX, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_fit, dim3(1), dim3(256), 0, s, grams, nmode, R, lambda, pin, nb, 1, normX, out);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

int normsq_device(const double* v, uint64_t n, double* part, double* out, hipStream_t s)
{
    const int nb = 256;
    hipLaunchKernelGGL(k_sumsq_partial, dim3(nb), dim3(256), 0, s, v, n, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1), 0, s, part, nb, 1, out);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

// ------------------------------------------------------------------ batched mode updates
// The mode-j update of several CP-ALS models (lockstep sessions of one rank
// and precision) in one launch per kernel: each kernel's table (up to
// DENSE_BATCH_MAX jobs, in the kernel arguments) gives job k the logical
// blocks [start[k], start[k + 1]) -- its own launch's grid -- and each block
// finds its job by a uniform scan.  The bodies are the single-model
// kernels', so every job computes what its own launches would.
template <typename J>
struct BatchTab {
    J a[DENSE_BATCH_MAX];
    uint32_t start[DENSE_BATCH_MAX + 1];
    int n;
};
template <typename J>
__device__ inline int batch_job(const BatchTab<J>& t, uint32_t b)
{
    int k = 0;
    while (k + 1 < t.n && b >= t.start[k + 1]) ++k;
    return k;
}

struct SpdJob {
    const double* grams;
    double *P, *wsA, *wsQ;
    int* status;
    int nmode, mode;
};
struct ApplyJob {
    const double* M;
    const double* P;
    void* U;
    double *psum, *pmax, *pin;
    uint64_t I;
};
struct LambdaJob {
    const double *psum, *pmax;
    double* lambda;
    int nb, which;
};
struct GramJob {
    void* U;
    void* S;
    const double* lambda;
    double* part;
    const double* Mi;
    double* pin;
    uint64_t I, rpb;
};
struct SumJob {
    const double* part;
    double* out;
    void* zero;
    uint64_t nz;
    int nparts, nsum;
};
struct FitJob {
    const double *grams, *lambda, *pin;
    double* out;
    double* gather;
    double normX;
    int nmode, nb;
};

template <int BT, int NPT>
__global__ void __launch_bounds__(BT) k_spd_inverse_batch(BatchTab<SpdJob> t, int R)
{
    const SpdJob& j = t.a[blockIdx.x];
    spd_inverse_body<BT, NPT>(j.grams, j.nmode, j.mode, R, j.P, j.status, j.wsA, j.wsQ);
}

template <typename MT, typename UT, int T16, bool LOWP>
__global__ void __launch_bounds__(256) k_apply_batch(BatchTab<ApplyJob> t, int R)
{
    const int k = batch_job(t, blockIdx.x);
    const ApplyJob& j = t.a[k];
    apply_body<MT, UT, T16, LOWP>(j.M, j.I, R, j.P, (UT*)j.U, j.psum, j.pmax, j.pin, blockIdx.x - t.start[k],
                                  t.start[k + 1] - t.start[k]);
}

__global__ void __launch_bounds__(256) k_lambda_batch(BatchTab<LambdaJob> t, int R)
{
    const LambdaJob& j = t.a[blockIdx.x / R];
    lambda_body(j.psum, j.pmax, j.nb, R, j.which, j.lambda, blockIdx.x % R);
}

template <typename T, typename ST, bool F64, int T16, bool INNER>
__global__ void __launch_bounds__(256) k_gram_rows_batch(BatchTab<GramJob> t, int R)
{
    const int k = batch_job(t, blockIdx.x);
    const GramJob& j = t.a[k];
    gram_rows_body<T, ST, F64, T16, true, INNER>((T*)j.U, j.I, R, j.rpb, j.part, j.lambda, (ST*)j.S, j.Mi, j.pin,
                                                 blockIdx.x - t.start[k]);
}

__global__ void __launch_bounds__(256) k_sum_cols_batch(BatchTab<SumJob> t, int n)
{
    const int k = batch_job(t, blockIdx.x);
    const SumJob& j = t.a[k];
    sum_cols_body(j.part, j.nparts, n, j.out, j.nsum, j.zero, j.nz, blockIdx.x - t.start[k],
                  t.start[k + 1] - t.start[k]);
}

__global__ void __launch_bounds__(256) k_fit_batch(BatchTab<FitJob> t, int R)
{
    const FitJob& j = t.a[blockIdx.x];
    fit_body(j.grams, j.nmode, R, j.lambda, j.pin, j.nb, 1, j.normX, j.out);
    if (threadIdx.x == 0) j.gather[0] = j.out[0];
}

template <typename UT, typename ST>
int dense_mode_batch(int n, const DenseModeJob* m, int R, bool f64, hipStream_t s)
{
    if (n < 1 || n > DENSE_BATCH_MAX || R < 1 || R > 64) {
        set_error("dense_mode_batch: %d jobs of rank %d (1..%d jobs, rank 1..64)", n, R, DENSE_BATCH_MAX);
        return ALTO_ERR_ARG;
    }
    constexpr bool same = std::is_same<UT, ST>::value;
    BatchTab<SpdJob> sp{};
    BatchTab<ApplyJob> ap{};
    BatchTab<LambdaJob> la{};
    BatchTab<GramJob> gr{};
    BatchTab<SumJob> su{};
    sp.n = ap.n = la.n = gr.n = su.n = n;
    ap.start[0] = gr.start[0] = su.start[0] = 0;
    for (int k = 0; k < n; ++k) {
        const DenseModeJob& x = m[k];
        sp.a[k] = SpdJob{x.grams, x.P, x.wsA, x.wsQ, x.status, x.nmode, x.mode};
        ap.a[k] = ApplyJob{x.M, x.P, x.U, x.psum, x.pmax, x.pin, x.I};
        ap.start[k + 1] = ap.start[k] + (uint32_t)x.nb;
        la.a[k] = LambdaJob{x.psum, x.pmax, x.lambda, x.nb, x.which};
        const int gb = gram_blocks(x.I);
        const uint64_t rpb = ((x.I + gb - 1) / gb + 15) / 16 * 16;
        gr.a[k] = GramJob{x.U, same ? nullptr : x.S, x.lambda, x.gpart, x.Mgram, x.pin_gram, x.I, rpb};
        gr.start[k + 1] = gr.start[k] + (uint32_t)gb;
        // the sums and the clearing of the next mode's M, as sum_cols_device
        const int nsum = (R * R + 31) / 32;
        uint64_t nzb = (x.zero_bytes / 16 + 255) / 256;
        if (nzb > 1024) nzb = 1024;
        if (x.zero_bytes && nzb == 0) nzb = 1;
        su.a[k] = SumJob{x.gpart, x.G, x.zero, x.zero_bytes, gb, nsum};
        su.start[k + 1] = su.start[k] + (uint32_t)(nsum + nzb);
    }
#define SPD_(BT, NPT) hipLaunchKernelGGL((k_spd_inverse_batch<BT, NPT>), dim3(n), dim3(BT), 0, s, sp, R)
    if (R <= 16) SPD_(256, 1);
    else if (R <= 32) SPD_(1024, 1);
    else SPD_(1024, 4);
#undef SPD_
    const size_t lds_ap = (size_t)R * (R + 2) * 8;
    const bool lowp = std::is_same<UT, float>::value && !dense_f64_enabled();
#define AP_(TT)                                                                                                  \
    do {                                                                                                         \
        if (lowp)                                                                                                \
            hipLaunchKernelGGL((k_apply_batch<double, UT, TT, true>), dim3(ap.start[n]), dim3(256), lds_ap, s, ap, R); \
        else                                                                                                     \
            hipLaunchKernelGGL((k_apply_batch<double, UT, TT, false>), dim3(ap.start[n]), dim3(256), lds_ap, s, ap, \
                               R);                                                                               \
    } while (0)
    switch ((R + 15) / 16) {
    case 1: AP_(1); break;
    case 2: AP_(2); break;
    case 3: AP_(3); break;
    default: AP_(4); break;
    }
#undef AP_
    hipLaunchKernelGGL(k_lambda_batch, dim3(n * R), dim3(256), 0, s, la, R);
    const size_t lds_gr = (size_t)R * R * 8;
    // storage that rounds (BF16): the last modes' fit inner products from
    // the stored values, in the same launch
#define GS_(TT)                                                                                                     \
    do {                                                                                                            \
        if (f64)                                                                                                    \
            hipLaunchKernelGGL((k_gram_rows_batch<UT, ST, true, TT, !same>), dim3(gr.start[n]), dim3(256), lds_gr, s, \
                               gr, R);                                                                              \
        else                                                                                                        \
            hipLaunchKernelGGL((k_gram_rows_batch<UT, ST, false, TT, !same>), dim3(gr.start[n]), dim3(256), lds_gr, \
                               s, gr, R);                                                                           \
    } while (0)
    switch ((R + 15) / 16) {
    case 1: GS_(1); break;
    case 2: GS_(2); break;
    case 3: GS_(3); break;
    default: GS_(4); break;
    }
#undef GS_
    hipLaunchKernelGGL(k_sum_cols_batch, dim3(su.start[n]), dim3(256), 0, s, su, R * R);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}
template int dense_mode_batch<double, double>(int, const DenseModeJob*, int, bool, hipStream_t);
template int dense_mode_batch<float, float>(int, const DenseModeJob*, int, bool, hipStream_t);
template int dense_mode_batch<float, bf16_t>(int, const DenseModeJob*, int, bool, hipStream_t);

int fit_batch(int n, const DenseFitJob* f, int R, hipStream_t s)
{
    if (n < 1 || n > DENSE_BATCH_MAX) {
        set_error("fit_batch: %d jobs (1..%d)", n, DENSE_BATCH_MAX);
        return ALTO_ERR_ARG;
    }
    BatchTab<FitJob> t{};
    t.n = n;
    for (int k = 0; k < n; ++k)
        t.a[k] = FitJob{f[k].grams, f[k].lambda, f[k].pin, f[k].out, f[k].gather, f[k].normX, f[k].nmode, f[k].nb};
    hipLaunchKernelGGL(k_fit_batch, dim3(n), dim3(256), 0, s, t, R);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

}  // namespace alto
