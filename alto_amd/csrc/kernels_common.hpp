// Device helpers shared by the stream kernels (MTTKRP, Phi, log-likelihood).
#pragma once

#include "alto_internal.hpp"

namespace alto {

// bf16 storage: the upper 16 bits of an IEEE binary32
struct bf16_t {
    uint16_t x;
};

template <typename AT> __device__ inline AT to_acc(double v) { return (AT)v; }
template <typename AT> __device__ inline AT to_acc(float v) { return (AT)v; }
template <typename AT> __device__ inline AT to_acc(bf16_t v) { return (AT)__uint_as_float((uint32_t)v.x << 16); }

// round-to-nearest-even float -> bf16 (NaN kept quiet)
__device__ inline bf16_t to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    bf16_t r;
    if ((u & 0x7fffffffu) > 0x7f800000u) {
        r.x = (uint16_t)((u >> 16) | 0x40u);
        return r;
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    r.x = (uint16_t)(u >> 16);
    return r;
}

template <typename ST> __device__ inline ST store_as(double v);
template <> __device__ inline double store_as<double>(double v) { return v; }
template <> __device__ inline float store_as<float>(double v) { return (float)v; }
template <> __device__ inline bf16_t store_as<bf16_t>(double v) { return to_bf16((float)v); }

template <int PW>
__device__ inline uint64_t field(const uint64_t (&pk)[PW], int off, int bits)
{
    if (bits == 0) return 0;
    uint64_t msk = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    if (PW == 1) return (pk[0] >> off) & msk;
    uint64_t v;
    if (off >= 64) v = pk[PW - 1] >> (off - 64);
    else if (off == 0) v = pk[0];
    else v = (pk[0] >> off) | (pk[PW - 1] << (64 - off));
    return v & msk;
}

template <int PW>
__device__ inline void set_field(uint64_t (&pk)[PW], int off, int bits, uint64_t v)
{
    if (bits == 0) return;
    const uint64_t msk = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    v &= msk;
    if (PW == 1) {
        pk[0] = (pk[0] & ~(msk << off)) | (v << off);
        return;
    }
    if (off >= 64) {
        pk[PW - 1] = (pk[PW - 1] & ~(msk << (off - 64))) | (v << (off - 64));
        return;
    }
    pk[0] = (pk[0] & ~(msk << off)) | (v << off);
    if (off + bits > 64) {  // the high part continues in word 1 (off > 0 here)
        const uint64_t hm = (1ull << (off + bits - 64)) - 1;
        pk[PW - 1] = (pk[PW - 1] & ~hm) | (v >> (64 - off));
    }
}

__device__ inline uint64_t shfl_u64(uint64_t x, int src, int width)
{
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    lo = (uint32_t)__shfl((int)lo, src, width);
    hi = (uint32_t)__shfl((int)hi, src, width);
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline double shfl_val(double x, int src, int width) { return __shfl(x, src, width); }
__device__ inline float shfl_val(float x, int src, int width) { return __shfl(x, src, width); }

template <int G, typename T>
__device__ inline T group_sum(T v)
{
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}

__device__ inline void atomic_add_hw(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ inline void atomic_add_hw(float* p, float v) { unsafeAtomicAdd(p, v); }

template <typename AT, typename OT, int G, int E, bool CONTIG = false>
__device__ inline void flush_row(OT* out, uint32_t row, int R, int gl, const AT (&acc)[E])
{
    OT* o = out + (uint64_t)row * R;
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int c = CONTIG ? gl * E + q : gl + q * G;
        if (c < R) atomic_add_hw(&o[c], (OT)acc[q]);
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that XCD x
// runs logical blocks [x*nb8, (x+1)*nb8), a contiguous slice of the stream.
__device__ inline uint64_t xcd_block(uint64_t nblocks_real)
{
    const uint64_t nb8 = gridDim.x >> 3;
    const uint64_t b = blockIdx.x;
    const uint64_t logical = (b & 7) * nb8 + (b >> 3);
    return logical < nblocks_real ? logical : ~0ull;
}

template <typename FT, typename VT, typename OT>
struct StreamArgs {
    const uint64_t* pk;
    const VT* val;
    const FT* Uo[ALTO_MAXN - 1];      // factors of the other modes
    int offo[ALTO_MAXN - 1], bitso[ALTO_MAXN - 1];
    int nother;
    int offn, bitsn;                   // the target mode's field
    OT* out;
    uint64_t nnz;
    uint64_t seg;
    uint64_t nblocks_real;
    int nmode, mode, rank;
    // LDS staging of hot rows (ranked streams): elements of Uo[o] staged and
    // their byte offset in the workgroup's LDS
    uint32_t hot_n[ALTO_MAXN - 1];
    uint32_t hot_base[ALTO_MAXN - 1];
};

template <typename FT, typename VT, typename OT>
inline void fill_stream_args(const alto_tensor* t, int mode, int R, int seg, StreamArgs<FT, VT, OT>& a)
{
    a.pk = t->ms[mode].pk.as<uint64_t>();
    a.val = nullptr;
    a.out = nullptr;
    for (int o = 0; o < ALTO_MAXN - 1; ++o) {
        a.Uo[o] = nullptr;
        a.offo[o] = a.bitso[o] = 0;
        a.hot_n[o] = a.hot_base[o] = 0;
    }
    a.nother = 0;
    a.offn = t->pack.off[mode];
    a.bitsn = t->pack.bits[mode];
    a.nnz = t->nnz;
    a.seg = (uint64_t)seg;
    a.nmode = t->nmode;
    a.mode = mode;
    a.rank = R;
}

// Factor rows in rank order for a ranked stream: dst row k = src row rank_row[k].
template <typename FT>
__global__ void k_permute_rows(const FT* __restrict__ src, const uint32_t* __restrict__ rank_row, uint64_t d, int R,
                               FT* __restrict__ dst)
{
    const uint64_t n = d * (uint64_t)R;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = i / (uint64_t)R;
        dst[i] = src[(uint64_t)rank_row[k] * R + (i - k * R)];
    }
}

// Points the stream arguments at the other modes' factors (natural mode
// order).  For a ranked stream the kernels index the factors by rank: the
// rows are first copied into rank order (into buffers reserved by
// reserve_rank_copies on this stream's tensor).
template <typename FT, typename VT, typename OT>
inline int set_factors(const alto_tensor* t, int mode, const FT* const* U, StreamArgs<FT, VT, OT>& a, hipStream_t s)
{
    const bool ranked = t->ms[mode].ranked != 0;
    int o = 0;
    for (int m = 0; m < t->nmode; ++m) {
        if (m == mode) continue;
        a.Uo[o] = U[m];
        if (ranked) {
            const uint64_t n = t->dims[m] * (uint64_t)a.rank;
            if (t->pcopy[m].bytes < n * sizeof(FT) + 64) {
                set_error("rank-ordered factor copies not reserved (mode %d)", m);
                return ALTO_ERR_UNSUPPORTED;
            }
            uint64_t b = (n + 255) / 256;
            if (b > 8192) b = 8192;
            hipLaunchKernelGGL(k_permute_rows<FT>, dim3((unsigned)b), dim3(256), 0, s, U[m],
                               t->rank_row[m].as<uint32_t>(), t->dims[m], a.rank, t->pcopy[m].as<FT>());
            HIP_TRY(hipGetLastError());
            a.Uo[o] = t->pcopy[m].as<FT>();
        }
        a.offo[o] = t->pack.off[m];
        a.bitso[o] = t->pack.bits[m];
        ++o;
    }
    a.nother = o;
    return ALTO_OK;
}

// LDS bytes of hot rows per workgroup of the ranked-stream kernels, split
// evenly over the other modes; returns the bytes used.
template <typename FT, typename VT, typename OT>
inline uint32_t plan_hot_rows(const alto_tensor* t, int mode, StreamArgs<FT, VT, OT>& a, uint32_t lds_bytes)
{
    if (!t->ms[mode].ranked || a.nother == 0) return 0;
    const uint64_t per = (lds_bytes / a.nother) & ~15ull;  // 16-byte aligned regions
    uint32_t base = 0;
    int o = 0;
    for (int m = 0; m < t->nmode; ++m) {
        if (m == mode) continue;
        uint64_t rows = per / ((uint64_t)a.rank * sizeof(FT));
        if (rows > t->dims[m]) rows = t->dims[m];
        a.hot_n[o] = (uint32_t)(rows * a.rank);
        a.hot_base[o] = base;
        base += (uint32_t)((rows * a.rank * sizeof(FT) + 15) / 16 * 16);
        ++o;
    }
    return base;
}

// lanes per nonzero: 2 strided columns per lane, 4..64 lanes
inline int group_lanes(int R) { return R <= 8 ? 4 : R <= 16 ? 8 : R <= 32 ? 16 : R <= 64 ? 32 : 64; }
int tuned_seg(uint64_t nnz, int G, int acc_bytes);

template <typename A>
inline dim3 stream_grid(uint64_t nnz, int seg, int G, A& a, int block_threads = 256)
{
    uint64_t groups = (nnz + seg - 1) / seg;
    uint64_t threads = groups * G;
    uint64_t blocks = (threads + block_threads - 1) / block_threads;
    if (blocks < 1) blocks = 1;
    a.nblocks_real = blocks;
    uint64_t nb8 = (blocks + 7) / 8;
    return dim3((unsigned)(nb8 * 8));
}

#define RANK_DISPATCH_PW(PW_, ...)                                                    \
    do {                                                                               \
        if ((PW_) == 1) { constexpr int PW = 1; __VA_ARGS__; }                         \
        else { constexpr int PW = 2; __VA_ARGS__; }                                    \
    } while (0)

#define RANK_DISPATCH(R_, PW_, ...)                                                    \
    do {                                                                               \
        const int _r = (R_);                                                           \
        constexpr int E = 2;                                                           \
        if (_r <= 8) { constexpr int G = 4; RANK_DISPATCH_PW(PW_, __VA_ARGS__); }      \
        else if (_r <= 16) { constexpr int G = 8; RANK_DISPATCH_PW(PW_, __VA_ARGS__); }\
        else if (_r <= 32) { constexpr int G = 16; RANK_DISPATCH_PW(PW_, __VA_ARGS__); }\
        else if (_r <= 64) { constexpr int G = 32; RANK_DISPATCH_PW(PW_, __VA_ARGS__); }\
        else { constexpr int G = 64; RANK_DISPATCH_PW(PW_, __VA_ARGS__); }             \
    } while (0)

template <typename T>
__global__ void k_fill(T* __restrict__ a, uint64_t n, T v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) a[i] = v;
}

template <typename T>
inline int fill_value(T* p, uint64_t n, T v, hipStream_t s)
{
    if (n == 0) return ALTO_OK;
    uint64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)b), dim3(256), 0, s, p, n, v);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

template <typename T>
inline int fill_zero(T* p, uint64_t n, hipStream_t s) { return fill_value<T>(p, n, (T)0, s); }

// entry points implemented in mttkrp.hip
int mttkrp_device(alto_tensor* t, int mode, int R, const void* const* U, void* out, int precision, int accumulate,
                  hipStream_t s, int seg);
int phi_device(alto_tensor* t, int mode, int R, const double* const* U, double* phi, double eps, hipStream_t s, int seg);
int phi_device_cached(alto_tensor* t, int mode, int R, const double* const* U, double* phi, double eps, hipStream_t s,
                      int seg, double* pc, int pim);
int mttkrp_batched_impl(int njobs, alto_tensor_t* const* tensors, const int* modes, uint64_t rank,
                        const void* const* const* factors, void* const* outs, int precision, int accumulate,
                        hipStream_t s);

}  // namespace alto
