// MTTKRP stream kernel and its launcher, instantiated per factor element
// type in mttkrp_f64.hip / mttkrp_f32.hip / mttkrp_bf16.hip (one translation
// unit each, so that they build in parallel).
//
// Reference semantics: mttkrp_alto_par / mttkrp_alto_da_mem_pull
// (include/alto.hpp:617-623, :1227-1318).
#pragma once

#include <algorithm>
#include <type_traits>
#include <vector>

#include "alto_internal.hpp"
#include "kernels_common.hpp"

namespace alto {

// launch configuration chosen by the host side (tunables)
struct MttkrpTune {
    int variant = -1;       // lane mapping 0..3, -1 = per element type
    int nt = 0;             // non-temporal stream loads
    int hot_block = 512;    // threads per workgroup of the ranked-stream kernel
    int hot_lds = 49152;    // LDS bytes per workgroup for hot rows
    int narrow_max = 32;    // FP32 / BF16 ranks that use two columns per lane
    int narrow_small = 1;   // ... and every rank below 4M nonzeros
};

// Column mapping of a lane group: strided (lane gl holds c = gl + q*G, so
// each gather / atomic instruction touches G contiguous elements) or
// contiguous (lane gl holds c = gl*E + q, loaded with one 16-byte load where
// the row allows it).
template <bool CONTIG, int G, int E>
__device__ inline int col_of(int gl, int q) { return CONTIG ? gl * E + q : gl + q * G; }

// Row loads from global memory (plain pointers) or from LDS (address space
// 3 pointers, so that the compiler emits ds_read rather than flat loads).
template <typename T> using lds_ptr = const __attribute__((address_space(3))) T*;
typedef double v2f64 __attribute__((ext_vector_type(2)));
typedef float v4f32 __attribute__((ext_vector_type(4)));
typedef float v2f32 __attribute__((ext_vector_type(2)));
typedef uint32_t v2u32 __attribute__((ext_vector_type(2)));

template <typename V, typename T> __device__ inline const V* vcast(const T* p) { return reinterpret_cast<const V*>(p); }
template <typename V, typename T> __device__ inline lds_ptr<V> vcast(lds_ptr<T> p)
{
    return reinterpret_cast<lds_ptr<V>>(p);
}

template <typename AT, typename P> __device__ inline AT elem_acc(P p, int c, double*) { return (AT)p[c]; }
template <typename AT, typename P> __device__ inline AT elem_acc(P p, int c, float*) { return (AT)p[c]; }
template <typename AT, typename P> __device__ inline AT elem_acc(P p, int c, bf16_t*)
{
    return (AT)__uint_as_float((uint32_t)p[c].x << 16);
}

template <typename FT, int E> struct VecLoad {
    static constexpr bool ok = false;
};
template <> struct VecLoad<double, 2> {
    static constexpr bool ok = true;
    template <typename AT, typename P> __device__ static void load(P p, AT (&x)[2])
    {
        const v2f64 a = *vcast<v2f64>(p);
        x[0] = (AT)a.x;
        x[1] = (AT)a.y;
    }
};
template <> struct VecLoad<double, 4> {
    static constexpr bool ok = true;
    template <typename AT, typename P> __device__ static void load(P p, AT (&x)[4])
    {
        const v2f64 a = *vcast<v2f64>(p);
        const v2f64 b = *vcast<v2f64>(p + 2);
        x[0] = (AT)a.x; x[1] = (AT)a.y; x[2] = (AT)b.x; x[3] = (AT)b.y;
    }
};
template <> struct VecLoad<float, 4> {
    static constexpr bool ok = true;
    template <typename AT, typename P> __device__ static void load(P p, AT (&x)[4])
    {
        const v4f32 a = *vcast<v4f32>(p);
        x[0] = (AT)a.x; x[1] = (AT)a.y; x[2] = (AT)a.z; x[3] = (AT)a.w;
    }
};
template <> struct VecLoad<float, 2> {
    static constexpr bool ok = true;
    template <typename AT, typename P> __device__ static void load(P p, AT (&x)[2])
    {
        const v2f32 a = *vcast<v2f32>(p);
        x[0] = (AT)a.x; x[1] = (AT)a.y;
    }
};
template <> struct VecLoad<bf16_t, 4> {
    static constexpr bool ok = true;
    template <typename AT, typename P> __device__ static void load(P p, AT (&x)[4])
    {
        const v2u32 a = *vcast<v2u32>(p);
        x[0] = (AT)__uint_as_float(a.x << 16); x[1] = (AT)__uint_as_float(a.x & 0xffff0000u);
        x[2] = (AT)__uint_as_float(a.y << 16); x[3] = (AT)__uint_as_float(a.y & 0xffff0000u);
    }
};
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));
template <> struct VecLoad<bf16_t, 8> {
    static constexpr bool ok = true;
    template <typename AT, typename P> __device__ static void load(P p, AT (&x)[8])
    {
        const v4u32 a = *vcast<v4u32>(p);
        x[0] = (AT)__uint_as_float(a.x << 16); x[1] = (AT)__uint_as_float(a.x & 0xffff0000u);
        x[2] = (AT)__uint_as_float(a.y << 16); x[3] = (AT)__uint_as_float(a.y & 0xffff0000u);
        x[4] = (AT)__uint_as_float(a.z << 16); x[5] = (AT)__uint_as_float(a.z & 0xffff0000u);
        x[6] = (AT)__uint_as_float(a.w << 16); x[7] = (AT)__uint_as_float(a.w & 0xffff0000u);
    }
};
template <> struct VecLoad<bf16_t, 2> {
    static constexpr bool ok = true;
    template <typename AT, typename P> __device__ static void load(P p, AT (&x)[2])
    {
        const uint32_t a = *vcast<uint32_t>(p);
        x[0] = (AT)__uint_as_float(a << 16); x[1] = (AT)__uint_as_float(a & 0xffff0000u);
    }
};

template <typename FT, typename AT, int G, int E, bool CONTIG = false, typename P>
__device__ inline void load_cols(P row, int gl, int R, AT (&x)[E])
{
    if constexpr (CONTIG && VecLoad<FT, E>::ok) {
        const int c0 = gl * E;
        // bf16 rows of odd rank are only 2-byte aligned: element loads there
        if (c0 + E <= R && (sizeof(FT) >= 4 || (R & 1) == 0)) {
            VecLoad<FT, E>::load(row + c0, x);
            return;
        }
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int c = col_of<CONTIG, G, E>(gl, q);
        x[q] = c < R ? elem_acc<AT>(row, c, (FT*)nullptr) : (AT)0;
    }
}

// factor row `off` (element offset) of other mode o: from the LDS copy of the
// hot rows when it is there, else from global memory
template <typename FT, typename AT, int G, int E, bool CONTIG, bool HOT>
__device__ inline void load_row(const FT* __restrict__ U, lds_ptr<FT> hot, uint32_t hot_n, uint32_t off, int gl, int R,
                                AT (&x)[E])
{
    if (HOT && off < hot_n) load_cols<FT, AT, G, E, CONTIG>(hot + off, gl, R, x);
    else load_cols<FT, AT, G, E, CONTIG>(U + off, gl, R, x);
}

template <bool NT, typename T>
__device__ inline T stream_ld(const T* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// Copies the leading hot_n[o] elements (the hot rows, in rank order) of each
// other mode's rank-ordered factor into this workgroup's LDS, 16 bytes per
// lane and step; every thread of the block takes part (called before any
// thread returns).
template <typename FT, typename VT, typename OT>
__device__ inline void stage_hot_rows(const StreamArgs<FT, VT, OT>& a, unsigned char* smem)
{
    for (int o = 0; o < a.nother; ++o) {
        const uint32_t words = (uint32_t)((a.hot_n[o] * sizeof(FT) + 15) / 16);
        const uint4* __restrict__ src = reinterpret_cast<const uint4*>(a.Uo[o]);
        uint4* dst = reinterpret_cast<uint4*>(smem + a.hot_base[o]);
        for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
}

// HOT: the stream holds frequency ranks for the other modes and Uo are the
// rank-ordered factor copies; the hot_n[o] / R most frequent rows of each are
// read from LDS (staged once per workgroup), the rest from global memory.
template <typename FT, typename VT, typename AT, typename OT, int G, int E, int PW, bool NT, bool CONTIG, bool HOT>
__global__ void __launch_bounds__(1024) k_mttkrp(StreamArgs<FT, VT, OT> a)
{
    constexpr int J = 4;  // nonzeros in flight per group
    extern __shared__ __align__(16) unsigned char hot_smem[];
    if constexpr (HOT) stage_hot_rows(a, hot_smem);
    const int lane = threadIdx.x & 63;
    const int gl = lane & (G - 1);
    const uint64_t block = xcd_block(a.nblocks_real);
    if (block == ~0ull) return;
    const uint64_t group = (block * blockDim.x + threadIdx.x) / G;
    const uint64_t s = group * a.seg;
    if (s >= a.nnz) return;
    const uint64_t e = s + a.seg < a.nnz ? s + a.seg : a.nnz;
    const int R = a.rank;
    const int nother = a.nother;

    AT acc[E];
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = 0;
    uint32_t cur = 0xffffffffu;

    for (uint64_t p = s; p < e; p += G) {
        // phase 0: lane gl decodes nonzero p+gl once (coalesced load)
        uint32_t offs[ALTO_MAXN - 1];
        uint32_t rown = 0;
        VT myv = 0;
        {
            uint64_t k[PW];
            if (p + gl < e) {
#pragma unroll
                for (int w = 0; w < PW; ++w) k[w] = stream_ld<NT>(&a.pk[(p + gl) * PW + w]);
                myv = stream_ld<NT>(&a.val[p + gl]);
            } else {
#pragma unroll
                for (int w = 0; w < PW; ++w) k[w] = 0;
            }
#pragma unroll
            for (int o = 0; o < ALTO_MAXN - 1; ++o)
                offs[o] = o < nother ? (uint32_t)(field<PW>(k, a.offo[o], a.bitso[o]) * (uint64_t)R) : 0u;
            rown = (uint32_t)field<PW>(k, a.offn, a.bitsn);
        }
        const int cnt = (int)(e - p < (uint64_t)G ? e - p : (uint64_t)G);
        for (int j0 = 0; j0 < cnt; j0 += J) {
            // phase 1: broadcast J nonzeros, issue all their row gathers
            AT prod[J][E];
            uint32_t rj[J];
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                const int j = j0 + jj < cnt ? j0 + jj : cnt - 1;
                rj[jj] = (uint32_t)__shfl((int)rown, j, G);
                const AT v = (AT)shfl_val(myv, j, G);
#pragma unroll
                for (int q = 0; q < E; ++q) prod[jj][q] = v;
            }
#pragma unroll
            for (int o = 0; o < ALTO_MAXN - 1; ++o) {
                if (o < nother) {
                    const FT* __restrict__ U = a.Uo[o];
                    const lds_ptr<FT> H = (lds_ptr<FT>)(hot_smem + a.hot_base[o]);
#pragma unroll
                    for (int jj = 0; jj < J; ++jj) {
                        const int j = j0 + jj < cnt ? j0 + jj : cnt - 1;
                        const uint32_t off = (uint32_t)__shfl((int)offs[o], j, G);
                        AT x[E];
                        load_row<FT, AT, G, E, CONTIG, HOT>(U, H, a.hot_n[o], off, gl, R, x);
#pragma unroll
                        for (int q = 0; q < E; ++q) prod[jj][q] *= x[q];
                    }
                }
            }
            // phase 2: register reduction while the output row repeats
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                if (j0 + jj < cnt) {
                    if (rj[jj] != cur) {
                        if (cur != 0xffffffffu) flush_row<AT, OT, G, E, CONTIG>(a.out, cur, R, gl, acc);
                        cur = rj[jj];
#pragma unroll
                        for (int q = 0; q < E; ++q) acc[q] = 0;
                    }
#pragma unroll
                    for (int q = 0; q < E; ++q) acc[q] += prod[jj][q];
                }
            }
        }
    }
    if (cur != 0xffffffffu) flush_row<AT, OT, G, E, CONTIG>(a.out, cur, R, gl, acc);
}

// ------------------------------------------------ 16-byte lanes, VALU broadcasts
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// value of lane L (a constant) of each G-lane group, through VALU paths:
// DPP quad_perm (G = 2, 4), DPP row_newbcast (G = 8, 16: lane L of each
// 16-lane row), readlane (G = 32, 64); G = 1 is the lane itself
template <int G, int L>
__device__ __forceinline__ uint32_t bcast32(uint32_t v)
{
    static_assert(L >= 0 && L < G, "broadcast lane outside the group");
    if constexpr (G == 64) {
        return __builtin_amdgcn_readlane(v, L);
    } else if constexpr (G == 32) {
        const uint32_t lo = __builtin_amdgcn_readlane(v, L), hi = __builtin_amdgcn_readlane(v, 32 + L);
        return (threadIdx.x & 32) ? hi : lo;
    } else if constexpr (G == 16) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x150 + L, 0xf, 0xf, false);
    } else if constexpr (G == 8) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x150 + L, 0xf, 0xf, false);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x150 + 8 + L, 0xf, 0xf, false);
        return (threadIdx.x & 8) ? hi : lo;
    } else if constexpr (G == 4) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, L | (L << 2) | (L << 4) | (L << 6), 0xf, 0xf, false);
    } else if constexpr (G == 2) {  // quad_perm [L, L, 2+L, 2+L]
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, L | (L << 2) | ((L + 2) << 4) | ((L + 2) << 6), 0xf, 0xf,
                                                     false);
    } else {
        static_assert(G == 1, "group of 1..64 lanes");
        return v;
    }
}
template <int G, int L> __device__ __forceinline__ double bcast_val(double v)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint64_t r = ((uint64_t)bcast32<G, L>((uint32_t)(u >> 32)) << 32) | bcast32<G, L>((uint32_t)u);
    return __longlong_as_double((long long)r);
}
template <int G, int L> __device__ __forceinline__ float bcast_val(float v)
{
    return __uint_as_float(bcast32<G, L>(__float_as_uint(v)));
}

// E = 16 / sizeof(FT) consecutive columns per lane in one 16-byte load; the
// host guarantees 16-byte aligned rows (R * sizeof(FT) % 16 == 0), so a lane
// either loads a whole vector or lies beyond the rank
template <typename FT, typename AT, int E, typename P>
__device__ __forceinline__ void load16(P row, int gl, int R, AT (&x)[E])
{
    // branch-free: lanes beyond the rank read the row's first vector and
    // drop it
    const int c0 = gl * E;
    const bool in = c0 < R;
    VecLoad<FT, E>::load(row + (in ? c0 : 0), x);
#pragma unroll
    for (int q = 0; q < E; ++q) x[q] = in ? x[q] : (AT)0;
}

// The MTTKRP stream kernel with 16-byte lanes: per batch of G nonzeros each
// lane decodes one, then the group walks them J at a time; the decoded
// fields reach the group through VALU broadcasts (no LDS permutes) and each
// factor row arrives with one 16-byte load per lane.  The next batch's
// nonzeros are loaded while the current batch's rows are gathered.
// NO = number of other modes (a template constant, so that the gathers of
// all modes are issued back to back with no branch between them).  HM = hot
// mask of a hot-pattern group of a split stream: for other mode o with bit
// o set every nonzero's rank is below hot_n[o] / R, and the row is read from
// the LDS copy of those rows (staged once per workgroup).
// RTNO: NO is an upper bound and modes o >= a.nother are skipped at run time
// (the multi-tensor launch, whose tensors differ in their number of modes).
template <typename FT, typename VT, typename AT, typename OT, int G, int E, int PW, int NO, int HM, bool RTNO>
__device__ __forceinline__ void mttkrp16_body(const StreamArgs<FT, VT, OT>& a, uint64_t block,
                                              unsigned char* hot_smem)
{
    // a batch is B = max(G, 4) nonzeros: each lane decodes NS = B / G of them
    // (slot k of lane l holds nonzero p + k*G + l); J = 4 of them are in
    // flight per group at a time
    constexpr int NS = G < 4 ? 4 / G : 1;
    constexpr int B = G * NS;
    constexpr int J = 4;
    const int lane = threadIdx.x & 63;
    const int gl = lane & (G - 1);
    const uint64_t group = (block * 256 + threadIdx.x) / G;
    const uint64_t s = group * a.seg;
    if (s >= a.nnz) return;
    const uint64_t e = s + a.seg < a.nnz ? s + a.seg : a.nnz;
    const int R = a.rank;

    AT acc[E];
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = 0;
    uint32_t cur = 0xffffffffu;

    // this lane's nonzeros of the next batch (past the segment: row 0 of
    // every mode, value 0 -- valid memory, and never accumulated)
    uint64_t kn[NS][PW];
    VT vn[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const uint64_t i = s + k * G + gl;
        vn[k] = 0;
#pragma unroll
        for (int w = 0; w < PW; ++w) kn[k][w] = 0;
        if (i < e) {
#pragma unroll
            for (int w = 0; w < PW; ++w) kn[k][w] = a.pk[i * PW + w];
            vn[k] = a.val[i];
        }
    }

    for (uint64_t p = s; p < e; p += B) {
        uint32_t offs[NS][NO > 0 ? NO : 1];
        uint32_t rown[NS];
        VT myv[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            myv[k] = vn[k];
#pragma unroll
            for (int o = 0; o < NO; ++o) offs[k][o] = (uint32_t)(field<PW>(kn[k], a.offo[o], a.bitso[o]) * (uint64_t)R);
            rown[k] = (uint32_t)field<PW>(kn[k], a.offn, a.bitsn);
        }
        const int cnt = (int)(e - p < (uint64_t)B ? e - p : (uint64_t)B);
        static_for<0, B / J>([&](auto bi) {
            constexpr int j0 = decltype(bi)::value * J;
            if (j0 >= cnt) return;
            AT prod[J][E];
            uint32_t rj[J];
            static_for<0, J>([&](auto ji) {
                constexpr int jj = decltype(ji)::value;
                constexpr int slot = (j0 + jj) / G, src = (j0 + jj) % G;
                rj[jj] = bcast32<G, src>(rown[slot]);
                const AT v = (AT)bcast_val<G, src>(myv[slot]);
#pragma unroll
                for (int q = 0; q < E; ++q) prod[jj][q] = v;
            });
            static_for<0, NO>([&](auto oi) {
                constexpr int o = decltype(oi)::value;
                if constexpr (RTNO)
                    if (o >= a.nother) return;  // uniform
                const FT* __restrict__ U = a.Uo[o];
                const lds_ptr<FT> H = (lds_ptr<FT>)(hot_smem + a.hot_base[o]);
                static_for<0, J>([&](auto ji) {
                    constexpr int jj = decltype(ji)::value;
                    constexpr int slot = (j0 + jj) / G, src = (j0 + jj) % G;
                    const uint32_t off = bcast32<G, src>(offs[slot][o]);
                    AT x[E];
                    if constexpr ((HM >> o) & 1) load16<FT, AT, E>(H + off, gl, R, x);
                    else load16<FT, AT, E>(U + off, gl, R, x);
#pragma unroll
                    for (int q = 0; q < E; ++q) prod[jj][q] *= x[q];
                });
            });
            if constexpr (j0 == 0) {
                // the next batch's nonzeros, issued behind this block's gathers
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    const uint64_t i = p + B + k * G + gl;
                    vn[k] = 0;
#pragma unroll
                    for (int w = 0; w < PW; ++w) kn[k][w] = 0;
                    if (i < e) {
#pragma unroll
                        for (int w = 0; w < PW; ++w) kn[k][w] = a.pk[i * PW + w];
                        vn[k] = a.val[i];
                    }
                }
            }
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                if (j0 + jj < cnt) {
                    if (rj[jj] != cur) {
                        if (cur != 0xffffffffu) flush_row<AT, OT, G, E, true>(a.out, cur, R, gl, acc);
                        cur = rj[jj];
#pragma unroll
                        for (int q = 0; q < E; ++q) acc[q] = 0;
                    }
#pragma unroll
                    for (int q = 0; q < E; ++q) acc[q] += prod[jj][q];
                }
            }
        });
    }
    if (cur != 0xffffffffu) flush_row<AT, OT, G, E, true>(a.out, cur, R, gl, acc);
}

template <typename FT, typename VT, typename AT, typename OT, int G, int E, int PW, int NO, int HM>
__global__ void __launch_bounds__(256) k_mttkrp16(StreamArgs<FT, VT, OT> a)
{
    extern __shared__ __align__(16) unsigned char hot_smem[];
    if constexpr (HM != 0) stage_hot_rows(a, hot_smem);
    const uint64_t block = xcd_block(a.nblocks_real);
    if (block == ~0ull) return;
    mttkrp16_body<FT, VT, AT, OT, G, E, PW, NO, HM, false>(a, block, hot_smem);
}

// MTTKRP over a fiber-ordered stream (nonzeros ordered by (row, row of the
// fiber mode a)): as k_mttkrp16, but the fiber-mode row U_a[row_a] is read
// only where a fiber starts.  In each block of J nonzeros of a group, a
// position whose fiber differs from the previous position's (for the first
// position: from the previous block's last) is a fiber head; the wave loads
// U_a rows at the positions where any of its groups has a head -- a
// wave-uniform branch, so every lane loads (groups without a head reload
// the row they already hold) and nothing is merged into live registers
// between the issues -- and every other position reuses the row of the
// last loaded one.  On long fibers most blocks need no U_a load: a
// 3-mode MTTKRP then gathers one row per nonzero instead of two.  The other
// NB = N - 2 modes' rows are gathered per nonzero; products are summed per
// output row and flushed as in k_mttkrp16.
template <typename FT, typename VT, typename AT, typename OT, int G, int E, int PW, int NB>
__global__ void __launch_bounds__(256) k_mttkrp_fib(StreamArgs<FT, VT, OT> a, const FT* __restrict__ Ua, int offa,
                                                   int bitsa)
{
    constexpr int NS = G < 4 ? 4 / G : 1;
    constexpr int B = G * NS;
    constexpr int J = 4;
    const uint64_t block = xcd_block(a.nblocks_real);
    if (block == ~0ull) return;
    const int lane = threadIdx.x & 63;
    const int gl = lane & (G - 1);
    const uint64_t group = (block * 256 + threadIdx.x) / G;
    const uint64_t s = group * a.seg;
    if (s >= a.nnz) return;
    const uint64_t e = s + a.seg < a.nnz ? s + a.seg : a.nnz;
    const int R = a.rank;

    AT acc[E], ua_last[E];
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = ua_last[q] = 0;
    uint32_t cur = 0xffffffffu;
    uint32_t prev_a = 0;  // fiber row offset of the previous position
    bool have_a = false;  // ua_last holds U_a at prev_a

    // this lane's nonzeros of the next batch (past the segment: row 0 of
    // every mode, value 0 -- valid memory, and never accumulated)
    uint64_t kn[NS][PW];
    VT vn[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const uint64_t i = s + k * G + gl;
        vn[k] = 0;
#pragma unroll
        for (int w = 0; w < PW; ++w) kn[k][w] = 0;
        if (i < e) {
#pragma unroll
            for (int w = 0; w < PW; ++w) kn[k][w] = a.pk[i * PW + w];
            vn[k] = a.val[i];
        }
    }

    for (uint64_t p = s; p < e; p += B) {
        uint32_t offs[NS][NB];
        uint32_t aoff[NS], rown[NS];
        VT myv[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            myv[k] = vn[k];
#pragma unroll
            for (int o = 0; o < NB; ++o) offs[k][o] = (uint32_t)(field<PW>(kn[k], a.offo[o], a.bitso[o]) * (uint64_t)R);
            aoff[k] = (uint32_t)(field<PW>(kn[k], offa, bitsa) * (uint64_t)R);
            rown[k] = (uint32_t)field<PW>(kn[k], a.offn, a.bitsn);
        }
        const int cnt = (int)(e - p < (uint64_t)B ? e - p : (uint64_t)B);
        static_for<0, B / J>([&](auto bi) {
            constexpr int j0 = decltype(bi)::value * J;
            if (j0 >= cnt) return;
            AT prod[J][E];
            uint32_t rj[J], aj[J];
            static_for<0, J>([&](auto ji) {
                constexpr int jj = decltype(ji)::value;
                constexpr int slot = (j0 + jj) / G, src = (j0 + jj) % G;
                rj[jj] = bcast32<G, src>(rown[slot]);
                aj[jj] = bcast32<G, src>(aoff[slot]);
                const AT v = (AT)bcast_val<G, src>(myv[slot]);
#pragma unroll
                for (int q = 0; q < E; ++q) prod[jj][q] = v;
            });
            // issue: the other modes' rows of the J nonzeros, then U_a rows
            // at the fiber heads
            static_for<0, NB>([&](auto oi) {
                constexpr int o = decltype(oi)::value;
                const FT* __restrict__ U = a.Uo[o];
                static_for<0, J>([&](auto ji) {
                    constexpr int jj = decltype(ji)::value;
                    constexpr int slot = (j0 + jj) / G, src = (j0 + jj) % G;
                    const uint32_t off = bcast32<G, src>(offs[slot][o]);
                    AT x[E];
                    load16<FT, AT, E>(U + off, gl, R, x);
#pragma unroll
                    for (int q = 0; q < E; ++q) prod[jj][q] *= x[q];
                });
            });
            // lanes beyond the rank load the row's first vector and keep it:
            // they never flush, so no select follows the load
            AT ra[J][E];
            uint32_t ldm = 0;  // positions loaded (uniform)
            const int c0 = gl * E < R ? gl * E : 0;
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                const bool head = jj == 0 ? (!have_a || aj[0] != prev_a) : aj[jj] != aj[jj - 1];
                if (__any(head)) {
                    VecLoad<FT, E>::load(Ua + aj[jj] + c0, ra[jj]);
                    ldm |= 1u << jj;
                }
            }
            // the mask reaches the selection below through an opaque copy, so
            // that the selection is not folded into the branches above (which
            // would chain every load to the next one's registers)
            asm volatile("" : "+s"(ldm));
            if constexpr (j0 == 0) {
                // the next batch's nonzeros, issued behind this block's gathers
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    const uint64_t i = p + B + k * G + gl;
                    vn[k] = 0;
#pragma unroll
                    for (int w = 0; w < PW; ++w) kn[k][w] = 0;
                    if (i < e) {
#pragma unroll
                        for (int w = 0; w < PW; ++w) kn[k][w] = a.pk[i * PW + w];
                        vn[k] = a.val[i];
                    }
                }
            }
            // math: each position's U_a row is that of the last load at or
            // before it
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                const bool use = (ldm >> jj) & 1u;
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    ua_last[q] = use ? ra[jj][q] : ua_last[q];
                    prod[jj][q] *= ua_last[q];
                }
            }
            prev_a = aj[J - 1];
            have_a = true;
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                if (j0 + jj < cnt) {
                    if (rj[jj] != cur) {
                        if (cur != 0xffffffffu) flush_row<AT, OT, G, E, true>(a.out, cur, R, gl, acc);
                        cur = rj[jj];
#pragma unroll
                        for (int q = 0; q < E; ++q) acc[q] = 0;
                    }
#pragma unroll
                    for (int q = 0; q < E; ++q) acc[q] += prod[jj][q];
                }
            }
        });
    }
    if (cur != 0xffffffffu) flush_row<AT, OT, G, E, true>(a.out, cur, R, gl, acc);
}

// Many MTTKRPs (different tensors, modes, sizes; one rank and lane shape) in
// one launch: the table travels in the kernel arguments, and job j owns the
// logical workgroups [start[j], start[j+1]) -- its own launch's grid, laid
// end to end with the other jobs' and mapped XCD-aware as a whole.
constexpr int MTTKRP_BATCH_MAX = 16;
template <typename FT, typename VT, typename OT>
struct BatchTable {
    StreamArgs<FT, VT, OT> a[MTTKRP_BATCH_MAX];
    uint32_t start[MTTKRP_BATCH_MAX + 1];
    int n;
};

template <typename FT, typename VT, typename AT, typename OT, int G, int E, int PW>
__global__ void __launch_bounds__(256) k_mttkrp16_batch(BatchTable<FT, VT, OT> tab)
{
    extern __shared__ __align__(16) unsigned char hot_smem[];
    const uint64_t block = xcd_block(tab.start[tab.n]);
    if (block == ~0ull) return;
    int j = 0;
    while (j + 1 < tab.n && block >= tab.start[j + 1]) ++j;  // uniform
    mttkrp16_body<FT, VT, AT, OT, G, E, PW, ALTO_MAXN - 1, 0, true>(tab.a[j], block - tab.start[j], hot_smem);
}

// the vector-lane kernel needs factor rows aligned to the lane's load width
template <typename FT, typename VT, typename OT>
inline bool rows_aligned(const StreamArgs<FT, VT, OT>& a, int R, int bytes)
{
    if (((uint64_t)R * sizeof(FT)) % bytes != 0) return false;
    for (int o = 0; o < a.nother; ++o)
        if ((uintptr_t)a.Uo[o] % bytes != 0) return false;
    return true;
}

// The fiber parts of a fiber-ordered stream through k_mttkrp_fib (G lanes
// per nonzero, E columns per lane): [0, nlong) with fiber mode fib_mode,
// then [nlong, nlong + nlong2) with fib_mode2; `base` carries the stream,
// value, output and target-mode fields.
template <typename FT, typename VT, typename AT, typename OT, int E>
int launch_fiber_tier(const alto_tensor* t, int mode, int fmode, uint64_t first, uint64_t count, int R,
                      const FT* const* U, const StreamArgs<FT, VT, OT>& base, int G, int seg, hipStream_t s)
{
    StreamArgs<FT, VT, OT> f = base;
    int o = 0;
    for (int m = 0; m < t->nmode; ++m) {
        if (m == mode || m == fmode) continue;
        f.Uo[o] = U[m];
        f.offo[o] = t->pack.off[m];
        f.bitso[o] = t->pack.bits[m];
        ++o;
    }
    for (int q = o; q < ALTO_MAXN - 1; ++q) {
        f.Uo[q] = nullptr;
        f.offo[q] = f.bitso[q] = 0;
    }
    f.nother = o;
    f.pk = base.pk + first * t->pack.pw;
    f.val = base.val + first;
    f.nnz = count;
    f.seg = (uint64_t)seg;
    const dim3 grid = stream_grid(f.nnz, seg, G, f);
    const FT* Ua = U[fmode];
    const int offa = t->pack.off[fmode], bitsa = t->pack.bits[fmode];
#define KF_(GG, PWW, NBB) \
    hipLaunchKernelGGL((k_mttkrp_fib<FT, VT, AT, OT, GG, E, PWW, NBB>), grid, dim3(256), 0, s, f, Ua, offa, bitsa)
#define KFN_(GG, PWW)                          \
    do {                                       \
        if (o == 1) KF_(GG, PWW, 1);           \
        else if (o == 2) KF_(GG, PWW, 2);      \
        else KF_(GG, PWW, 3);                  \
    } while (0)
#define KFG_(GG)                               \
    do {                                       \
        if (G == GG) {                         \
            if (t->pack.pw == 1) KFN_(GG, 1);  \
            else KFN_(GG, 2);                  \
        }                                      \
    } while (0)
    KFG_(1);
    KFG_(2);
    KFG_(4);
    KFG_(8);
    KFG_(16);
    KFG_(32);
    KFG_(64);
#undef KFG_
#undef KFN_
#undef KF_
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

template <typename FT, typename VT, typename AT, typename OT, int E>
int launch_fiber_part(const alto_tensor* t, int mode, int R, const FT* const* U, const StreamArgs<FT, VT, OT>& base,
                      int G, int seg, hipStream_t s)
{
    const ModeStream& M = t->ms[mode];
    if (M.nlong)
        ALTO_TRY((launch_fiber_tier<FT, VT, AT, OT, E>(t, mode, M.fib_mode, 0, M.nlong, R, U, base, G, seg, s)));
    if (M.nlong2)
        ALTO_TRY((launch_fiber_tier<FT, VT, AT, OT, E>(t, mode, M.fib_mode2, M.nlong, M.nlong2, R, U, base, G, seg,
                                                       s)));
    return ALTO_OK;
}

// launch_fiber_part instantiated per element type and columns per lane in
// mttkrp_fib_{f64,f32,bf16}.hip (translation units of their own, built in
// parallel with the others)
int fiber_part(const alto_tensor* t, int mode, int R, const double* const* U,
               const StreamArgs<double, double, double>& a, int G, int E, int seg, hipStream_t s);
int fiber_part(const alto_tensor* t, int mode, int R, const float* const* U, const StreamArgs<float, float, double>& a,
               int G, int E, int seg, hipStream_t s);
int fiber_part(const alto_tensor* t, int mode, int R, const bf16_t* const* U,
               const StreamArgs<bf16_t, float, double>& a, int G, int E, int seg, hipStream_t s);

// the stream arguments narrowed to the part after the fiber parts
template <typename FT, typename VT, typename OT>
inline void skip_long_part(const alto_tensor* t, uint64_t nlong, StreamArgs<FT, VT, OT>& a)
{
    a.pk += nlong * t->pack.pw;
    a.val += nlong;
    a.nnz -= nlong;
}

// Lane mappings: 0 = 2 strided columns per lane (each gather instruction
// reads G consecutive elements of a row); 1 = 4 contiguous columns per lane;
// 2 = contiguous columns per lane in one vector load (16 bytes; 8 for BF16),
// broadcasts through DPP / readlane.  -1 = 2 (measured fastest on MI355X for
// every element type).  Ranked streams without the hot-pattern split use
// mapping 0 or 1.
template <typename FT, typename VT, typename AT, typename OT>
int launch_mttkrp(const alto_tensor* t, int mode, int R, const FT* const* U, const VT* val, OT* out, int seg,
                  const MttkrpTune& k, hipStream_t s)
{
    StreamArgs<FT, VT, OT> a;
    fill_stream_args(t, mode, R, seg, a);
    ALTO_TRY(set_factors(t, mode, U, a, s));
    a.val = val;
    a.out = out;
    int v = k.variant >= 0 ? k.variant : 2;
    const bool hot = t->ms[mode].ranked != 0 && a.nother > 0;
    // elements per lane of the vector-lane mapping: one 16-byte load (FP64,
    // FP32), one 8-byte load for BF16 (four columns: eight would need ~110
    // VGPRs and halve the occupancy)
    constexpr int E16 = std::is_same<FT, bf16_t>::value ? 4 : 16 / sizeof(FT);
    const ModeStream& M = t->ms[mode];
    const bool split = M.split && hot && a.nother <= 2;
    if (v == 2 && ((hot && !split) || !rows_aligned(a, R, E16 * (int)sizeof(FT)) || R > 64 * E16 || a.nother < 1)) v = 0;
    // FP32 / BF16 rows of rank <= 16: two columns per lane, so that a wave
    // holds at most 16 groups (with 4 columns per lane it would hold up to
    // 32, and each row change of any group costs the wave a round of
    // atomics); measured 1.3x faster at ranks 8 and 16
    // Below 4M nonzeros (flushes dominate, launches are latency-bound) every
    // rank takes two columns per lane, as FP64 does with its 16-byte lanes:
    // 4 groups per wave at rank 64 instead of 2 made FP32 ~40% slower than
    // FP64 there.
    constexpr bool NARROW_OK = !std::is_same<FT, double>::value;
    const bool narrow_rank = R <= k.narrow_max || (k.narrow_small && t->nnz < (4ull << 20) && R <= 128);
    const uint64_t nfib = M.nlong + M.nlong2;
    const bool fib = M.fib_mode >= 0 && nfib > 0 && !hot && a.nother >= 2;
    if (v == 2 && NARROW_OK && narrow_rank && !split && a.nother >= 1 && rows_aligned(a, R, 2 * (int)sizeof(FT))) {
        int G = 1;
        while (G * 2 < R) G <<= 1;
        if (seg <= 0) seg = tuned_seg(t->nnz, G, (int)sizeof(AT));
        a.seg = (uint64_t)seg;
        if (fib) {
            ALTO_TRY(fiber_part(t, mode, R, U, a, G, 2, seg, s));
            skip_long_part(t, nfib, a);
            if (a.nnz == 0) return ALTO_OK;
        }
        dim3 grid = stream_grid(a.nnz, seg, G, a);
#define KN_(GG, PWW)                                                                                            \
    do {                                                                                                        \
        switch (a.nother) {                                                                                     \
        case 1: hipLaunchKernelGGL((k_mttkrp16<FT, VT, AT, OT, GG, 2, PWW, 1, 0>), grid, dim3(256), 0, s, a); break; \
        case 2: hipLaunchKernelGGL((k_mttkrp16<FT, VT, AT, OT, GG, 2, PWW, 2, 0>), grid, dim3(256), 0, s, a); break; \
        case 3: hipLaunchKernelGGL((k_mttkrp16<FT, VT, AT, OT, GG, 2, PWW, 3, 0>), grid, dim3(256), 0, s, a); break; \
        default: hipLaunchKernelGGL((k_mttkrp16<FT, VT, AT, OT, GG, 2, PWW, 4, 0>), grid, dim3(256), 0, s, a); break; \
        }                                                                                                       \
    } while (0)
#define KNG_(GG)                                 \
    do {                                         \
        if (G == GG) {                           \
            if (t->pack.pw == 1) KN_(GG, 1);     \
            else KN_(GG, 2);                     \
        }                                        \
    } while (0)
        KNG_(1);
        KNG_(2);
        KNG_(4);
        KNG_(8);
        KNG_(16);
        KNG_(32);
        KNG_(64);
#undef KNG_
#undef KN_
        HIP_TRY(hipGetLastError());
        return ALTO_OK;
    }
    if (v == 2) {
        int G = 1;
        while (G * E16 < R) G <<= 1;
        if (seg <= 0) seg = tuned_seg(t->nnz, G, (int)sizeof(AT));
        a.seg = (uint64_t)seg;
        if (fib && !split) {
            ALTO_TRY(fiber_part(t, mode, R, U, a, G, E16, seg, s));
            skip_long_part(t, nfib, a);
            if (a.nnz == 0) return ALTO_OK;
        }
        // one launch per hot-pattern group (one launch over the whole stream
        // when it is not split)
        const int npat = split ? 1 << a.nother : 1;
        for (int p = 0; p < npat; ++p) {
            StreamArgs<FT, VT, OT> b = a;
            uint32_t lds = 0;
            if (split) {
                const uint64_t s0 = M.pat_start[p], cnt = M.pat_start[p + 1] - s0;
                if (cnt == 0) continue;
                b.pk = a.pk + s0 * t->pack.pw;
                b.val = a.val + s0;
                b.nnz = cnt;
                for (int o = 0; o < a.nother; ++o) {
                    b.hot_n[o] = (p >> o) & 1 ? (uint32_t)(M.hot_k[o] * R) : 0u;
                    b.hot_base[o] = lds;
                    lds += (uint32_t)((b.hot_n[o] * sizeof(FT) + 15) / 16 * 16);
                }
            }
            if (lds > 160u * 1024u) {
                set_error("mttkrp: %u bytes of staged rows exceed the 160 KiB of LDS (rank %d)", lds, R);
                return ALTO_ERR_UNSUPPORTED;
            }
            dim3 grid = stream_grid(b.nnz, seg, G, b);
#define K16L_(GG, PWW, NOO, HMM) \
    hipLaunchKernelGGL((k_mttkrp16<FT, VT, AT, OT, GG, E16, PWW, NOO, HMM>), grid, dim3(256), lds, s, b)
#define K16N_(GG, PWW)                                     \
    do {                                                   \
        switch (b.nother * 16 + p) {                       \
        case 16: K16L_(GG, PWW, 1, 0); break;              \
        case 17: K16L_(GG, PWW, 1, 1); break;              \
        case 32: K16L_(GG, PWW, 2, 0); break;              \
        case 33: K16L_(GG, PWW, 2, 1); break;              \
        case 34: K16L_(GG, PWW, 2, 2); break;              \
        case 35: K16L_(GG, PWW, 2, 3); break;              \
        case 48: K16L_(GG, PWW, 3, 0); break;              \
        default: K16L_(GG, PWW, 4, 0); break;              \
        }                                                  \
    } while (0)
#define K16_(GG)                                 \
    do {                                         \
        if (G == GG) {                           \
            if (t->pack.pw == 1) K16N_(GG, 1);   \
            else K16N_(GG, 2);                   \
        }                                        \
    } while (0)
            K16_(1);
            K16_(2);
            K16_(4);
            K16_(8);
            K16_(16);
            K16_(32);
            K16_(64);
#undef K16_
#undef K16N_
#undef K16L_
            HIP_TRY(hipGetLastError());
        }
        return ALTO_OK;
    }
    const int E = v == 1 ? 4 : 2;
    int G = 4;
    while (G * E < R && G < 64) G <<= 1;
    if (seg <= 0) seg = tuned_seg(t->nnz, G, (int)sizeof(AT));
    a.seg = (uint64_t)seg;
    const int bt = hot ? k.hot_block : 256;
    const uint32_t lds = hot ? plan_hot_rows(t, mode, a, (uint32_t)k.hot_lds) : 0;
    dim3 grid = stream_grid(t->nnz, seg, G, a, bt);
#define K_(PWW, NTT, HOTT) \
    hipLaunchKernelGGL((k_mttkrp<FT, VT, AT, OT, GG, EE, PWW, NTT, CC, HOTT>), grid, dim3(bt), lds, s, a)
#define L_(GG_, EE_, CC_)                            \
    do {                                             \
        constexpr int GG = GG_, EE = EE_;            \
        constexpr bool CC = CC_;                     \
        if (t->pack.pw == 1) {                       \
            if (hot) K_(1, false, true);             \
            else if (k.nt) K_(1, true, false);       \
            else K_(1, false, false);                \
        } else {                                     \
            if (hot) K_(2, false, true);             \
            else K_(2, false, false);                \
        }                                            \
    } while (0)
#define LG_(EE, CC)                         \
    do {                                    \
        switch (G) {                        \
        case 4: L_(4, EE, CC); break;       \
        case 8: L_(8, EE, CC); break;       \
        case 16: L_(16, EE, CC); break;     \
        case 32: L_(32, EE, CC); break;     \
        default: L_(64, EE, CC); break;     \
        }                                   \
    } while (0)
    if (v == 1) LG_(4, true);
    else LG_(2, false);
#undef LG_
#undef L_
#undef K_
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

// lanes per nonzero (G) and elements per lane (E) of the vector-lane kernel
// for this tensor and mode, as launch_mttkrp picks them; false when the
// launch would take another kernel (ranked streams, unaligned rows, other
// mappings)
template <typename FT, typename VT, typename OT>
inline bool vector_lane_shape(const alto_tensor* t, int mode, int R, const StreamArgs<FT, VT, OT>& a,
                              const MttkrpTune& k, int& G, int& E)
{
    constexpr int E16 = std::is_same<FT, bf16_t>::value ? 4 : 16 / sizeof(FT);
    constexpr bool NARROW_OK = !std::is_same<FT, double>::value;
    if (!(k.variant < 0 || k.variant == 2) || a.nother < 1 || t->ms[mode].ranked != 0) return false;
    const bool narrow_rank = R <= k.narrow_max || (k.narrow_small && t->nnz < (4ull << 20) && R <= 128);
    if (NARROW_OK && narrow_rank && rows_aligned(a, R, 2 * (int)sizeof(FT))) E = 2;
    else if (rows_aligned(a, R, E16 * (int)sizeof(FT)) && R <= 64 * E16) E = E16;
    else return false;
    G = 1;
    while (G * E < R) G <<= 1;
    return true;
}

template <typename FT, typename VT, typename AT, typename OT, int E>
inline void launch_batch_table(const BatchTable<FT, VT, OT>& tab, int G, int pw, dim3 grid, hipStream_t s)
{
#define KB_(GG)                                                                                              \
    do {                                                                                                     \
        if (pw == 1) hipLaunchKernelGGL((k_mttkrp16_batch<FT, VT, AT, OT, GG, E, 1>), grid, dim3(256), 0, s, tab); \
        else hipLaunchKernelGGL((k_mttkrp16_batch<FT, VT, AT, OT, GG, E, 2>), grid, dim3(256), 0, s, tab);        \
    } while (0)
    switch (G) {
    case 2: KB_(2); break;
    case 4: KB_(4); break;
    case 8: KB_(8); break;
    case 16: KB_(16); break;
    case 32: KB_(32); break;
    default: KB_(64); break;
    }
#undef KB_
}

// Job i: the MTTKRP of ts[i] along modes[i] with factors U[i] and values
// vals[i] into outs[i] (cleared or accumulated by the caller).  Jobs of one
// lane shape (G in 2..64, E, key width) share launches of up to
// MTTKRP_BATCH_MAX jobs; the others are launched one by one through
// launch_mttkrp.  Segment lengths: `seg` when positive, else the size rule
// of each tensor.
template <typename FT, typename VT, typename AT, typename OT>
int launch_mttkrp_batch(int n, alto_tensor* const* ts, const int* modes, int R, const FT* const* const* U,
                        const VT* const* vals, OT* const* outs, int seg, const MttkrpTune& k, hipStream_t s)
{
    constexpr int E16 = std::is_same<FT, bf16_t>::value ? 4 : 16 / sizeof(FT);
    struct Group {
        int G, E, pw;
        std::vector<StreamArgs<FT, VT, OT>> jobs;
    };
    std::vector<Group> groups;
    for (int i = 0; i < n; ++i) {
        const alto_tensor* t = ts[i];
        if (t->ms[modes[i]].ranked) {  // rank-ordered rows: the hot-row kernels
            ALTO_TRY((launch_mttkrp<FT, VT, AT, OT>(t, modes[i], R, U[i], vals[i], outs[i], seg, k, s)));
            continue;
        }
        StreamArgs<FT, VT, OT> a;
        fill_stream_args(t, modes[i], R, 0, a);
        ALTO_TRY(set_factors(t, modes[i], U[i], a, s));
        a.val = vals[i];
        a.out = outs[i];
        int G = 0, E = 0;
        if (!vector_lane_shape(t, modes[i], R, a, k, G, E) || G < 2 || G > 64) {
            ALTO_TRY((launch_mttkrp<FT, VT, AT, OT>(t, modes[i], R, U[i], vals[i], outs[i], seg, k, s)));
            continue;
        }
        a.seg = (uint64_t)(seg > 0 ? seg : tuned_seg(t->nnz, G, (int)sizeof(AT)));
        (void)stream_grid(t->nnz, (int)a.seg, G, a);
        Group* g = nullptr;
        for (auto& x : groups)
            if (x.G == G && x.E == E && x.pw == t->pack.pw) g = &x;
        if (!g) {
            groups.push_back(Group{G, E, t->pack.pw, {}});
            g = &groups.back();
        }
        g->jobs.push_back(a);
    }
    for (const auto& g : groups) {
        for (size_t j0 = 0; j0 < g.jobs.size(); j0 += MTTKRP_BATCH_MAX) {
            BatchTable<FT, VT, OT> tab;
            tab.n = (int)std::min<size_t>(MTTKRP_BATCH_MAX, g.jobs.size() - j0);
            uint64_t total = 0;
            for (int j = 0; j < MTTKRP_BATCH_MAX; ++j) {
                if (j < tab.n) tab.a[j] = g.jobs[j0 + j];
                else tab.a[j] = g.jobs[j0];  // unused entries
                tab.start[j] = (uint32_t)total;
                if (j < tab.n) total += g.jobs[j0 + j].nblocks_real;
            }
            tab.start[MTTKRP_BATCH_MAX] = (uint32_t)total;
            for (int j = tab.n; j <= MTTKRP_BATCH_MAX; ++j) tab.start[j] = (uint32_t)total;
            if (total >= (1ull << 31)) {
                set_error("mttkrp_batched: %llu workgroups in one launch", (unsigned long long)total);
                return ALTO_ERR_UNSUPPORTED;
            }
            const dim3 grid((unsigned)((total + 7) / 8 * 8));
            if (g.E == 2) launch_batch_table<FT, VT, AT, OT, 2>(tab, g.G, g.pw, grid, s);
            else if constexpr (E16 != 2) launch_batch_table<FT, VT, AT, OT, E16>(tab, g.G, g.pw, grid, s);
            HIP_TRY(hipGetLastError());
        }
    }
    return ALTO_OK;
}

int launch_mttkrp_f64(const alto_tensor* t, int mode, int R, const double* const* U, const double* val, double* out,
                      int seg, const MttkrpTune& k, hipStream_t s);
int launch_mttkrp_f32(const alto_tensor* t, int mode, int R, const float* const* U, const float* val, double* out,
                      int seg, const MttkrpTune& k, hipStream_t s);
int launch_mttkrp_bf16(const alto_tensor* t, int mode, int R, const bf16_t* const* U, const float* val, double* out,
                       int seg, const MttkrpTune& k, hipStream_t s);
int launch_mttkrp_batch_f64(int n, alto_tensor* const* ts, const int* modes, int R, const double* const* const* U,
                            const double* const* vals, double* const* outs, int seg, const MttkrpTune& k, hipStream_t s);
int launch_mttkrp_batch_f32(int n, alto_tensor* const* ts, const int* modes, int R, const float* const* const* U,
                            const float* const* vals, double* const* outs, int seg, const MttkrpTune& k, hipStream_t s);
int launch_mttkrp_batch_bf16(int n, alto_tensor* const* ts, const int* modes, int R, const bf16_t* const* const* U,
                            const float* const* vals, double* const* outs, int seg, const MttkrpTune& k, hipStream_t s);

}  // namespace alto
