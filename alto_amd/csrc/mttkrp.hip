// MTTKRP and CP-APR Phi over the per-mode chunk-ordered nonzero streams.
//
// Layout: the nonzeros of each ALTO chunk (2^chunk_log2 consecutive nonzeros
// in ALTO order) are stored, per target mode, sorted by that mode's row
// (stable), as packed coordinates + value.  A lane group of G lanes (G =
// power of two >= rank, up to 64; E elements per lane) walks a contiguous
// range of the stream: batches of G nonzeros are loaded coalesced (one per
// lane) and broadcast with ds_bpermute; the Hadamard product of the input
// factor rows is formed with row-contiguous gathers (lane r reads column r);
// contributions are reduced in registers while the output row stays the same
// and flushed with one hardware FP atomic per column when it changes.
// Workgroups are mapped XCD-aware: the 8 XCDs each walk one contiguous eighth
// of the stream so that each XCD's L2 holds the factor rows of its region.
//
// Reference semantics: mttkrp_alto_par / mttkrp_alto_da_mem_pull
// (include/alto.hpp:617-623, :1227-1318) and CalculatePhi_alto_par
// (include/apr.hpp:295-303, :383-501).
#include "alto_internal.hpp"
#include "kernels_common.hpp"

namespace alto {

// nonzeros per lane group; set by alto_tune_set (ALTO_TUNE_SEG) or the
// autotuner, else 256
static int g_seg = 0;
int tuned_seg(int R)
{
    (void)R;
    return g_seg > 0 ? g_seg : 256;
}
static int g_chunk = 0;
static int g_nt = 0;
// -1: per element type, from measurements on MI355X at rank 32 (FP64: 2
// strided columns / 4 in flight; FP32: 8 in flight; BF16: 4 contiguous
// columns, 8-byte loads)
static int g_variant = -1;
static int g_graphs = 0;
static int g_fiber = 1;
int graphs_enabled() { return g_graphs; }
int fiber_order_enabled() { return g_fiber; }
int tuned_chunk_log2() { return g_chunk; }
int set_chunk(int v)
{
    if (v < 0 || v > 31) return ALTO_ERR_ARG;
    g_chunk = v;
    return ALTO_OK;
}
int set_seg(int v)
{
    if (v < 0 || v > (1 << 20)) return ALTO_ERR_ARG;
    g_seg = v;
    return ALTO_OK;
}

// Column mapping of a lane group: strided (lane gl holds c = gl + q*G, so
// each gather / atomic instruction touches G contiguous elements) or
// contiguous (lane gl holds c = gl*E + q, loaded with one 16-byte load where
// the row allows it).
template <bool CONTIG, int G, int E>
__device__ inline int col_of(int gl, int q) { return CONTIG ? gl * E + q : gl + q * G; }

template <typename FT, int E> struct VecLoad {
    static constexpr bool ok = false;
};
template <> struct VecLoad<double, 2> {
    static constexpr bool ok = true;
    template <typename AT> __device__ static void load(const double* p, AT (&x)[2])
    {
        const double2 a = *reinterpret_cast<const double2*>(p);
        x[0] = (AT)a.x;
        x[1] = (AT)a.y;
    }
};
template <> struct VecLoad<double, 4> {
    static constexpr bool ok = true;
    template <typename AT> __device__ static void load(const double* p, AT (&x)[4])
    {
        const double2 a = *reinterpret_cast<const double2*>(p);
        const double2 b = *reinterpret_cast<const double2*>(p + 2);
        x[0] = (AT)a.x; x[1] = (AT)a.y; x[2] = (AT)b.x; x[3] = (AT)b.y;
    }
};
template <> struct VecLoad<float, 4> {
    static constexpr bool ok = true;
    template <typename AT> __device__ static void load(const float* p, AT (&x)[4])
    {
        const float4 a = *reinterpret_cast<const float4*>(p);
        x[0] = (AT)a.x; x[1] = (AT)a.y; x[2] = (AT)a.z; x[3] = (AT)a.w;
    }
};
template <> struct VecLoad<float, 2> {
    static constexpr bool ok = true;
    template <typename AT> __device__ static void load(const float* p, AT (&x)[2])
    {
        const float2 a = *reinterpret_cast<const float2*>(p);
        x[0] = (AT)a.x; x[1] = (AT)a.y;
    }
};
template <> struct VecLoad<bf16_t, 4> {
    static constexpr bool ok = true;
    template <typename AT> __device__ static void load(const bf16_t* p, AT (&x)[4])
    {
        const uint2 a = *reinterpret_cast<const uint2*>(p);
        x[0] = (AT)__uint_as_float(a.x << 16); x[1] = (AT)__uint_as_float(a.x & 0xffff0000u);
        x[2] = (AT)__uint_as_float(a.y << 16); x[3] = (AT)__uint_as_float(a.y & 0xffff0000u);
    }
};
template <> struct VecLoad<bf16_t, 2> {
    static constexpr bool ok = true;
    template <typename AT> __device__ static void load(const bf16_t* p, AT (&x)[2])
    {
        const uint32_t a = *reinterpret_cast<const uint32_t*>(p);
        x[0] = (AT)__uint_as_float(a << 16); x[1] = (AT)__uint_as_float(a & 0xffff0000u);
    }
};

template <typename FT, typename AT, int G, int E, bool CONTIG = false>
__device__ inline void load_cols(const FT* __restrict__ row, int gl, int R, AT (&x)[E])
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
        x[q] = c < R ? to_acc<AT>(row[c]) : (AT)0;
    }
}

template <bool NT, typename T>
__device__ inline T stream_ld(const T* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// FIB: the first other mode is the stream's fiber mode; a nonzero whose row
// there equals the previous nonzero's reuses the row already in registers
// instead of gathering it again.
template <typename FT, typename VT, typename AT, typename OT, int G, int E, int PW, int J, bool NT, bool CONTIG,
          bool FIB>
__global__ void __launch_bounds__(256) k_mttkrp(StreamArgs<FT, VT, OT> a)
{
    const int lane = threadIdx.x & 63;
    const int gl = lane & (G - 1);
    const uint64_t block = xcd_block(a.nblocks_real);
    if (block == ~0ull) return;
    const uint64_t group = (block * 256 + threadIdx.x) / G;
    const uint64_t s = group * a.seg;
    if (s >= a.nnz) return;
    const uint64_t e = s + a.seg < a.nnz ? s + a.seg : a.nnz;
    const int R = a.rank;
    const int nother = a.nother;

    AT acc[E];
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = 0;
    uint32_t cur = 0xffffffffu;
    // last fiber-mode row offset and its values (row offsets are < 2^32 - R,
    // so the initial value never matches)
    uint32_t foff = 0xffffffffu;
    AT fx[E];
#pragma unroll
    for (int q = 0; q < E; ++q) fx[q] = 0;

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
            if constexpr (FIB) {
                // fiber mode: gather only rows that differ from the previous
                // nonzero's (the predicate is uniform across the group)
                const FT* __restrict__ U = a.Uo[0];
                AT fxj[J][E];
                uint32_t prev = foff;
#pragma unroll
                for (int jj = 0; jj < J; ++jj) {
                    const int j = j0 + jj < cnt ? j0 + jj : cnt - 1;
                    const uint32_t off = (uint32_t)__shfl((int)offs[0], j, G);
                    if (off != prev) {
                        load_cols<FT, AT, G, E, CONTIG>(U + off, gl, R, fxj[jj]);
                    } else {
#pragma unroll
                        for (int q = 0; q < E; ++q) fxj[jj][q] = jj == 0 ? fx[q] : fxj[jj - 1][q];
                    }
                    prev = off;
                }
                foff = prev;
#pragma unroll
                for (int q = 0; q < E; ++q) fx[q] = fxj[J - 1][q];
#pragma unroll
                for (int jj = 0; jj < J; ++jj)
#pragma unroll
                    for (int q = 0; q < E; ++q) prod[jj][q] *= fxj[jj][q];
            }
#pragma unroll
            for (int o = FIB ? 1 : 0; o < ALTO_MAXN - 1; ++o) {
                if (o < nother) {
                    const FT* __restrict__ U = a.Uo[o];
#pragma unroll
                    for (int jj = 0; jj < J; ++jj) {
                        const int j = j0 + jj < cnt ? j0 + jj : cnt - 1;
                        const uint32_t off = (uint32_t)__shfl((int)offs[o], j, G);
                        AT x[E];
                        load_cols<FT, AT, G, E, CONTIG>(U + off, gl, R, x);
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

// Phi: per nonzero pi = prod_{m != n} U_m[row_m,:], d = <B[row,:], pi>,
// contribution (x / max(d, eps)) * pi.
template <int G, int E, int PW, int J>
__global__ void __launch_bounds__(256) k_phi(StreamArgs<double, double, double> a, const double* __restrict__ B, double eps)
{
    const int lane = threadIdx.x & 63;
    const int gl = lane & (G - 1);
    const uint64_t block = xcd_block(a.nblocks_real);
    if (block == ~0ull) return;
    const uint64_t group = (block * 256 + threadIdx.x) / G;
    const uint64_t s = group * a.seg;
    if (s >= a.nnz) return;
    const uint64_t e = s + a.seg < a.nnz ? s + a.seg : a.nnz;
    const int R = a.rank;
    const int nother = a.nother;

    double acc[E];
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = 0;
    uint32_t cur = 0xffffffffu;
    for (uint64_t p = s; p < e; p += G) {
        uint32_t offs[ALTO_MAXN - 1];
        uint32_t rown = 0;
        double myv = 0;
        {
            uint64_t k[PW];
            if (p + gl < e) {
#pragma unroll
                for (int w = 0; w < PW; ++w) k[w] = a.pk[(p + gl) * PW + w];
                myv = a.val[p + gl];
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
            double pi[J][E], b[J][E], v[J];
            uint32_t rj[J];
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                const int j = j0 + jj < cnt ? j0 + jj : cnt - 1;
                rj[jj] = (uint32_t)__shfl((int)rown, j, G);
                v[jj] = shfl_val(myv, j, G);
                load_cols<double, double, G, E>(B + (uint64_t)rj[jj] * R, gl, R, b[jj]);
#pragma unroll
                for (int q = 0; q < E; ++q) pi[jj][q] = 1.0;
            }
#pragma unroll
            for (int o = 0; o < ALTO_MAXN - 1; ++o) {
                if (o < nother) {
                    const double* __restrict__ U = a.Uo[o];
#pragma unroll
                    for (int jj = 0; jj < J; ++jj) {
                        const int j = j0 + jj < cnt ? j0 + jj : cnt - 1;
                        const uint32_t off = (uint32_t)__shfl((int)offs[o], j, G);
                        double x[E];
                        load_cols<double, double, G, E>(U + off, gl, R, x);
#pragma unroll
                        for (int q = 0; q < E; ++q) pi[jj][q] *= x[q];
                    }
                }
            }
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                double d = 0;
#pragma unroll
                for (int q = 0; q < E; ++q) d += b[jj][q] * pi[jj][q];
                d = group_sum<G>(d);
                const double w = v[jj] / fmax(d, eps);
                if (j0 + jj < cnt) {
                    if (rj[jj] != cur) {
                        if (cur != 0xffffffffu) flush_row<double, double, G, E>(a.out, cur, R, gl, acc);
                        cur = rj[jj];
#pragma unroll
                        for (int q = 0; q < E; ++q) acc[q] = 0;
                    }
#pragma unroll
                    for (int q = 0; q < E; ++q) acc[q] += w * pi[jj][q];
                }
            }
        }
    }
    if (cur != 0xffffffffu) flush_row<double, double, G, E>(a.out, cur, R, gl, acc);
}

// ------------------------------------------------------------------ host side
template <typename FT, typename VT, typename AT, typename OT>
static int launch_mttkrp(const alto_tensor* t, int mode, int R, const FT* const* U, const VT* val, OT* out, int seg,
                         hipStream_t s)
{
    StreamArgs<FT, VT, OT> a;
    fill_stream_args(t, mode, R, seg, a);
    set_factors(t, mode, U, a);
    a.val = val;
    a.out = out;
    // variants: 0 = 2 strided columns per lane, 4 nonzeros in flight per group;
    // 1 = same with 8 in flight; 2 = 4 contiguous columns per lane (16-byte
    // loads), 4 in flight; 3 = 4 contiguous columns, 8 in flight
    const int v = g_variant >= 0 ? g_variant : std::is_same<FT, double>::value ? 0 : std::is_same<FT, float>::value ? 1 : 2;
    const int E = v >= 2 ? 4 : 2;
    int G = 4;
    while (G * E < R && G < 64) G <<= 1;
    dim3 grid = stream_grid(t->nnz, seg, G, a);
    const bool fib = t->ms[mode].fib >= 0 && a.nother > 0;
#define L_(GG, EE, JJ, CC)                                                                                      \
    do {                                                                                                         \
        if (t->pack.pw == 1) {                                                                                   \
            if (g_nt) hipLaunchKernelGGL((k_mttkrp<FT, VT, AT, OT, GG, EE, 1, JJ, true, CC, false>), grid, dim3(256), 0, s, a); \
            else if (fib) hipLaunchKernelGGL((k_mttkrp<FT, VT, AT, OT, GG, EE, 1, JJ, false, CC, true>), grid, dim3(256), 0, s, a); \
            else hipLaunchKernelGGL((k_mttkrp<FT, VT, AT, OT, GG, EE, 1, JJ, false, CC, false>), grid, dim3(256), 0, s, a); \
        } else {                                                                                                 \
            if (fib) hipLaunchKernelGGL((k_mttkrp<FT, VT, AT, OT, GG, EE, 2, JJ, false, CC, true>), grid, dim3(256), 0, s, a); \
            else hipLaunchKernelGGL((k_mttkrp<FT, VT, AT, OT, GG, EE, 2, JJ, false, CC, false>), grid, dim3(256), 0, s, a); \
        }                                                                                                        \
    } while (0)
#define LG_(EE, JJ, CC)                                                                                          \
    do {                                                                                                         \
        switch (G) {                                                                                             \
        case 4: L_(4, EE, JJ, CC); break;                                                                        \
        case 8: L_(8, EE, JJ, CC); break;                                                                        \
        case 16: L_(16, EE, JJ, CC); break;                                                                      \
        case 32: L_(32, EE, JJ, CC); break;                                                                      \
        default: L_(64, EE, JJ, CC); break;                                                                      \
        }                                                                                                        \
    } while (0)
    switch (v) {
    case 1: LG_(2, 8, false); break;
    case 2: LG_(4, 4, true); break;
    case 3: LG_(4, 8, true); break;
    default: LG_(2, 4, false); break;
    }
#undef LG_
#undef L_
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

// factor row offsets (row * rank) are carried as 32-bit values
static int check_offsets(const alto_tensor* t, int R)
{
    for (int m = 0; m < t->nmode; ++m)
        if (t->dims[m] * (uint64_t)R >= (1ull << 32)) {
            set_error("mode %d: dims*rank = %llu exceeds 2^32 (32-bit row offsets)", m,
                      (unsigned long long)(t->dims[m] * (uint64_t)R));
            return ALTO_ERR_UNSUPPORTED;
        }
    return ALTO_OK;
}

int mttkrp_device(alto_tensor* t, int mode, int R, const void* const* U, void* out, int precision, int accumulate,
                  hipStream_t s, int seg)
{
    if (mode < 0 || mode >= t->nmode || R < 1 || R > 128) {
        set_error("mttkrp: mode %d / rank %d out of range (rank 1..128)", mode, R);
        return ALTO_ERR_ARG;
    }
    ALTO_TRY(check_offsets(t, R));
    ALTO_TRY(prepare_mode_streams(t, precision == ALTO_PREC_F64, precision != ALTO_PREC_F64, s));
    uint64_t nout = t->dims[mode] * (uint64_t)R;
    if (seg <= 0) seg = default_seg(R);
    if (precision == ALTO_PREC_F64) {
        if (!accumulate) ALTO_TRY(fill_zero((double*)out, nout, s));
        return launch_mttkrp<double, double, double, double>(t, mode, R, (const double* const*)U,
                                                             t->ms[mode].val64.as<double>(), (double*)out, seg, s);
    }
    if (!accumulate) ALTO_TRY(fill_zero((float*)out, nout, s));
    if (precision == ALTO_PREC_F32)
        return launch_mttkrp<float, float, float, float>(t, mode, R, (const float* const*)U, t->ms[mode].val32.as<float>(),
                                                         (float*)out, seg, s);
    if (precision == ALTO_PREC_BF16)
        return launch_mttkrp<bf16_t, float, float, float>(t, mode, R, (const bf16_t* const*)U,
                                                          t->ms[mode].val32.as<float>(), (float*)out, seg, s);
    set_error("mttkrp: unknown precision %d", precision);
    return ALTO_ERR_ARG;
}

int phi_device(alto_tensor* t, int mode, int R, const double* const* U, double* phi, double eps, hipStream_t s, int seg)
{
    if (mode < 0 || mode >= t->nmode || R < 1 || R > 128) {
        set_error("phi: mode %d / rank %d out of range (rank 1..128)", mode, R);
        return ALTO_ERR_ARG;
    }
    ALTO_TRY(check_offsets(t, R));
    ALTO_TRY(prepare_mode_streams(t, 1, 0, s));
    if (seg <= 0) seg = default_seg(R);
    ALTO_TRY(fill_zero(phi, t->dims[mode] * (uint64_t)R, s));
    StreamArgs<double, double, double> a;
    fill_stream_args(t, mode, R, seg, a);
    set_factors(t, mode, U, a);
    a.val = t->ms[mode].val64.as<double>();
    a.out = phi;
    dim3 grid = stream_grid(t->nnz, seg, group_lanes(R), a);
    RANK_DISPATCH(R, t->pack.pw, { hipLaunchKernelGGL((k_phi<G, E, PW, 4>), grid, dim3(256), 0, s, a, U[mode], eps); });
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

}  // namespace alto

using namespace alto;

extern "C" {

int alto_tune_set(int param, int value)
{
    if (param == ALTO_TUNE_SEG) return set_seg(value);
    if (param == ALTO_TUNE_CHUNK_LOG2) return set_chunk(value);
    if (param == ALTO_TUNE_FIBER_ORDER) {
        g_fiber = value ? 1 : 0;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_GRAPHS) {
        g_graphs = value ? 1 : 0;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_MTTKRP_VARIANT) {
        if (value < -1 || value > 3) return ALTO_ERR_ARG;
        g_variant = value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_NONTEMPORAL) {
        g_nt = value ? 1 : 0;
        return ALTO_OK;
    }
    set_error("alto_tune_set: unknown parameter %d", param);
    return ALTO_ERR_ARG;
}

int alto_mttkrp(alto_tensor_t* t, int mode, uint64_t rank, const void* const* factors, void* out, int precision,
                int accumulate, alto_stream_t stream)
{
    if (!t || !factors || !out) {
        set_error("alto_mttkrp: null argument");
        return ALTO_ERR_ARG;
    }
    return mttkrp_device(t, mode, (int)rank, factors, out, precision, accumulate, (hipStream_t)stream, 0);
}

int alto_phi(alto_tensor_t* t, int mode, uint64_t rank, const double* const* factors, double* phi, double eps_div_zero,
             alto_stream_t stream)
{
    if (!t || !factors || !phi) {
        set_error("alto_phi: null argument");
        return ALTO_ERR_ARG;
    }
    return phi_device(t, mode, (int)rank, factors, phi, eps_div_zero, (hipStream_t)stream, 0);
}

}  // extern "C"
