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

template <typename FT, typename VT, typename AT, typename OT, int G, int E, int PW>
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
    const int mode = a.mode;

    AT acc[E];
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = 0;
    uint64_t cur = ~0ull;

    for (uint64_t p = s; p < e; p += G) {
        uint64_t myk[PW];
        VT myv = 0;
        if (p + gl < e) {
#pragma unroll
            for (int w = 0; w < PW; ++w) myk[w] = a.pk[(p + gl) * PW + w];
            myv = a.val[p + gl];
        } else {
#pragma unroll
            for (int w = 0; w < PW; ++w) myk[w] = 0;
        }
        const int cnt = (int)(e - p < (uint64_t)G ? e - p : (uint64_t)G);
        for (int j = 0; j < cnt; ++j) {
            uint64_t k[PW];
#pragma unroll
            for (int w = 0; w < PW; ++w) k[w] = shfl_u64(myk[w], j, G);
            const AT v = (AT)shfl_val(myv, j, G);
            const uint64_t row = field<PW>(k, a.off[mode], a.bits[mode]);
            if (row != cur) {
                if (cur != ~0ull) flush_row<AT, OT, E, G>(a.out, cur, R, gl, acc);
                cur = row;
#pragma unroll
                for (int q = 0; q < E; ++q) acc[q] = 0;
            }
            AT prod[E];
#pragma unroll
            for (int q = 0; q < E; ++q) prod[q] = v;
            for (int m = 0; m < a.nmode; ++m) {
                if (m == mode) continue;
                const uint64_t r = field<PW>(k, a.off[m], a.bits[m]);
                const FT* u = a.U[m] + r * (uint64_t)R;
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    const int c = gl + q * G;
                    prod[q] *= c < R ? to_acc<AT>(u[c]) : (AT)0;
                }
            }
#pragma unroll
            for (int q = 0; q < E; ++q) acc[q] += prod[q];
        }
    }
    if (cur != ~0ull) flush_row<AT, OT, E, G>(a.out, cur, R, gl, acc);
}

// Phi: per nonzero pi = prod_{m != n} U_m[row_m,:], d = <B[row,:], pi>,
// contribution (x / max(d, eps)) * pi.  B's row is reloaded on row change.
template <int G, int E, int PW>
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
    const int mode = a.mode;

    double acc[E], b[E];
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = b[q] = 0;
    uint64_t cur = ~0ull;
    for (uint64_t p = s; p < e; p += G) {
        uint64_t myk[PW];
        double myv = 0;
        if (p + gl < e) {
#pragma unroll
            for (int w = 0; w < PW; ++w) myk[w] = a.pk[(p + gl) * PW + w];
            myv = a.val[p + gl];
        } else {
#pragma unroll
            for (int w = 0; w < PW; ++w) myk[w] = 0;
        }
        const int cnt = (int)(e - p < (uint64_t)G ? e - p : (uint64_t)G);
        for (int j = 0; j < cnt; ++j) {
            uint64_t k[PW];
#pragma unroll
            for (int w = 0; w < PW; ++w) k[w] = shfl_u64(myk[w], j, G);
            const double v = shfl_val(myv, j, G);
            const uint64_t row = field<PW>(k, a.off[mode], a.bits[mode]);
            if (row != cur) {
                if (cur != ~0ull) flush_row<double, double, E, G>(a.out, cur, R, gl, acc);
                cur = row;
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    acc[q] = 0;
                    const int c = gl + q * G;
                    b[q] = c < R ? B[row * (uint64_t)R + c] : 0.0;
                }
            }
            double pi[E];
#pragma unroll
            for (int q = 0; q < E; ++q) pi[q] = 1.0;
            for (int m = 0; m < a.nmode; ++m) {
                if (m == mode) continue;
                const uint64_t r = field<PW>(k, a.off[m], a.bits[m]);
                const double* u = a.U[m] + r * (uint64_t)R;
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    const int c = gl + q * G;
                    pi[q] *= c < R ? u[c] : 0.0;
                }
            }
            double d = 0;
#pragma unroll
            for (int q = 0; q < E; ++q) d += b[q] * pi[q];
            d = group_sum<G>(d);
            const double w = v / fmax(d, eps);
#pragma unroll
            for (int q = 0; q < E; ++q) acc[q] += w * pi[q];
        }
    }
    if (cur != ~0ull) flush_row<double, double, E, G>(a.out, cur, R, gl, acc);
}

// ------------------------------------------------------------------ host side
template <typename FT, typename VT, typename AT, typename OT>
static int launch_mttkrp(const alto_tensor* t, int mode, int R, const FT* const* U, const VT* val, OT* out, int seg,
                         hipStream_t s)
{
    StreamArgs<FT, VT, OT> a;
    fill_stream_args(t, mode, R, seg, a);
    for (int m = 0; m < t->nmode; ++m) a.U[m] = U[m];
    a.val = val;
    a.out = out;
    dim3 grid = stream_grid(t->nnz, seg, group_lanes(R), a);
    RANK_DISPATCH(R, t->pack.pw, {
        hipLaunchKernelGGL((k_mttkrp<FT, VT, AT, OT, G, E, PW>), grid, dim3(256), 0, s, a);
    });
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

int mttkrp_device(alto_tensor* t, int mode, int R, const void* const* U, void* out, int precision, int accumulate,
                  hipStream_t s, int seg)
{
    if (mode < 0 || mode >= t->nmode || R < 1 || R > 128) {
        set_error("mttkrp: mode %d / rank %d out of range (rank 1..128)", mode, R);
        return ALTO_ERR_ARG;
    }
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
    ALTO_TRY(prepare_mode_streams(t, 1, 0, s));
    if (seg <= 0) seg = default_seg(R);
    ALTO_TRY(fill_zero(phi, t->dims[mode] * (uint64_t)R, s));
    StreamArgs<double, double, double> a;
    fill_stream_args(t, mode, R, seg, a);
    for (int m = 0; m < t->nmode; ++m) a.U[m] = U[m];
    a.val = t->ms[mode].val64.as<double>();
    a.out = phi;
    dim3 grid = stream_grid(t->nnz, seg, group_lanes(R), a);
    RANK_DISPATCH(R, t->pack.pw, { hipLaunchKernelGGL((k_phi<G, E, PW>), grid, dim3(256), 0, s, a, U[mode], eps); });
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

}  // namespace alto

using namespace alto;

extern "C" {

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
