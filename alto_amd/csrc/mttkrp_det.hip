// Deterministic-order MTTKRP and CP-APR Phi (ALTO_TUNE_DETERMINISTIC): every
// sum is taken in one fixed order, so a run gives bit-identical results
// whatever the scheduling of the launch, for parity runs that must not
// depend on the order of floating-point atomics.  The reference's pull
// phase of mttkrp_alto_da_mem_pull (include/alto.hpp:1290-1317) has such a
// fixed order; this is the stream-walk form of the same property.
//
// The mode's stream is fully sorted by output row (one chunk).  Lane group g
// walks segment [g seg, (g + 1) seg) one nonzero at a time, its lanes
// holding strided columns of the row's running sum.  Rows that start and end
// inside the segment belong to no other segment: they are added into the
// output by this group alone.  The first and the last row of each segment
// may continue in the neighbouring segments: their partial sums go to two
// slots per segment, and a second launch adds the slots of each such row in
// segment order (one workgroup per row, columns across its lanes).
#include "mttkrp_kernels.hpp"

namespace alto {

static constexpr uint32_t DET_EMPTY = 0xffffffffu;
static constexpr uint64_t DET_SEG = 1024;

template <typename FT, typename VT, typename AT, int G, int E, int PW, bool PHI>
__global__ void __launch_bounds__(256) k_det_stream(StreamArgs<FT, VT, double> a, const FT* __restrict__ Bm,
                                                    double eps, uint32_t* __restrict__ brow,
                                                    double* __restrict__ bpart, uint64_t ngroups)
{
    const int gl = (threadIdx.x & 63) & (G - 1);
    const uint64_t g = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (g >= ngroups) return;
    const int R = a.rank;
    const uint64_t s = g * a.seg;
    if (s >= a.nnz) {
        if (gl == 0) brow[2 * g] = brow[2 * g + 1] = DET_EMPTY;
        return;
    }
    const uint64_t e = s + a.seg < a.nnz ? s + a.seg : a.nnz;
    AT acc[E];
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = 0;
    uint32_t cur = DET_EMPTY;
    bool head = false;  // the segment's first row went to its slot
    for (uint64_t p = s; p < e; ++p) {
        uint64_t k[PW];
#pragma unroll
        for (int w = 0; w < PW; ++w) k[w] = a.pk[p * PW + w];
        const uint32_t row = (uint32_t)field<PW>(k, a.offn, a.bitsn);
        AT pi[E];
#pragma unroll
        for (int q = 0; q < E; ++q) pi[q] = PHI ? (AT)1 : (AT)a.val[p];
        for (int o = 0; o < a.nother; ++o) {
            const uint32_t off = (uint32_t)(field<PW>(k, a.offo[o], a.bitso[o]) * (uint64_t)R);
            AT x[E];
            load_cols<FT, AT, G, E>(a.Uo[o] + off, gl, R, x);
#pragma unroll
            for (int q = 0; q < E; ++q) pi[q] *= x[q];
        }
        if constexpr (PHI) {
            AT b[E];
            load_cols<FT, AT, G, E>(Bm + (uint64_t)row * R, gl, R, b);
            AT d = 0;
#pragma unroll
            for (int q = 0; q < E; ++q) d += b[q] * pi[q];
            d = group_sum<G>(d);
            const AT w = (AT)a.val[p] / fmax(d, (AT)eps);
#pragma unroll
            for (int q = 0; q < E; ++q) pi[q] *= w;
        }
        if (row != cur) {
            if (cur != DET_EMPTY) {
                if (!head) {
                    if (gl == 0) brow[2 * g] = cur;
#pragma unroll
                    for (int q = 0; q < E; ++q) {
                        const int c = gl + q * G;
                        if (c < R) bpart[(2 * g) * R + c] = (double)acc[q];
                    }
                    head = true;
                } else {
#pragma unroll
                    for (int q = 0; q < E; ++q) {
                        const int c = gl + q * G;
                        if (c < R) a.out[(uint64_t)cur * R + c] += (double)acc[q];
                    }
                }
            }
            cur = row;
#pragma unroll
            for (int q = 0; q < E; ++q) acc[q] = 0;
        }
#pragma unroll
        for (int q = 0; q < E; ++q) acc[q] += pi[q];
    }
    if (gl == 0) {
        if (!head) brow[2 * g] = DET_EMPTY;
        brow[2 * g + 1] = cur;
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int c = gl + q * G;
        if (c < R) bpart[(2 * g + 1) * R + c] = (double)acc[q];
    }
}

// Slots j (2 per segment, in segment order; the rows of the non-empty ones
// never decrease): the first slot of each row adds the row's slots, in
// order, into the output.  Only a segment's first slot can be empty when the
// segment holds data, so the previous non-empty slot is at most two back.
__global__ void __launch_bounds__(64) k_det_combine(const uint32_t* __restrict__ brow, const double* __restrict__ bpart,
                                                    uint64_t nslots, int R, double* __restrict__ out)
{
    const uint64_t j = blockIdx.x;
    const uint32_t row = brow[j];
    if (row == DET_EMPTY) return;
    int64_t q = (int64_t)j - 1;
    while (q >= 0 && brow[q] == DET_EMPTY) --q;
    if (q >= 0 && brow[q] == row) return;  // not the row's first slot
    for (int c = threadIdx.x; c < R; c += blockDim.x) {
        double sum = 0;
        for (uint64_t k = j; k < nslots; ++k) {
            const uint32_t rk = brow[k];
            if (rk == DET_EMPTY) continue;
            if (rk != row) break;
            sum += bpart[k * R + c];
        }
        out[(uint64_t)row * R + c] += sum;
    }
}

// slots of the tensor's deterministic walk for rank R (kept on the tensor,
// grown when needed, after the work that used them has drained)
static int det_slots(alto_tensor* t, int R, uint64_t ngroups)
{
    const uint64_t rows_b = 2 * ngroups * sizeof(uint32_t), part_b = 2 * ngroups * (uint64_t)R * sizeof(double);
    if (t->det_rows.bytes < rows_b || t->det_part.bytes < part_b) {
        ALTO_TRY(t->wait_uses());
        if (t->det_rows.bytes < rows_b) {
            t->det_rows.release();
            ALTO_TRY(t->det_rows.alloc(rows_b));
        }
        if (t->det_part.bytes < part_b) {
            t->det_part.release();
            ALTO_TRY(t->det_part.alloc(part_b));
        }
    }
    return ALTO_OK;
}

template <typename FT, typename VT, typename AT, bool PHI>
static int det_launch(alto_tensor* t, int mode, int R, const FT* const* U, const VT* val, double* out, double eps,
                      hipStream_t s)
{
    StreamArgs<FT, VT, double> a;
    fill_stream_args(t, mode, R, (int)DET_SEG, a);
    ALTO_TRY(set_factors(t, mode, U, a, s));
    a.val = val;
    a.out = out;
    const uint64_t ngroups = (t->nnz + DET_SEG - 1) / DET_SEG;
    ALTO_TRY(det_slots(t, R, ngroups));
    uint32_t* brow = t->det_rows.as<uint32_t>();
    double* bpart = t->det_part.as<double>();
    RANK_DISPATCH(R, t->pack.pw, {
        const uint64_t threads = ngroups * G;
        const dim3 grid((unsigned)((threads + 255) / 256));
        hipLaunchKernelGGL((k_det_stream<FT, VT, AT, G, E, PW, PHI>), grid, dim3(256), 0, s, a, U[mode], eps, brow,
                           bpart, ngroups);
    });
    HIP_TRY(hipGetLastError());
    if (ngroups) hipLaunchKernelGGL(k_det_combine, dim3((unsigned)(2 * ngroups)), dim3(64), 0, s, brow, bpart,
                                    2 * ngroups, R, out);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

int mttkrp_det_device(alto_tensor* t, int mode, int R, const void* const* U, double* out, int precision, hipStream_t s)
{
    if (t->nnz >= (1ull << 40)) {
        set_error("deterministic MTTKRP: too many nonzeros");
        return ALTO_ERR_UNSUPPORTED;
    }
    if (precision == ALTO_PREC_F64)
        return det_launch<double, double, double, false>(t, mode, R, (const double* const*)U,
                                                          t->ms[mode].val64.as<double>(), out, 0.0, s);
    if (precision == ALTO_PREC_F32)
        return det_launch<float, float, float, false>(t, mode, R, (const float* const*)U,
                                                       t->ms[mode].val32.as<float>(), out, 0.0, s);
    if (precision == ALTO_PREC_BF16)
        return det_launch<bf16_t, float, float, false>(t, mode, R, (const bf16_t* const*)U,
                                                        t->ms[mode].val32.as<float>(), out, 0.0, s);
    set_error("mttkrp: unknown precision %d", precision);
    return ALTO_ERR_ARG;
}

int phi_det_device(alto_tensor* t, int mode, int R, const double* const* U, double* phi, double eps, hipStream_t s)
{
    return det_launch<double, double, double, true>(t, mode, R, U, t->ms[mode].val64.as<double>(), phi, eps, s);
}

}  // namespace alto
