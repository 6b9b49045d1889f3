// MTTKRP (host side; kernels in mttkrp_kernels.hpp) and CP-APR Phi over the
// per-mode chunk-ordered nonzero streams.
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
#include "mttkrp_kernels.hpp"

namespace alto {

// Nonzeros walked by one lane group: ALTO_TUNE_SEG when set, else a power
// of two in [512, 4096] (from 256 below 4M nonzeros) sized so that a launch
// has about 2^21 threads (G lanes per group), doubled for FP32 accumulation.  Long segments cut the
// flushes -- and with them the atomic traffic on heavily used output rows
// -- while enough groups remain to fill the chip (measured on MI355X: 512
// for the 77M-nonzero 3-mode tensor in FP64, 1024 in FP32; 1024-4096 for the
// 200M-nonzero 5-mode one; 512 for 14M-nonzero rank-8 tensors).
static int g_seg = 0;
int tuned_seg(uint64_t nnz, int G, int acc_bytes)
{
    if (g_seg > 0) return g_seg;
    if (nnz < (4ull << 20)) {
        // below 4M nonzeros a launch is latency-bound: short segments give
        // more lane groups.  Measured optimum on MI355X from 10K to 4M
        // nonzeros (3-5 modes, ranks 8-64, FP64 and FP32): ~0.16 sqrt(nnz)
        // nonzeros per group, as a power of two from 16
        const double target = 0.16 * sqrt((double)nnz);
        int seg = 16;
        while (seg < 256 && seg * 1.5 < target) seg <<= 1;
        return seg;
    }
    // large tensors: about 2^21 threads per launch (flushes cost more than
    // the lost parallelism), twice as long segments with FP32 accumulation
    const uint64_t want = (nnz * (uint64_t)G >> 21) * (acc_bytes < 8 ? 2 : 1);
    int seg = 512;
    while (seg < 4096 && (uint64_t)seg * 3 / 2 < want) seg <<= 1;
    return seg;
}
static int g_chunk = 0;
static int g_nt = 0;
// lane mapping of the MTTKRP kernel; -1 = per element type
static int g_variant = -1;
static int g_graphs = 0;
static int g_hot = 0;
static int g_hot_block = 512;
static int g_hot_lds = 32768;
static int g_narrow_max = 32;
static int g_narrow_small = 1;
// Phi kernel: 0 = runtime batch loop with LDS permutes, 1 = compile-time
// batch with DPP broadcasts (5% faster on the 120M-count CP-APR workload)
static int g_phi_variant = 1;
// Fiber layout of the mode streams (tensors from 4M nonzeros, 3 or more
// modes): fibers of at least g_fiber_min nonzeros are walked with the
// fiber-mode factor row held in registers.  g_fibers = number of fiber
// modes per stream (0 = off; 2: the nonzeros outside the first mode's
// fibers are searched for fibers of a second mode -- measured slower on the
// 76.9M-nonzero rank-32 tensor: all-mode FP64 4.48 -> 4.53 ms, FP32 3.27 ->
// 3.65 ms, although it moves ~10% of the nonzeros into fibers)
static int g_fibers = 1;
// 3 (FP32 all-mode MTTKRP of the 76.9M-nonzero rank-32 tensor 3.27 ms
// against 3.44 at 4; FP64 equal within noise)
static int g_fiber_min = 3;
static int g_fiber_chunk = 30;
static uint64_t g_fiber_min_nnz = 4ull << 20;
static int g_fiber_rule = 0;  // fiber mode: 0 = the other mode with the fewest rows, 1 = the most
int fiber_rule() { return g_fiber_rule; }
int fiber_settings(int* min_len, int* chunk_log2, uint64_t* min_nnz)
{
    *min_len = g_fiber_min;
    *chunk_log2 = g_fiber_chunk;
    *min_nnz = g_fiber_min_nnz;
    return g_fibers;
}
// deterministic-order MTTKRP / Phi (fixed summation order, no float atomics)
static int g_det = 0;
int det_enabled() { return g_det; }
// CP-APR: products of the other modes' rows kept per nonzero across a
// mode's inner iterations (when they fit)
static int g_apr_pi_cache = 1;
int apr_pi_cache_enabled() { return g_apr_pi_cache; }
int graphs_enabled() { return g_graphs; }
// lockstep CP-ALS: the sessions' mode updates and fits in batched launches
static int g_lockstep_dense = 1;
int lockstep_dense_batch_enabled() { return g_lockstep_dense; }
// lockstep CP-ALS: repeated iterations replayed as a HIP graph
static int g_lockstep_graphs = 0;
int lockstep_graphs_enabled() { return g_lockstep_graphs; }
// FP32 / BF16 CP-ALS: Grams and row solves on FP64 MFMA
static int g_dense_f64 = 1;
int dense_f64_enabled() { return g_dense_f64; }
// server: small jobs iterated in lockstep groups
static uint64_t g_serve_lock_nnz = 0;
static int g_serve_lock_sessions = 64;
void serve_lockstep_settings(uint64_t* max_nnz, int* max_sessions)
{
    *max_nnz = g_serve_lock_nnz;
    *max_sessions = g_serve_lock_sessions;
}
int hot_rows_enabled() { return g_hot; }
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

// The same sum with the batch loop unrolled at compile time: each lane
// decodes one nonzero of a batch of G, and the group takes each nonzero's
// row offsets and value from the decoding lane through DPP / readlane
// broadcasts (no LDS permutes); the NO other modes' gathers of J = 4
// nonzeros are issued back to back, and the next batch is loaded behind the
// first gathers.  Lanes still read strided 8-byte columns (16-byte lanes
// raise the register count past the point where occupancy pays for them;
// holding this kernel to 128 VGPRs for a fourth wave per SIMD measured
// slower too).
// PIM: the products pi of the other modes' rows are computed from the
// factors (0), computed and stored per nonzero into Pc in stream order (1),
// or read back from Pc (2): within CP-APR's inner iterations of one mode the
// other factors do not change, so the later iterations stream pi instead of
// gathering NO rows per nonzero (the reference's precomputed rows,
// include/apr.hpp:1049-1097 and 1128-1165).
template <int G, int E, int PW, int NO, int PIM>
__global__ void __launch_bounds__(256)
    k_phi_b(StreamArgs<double, double, double> a, const double* __restrict__ B, double eps, double* __restrict__ Pc)
{
    constexpr int J = G < 4 ? G : 4;
    const int lane = threadIdx.x & 63;
    const int gl = lane & (G - 1);
    const uint64_t block = xcd_block(a.nblocks_real);
    if (block == ~0ull) return;
    const uint64_t group = (block * 256 + threadIdx.x) / G;
    const uint64_t s = group * a.seg;
    if (s >= a.nnz) return;
    const uint64_t e = s + a.seg < a.nnz ? s + a.seg : a.nnz;
    const int R = a.rank;

    double acc[E], b_last[E];
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = b_last[q] = 0;
    uint32_t cur = 0xffffffffu;
    uint32_t b_row = 0xffffffffu;  // the model row b_last holds (this group's)
    // this lane's nonzero of the next batch (past the segment: row 0 of
    // every mode, value 0 -- valid memory, never accumulated)
    uint64_t kn[PW];
    double vn = 0;
#pragma unroll
    for (int w = 0; w < PW; ++w) kn[w] = 0;
    if (s + gl < e) {
#pragma unroll
        for (int w = 0; w < PW; ++w) kn[w] = a.pk[(s + gl) * PW + w];
        vn = a.val[s + gl];
    }
    for (uint64_t p = s; p < e; p += G) {
        uint32_t offs[NO];
#pragma unroll
        for (int o = 0; o < NO; ++o) offs[o] = (uint32_t)(field<PW>(kn, a.offo[o], a.bitso[o]) * (uint64_t)R);
        const uint32_t rown = (uint32_t)field<PW>(kn, a.offn, a.bitsn);
        const double myv = vn;
        const int cnt = (int)(e - p < (uint64_t)G ? e - p : (uint64_t)G);
        static_for<0, G / J>([&](auto bi) {
            constexpr int j0 = decltype(bi)::value * J;
            if (j0 >= cnt) return;
            double pi[J][E], b[J][E], v[J];
            uint32_t rj[J];
            static_for<0, J>([&](auto ji) {
                constexpr int jj = decltype(ji)::value;
                rj[jj] = bcast32<G, j0 + jj>(rown);
                v[jj] = bcast_val<G, j0 + jj>(myv);
#pragma unroll
                for (int q = 0; q < E; ++q) pi[jj][q] = 1.0;
            });
            // the model row only at the positions where some group's row
            // changes: a wave-uniform branch, so every lane loads (groups
            // whose row did not change reload the row they hold), and the
            // stream's runs of one row skip the load
            double rb[J][E];
            uint32_t ldm = 0;  // positions loaded (uniform)
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                const bool head = jj == 0 ? rj[0] != b_row : rj[jj] != rj[jj - 1];
                if (__any(head)) {
                    load_cols<double, double, G, E>(B + (uint64_t)rj[jj] * R, gl, R, rb[jj]);
                    ldm |= 1u << jj;
                }
            }
            // the mask reaches the selection below through an opaque copy, so
            // that the selection is not folded into the branches above
            asm volatile("" : "+s"(ldm));
            if constexpr (PIM == 2) {
#pragma unroll
                for (int jj = 0; jj < J; ++jj) {
                    const uint64_t i = p + j0 + jj < e ? p + j0 + jj : e - 1;
#pragma unroll
                    for (int q = 0; q < E; ++q) {
                        const int c = gl + q * G;
                        pi[jj][q] = c < R ? Pc[i * R + c] : 0.0;
                    }
                }
            } else {
                static_for<0, NO>([&](auto oi) {
                    constexpr int o = decltype(oi)::value;
                    const double* __restrict__ U = a.Uo[o];
                    static_for<0, J>([&](auto ji) {
                        constexpr int jj = decltype(ji)::value;
                        const uint32_t off = bcast32<G, j0 + jj>(offs[o]);
                        double x[E];
                        load_cols<double, double, G, E>(U + off, gl, R, x);
#pragma unroll
                        for (int q = 0; q < E; ++q) pi[jj][q] *= x[q];
                    });
                });
            }
            if constexpr (PIM == 1) {
#pragma unroll
                for (int jj = 0; jj < J; ++jj)
                    if (j0 + jj < cnt)
#pragma unroll
                        for (int q = 0; q < E; ++q) {
                            const int c = gl + q * G;
                            if (c < R) Pc[(p + j0 + jj) * R + c] = pi[jj][q];
                        }
            }
            if constexpr (j0 == 0) {
                const uint64_t i = p + G + gl;
                vn = 0;
#pragma unroll
                for (int w = 0; w < PW; ++w) kn[w] = 0;
                if (i < e) {
#pragma unroll
                    for (int w = 0; w < PW; ++w) kn[w] = a.pk[i * PW + w];
                    vn = a.val[i];
                }
            }
            // each position's model row is that of the last load at or before it
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                const bool use = (ldm >> jj) & 1u;
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    b_last[q] = use ? rb[jj][q] : b_last[q];
                    b[jj][q] = b_last[q];
                }
            }
            b_row = rj[J - 1];
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
        });
    }
    if (cur != 0xffffffffu) flush_row<double, double, G, E>(a.out, cur, R, gl, acc);
}

// Phi with 16-byte lanes (the k_mttkrp16 lane mapping): lane gl of a group
// of G holds columns 2 gl and 2 gl + 1, so every factor row (the model row
// B[row] and the NO other modes' rows) arrives with one 16-byte load per
// lane, half the gather instructions of the strided 8-byte lanes.  Needs an
// even rank and 16-byte aligned factors (rows of 16 R / 2 bytes).  Per
// nonzero: Pi = the product of the other rows, d = <B[row], Pi> summed over
// the group, and Pi x / max(d, eps) added into the row's run.  The model
// row is loaded only where the row changes inside a block of J nonzeros or
// from the previous block, with the decision uniform over the wave (the
// stream is ordered by row inside each chunk, so long runs skip most
// model-row loads).  PIM as in k_phi_b: 0 gather the other rows, 1 gather
// and store their products per nonzero into Pc, 2 read the products back
// (one 16-byte load per lane instead of NO).
template <int G, int PW, int NO, int PIM>
__global__ void __launch_bounds__(256)
    k_phi16(StreamArgs<double, double, double> a, const double* __restrict__ Bm, double eps, double* __restrict__ Pc)
{
    constexpr int E = 2;
    constexpr int NS = G < 4 ? 4 / G : 1;
    constexpr int BB = G * NS;
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
    const int c0 = gl * E < R ? gl * E : 0;
    const bool inr = gl * E < R;

    double acc[E], blast[E];
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = blast[q] = 0;
    uint32_t cur = 0xffffffffu;
    uint32_t prev_r = 0xffffffffu;  // row whose model row blast holds

    uint64_t kn[NS][PW];
    double vn[NS];
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
    for (uint64_t p = s; p < e; p += BB) {
        uint32_t offs[NS][NO];
        uint32_t rown[NS];
        double myv[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            myv[k] = vn[k];
#pragma unroll
            for (int o = 0; o < NO; ++o) offs[k][o] = (uint32_t)(field<PW>(kn[k], a.offo[o], a.bitso[o]) * (uint64_t)R);
            rown[k] = (uint32_t)field<PW>(kn[k], a.offn, a.bitsn);
        }
        const int cnt = (int)(e - p < (uint64_t)BB ? e - p : (uint64_t)BB);
        static_for<0, BB / J>([&](auto bi) {
            constexpr int j0 = decltype(bi)::value * J;
            if (j0 >= cnt) return;
            double pi[J][E], v[J];
            uint32_t rj[J];
            static_for<0, J>([&](auto ji) {
                constexpr int jj = decltype(ji)::value;
                constexpr int slot = (j0 + jj) / G, src = (j0 + jj) % G;
                rj[jj] = bcast32<G, src>(rown[slot]);
                v[jj] = bcast_val<G, src>(myv[slot]);
#pragma unroll
                for (int q = 0; q < E; ++q) pi[jj][q] = 1.0;
            });
            if constexpr (PIM == 2) {
                // lanes past the rank read the row's first pair (never used)
#pragma unroll
                for (int jj = 0; jj < J; ++jj) {
                    const uint64_t i = p + j0 + jj < e ? p + j0 + jj : e - 1;
                    VecLoad<double, E>::load(Pc + i * R + c0, pi[jj]);
                }
            } else {
                static_for<0, NO>([&](auto oi) {
                    constexpr int o = decltype(oi)::value;
                    const double* __restrict__ U = a.Uo[o];
                    static_for<0, J>([&](auto ji) {
                        constexpr int jj = decltype(ji)::value;
                        constexpr int slot = (j0 + jj) / G, src = (j0 + jj) % G;
                        const uint32_t off = bcast32<G, src>(offs[slot][o]);
                        double x[E];
                        VecLoad<double, E>::load(U + off + c0, x);
#pragma unroll
                        for (int q = 0; q < E; ++q) pi[jj][q] *= x[q];
                    });
                });
            }
            if constexpr (PIM == 1) {
#pragma unroll
                for (int jj = 0; jj < J; ++jj)
                    if (inr && j0 + jj < cnt)
                        *reinterpret_cast<v2f64*>(Pc + (p + j0 + jj) * R + c0) = v2f64{pi[jj][0], pi[jj][1]};
            }
            // model rows where the row changes (uniform over the wave)
            double br[J][E];
            uint32_t ldm = 0;
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                const bool chg = jj == 0 ? rj[0] != prev_r : rj[jj] != rj[jj - 1];
                if (__any(chg)) {
                    VecLoad<double, E>::load(Bm + (uint64_t)rj[jj] * R + c0, br[jj]);
                    ldm |= 1u << jj;
                }
            }
            // the mask reaches the selection below through an opaque copy, so
            // that the selection is not folded into the branches above
            asm volatile("" : "+s"(ldm));
            if constexpr (j0 == 0) {
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    const uint64_t i = p + BB + k * G + gl;
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
                const bool use = (ldm >> jj) & 1u;
                double d = 0;
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    blast[q] = use ? br[jj][q] : blast[q];
                    d += (inr ? blast[q] : 0.0) * pi[jj][q];
                }
                d = group_sum<G>(d);
                const double w = v[jj] / fmax(d, eps);
                if (j0 + jj < cnt) {
                    if (rj[jj] != cur) {
                        if (cur != 0xffffffffu) flush_row<double, double, G, E, true>(a.out, cur, R, gl, acc);
                        cur = rj[jj];
#pragma unroll
                        for (int q = 0; q < E; ++q) acc[q] = 0;
                    }
#pragma unroll
                    for (int q = 0; q < E; ++q) acc[q] += w * pi[jj][q];
                }
            }
            prev_r = rj[J - 1];
        });
    }
    if (cur != 0xffffffffu) flush_row<double, double, G, E, true>(a.out, cur, R, gl, acc);
}

// ------------------------------------------------------------------ host side
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

static size_t elem_bytes(int precision) { return precision == ALTO_PREC_F64 ? 8 : precision == ALTO_PREC_F32 ? 4 : 2; }

// Stream layout for MTTKRPs of rank R in `precision`: per-mode streams,
// rank-ordered factor buffers and, with hot rows on, each stream grouped by
// hot pattern for the LDS budget.  A no-op once in place (so it may run
// inside a graph capture after a first call outside it).
int mttkrp_prepare(alto_tensor* t, int R, int precision, hipStream_t s)
{
    ALTO_TRY(check_offsets(t, R));
    ALTO_TRY(prepare_mode_streams(t, precision == ALTO_PREC_F64, precision != ALTO_PREC_F64, s));
    const size_t eb = elem_bytes(precision);
    ALTO_TRY(reserve_rank_copies(t, (uint64_t)R, eb, s));
    const int no = t->nmode - 1;
    const int v = g_variant >= 0 ? g_variant : 2;
    const uint64_t lane_bytes = precision == ALTO_PREC_BF16 ? 8 : 16;
    if (!g_hot || v != 2 || no < 1 || no > 2 || ((uint64_t)R * eb) % lane_bytes != 0) return ALTO_OK;
    for (int m = 0; m < t->nmode; ++m) {
        if (!t->ms[m].ranked) continue;
        int want[ALTO_MAXN - 1] = {0};
        bool any = false;
        int o = 0;
        for (int q = 0; q < t->nmode; ++q) {
            if (q == m) continue;
            uint64_t rows = (uint64_t)g_hot_lds / no / ((uint64_t)R * eb);
            if (rows > t->dims[q]) rows = t->dims[q];
            want[o] = (int)rows;
            any = any || rows > 0;
            ++o;
        }
        if (!any) continue;
        bool same = t->ms[m].split != 0;
        for (int q = 0; q < no && same; ++q) same = t->ms[m].hot_k[q] == want[q];
        if (!same) ALTO_TRY(split_stream(t, m, want, s));
    }
    return ALTO_OK;
}

// variant >= -1: that lane mapping (and `seg` as given), for the autotuner;
// -2: the process-wide knobs, else the tensor's tuned shape, else defaults
static int mttkrp_device_shape(alto_tensor* t, int mode, int R, const void* const* U, void* out, int precision,
                               int accumulate, hipStream_t s, int seg, int variant)
{
    if (mode < 0 || mode >= t->nmode || R < 1 || R > 128) {
        set_error("mttkrp: mode %d / rank %d out of range (rank 1..128)", mode, R);
        return ALTO_ERR_ARG;
    }
    ALTO_TRY(mttkrp_prepare(t, R, precision, s));
    uint64_t nout = t->dims[mode] * (uint64_t)R;
    if (g_det) {
        if (!accumulate) ALTO_TRY(fill_zero((double*)out, nout, s));
        return mttkrp_det_device(t, mode, R, U, (double*)out, precision, s);
    }
    MttkrpTune k;
    k.variant = variant >= -1 ? variant : g_variant;
    if (variant == -2 && seg <= 0 && g_seg == 0 && g_variant < 0) {  // shape from the autotuner, if it ran
        for (const auto& c : t->tuned)
            if (c.mode == mode && c.rank == R && c.precision == precision) {
                k.variant = c.variant;
                seg = c.seg;
                break;
            }
    }
    k.nt = g_nt;
    k.hot_block = g_hot_block;
    k.hot_lds = g_hot_lds;
    k.narrow_max = g_narrow_max;
    k.narrow_small = g_narrow_small;
    if (precision == ALTO_PREC_F64) {
        if (!accumulate) ALTO_TRY(fill_zero((double*)out, nout, s));
        return launch_mttkrp_f64(t, mode, R, (const double* const*)U, t->ms[mode].val64.as<double>(), (double*)out, seg,
                                 k, s);
    }
    // FP32 / BF16: products and row sums in FP32 registers, flushed into an
    // FP64 output (FP32 atomics summing thousands of row partials lost
    // ~5e-4 of the heavy rows of a 200M-nonzero Zipf(1.2) tensor)
    if (!accumulate) ALTO_TRY(fill_zero((double*)out, nout, s));
    if (precision == ALTO_PREC_F32)
        return launch_mttkrp_f32(t, mode, R, (const float* const*)U, t->ms[mode].val32.as<float>(), (double*)out, seg,
                                 k, s);
    if (precision == ALTO_PREC_BF16)
        return launch_mttkrp_bf16(t, mode, R, (const bf16_t* const*)U, t->ms[mode].val32.as<float>(), (double*)out,
                                  seg, k, s);
    set_error("mttkrp: unknown precision %d", precision);
    return ALTO_ERR_ARG;
}

int mttkrp_device(alto_tensor* t, int mode, int R, const void* const* U, void* out, int precision, int accumulate,
                  hipStream_t s, int seg)
{
    return mttkrp_device_shape(t, mode, R, U, out, precision, accumulate, s, seg, -2);
}

int phi_device(alto_tensor* t, int mode, int R, const double* const* U, double* phi, double eps, hipStream_t s, int seg)
{
    return phi_device_cached(t, mode, R, U, phi, eps, s, seg, nullptr, 0);
}

// pim: 0 = no cache; 1 = store the products of the other modes' rows per
// nonzero into pc (nnz x R doubles, stream order); 2 = read them from pc
// (stored by a pim 1 call for the same mode and factors)
int phi_device_cached(alto_tensor* t, int mode, int R, const double* const* U, double* phi, double eps, hipStream_t s,
                      int seg, double* pc, int pim)
{
    if (mode < 0 || mode >= t->nmode || R < 1 || R > 128) {
        set_error("phi: mode %d / rank %d out of range (rank 1..128)", mode, R);
        return ALTO_ERR_ARG;
    }
    ALTO_TRY(check_offsets(t, R));
    ALTO_TRY(prepare_mode_streams(t, 1, 0, s));
    if (seg <= 0) seg = tuned_seg(t->nnz, group_lanes(R), 8);
    ALTO_TRY(fill_zero(phi, t->dims[mode] * (uint64_t)R, s));
    ALTO_TRY(reserve_rank_copies(t, (uint64_t)R, 8, s));
    if (g_det) return phi_det_device(t, mode, R, U, phi, eps, s);
    StreamArgs<double, double, double> a;
    fill_stream_args(t, mode, R, seg, a);
    ALTO_TRY(set_factors(t, mode, U, a, s));
    a.val = t->ms[mode].val64.as<double>();
    a.out = phi;
    dim3 grid = stream_grid(t->nnz, seg, group_lanes(R), a);
    bool al16 = R % 2 == 0 && ((uintptr_t)U[mode] & 15) == 0;
    for (int o = 0; o < a.nother; ++o) al16 = al16 && ((uintptr_t)a.Uo[o] & 15) == 0;
    if (g_phi_variant == 2 && al16 && (!pim || ((uintptr_t)pc & 15) == 0)) {
        const int no = t->nmode - 1;
#define PHI16_P_(N_, P_) hipLaunchKernelGGL((k_phi16<G, PW, N_, P_>), grid, dim3(256), 0, s, a, U[mode], eps, pc)
#define PHI16_NO_(N_)                      \
    do {                                   \
        if (pim == 1) PHI16_P_(N_, 1);     \
        else if (pim == 2) PHI16_P_(N_, 2); \
        else PHI16_P_(N_, 0);              \
    } while (0)
        RANK_DISPATCH(R, t->pack.pw, {
            if (no == 1) PHI16_NO_(1);
            else if (no == 2) PHI16_NO_(2);
            else if (no == 3) PHI16_NO_(3);
            else PHI16_NO_(4);
        });
#undef PHI16_NO_
#undef PHI16_P_
    } else if (g_phi_variant >= 1 || pim) {
        const int no = t->nmode - 1;
#define PHI_NO_(N_)                                                                                              \
    do {                                                                                                         \
        if (pim == 1) hipLaunchKernelGGL((k_phi_b<G, E, PW, N_, 1>), grid, dim3(256), 0, s, a, U[mode], eps, pc); \
        else if (pim == 2)                                                                                       \
            hipLaunchKernelGGL((k_phi_b<G, E, PW, N_, 2>), grid, dim3(256), 0, s, a, U[mode], eps, pc);          \
        else hipLaunchKernelGGL((k_phi_b<G, E, PW, N_, 0>), grid, dim3(256), 0, s, a, U[mode], eps, pc);          \
    } while (0)
        RANK_DISPATCH(R, t->pack.pw, {
            if (no == 1) PHI_NO_(1);
            else if (no == 2) PHI_NO_(2);
            else if (no == 3) PHI_NO_(3);
            else PHI_NO_(4);
        });
#undef PHI_NO_
    } else {
        RANK_DISPATCH(R, t->pack.pw,
                      { hipLaunchKernelGGL((k_phi<G, E, PW, 4>), grid, dim3(256), 0, s, a, U[mode], eps); });
    }
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
    if (param == ALTO_TUNE_NARROW_RANK) {
        if (value < 0 || value > 128) return ALTO_ERR_ARG;
        g_narrow_max = value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_HOT_ROWS) {
        g_hot = value ? 1 : 0;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_HOT_BLOCK) {
        if (value != 256 && value != 512 && value != 1024) return ALTO_ERR_ARG;
        g_hot_block = value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_HOT_LDS) {
        if (value < 0 || value > 160 * 1024) return ALTO_ERR_ARG;
        g_hot_lds = value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_GRAPHS) {
        g_graphs = value ? 1 : 0;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_MTTKRP_VARIANT) {
        if (value < -1 || value > 2) return ALTO_ERR_ARG;
        g_variant = value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_NARROW_SMALL) {
        g_narrow_small = value ? 1 : 0;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_PHI_VARIANT) {
        if (value < 0 || value > 2) return ALTO_ERR_ARG;
        g_phi_variant = value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_FIBERS) {
        if (value < 0 || value > 2) return ALTO_ERR_ARG;
        g_fibers = value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_DETERMINISTIC) {
        g_det = value ? 1 : 0;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_DENSE_F64) {
        g_dense_f64 = value ? 1 : 0;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_SERVE_LOCKSTEP_NNZ) {
        if (value < 0) return ALTO_ERR_ARG;
        g_serve_lock_nnz = (uint64_t)value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_SERVE_LOCKSTEP_SESSIONS) {
        if (value < 1 || value > 4096) return ALTO_ERR_ARG;
        g_serve_lock_sessions = value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_LOCKSTEP_GRAPHS) {
        g_lockstep_graphs = value ? 1 : 0;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_LOCKSTEP_DENSE_BATCH) {
        g_lockstep_dense = value ? 1 : 0;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_APR_PI_CACHE) {
        g_apr_pi_cache = value ? 1 : 0;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_FIBER_MIN) {
        if (value < 2 || value > (1 << 20)) return ALTO_ERR_ARG;
        g_fiber_min = value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_FIBER_CHUNK_LOG2) {
        if (value < 10 || value > 31) return ALTO_ERR_ARG;
        g_fiber_chunk = value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_FIBER_MODE_RULE) {
        if (value < 0 || value > 1) return ALTO_ERR_ARG;
        g_fiber_rule = value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_FIBER_MIN_NNZ) {
        if (value < 0) return ALTO_ERR_ARG;
        g_fiber_min_nnz = (uint64_t)value;
        return ALTO_OK;
    }
    if (param == ALTO_TUNE_NONTEMPORAL) {
        g_nt = value ? 1 : 0;
        return ALTO_OK;
    }
    set_error("alto_tune_set: unknown parameter %d", param);
    return ALTO_ERR_ARG;
}

int alto_mttkrp_autotune(alto_tensor_t* t, uint64_t rank, int precision, const void* const* factors, void* out,
                         int reps, alto_stream_t stream, alto_tune_entry_t* result)
{
    if (!t || !factors || !out || rank < 1 || rank > 128 || reps < 1) {
        set_error("alto_mttkrp_autotune: invalid argument");
        return ALTO_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    const int R = (int)rank;
    static const int variants[] = {2, 0, 1};
    static const int segs[] = {16, 32, 64, 128, 256, 512, 1024, 2048, 4096};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int st = ALTO_OK;
    // times `reps` accumulating launches of one shape (variant -1 / seg 0:
    // the defaults, without any tuned entry)
    auto time_cfg = [&](int mode, int v, int sg, double* ms) -> int {
        // warm-up first (it also prepares the streams)
        int r = mttkrp_device_shape(t, mode, R, factors, out, precision, 1, s, sg, v);
        if (r == ALTO_OK && hipEventRecord(e0, s) != hipSuccess) r = ALTO_ERR_HIP;
        for (int k = 0; k < reps && r == ALTO_OK; ++k)
            r = mttkrp_device_shape(t, mode, R, factors, out, precision, 1, s, sg, v);
        if (r == ALTO_OK && hipEventRecord(e1, s) != hipSuccess) r = ALTO_ERR_HIP;
        if (r == ALTO_OK && hipEventSynchronize(e1) != hipSuccess) r = ALTO_ERR_HIP;
        float f = 0;
        if (r == ALTO_OK && hipEventElapsedTime(&f, e0, e1) != hipSuccess) r = ALTO_ERR_HIP;
        *ms = (double)f / reps;
        return r;
    };
    // the defaults must not pick up an earlier tuned entry for this case
    std::vector<alto_tensor::Tuned> keep;
    for (const auto& c : t->tuned)
        if (!(c.rank == R && c.precision == precision)) keep.push_back(c);
    t->tuned = keep;
    for (int mode = 0; mode < t->nmode && st == ALTO_OK; ++mode) {
        double def_ms = 0, best_ms = 0;
        st = time_cfg(mode, -2, 0, &def_ms);
        int bv = -1, bs = 0;
        best_ms = def_ms;
        // segment lengths within 4x of the size rule's choice (the others
        // measured slower on every tensor tried, up to 10x on large ones)
        const int h = tuned_seg(t->nnz, group_lanes(R), precision == ALTO_PREC_F64 ? 8 : 4);
        for (int v : variants)
            for (int sg : segs) {
                if (sg * 4 < h || sg > h * 4) continue;
                if (st != ALTO_OK) break;
                double ms = 0;
                st = time_cfg(mode, v, sg, &ms);
                if (st == ALTO_OK && ms < best_ms) {
                    best_ms = ms;
                    bv = v;
                    bs = sg;
                }
            }
        if (st != ALTO_OK) break;
        if (bv >= 0) t->tuned.push_back({mode, R, precision, bv, bs});
        if (result) {
            result[mode].mode = mode;
            result[mode].variant = bv;
            result[mode].seg = bs;
            result[mode].ms = best_ms;
            result[mode].default_ms = def_ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

int alto_mttkrp(alto_tensor_t* t, int mode, uint64_t rank, const void* const* factors, void* out, int precision,
                int accumulate, alto_stream_t stream)
{
    if (!t || !factors || !out) {
        set_error("alto_mttkrp: null argument");
        return ALTO_ERR_ARG;
    }
    ALTO_TRY(mttkrp_device(t, mode, (int)rank, factors, out, precision, accumulate, (hipStream_t)stream, 0));
    return t->note_use((hipStream_t)stream);
}

int alto_mttkrp_batched(int njobs, alto_tensor_t* const* tensors, const int* modes, uint64_t rank,
                        const void* const* const* factors, void* const* outs, int precision, int accumulate,
                        alto_stream_t stream)
{
    ALTO_TRY(alto::mttkrp_batched_impl(njobs, tensors, modes, rank, factors, outs, precision, accumulate,
                                       (hipStream_t)stream));
    for (int i = 0; i < njobs; ++i) {
        bool seen = false;  // one event per tensor listed several times
        for (int k = 0; k < i && !seen; ++k) seen = tensors[k] == tensors[i];
        if (!seen) ALTO_TRY(tensors[i]->note_use((hipStream_t)stream));
    }
    return ALTO_OK;
}

}  // extern "C"

namespace alto {

// alto_mttkrp_batched without the tensors' use events (for callers that
// synchronize the stream themselves, such as lockstep CP-ALS)
int mttkrp_batched_impl(int njobs, alto_tensor_t* const* tensors, const int* modes, uint64_t rank,
                        const void* const* const* factors, void* const* outs, int precision, int accumulate,
                        hipStream_t s)
{
    if (njobs < 0 || (njobs > 0 && (!tensors || !modes || !factors || !outs))) {
        set_error("alto_mttkrp_batched: null argument");
        return ALTO_ERR_ARG;
    }
    if (precision != ALTO_PREC_F64 && precision != ALTO_PREC_F32 && precision != ALTO_PREC_BF16) {
        set_error("alto_mttkrp_batched: unknown precision %d", precision);
        return ALTO_ERR_ARG;
    }
    const int R = (int)rank;
    for (int i = 0; i < njobs; ++i) {
        alto_tensor* t = tensors[i];
        if (!t || !factors[i] || !outs[i]) {
            set_error("alto_mttkrp_batched: null argument in job %d", i);
            return ALTO_ERR_ARG;
        }
        if (modes[i] < 0 || modes[i] >= t->nmode || rank < 1 || rank > 128) {
            set_error("alto_mttkrp_batched: job %d: mode %d / rank %llu out of range (rank 1..128)", i, modes[i],
                      (unsigned long long)rank);
            return ALTO_ERR_ARG;
        }
        for (int m = 0; m < t->nmode; ++m)
            if (!factors[i][m]) {
                set_error("alto_mttkrp_batched: job %d: null factor %d", i, m);
                return ALTO_ERR_ARG;
            }
    }
    if (g_det) {  // one deterministic MTTKRP per job
        for (int i = 0; i < njobs; ++i)
            ALTO_TRY(mttkrp_device(tensors[i], modes[i], R, factors[i], outs[i], precision, accumulate, s, 0));
        return ALTO_OK;
    }
    MttkrpTune k;
    k.variant = g_variant;
    k.nt = g_nt;
    k.hot_block = g_hot_block;
    k.hot_lds = g_hot_lds;
    k.narrow_max = g_narrow_max;
    k.narrow_small = g_narrow_small;
    std::vector<const void*> vals(njobs);
    for (int i = 0; i < njobs; ++i) {
        alto_tensor* t = tensors[i];
        ALTO_TRY(mttkrp_prepare(t, R, precision, s));
        const uint64_t nout = t->dims[modes[i]] * (uint64_t)R;
        if (precision == ALTO_PREC_F64) {
            vals[i] = t->ms[modes[i]].val64.as<double>();
            if (!accumulate) ALTO_TRY(fill_zero((double*)outs[i], nout, s));
        } else {
            vals[i] = t->ms[modes[i]].val32.as<float>();
            if (!accumulate) ALTO_TRY(fill_zero((double*)outs[i], nout, s));
        }
    }
    if (njobs == 0) return ALTO_OK;
    if (precision == ALTO_PREC_F64)
        return launch_mttkrp_batch_f64(njobs, tensors, modes, R, (const double* const* const*)factors,
                                       (const double* const*)vals.data(), (double* const*)outs, g_seg, k, s);
    if (precision == ALTO_PREC_F32)
        return launch_mttkrp_batch_f32(njobs, tensors, modes, R, (const float* const* const*)factors,
                                       (const float* const*)vals.data(), (double* const*)outs, g_seg, k, s);
    return launch_mttkrp_batch_bf16(njobs, tensors, modes, R, (const bf16_t* const* const*)factors,
                                    (const float* const*)vals.data(), (double* const*)outs, g_seg, k, s);
}

}  // namespace alto

extern "C" {

int alto_phi(alto_tensor_t* t, int mode, uint64_t rank, const double* const* factors, double* phi, double eps_div_zero,
             alto_stream_t stream)
{
    if (!t || !factors || !phi) {
        set_error("alto_phi: null argument");
        return ALTO_ERR_ARG;
    }
    ALTO_TRY(phi_device(t, mode, (int)rank, factors, phi, eps_div_zero, (hipStream_t)stream, 0));
    return t->note_use((hipStream_t)stream);
}

}  // extern "C"
