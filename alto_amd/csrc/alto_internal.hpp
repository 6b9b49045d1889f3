// Internal definitions shared by the HIP translation units of libalto_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "alto_hip.h"

#define ALTO_MAXN 5

namespace alto {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
void set_error_hip(hipError_t e, const char* what);

#define HIP_TRY(expr)                                             \
    do {                                                          \
        hipError_t _e = (expr);                                   \
        if (_e != hipSuccess) {                                   \
            ::alto::set_error_hip(_e, #expr);                     \
            return ALTO_ERR_HIP;                                  \
        }                                                         \
    } while (0)

#define ALTO_TRY(expr)                       \
    do {                                     \
        int _s = (expr);                     \
        if (_s != ALTO_OK) return _s;        \
    } while (0)

// ---------------------------------------------------------------- device memory
// Thin owning wrapper; all device allocations of the library go through it.
// Blocks go back to a cache shared by all streams on release, so a buffer
// must be released only after the work that uses it has completed.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    size_t cap = 0;   // size class actually reserved
    int device = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int alloc(size_t n);
    void release();
    void swap(DevBuf& o)
    {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(cap, o.cap);
        std::swap(device, o.device);
    }
    template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

// ---------------------------------------------------------------- tensor layout
// Packed-coordinate layout used by the per-mode streams: mode n occupies bits
// [off[n], off[n]+bits[n]) of a 64-bit word (pw==1) or of a 128-bit pair (pw==2).
struct PackLayout {
    int nmode;
    int pw;                   // words per packed coordinate (1 or 2)
    int bits[ALTO_MAXN];
    int off[ALTO_MAXN];
};

// Byte lookup table that applies a fixed bit permutation to a W-word key:
// out = OR_b table[b][byte_b(in)], each entry W words.
struct BitPermLut {
    int in_words = 1, out_words = 1;
    int nb = 0;   // input bytes that carry index bits
    std::vector<uint64_t> host;   // [in_words*8][256][out_words]
    DevBuf dev;
};

struct ModeStream {
    // Nonzeros of one ALTO chunk sorted by this mode's row (stable), per chunk.
    DevBuf pk;      // nnz * pw uint64
    DevBuf val64;   // nnz double (present when FP64 arithmetic was prepared)
    DevBuf val32;   // nnz float  (present when FP32/BF16 arithmetic was prepared)
    int chunk_log2 = -1;  // chunk size the stream was ordered with
    // 1: the fields of the other modes hold frequency ranks (rank 0 = the
    // row with the most nonzeros) instead of rows; the kernels then read
    // the other factors through row copies permuted into rank order
    int ranked = 0;
    // Hot-pattern split of a ranked stream: the nonzeros are grouped by their
    // pattern (bit o set when the rank of the o-th other mode is below
    // hot_k[o]), then by chunk and row; pattern p occupies
    // [pat_start[p], pat_start[p+1]).  split = 0: not grouped.
    int split = 0;
    int hot_k[ALTO_MAXN - 1] = {0};
    std::vector<uint64_t> pat_start;
    // Fiber layout (fib_mode >= 0): the stream is ordered by (chunk, row of
    // this mode, row of mode fib_mode) and split in two parts that each keep
    // that order: first the nonzeros of fibers -- runs of equal (chunk, row,
    // fiber-mode row) -- of at least fib_min nonzeros (nlong of them), then
    // the others.  With a second fiber mode (fib_mode2 >= 0) the others are
    // ordered and split again in the same way by (chunk, row, row of
    // fib_mode2): its fibers (nlong2 nonzeros) follow the first part, the
    // rest comes last.  MTTKRP reads the fiber mode's row once per fiber in
    // the fiber parts.
    int fib_mode = -1;
    int fib_mode2 = -1;
    int fib_min = 0;
    uint64_t nlong = 0;
    uint64_t nlong2 = 0;
};

}  // namespace alto

struct alto_tensor {
    int nmode = 0;
    int W = 1;              // words per ALTO key (LIT width / 64)
    int nprtn = 1;          // reference partition count
    int alto_bits = 0;
    int val_kind = ALTO_VAL_F64;
    uint64_t nnz = 0;
    uint64_t dims[ALTO_MAXN] = {0};
    uint64_t masks[ALTO_MAXN * 2] = {0};
    uint64_t alto_mask[2] = {0};
    uint64_t cr_masks[ALTO_MAXN * 2] = {0};
    uint64_t alto_cr_mask[2] = {0};
    int is_oidx = 0;
    std::vector<uint64_t> prtn_ptr;    // nprtn+1 (host copy)
    std::vector<uint64_t> intervals;   // nprtn*nmode*2 (host copy)

    alto::PackLayout pack;
    alto::BitPermLut lin;     // packed -> ALTO key
    alto::BitPermLut delin;   // ALTO key -> packed

    alto::DevBuf key;         // nnz * W uint64, ALTO order, CR bits included
    alto::DevBuf vals;        // nnz double, ALTO order (ValType cast applied)

    int chunk_log2 = 20;      // log2 nonzeros per chunk of the per-mode streams
    alto::ModeStream ms[ALTO_MAXN];
    int device = 0;

    // Frequency ranking of each mode's rows (descending nonzero count, ties
    // by row): rank_row[m][k] = row of rank k, row_rank[m][r] = rank of row r.
    // Present when the streams are ranked.
    alto::DevBuf rank_row[ALTO_MAXN];
    alto::DevBuf row_rank[ALTO_MAXN];
    // factor rows permuted into rank order for the stream kernels, one buffer
    // per mode (kept, grown on demand)
    alto::DevBuf pcopy[ALTO_MAXN];
    // per-segment row slots of the deterministic-order walk (kept, grown on
    // demand): the first / last row of each segment and their partial sums
    alto::DevBuf det_rows, det_part;

    // MTTKRP launch shapes chosen by alto_mttkrp_autotune for (mode, rank,
    // precision); used when the process-wide knobs are at their defaults
    struct Tuned {
        int mode, rank, precision, variant, seg;
    };
    std::vector<Tuned> tuned;

    // Streams that asynchronous calls enqueued work on for this tensor, each
    // with an event recorded behind that work: the tensor's buffers go back
    // to the shared device cache (destroy, re-ordered streams, grown copies)
    // only after every such event has completed.
    std::mutex use_mu;
    std::vector<std::pair<hipStream_t, hipEvent_t>> uses;
    int note_use(hipStream_t s);  // event behind the work enqueued so far on s
    int wait_uses();              // host wait for every recorded event
    ~alto_tensor();
};

namespace alto {

// create.hip
int build_luts(alto_tensor* t, hipStream_t s);
size_t pool_in_use(int dev);
int compute_masks(int nmode, const uint64_t* dims, int W, uint64_t* masks, uint64_t* alto_mask);
int radix_sort_pairs(uint64_t* keys, int W, uint32_t* payload, uint64_t n, int bits, hipStream_t s);
int delinearize_all(const alto_tensor* t, uint64_t* coords_dev, hipStream_t s);
int prepare_mode_streams(alto_tensor* t, int need64, int need32, hipStream_t s);
int tuned_chunk_log2();
int hot_rows_enabled();
int reserve_rank_copies(alto_tensor* t, uint64_t rank, size_t elem_bytes, hipStream_t s);
int split_stream(alto_tensor* t, int mode, const int* hot_k, hipStream_t s);
int mttkrp_prepare(alto_tensor* t, int R, int precision, hipStream_t s);
// fiber layout settings (mttkrp.hip): 0 when off, else the number of fiber
// modes (1 or 2); the minimum fiber length of the fiber parts, the chunk
// size and the smallest tensor that gets the layout
int fiber_settings(int* min_len, int* chunk_log2, uint64_t* min_nnz);
int fiber_rule();
// deterministic-order reductions (ALTO_TUNE_DETERMINISTIC): streams fully
// sorted by row, MTTKRP / Phi sums in a fixed order
int det_enabled();
// lockstep CP-ALS sessions' dense steps batched (ALTO_TUNE_LOCKSTEP_DENSE_BATCH)
int lockstep_dense_batch_enabled();
// FP32 / BF16 CP-ALS Grams and row solves on FP64 MFMA (ALTO_TUNE_DENSE_F64)
int dense_f64_enabled();
// ... and replayed as a HIP graph while the active set repeats (ALTO_TUNE_LOCKSTEP_GRAPHS)
int lockstep_graphs_enabled();
int apr_pi_cache_enabled();
int mttkrp_det_device(alto_tensor* t, int mode, int R, const void* const* U, double* out, int precision, hipStream_t s);
int phi_det_device(alto_tensor* t, int mode, int R, const double* const* U, double* phi, double eps, hipStream_t s);

// cpd.hip: per-phase event timing of a session's eager iterations (default on)
int cpd_session_set_timing(alto_cpd_session_t* sess, int on);
// the stream a session's later work goes to; waits for the session's
// current stream first
int cpd_session_set_stream(alto_cpd_session_t* sess, hipStream_t s);
// iterations run so far and whether the fit tolerance stopped the session
int cpd_session_state(const alto_cpd_session_t* sess, int* iterations, int* converged);
// server: jobs of at most *max_nnz nonzeros (0: none) iterated in lockstep
// groups, at most *max_sessions of them at once (ALTO_TUNE_SERVE_LOCKSTEP_*)
void serve_lockstep_settings(uint64_t* max_nnz, int* max_sessions);

}  // namespace alto
