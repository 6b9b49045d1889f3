/*
 * libalto_hip — C ABI of the MI355X (gfx950) ALTO engine.
 *
 * Drop-in boundary for the reference's header-only templates (IntelLabs/ALTO):
 * each entry point names the reference interface it replaces.  Plain pointers
 * and sizes only; every call returns an alto status code (0 = success) and
 * alto_last_error() describes the last failure of the calling thread.
 *
 * Pointer kinds: "host" buffers are ordinary CPU memory (the call
 * synchronizes before returning, as the reference's synchronous OpenMP calls
 * do); "device" buffers are HIP device memory of the current device.
 */
#ifndef ALTO_HIP_H_
#define ALTO_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct alto_tensor alto_tensor_t;
typedef void* alto_stream_t; /* hipStream_t; NULL = the default stream */

enum {
    ALTO_OK = 0,
    ALTO_ERR_ARG = -1,
    ALTO_ERR_HIP = -2,
    ALTO_ERR_OOM = -3,
    ALTO_ERR_UNSUPPORTED = -4,
    ALTO_ERR_BUSY = -5
};

/* ValType of the ALTO tensor (include/common.hpp:62-69) */
enum { ALTO_VAL_F64 = 0, ALTO_VAL_U64 = 1 };
/* arithmetic of MTTKRP / CP iterations */
enum { ALTO_PREC_F64 = 0, ALTO_PREC_F32 = 1, ALTO_PREC_BF16 = 2 };
enum { ALTO_PTR_HOST = 0, ALTO_PTR_DEVICE = 1 };

#define ALTO_MAX_MODES 5

const char* alto_last_error(void);
const char* alto_version(void);

/* ---- ALTO format -------------------------------------------------------- */

/* create_alto<LIT>(SparseTensor*, AltoTensor<LIT>**, int nprtn)
 * (include/alto.hpp:108-294).  cidx holds nmodes pointers to nnz zero-based
 * uint64 coordinates (SparseTensor::cidx, include/sptensor.hpp:8-26); vals
 * holds nnz doubles.  lit_bits selects the ALTO key width (64 or 128, 0 =
 * smallest that fits).  nprtn is the reference partition count (the OpenMP
 * thread count at src/main.cpp:353) used for prtn_ptr / intervals / CR bits.
 * Linearization, sort and partitioning run on the device. */
int alto_tensor_create(int nmodes, const uint64_t* dims, uint64_t nnz,
                       const uint64_t* const* cidx, const double* vals, int ptr_kind,
                       int val_kind, int lit_bits, int nprtn, alto_stream_t stream,
                       alto_tensor_t** out);

/* destroy_alto<LIT> (include/alto.hpp:296-322) */
int alto_tensor_destroy(alto_tensor_t* t);

typedef struct {
    int nmodes, lit_words, alto_bits, nprtn, is_oidx, val_kind;
    uint64_t nnz;
    uint64_t dims[ALTO_MAX_MODES];
    uint64_t mode_masks[ALTO_MAX_MODES * 2]; /* lit_words words per mode */
    uint64_t alto_mask[2];
    uint64_t cr_masks[ALTO_MAX_MODES * 2];
    uint64_t alto_cr_mask[2];
    uint64_t device_bytes;
} alto_tensor_info_t;

/* AltoTensor<LIT> scalar fields (include/alto.hpp:27-54) */
int alto_tensor_info(const alto_tensor_t* t, alto_tensor_info_t* info);

/* Copies AltoTensor<LIT>::idx (nnz*lit_words words, CR bits included),
 * vals (as double), prtn_ptr (nprtn+1) and prtn_intervals (nprtn*nmodes
 * {start,stop} pairs) into host buffers; any pointer may be NULL. */
int alto_tensor_export(const alto_tensor_t* t, uint64_t* idx, double* vals,
                       uint64_t* prtn_ptr, uint64_t* intervals);

/* Delinearized coordinates (pext of every key per mode, include/alto.hpp:431-455),
 * nmodes x nnz, into a host buffer. */
int alto_tensor_delinearize(const alto_tensor_t* t, uint64_t* coords);

/* AltoTensor<LIT>::oidx[mode] (include/alto.hpp:246-285): positions sorted by
 * the mode's row index; ties in ascending position.  Host buffer of nnz. */
int alto_tensor_export_oidx(const alto_tensor_t* t, int mode, uint64_t* oidx);

/* Device pdep / pext over arrays (include/bitops.hpp): keys of lit_words
 * words.  alto_pdep_keys: coords (nmodes x n, device) -> keys (n*lit_words,
 * device); alto_pext_keys: keys -> coords. */
int alto_pdep_keys(const alto_tensor_t* t, const uint64_t* coords, uint64_t n, uint64_t* keys,
                   alto_stream_t stream);
int alto_pext_keys(const alto_tensor_t* t, const uint64_t* keys, uint64_t n, uint64_t* coords,
                   alto_stream_t stream);

/* ---- MTTKRP --------------------------------------------------------------- */

/* mttkrp_alto_par(target_mode, factors, rank, at, wlocks, ofibs)
 * (include/alto.hpp:617-623).  factors: nmodes device pointers to row-major
 * I_n x rank matrices of the precision's storage type (double / float /
 * bf16).  out: device I_mode x rank doubles in every precision (FP32 / BF16:
 * products and runs of one output row summed in FP32 registers, the row
 * partials added into the FP64 output).  accumulate != 0 adds into out as
 * the reference does; 0 overwrites it.  Asynchronous on `stream`. */
int alto_mttkrp(alto_tensor_t* t, int mode, uint64_t rank, const void* const* factors, void* out,
                int precision, int accumulate, alto_stream_t stream);

/* Many MTTKRPs in one launch (no reference counterpart: the reference runs
 * one mttkrp_alto_par per call, include/alto.hpp:617-623; this is the
 * multi-tensor form for batches of small decompositions).  Job i is
 * alto_mttkrp(tensors[i], modes[i], rank, factors[i], outs[i], precision,
 * accumulate): factors[i] holds that tensor's nmodes device pointers.  Jobs
 * that take the vector-lane kernel with the same lanes per nonzero, elements
 * per lane and key width share a launch of up to 16 jobs; each job keeps its
 * own segment length (ALTO_TUNE_SEG when set, else the size rule) and grid,
 * laid end to end with the others'.  Other jobs are launched one by one.
 * Shapes chosen by alto_mttkrp_autotune are not used here.  Each output equals that job's alto_mttkrp up to the order of
 * the floating-point atomic additions.  Asynchronous on `stream`. */
int alto_mttkrp_batched(int njobs, alto_tensor_t* const* tensors, const int* modes, uint64_t rank,
                        const void* const* const* factors, void* const* outs, int precision, int accumulate,
                        alto_stream_t stream);

/* mttkrp / mttkrp_par on a coordinate tensor (include/cpd.hpp:460-543),
 * without building an ALTO tensor: out (dims[mode] x rank, double) += the
 * sum over the nnz nonzeros of vals[i] times the Hadamard product of rows
 * cidx[m][i] of factors[m], m != mode (FP64, atomic adds in any order);
 * accumulate == 0 zeroes out first.  factors[mode] is not read (it may be
 * NULL or alias out).  cidx: nmodes arrays of zero-based coordinates.  All
 * pointers of ptr_kind: host coordinates are range-checked (ALTO_ERR_ARG),
 * device ones skip a nonzero with a coordinate out of range.  ALTO_PTR_HOST: synchronous;
 * ALTO_PTR_DEVICE: asynchronous on `stream`. */
int alto_coo_mttkrp(int nmodes, const uint64_t* dims, uint64_t nnz, const uint64_t* const* cidx,
                    const double* vals, int mode, uint64_t rank, const double* const* factors, double* out,
                    int accumulate, int ptr_kind, alto_stream_t stream);

/* Tuning knobs of the stream kernels (no reference counterpart; the
 * reference specializes at compile time, config.mk:25-26).  ALTO_TUNE_SEG:
 * nonzeros walked by one lane group (0 = default).  ALTO_TUNE_CHUNK_LOG2:
 * log2 of the ALTO chunk inside which each mode's stream is row-ordered
 * (0 = default); streams are re-ordered on their next use.
 * ALTO_TUNE_NONTEMPORAL: 1 = stream the nonzeros with non-temporal loads.
 * ALTO_TUNE_MTTKRP_VARIANT: lane mapping of the MTTKRP kernel: 0 = 2 strided
 * columns per lane, 1 = 4 contiguous columns per lane, 2 = one vector load
 * per lane and row, 16 bytes (8 for BF16) (needs rows aligned to that width,
 * else 0 is used); -1 (default) = 2.
 * ALTO_TUNE_GRAPHS: 1 = CP-ALS sessions capture one iteration as a HIP graph
 * and replay it (non-verbose iterate calls; phase timers are then not kept).
 * ALTO_TUNE_HOT_ROWS: 1 = the streams address the other modes' factor rows
 * by frequency rank and MTTKRP workgroups stage the most frequent rows of
 * each input factor in LDS: with the 16-byte mapping on tensors of up to 3
 * modes each stream is grouped by hot pattern (which of its rows are in
 * LDS), one launch per group; with the other mappings each row read tests
 * its rank.  0 (default) = rows from global memory.  Streams are re-ordered
 * on their next use.
 * ALTO_TUNE_HOT_BLOCK: threads per workgroup of the per-row-test kernel
 * (256/512/1024).
 * ALTO_TUNE_HOT_LDS: LDS bytes per workgroup for the staged rows (default
 * 32768).
 * ALTO_TUNE_NARROW_RANK: FP32 / BF16 ranks up to this value (0..128,
 * default 32) use two columns per lane in the vector-lane mapping (fewer
 * lane groups per wave, so fewer rounds of FP64 row flushes).
 * ALTO_TUNE_PHI_VARIANT: CP-APR Phi kernel, 0 = runtime batch loop with LDS
 * permutes, 1 = compile-time batch with DPP broadcasts (default), 2 = 16-byte
 * lanes with the model row reloaded only where the row changes (even ranks
 * and 16-byte aligned factors; others take 1).
 * ALTO_TUNE_NARROW_SMALL: 1 (default) = below 4M nonzeros every FP32 / BF16
 * rank uses two columns per lane.
 * ALTO_TUNE_FIBERS: 1 (default) = mode streams of tensors with at least
 * ALTO_TUNE_FIBER_MIN_NNZ nonzeros (default 4M) and 3+ modes are ordered by
 * (row, row of a second "fiber" mode -- the other mode with the fewest
 * rows) inside chunks of 2^ALTO_TUNE_FIBER_CHUNK_LOG2 nonzeros (default 30)
 * and split: the nonzeros of fibers of at least ALTO_TUNE_FIBER_MIN
 * (default 3) nonzeros first, walked by a kernel that reads the fiber-mode
 * factor row once per fiber, then the rest.  2 = the rest is ordered and
 * split again by (row, row of the other mode with the next fewest rows):
 * its fibers follow the first part.  0 = off.
 * ALTO_TUNE_FIBER_MODE_RULE: 0 = the fiber modes are the other modes with
 * the fewest rows, 1 = with the most.
 * ALTO_TUNE_DETERMINISTIC: 1 = MTTKRP (every precision, batched calls
 * included) and Phi sum in one fixed order, with no floating-point atomics:
 * the mode streams are fully sorted by row, each segment of 1024 nonzeros
 * is walked in order, rows inside a segment are written by it alone and the
 * rows split between segments are summed in segment order by a second
 * launch.  Results are then bit-identical from run to run (slower: for
 * parity runs).  Calls in this mode on one tensor share its slot buffer:
 * they must not overlap on different streams.
 * ALTO_TUNE_APR_PI_CACHE: 1 (default) = CP-APR stores, at the first inner
 * iteration of each mode, the product of the other modes' factor rows for
 * every nonzero (nnz x rank doubles, when that fits in half the free device
 * memory and 64 GiB) and streams it in the mode's later inner iterations
 * instead of gathering those rows again (the reference's precomputed rows,
 * include/apr.hpp:1049-1165); 0 = gather them every time.
 * ALTO_TUNE_LOCKSTEP_DENSE_BATCH: 1 (default) = alto_cpd_sessions_iterate
 * at rank <= 64 runs each mode's R x R solves, row solves, normalisations and
 * Gram updates of all its sessions in one launch per kernel, and their fits
 * in one launch, on the caller's stream; 0 = one chain of launches per
 * session over 4 internal streams.  ALTO_TUNE_LOCKSTEP_GRAPHS: 1 = with the
 * batched launches, an iteration whose active sessions repeat the previous
 * iteration's is captured once as a HIP graph and replayed (0, default:
 * eager launches).  ALTO_TUNE_SERVE_LOCKSTEP_NNZ: servers created afterwards
 * iterate jobs of at most this many nonzeros (and rank <= 64) in lockstep
 * groups of one rank, precision and tolerance, one iteration per group at a
 * time, members joining and leaving between iterations (0, default: every
 * job iterated by its worker); ALTO_TUNE_SERVE_LOCKSTEP_SESSIONS (default
 * 64): at most this many such jobs admitted at once.  ALTO_TUNE_DENSE_F64:
 * 1 (default) = FP32 / BF16 CP-ALS forms the Grams of its stored factors and
 * the row solves U = M V^-1 on FP64 MFMA (the MTTKRP stays in the model's
 * precision); 0 = on FP32 MFMA.
 * Streams are re-ordered on their next use. */
enum { ALTO_TUNE_SEG = 1, ALTO_TUNE_CHUNK_LOG2 = 2, ALTO_TUNE_NONTEMPORAL = 3, ALTO_TUNE_MTTKRP_VARIANT = 4,
       ALTO_TUNE_GRAPHS = 5, ALTO_TUNE_HOT_ROWS = 7, ALTO_TUNE_HOT_BLOCK = 8, ALTO_TUNE_HOT_LDS = 9,
       ALTO_TUNE_NARROW_RANK = 10, ALTO_TUNE_PHI_VARIANT = 11, ALTO_TUNE_NARROW_SMALL = 12,
       ALTO_TUNE_FIBERS = 13, ALTO_TUNE_FIBER_MIN = 14, ALTO_TUNE_FIBER_CHUNK_LOG2 = 15,
       ALTO_TUNE_FIBER_MIN_NNZ = 16, ALTO_TUNE_FIBER_MODE_RULE = 17, ALTO_TUNE_DETERMINISTIC = 18,
       ALTO_TUNE_APR_PI_CACHE = 19, ALTO_TUNE_LOCKSTEP_DENSE_BATCH = 20, ALTO_TUNE_LOCKSTEP_GRAPHS = 21,
       ALTO_TUNE_SERVE_LOCKSTEP_NNZ = 22, ALTO_TUNE_SERVE_LOCKSTEP_SESSIONS = 23, ALTO_TUNE_DENSE_F64 = 24 };
int alto_tune_set(int param, int value);

/* Autotuner (no reference counterpart; the reference fixes its shapes at
 * compile time, config.mk:25-26): for every mode of `t`, times `reps`
 * MTTKRPs of rank `rank` in `precision` for each lane mapping (0, 1, 2) and
 * segment length (powers of two from 16 to 4096, within 4x of the default
 * for the tensor's size) against the defaults, and keeps the fastest
 * shape on the tensor; later MTTKRPs of that (mode, rank, precision) on `t`
 * -- CP-ALS sessions included -- use it while ALTO_TUNE_SEG and
 * ALTO_TUNE_MTTKRP_VARIANT are at their defaults.  `factors` are device
 * factor matrices of that precision; `out` is device scratch of
 * max_n dims[n] x rank doubles, overwritten.  `result` (may be NULL) receives nmode entries
 * (variant -1 / seg 0: the defaults were fastest).  Synchronous. */
typedef struct {
    int mode;
    int variant;
    int seg;
    double ms;          /* mean time of the chosen shape */
    double default_ms;  /* mean time with the defaults */
} alto_tune_entry_t;
int alto_mttkrp_autotune(alto_tensor_t* t, uint64_t rank, int precision, const void* const* factors, void* out,
                         int reps, alto_stream_t stream, alto_tune_entry_t* result);

/* ---- CP-ALS --------------------------------------------------------------- */

typedef struct {
    double total, mttkrp, pseudoinv, memcopy, normalize, update, fit; /* seconds */
} alto_cpd_timers_t;

/* cpd_alto<LIT>(AltoTensor*, KruskalModel*, max_iters, epsilon)
 * (include/cpd.hpp:43-186).  U: nmodes host pointers to row-major
 * I_n x rank double factors (initial model in, result out); lambda: rank
 * doubles (out).  fit_trace (may be NULL) receives max_iters doubles.
 * precision selects factor storage / MTTKRP arithmetic; the R x R solve,
 * the fit and ||X||^2 stay FP64.  verbose != 0 prints the reference's
 * per-iteration and timer lines. */
int alto_cpd_als(alto_tensor_t* t, uint64_t rank, int max_iters, double epsilon, int precision,
                 double* const* U, double* lambda, double* fit_trace, int* iters,
                 alto_cpd_timers_t* timers, int verbose, alto_stream_t stream);

/* Device-resident CP-ALS model (no reference counterpart: the reference's
 * cpd_alto keeps its temporaries on the stack of one call).  create uploads
 * the initial factors (host, double) and builds the Grams; iterate runs up
 * to max_iters further iterations of cpd_alto's loop (stopping early when
 * |fit - previous fit| < epsilon after the first iteration) with all work on
 * `stream`; get downloads factors / lambda (any pointer may be NULL) and the
 * last fit.  One session per thread at a time. */
typedef struct alto_cpd_session alto_cpd_session_t;
int alto_cpd_session_create(alto_tensor_t* t, uint64_t rank, int precision, double* const* U,
                            alto_stream_t stream, alto_cpd_session_t** out);
/* As alto_cpd_session_create, with the initial factors drawn on the device:
 * U_n[i] = uniform (0, 1] from SplitMix64 of (seed, n, i) (not glibc rand). */
int alto_cpd_session_create_seeded(alto_tensor_t* t, uint64_t rank, int precision, uint64_t seed,
                                   alto_stream_t stream, alto_cpd_session_t** out);
int alto_cpd_session_iterate(alto_cpd_session_t* s, int max_iters, double epsilon, double* fit_trace,
                             int* iters, int* converged, int verbose);
/* Several sessions iterated in lockstep on one stream (no reference
 * counterpart: the reference runs one cpd_alto per process).  Each
 * iteration runs, for every mode index j, one alto_mttkrp_batched over the
 * sessions that have a mode j and have not converged, then their mode-j
 * updates -- at rank <= 64 each kernel of the update once for all of them
 * (R x R inverse, row solves, lambda, Gram + scaling; groups of 16), above
 * that one chain per session over 4 internal streams joined back to
 * `stream` -- then all their fits, read with one synchronization
 * (ALTO_TUNE_LOCKSTEP_DENSE_BATCH, ALTO_TUNE_LOCKSTEP_GRAPHS).  Every
 * session ends where alto_cpd_session_iterate(max_iters, epsilon) would leave
 * it, up to the order of the MTTKRP's floating-point atomic additions (bit
 * for bit with ALTO_TUNE_DETERMINISTIC).  All sessions share the rank and
 * the precision; earlier work on each session's own stream is waited for;
 * the sessions' graphs and phase timers are not used.  iters (may be null):
 * iterations run per session. */
int alto_cpd_sessions_iterate(int n, alto_cpd_session_t* const* sessions, int max_iters, double epsilon,
                              alto_stream_t stream, int* iters);
int alto_cpd_session_get(alto_cpd_session_t* s, double* const* U, double* lambda, double* fit,
                         alto_cpd_timers_t* timers);
int alto_cpd_session_destroy(alto_cpd_session_t* s);

/* cpd_fit_alto (include/cpd.hpp:189-271) for a model given in host memory */
int alto_cpd_fit(alto_tensor_t* t, uint64_t rank, double* const* U, const double* lambda,
                 double* fit);

/* cpd_fit_alto (include/cpd.hpp:189-271) from its arguments: the fit of the
 * model (lambda, last factor U_last: dims[nmodes-1] x rank) against a tensor
 * of squared norm normX, given the Grams (nmodes rank x rank matrices, only
 * the lower triangle read, as the reference defines it) and the last mode's
 * MTTKRP U_mttkrp.  Pointers of ptr_kind; the arithmetic runs on the
 * device. */
int alto_cpd_fit_parts(int nmodes, const uint64_t* dims, uint64_t rank, const double* const* grams,
                       const double* lambda, const double* U_last, const double* U_mttkrp, double normX, int ptr_kind,
                       double* fit);

/* pseudo_inverse (include/cpd.hpp:546-646): U (device dim x rank, double)
 * <- U V^-1 with V the Hadamard product of grams[m != mode] (device, nmodes
 * consecutive rank x rank matrices).  Cholesky; when it fails, the minimum-
 * norm least-squares solution (*used_fallback = 1). */
int alto_pseudo_inverse(const double* grams, int nmodes, int mode, uint64_t rank, double* U, uint64_t dim,
                        int* used_fallback, alto_stream_t stream);

/* update_gram (src/kruskal_model.cpp:11-24): gram (device, rank x rank
 * double, both triangles) = U^T U for a device factor U (dim x rank) stored
 * as double / float / bf16 per precision; computed on the matrix cores. */
int alto_update_gram(const void* U, uint64_t dim, uint64_t rank, int precision, double* gram,
                     alto_stream_t stream);

/* alto_update_gram without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-block partial Grams live in a static 16 MiB device
 * array that each GPU holds until the process ends; sizes whose partials
 * (one rank x rank FP64 matrix per 256 rows, at most 512 of them) exceed it
 * are rejected with ALTO_ERR_UNSUPPORTED (ranks up to 64 always fit; rank
 * 128 up to 32768 rows).  Asynchronous on `stream`; the result equals
 * alto_update_gram's bit for bit.  Every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU (different streams or threads)
 * race: order them on one stream or synchronize between them. */
int alto_update_gram_noalloc(const void* U, uint64_t dim, uint64_t rank, int precision, double* gram,
                             alto_stream_t stream);

/* ---- CP-APR --------------------------------------------------------------- */

typedef struct {
    double loglik, ls_fit, kkt;
    int outer_iters, total_inner;
    int inner_per_mode[ALTO_MAX_MODES];
    double t_pre, t_apr, t_phi, t_post; /* seconds */
} alto_apr_result_t;

/* cp_apr_alto<LIT>(AltoTensor*, KruskalModel*, max_iters, max_inner_iters,
 * epsilon, eps_div_zero, kappa_tol, kappa, dobench) (include/apr.hpp:984-1376).
 * U / lambda as for alto_cpd_als (FP64). */
int alto_cp_apr(alto_tensor_t* t, uint64_t rank, int max_iters, int max_inner, double epsilon,
                double eps_div_zero, double kappa_tol, double kappa, int dobench, double* const* U,
                double* lambda, alto_apr_result_t* res, int verbose, alto_stream_t stream);

/* cp_apr_alto with a choice of stop rule: ALTO_APR_STOP_ALTO (cp_apr_alto's
 * |KKT - previous KKT| <= epsilon^2, include/apr.hpp:1261), ALTO_APR_BENCH
 * (its dobench: every iteration runs), ALTO_APR_STOP_COO (the COO driver
 * cp_apr_mu_alto / cp_apr_mu_alto_otf, src/apr.cpp:239-779: stop after an
 * outer iteration in which every mode's first inner step met epsilon; that
 * iteration counts).  Otherwise as alto_cp_apr. */
enum { ALTO_APR_STOP_ALTO = 0, ALTO_APR_BENCH = 1, ALTO_APR_STOP_COO = 2 };
int alto_cp_apr_ex(alto_tensor_t* t, uint64_t rank, int max_iters, int max_inner, double epsilon,
                   double eps_div_zero, double kappa_tol, double kappa, int stop_rule, double* const* U,
                   double* lambda, alto_apr_result_t* res, int verbose, alto_stream_t stream);

/* CalculatePhi_alto_par (include/apr.hpp:295-303): Phi (device, I_mode x rank,
 * double) = sum over the mode's nonzeros of (x / max(<B[row,:], pi>, eps)) * pi,
 * B = factors[mode].  factors: nmodes device pointers (double). */
int alto_phi(alto_tensor_t* t, int mode, uint64_t rank, const double* const* factors, double* phi,
             double eps_div_zero, alto_stream_t stream);

/* Kruskal model utilities on the device (pointers of ptr_kind):
 * ALTO_KRUSKAL_L1 -- KruskalModelNormalizeMode (src/kruskal_model.cpp:178-222)
 *   on `mode`, or KruskalModelNormalize (:128-176) on every mode when mode
 *   is -1: columns divided by their L1 norms, lambda multiplied by them;
 * ALTO_KRUSKAL_NORM2 / ALTO_KRUSKAL_NORMMAX -- KruskalModelNorm (:307-341)
 *   MAT_NORM_2 / MAT_NORM_MAX on `mode`: lambda = the columns' 2-norms, or
 *   max(1, largest entry) (signed), columns divided by lambda;
 * ALTO_KRUSKAL_REDISTRIBUTE -- RedistributeLambda (:355-378) on `mode`:
 *   columns multiplied by lambda, lambda = 1.  Synchronous. */
enum { ALTO_KRUSKAL_L1 = 0, ALTO_KRUSKAL_NORM2 = 1, ALTO_KRUSKAL_NORMMAX = 2, ALTO_KRUSKAL_REDISTRIBUTE = 3 };
int alto_kruskal_normalize(int nmodes, const uint64_t* dims, uint64_t rank, double* const* U, double* lambda,
                           int mode, int kind, int ptr_kind, alto_stream_t stream);

/* KruskalModelRandomInit (src/kruskal_model.cpp:109-125) on host buffers:
 * glibc srand(seed) then U[n][i] = rand()/RAND_MAX row-major, lambda = 1. */
int alto_kruskal_random_init(int nmodes, const uint64_t* dims, uint64_t rank, unsigned seed,
                             double* const* U, double* lambda);

/* ---- serving: continuous batching of CP-ALS jobs ----------------------------
 * No reference counterpart (the reference runs one decomposition per
 * process).  A server owns max_concurrent workers, each with its own
 * non-blocking HIP stream.  Submitted jobs wait in arrival order and are
 * admitted as soon as a worker is free and their estimated HBM footprint fits
 * the budget (a smaller later job may pass a head job that does not fit);
 * admitted jobs build their ALTO tensor and run CP-ALS concurrently on the
 * device.  Inputs are copied at submit, so the caller may free them. */
typedef struct alto_server alto_server_t;
typedef struct {
    int nmodes;
    const uint64_t* dims;
    uint64_t nnz;
    const uint64_t* const* cidx; /* nmodes pointers, zero-based coordinates */
    const double* vals;
    int ptr_kind;    /* ALTO_PTR_HOST: copied at submit; ALTO_PTR_DEVICE: used in place,
                        the caller keeps them alive until the job finished */
    uint64_t rank;
    int max_iters;
    double epsilon;  /* stop when |fit delta| < epsilon (after the first iteration) */
    unsigned seed;   /* initial factors: device counter-based generator keyed by seed */
    int precision;   /* ALTO_PREC_* */
    int want_factors; /* keep the final factors for alto_server_result */
} alto_job_desc_t;
typedef struct {
    int status; /* 0 queued, 1 running, 2 done, < 0 error */
    int iters;
    double fit;
    double t_submit, t_start, t_end; /* seconds on the server's monotonic clock */
    uint64_t footprint;              /* estimated HBM bytes while running */
} alto_job_result_t;
typedef struct {
    int queued, running;
    uint64_t hbm_in_use, hbm_peak, hbm_budget;
} alto_server_stats_t;
int alto_server_create(uint64_t hbm_budget_bytes, int max_concurrent, alto_server_t** out);
int alto_server_submit(alto_server_t* s, const alto_job_desc_t* job, uint64_t* job_id);
/* wait != 0 blocks until the job finished; U (nmodes host pointers) and
 * lambda may be NULL. */
int alto_server_result(alto_server_t* s, uint64_t job_id, int wait, alto_job_result_t* res,
                       double* const* U, double* lambda);
int alto_server_release(alto_server_t* s, uint64_t job_id);
int alto_server_stats(alto_server_t* s, alto_server_stats_t* stats);
/* waits for every submitted job, then stops the workers */
int alto_server_destroy(alto_server_t* s);

/* ---- device memory helpers (for callers without their own allocator) ---- */
int alto_device_malloc(void** p, uint64_t bytes);
int alto_device_free(void* p);
int alto_memcpy_h2d(void* dst, const void* src, uint64_t bytes);
int alto_memcpy_d2h(void* dst, const void* src, uint64_t bytes);
/* device to device, asynchronous on `stream` */
int alto_memcpy_d2d(void* dst, const void* src, uint64_t bytes, alto_stream_t stream);
int alto_device_synchronize(void);
/* Returns the library's cached device blocks to the runtime (the library
 * reuses freed buffers internally so that concurrent work never blocks on
 * hipFree). */
int alto_device_trim(void);

#ifdef __cplusplus
}
#endif
#endif /* ALTO_HIP_H_ */
