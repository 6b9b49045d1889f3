// CP-ALS on the device: the per-iteration loop of cpd_alto
// (include/cpd.hpp:43-186) with every step on the GPU and only the fit
// read back per iteration (convergence test and progress line).
//
// Per mode j: MTTKRP into M (overwrite) -> [last mode: keep M for the fit]
// -> V = Hadamard(other Grams), Cholesky (pseudo-inverse fallback) ->
// U_j = M V^-1 with column statistics -> lambda (2-norm at iteration 0,
// max-norm after) -> U_j /= lambda (+ BF16 storage copy) -> Gram_j (MFMA).
// Precisions: FP64 (everything FP64); FP32 (factors, MTTKRP, Grams in FP32,
// Cholesky / solve arithmetic / fit in FP64); BF16 (FP32 master factors,
// BF16 copies feed MTTKRP and the Grams, FP32 accumulation).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <chrono>
#include <type_traits>
#include <vector>

#include "alto_internal.hpp"
#include "dense.hpp"
#include "kernels_common.hpp"

namespace alto {

// Per-phase event timing of eager iterations.  The events are created once
// and reused by every later iteration (a mark is one hipEventRecord); they
// are destroyed with the timer, whose owner has drained the stream first.
struct PhaseTimer {
    bool on = false;
    hipStream_t s = nullptr;
    std::vector<hipEvent_t> ev;  // created on demand, kept
    std::vector<int> phase;      // phase of mark k (-1: none)
    size_t used = 0;             // marks since the last collect
    int init(bool enable, hipStream_t st)
    {
        on = enable;
        s = st;
        used = 0;
        return ALTO_OK;
    }
    int mark(int ph)
    {
        if (!on) return ALTO_OK;
        if (used == ev.size()) {
            // timing only: the marks are read after the stream has been
            // synchronized, so no mark needs the system-scope release (an L2
            // write-back per mark, which stalls the kernels behind it)
            hipEvent_t e = nullptr;
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            ev.push_back(e);
            phase.push_back(-1);
        }
        HIP_TRY(hipEventRecord(ev[used], s));
        phase[used] = ph;
        ++used;
        return ALTO_OK;
    }
    // accumulate the time between consecutive marks into the phase of the
    // later mark (the stream has been synchronized by the caller)
    int collect(double* acc)
    {
        if (!on) return ALTO_OK;
        for (size_t k = 1; k < used; ++k) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ev[k - 1], ev[k]));
            if (phase[k] >= 0) acc[phase[k]] += ms * 1e-3;
        }
        used = 0;
        return ALTO_OK;
    }
    ~PhaseTimer()
    {
        for (auto e : ev) {
            (void)hipEventSynchronize(e);
            (void)hipEventDestroy(e);
        }
    }
};

enum { PH_MTTKRP = 0, PH_PINV, PH_COPY, PH_NORM, PH_UPDATE, PH_FIT, PH_N };

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename UT, typename ST, typename MT>
struct CpdState {
    int N = 0, R = 0, prec = 0;
    DevBuf U[ALTO_MAXN], S[ALTO_MAXN];
    DevBuf M, scratch, grams, gpart, L, P, wsA, wsQ, status, psum, pmax, pin, pin_gram, lambda, ipart, fitbuf, nx,
        nxpart;
    int nb_solve = 1;
    // storage that rounds (BF16 factors in FP32 models): the fit's inner
    // product is taken from the stored values in the last mode's Gram pass
    // (pin_gram), so that the fit is that of the model the MTTKRPs see
    static constexpr bool rounds = !std::is_same<UT, ST>::value;
    UT* u(int m) { return U[m].as<UT>(); }
    ST* st(int m) { return std::is_same<UT, ST>::value ? reinterpret_cast<ST*>(U[m].p) : S[m].as<ST>(); }
};

template <typename UT, typename ST, typename MT>
static int cpd_alloc(alto_tensor* t, int R, CpdState<UT, ST, MT>& c)
{
    c.N = t->nmode;
    c.R = R;
    uint64_t maxd = 0, maxrows_solve = 0;
    for (int m = 0; m < t->nmode; ++m) {
        maxd = std::max<uint64_t>(maxd, t->dims[m]);
        ALTO_TRY(c.U[m].alloc(t->dims[m] * R * sizeof(UT)));
        if (!std::is_same<UT, ST>::value) ALTO_TRY(c.S[m].alloc(t->dims[m] * R * sizeof(ST)));
    }
    maxrows_solve = maxd;
    c.nb_solve = solve_blocks(maxrows_solve, R);
    ALTO_TRY(c.M.alloc(maxd * R * sizeof(MT)));
    ALTO_TRY(c.grams.alloc((size_t)t->nmode * R * R * 8));
    ALTO_TRY(c.gpart.alloc((size_t)gram_blocks(maxd) * R * R * 8));
    ALTO_TRY(c.L.alloc((size_t)R * R * 8));
    ALTO_TRY(c.P.alloc((size_t)R * R * 8));
    ALTO_TRY(c.wsA.alloc((size_t)R * R * 8));
    ALTO_TRY(c.wsQ.alloc((size_t)R * R * 8));
    ALTO_TRY(c.status.alloc(8));
    ALTO_TRY(c.psum.alloc((size_t)c.nb_solve * R * 8));
    ALTO_TRY(c.pmax.alloc((size_t)c.nb_solve * R * 8));
    ALTO_TRY(c.pin.alloc((size_t)c.nb_solve * 8));
    ALTO_TRY(c.pin_gram.alloc((size_t)gram_blocks(maxd) * 8));
    ALTO_TRY(c.lambda.alloc((size_t)R * 8));
    ALTO_TRY(c.ipart.alloc((size_t)64 * R * 8));
    ALTO_TRY(c.fitbuf.alloc(4 * 8));
    ALTO_TRY(c.nx.alloc(8));
    ALTO_TRY(c.nxpart.alloc(256 * 8));
    return ALTO_OK;
}

// Grams on FP64 MFMA for FP64 models and, with ALTO_TUNE_DENSE_F64, for
// FP32 / BF16 models (from their stored factors)
static bool gram_f64(int prec) { return prec == ALTO_PREC_F64 || dense_f64_enabled(); }

template <typename UT, typename ST, typename MT>
static int cpd_upload(alto_tensor* t, CpdState<UT, ST, MT>& c, double* const* Uh, hipStream_t s,
                      uint64_t dev_seed = 0)
{
    const int R = c.R;
    std::vector<UT> tmp;
    DevBuf ones;
    ALTO_TRY(ones.alloc(R * 8));
    ALTO_TRY(fill_value<double>(ones.as<double>(), R, 1.0, s));
    for (int m = 0; m < c.N; ++m) {
        uint64_t n = t->dims[m] * R;
        if (!Uh) {  // seeded on the device (counter-based), then converted
            DevBuf d64;
            double* dst = std::is_same<UT, double>::value ? reinterpret_cast<double*>(c.U[m].p) : nullptr;
            if (!dst) {
                ALTO_TRY(d64.alloc(n * 8));
                dst = d64.as<double>();
            }
            ALTO_TRY(random_factor_device(dst, n, dev_seed, m, s));
            if (!std::is_same<UT, double>::value)
                ALTO_TRY((scale_device<double, UT>(dst, t->dims[m], R, ones.as<double>(), c.u(m), s)));
            HIP_TRY(hipStreamSynchronize(s));
        } else if (std::is_same<UT, double>::value) {
            HIP_TRY(hipMemcpyAsync(c.U[m].p, Uh[m], n * 8, hipMemcpyHostToDevice, s));
        } else {
            tmp.resize(n);
            for (uint64_t k = 0; k < n; ++k) tmp[k] = (UT)Uh[m][k];
            HIP_TRY(hipMemcpyAsync(c.U[m].p, tmp.data(), n * sizeof(UT), hipMemcpyHostToDevice, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        if (!std::is_same<UT, ST>::value)
            ALTO_TRY((scale_device<UT, ST>(c.u(m), t->dims[m], R, ones.as<double>(), c.st(m), s)));
        ALTO_TRY(gram_device<ST>(c.st(m), t->dims[m], R, c.grams.template as<double>() + (size_t)m * R * R,
                                 c.gpart.template as<double>(), s, gram_f64(c.prec)));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return ALTO_OK;
}

template <typename UT, typename ST, typename MT>
static int cpd_download(alto_tensor* t, CpdState<UT, ST, MT>& c, double* const* Uh, double* lambda_h, hipStream_t s)
{
    const int R = c.R;
    std::vector<UT> tmp;
    for (int m = 0; m < c.N; ++m) {
        uint64_t n = t->dims[m] * R;
        if (std::is_same<UT, double>::value) {
            HIP_TRY(hipMemcpyAsync(Uh[m], c.U[m].p, n * 8, hipMemcpyDeviceToHost, s));
        } else {
            tmp.resize(n);
            HIP_TRY(hipMemcpyAsync(tmp.data(), c.U[m].p, n * sizeof(UT), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            for (uint64_t k = 0; k < n; ++k) Uh[m][k] = (double)tmp[k];
        }
    }
    if (lambda_h) HIP_TRY(hipMemcpyAsync(lambda_h, c.lambda.p, R * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ALTO_OK;
}

template <typename UT, typename ST, typename MT>
static int cpd_mode_rest(alto_tensor* t, CpdState<UT, ST, MT>& c, int j, int it, PhaseTimer& pt, hipStream_t s);

// one mode update (cpd_alto's loop body, include/cpd.hpp:103-138)
template <typename UT, typename ST, typename MT>
static int cpd_mode_step(alto_tensor* t, CpdState<UT, ST, MT>& c, int j, int it, PhaseTimer& pt, hipStream_t s)
{
    const int R = c.R, N = c.N;
    const void* fptr[ALTO_MAXN];
    for (int m = 0; m < N; ++m) fptr[m] = c.st(m);
    // the modes of an iteration run back to back: mode j > 0 starts at the
    // previous mode's last mark
    if (j == 0) ALTO_TRY(pt.mark(-1));
    // M is zero here: cleared at set-up and by the previous mode's last launch
    ALTO_TRY(mttkrp_device(t, j, R, fptr, c.M.p, c.prec, 1, s, 0));
    ALTO_TRY(pt.mark(PH_MTTKRP));
    return cpd_mode_rest(t, c, j, it, pt, s);
}

// the mode update after its MTTKRP (M holds M_j): R x R solve, row solves,
// λ, Gram + scaling (which also clears M for the next mode)
template <typename UT, typename ST, typename MT>
static int cpd_mode_rest(alto_tensor* t, CpdState<UT, ST, MT>& c, int j, int it, PhaseTimer& pt, hipStream_t s)
{
    const int R = c.R, N = c.N;
    ALTO_TRY(chol_device(c.grams.template as<double>(), N, j, R, c.L.template as<double>(), c.P.template as<double>(),
                         c.wsA.template as<double>(), c.wsQ.template as<double>(), c.status.template as<int>(), s));
    // the last mode's solve also leaves the fit's inner-product partials
    // (storage that rounds: the Gram pass below does)
    const bool last = j == N - 1;
    ALTO_TRY((solve_device<MT, UT>(c.M.template as<MT>(), t->dims[j], R, c.L.template as<double>(),
                                   c.P.template as<double>(), c.status.template as<int>(), c.u(j),
                                   c.psum.template as<double>(), c.pmax.template as<double>(), c.nb_solve, s,
                                   last && !c.rounds ? c.pin.template as<double>() : nullptr)));
    ALTO_TRY(pt.mark(PH_PINV));
    ALTO_TRY(lambda_device(c.psum.template as<double>(), c.pmax.template as<double>(), c.nb_solve, R, it == 0 ? 0 : 1,
                           c.lambda.template as<double>(), s));
    ALTO_TRY(pt.mark(PH_NORM));
    // ... and clears M for the next mode's MTTKRP in the same launches
    const uint64_t next_rows = t->dims[(j + 1) % N];
    const bool inner = last && c.rounds;
    static_assert(std::is_same<MT, double>::value, "the Gram pass reads an FP64 M");
    ALTO_TRY((gram_scale_device<UT, ST>(
        c.u(j), t->dims[j], R, c.lambda.template as<double>(), std::is_same<UT, ST>::value ? nullptr : c.st(j),
        c.grams.template as<double>() + (size_t)j * R * R, c.gpart.template as<double>(), s, gram_f64(c.prec), c.M.p,
        next_rows * R * sizeof(MT), inner ? c.M.template as<double>() : nullptr,
        inner ? (R <= 64 ? c.pin_gram.template as<double>() : c.ipart.template as<double>()) : nullptr)));
    ALTO_TRY(pt.mark(PH_UPDATE));
    return ALTO_OK;
}

// fit after an iteration: the inner product came out of the last mode's
// solve (storage that rounds: out of its Gram pass)
template <typename UT, typename ST, typename MT>
static int cpd_fit_step(alto_tensor* t, CpdState<UT, ST, MT>& c, double normX, hipStream_t s)
{
    if (c.rounds && c.R > 64)
        return fit_device(c.grams.template as<double>(), c.N, c.R, c.lambda.template as<double>(),
                          c.ipart.template as<double>(), normX, c.fitbuf.template as<double>(), s);
    if (c.rounds)
        return fit_scalar_device(c.grams.template as<double>(), c.N, c.R, c.lambda.template as<double>(),
                                 c.pin_gram.template as<double>(), gram_blocks(t->dims[c.N - 1]), normX,
                                 c.fitbuf.template as<double>(), s);
    return fit_scalar_device(c.grams.template as<double>(), c.N, c.R, c.lambda.template as<double>(),
                             c.pin.template as<double>(), c.nb_solve, normX, c.fitbuf.template as<double>(), s);
}

// A CP-ALS model resident on the device (factors, Grams, workspaces) that
// can be iterated in several calls; the iteration counter and the previous
// fit persist, so iterate(a) then iterate(b) equals iterate(a + b).
struct CpdSessionBase {
    alto_tensor* t = nullptr;
    int R = 0, prec = 0, it = 0;
    double prev_fit = 0, fit = 0, normX = 0;
    bool converged = false;
    // per-phase event timing of the eager iterations (reported by
    // alto_cpd_session_get); off for callers that do not read it
    bool timed = true;
    hipStream_t s = nullptr;
    double acc[PH_N] = {0};
    double total = 0;
    PhaseTimer pt;  // events reused by every eager iteration
    // event behind the last work this session enqueued (on any stream):
    // destruction waits for it before the buffers go back to the cache
    hipEvent_t last = nullptr;
    int fence(hipStream_t st)
    {
        if (!last) HIP_TRY(hipEventCreateWithFlags(&last, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(last, st));
        return ALTO_OK;
    }
    virtual ~CpdSessionBase()
    {
        if (last) {
            (void)hipEventSynchronize(last);
            (void)hipEventDestroy(last);
        }
    }
    virtual int iterate(int max_iters, double eps, double* fit_trace, int* ran, int verbose) = 0;
    virtual int get(double* const* Uh, double* lambda_h) = 0;
    // pieces of one eager iteration, for iterating several sessions in
    // lockstep on one stream: the storage factors and M of mode j, the rest
    // of mode j after its MTTKRP, the fit (left in the session's fit buffer)
    virtual int nmodes() const = 0;
    virtual void mode_io(int j, const void** fptr, void** M) = 0;
    virtual int mode_rest(int j, hipStream_t st) = 0;
    virtual int fit_launch(hipStream_t st, double* fit_dst) = 0;
    // the same pieces as jobs of the batched dense launches (rank <= 64):
    // the rest of mode j, and the fit (its value also copied to *gather)
    virtual void mode_job(int j, DenseModeJob* x) = 0;
    virtual void fit_job(double* gather, DenseFitJob* f) = 0;
};

int graphs_enabled();

template <typename UT, typename ST, typename MT>
struct CpdSession : CpdSessionBase {
    CpdState<UT, ST, MT> c;
    // one CP-ALS iteration captured as a graph: [0] first iteration (2-norm),
    // [1] later iterations (max-norm); the fit lands in pinned host memory
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    double* fit_host = nullptr;
    // the graphs are recorded on a private stream (the caller's stream may
    // be the null stream, which cannot be captured) and launched on `s`
    hipStream_t cs = nullptr;
    ~CpdSession() override
    {
        if (last) (void)hipEventSynchronize(last);  // before the buffers (members) are released
        for (auto& e : exec)
            if (e) (void)hipGraphExecDestroy(e);
        if (fit_host) (void)hipHostFree(fit_host);
        if (cs) (void)hipStreamDestroy(cs);
    }
    int capture(int kind)
    {
        if (!fit_host) HIP_TRY(hipHostMalloc((void**)&fit_host, 32, hipHostMallocDefault));
        if (!cs) HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        PhaseTimer off;
        ALTO_TRY(off.init(false, cs));
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        int st = ALTO_OK;
        for (int j = 0; j < c.N && st == ALTO_OK; ++j) st = cpd_mode_step(t, c, j, kind, off, cs);
        if (st == ALTO_OK) st = cpd_fit_step(t, c, normX, cs);
        if (st == ALTO_OK) {
            hipError_t e = hipMemcpyAsync(fit_host, c.fitbuf.p, 24, hipMemcpyDeviceToHost, cs);
            if (e != hipSuccess) {
                set_error_hip(e, "hipMemcpyAsync (capture)");
                st = ALTO_ERR_HIP;
            }
        }
        hipError_t ec = hipStreamEndCapture(cs, &g);
        if (st != ALTO_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (ec != hipSuccess) {
            set_error_hip(ec, "hipStreamEndCapture");
            return ALTO_ERR_HIP;
        }
        hipError_t ei = hipGraphInstantiate(&exec[kind], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            set_error_hip(ei, "hipGraphInstantiate");
            return ALTO_ERR_HIP;
        }
        return ALTO_OK;
    }
    int init(alto_tensor* tt, int rank, int precision, double* const* Uh, hipStream_t st, uint64_t dev_seed = 0)
    {
        t = tt;
        R = rank;
        prec = precision;
        s = st;
        c.prec = prec;
        // stream layout and rank-ordered factor buffers are set up here,
        // outside any graph capture of the iterations
        ALTO_TRY(mttkrp_prepare(t, R, prec, s));
        ALTO_TRY(reserve_rank_copies(t, (uint64_t)R, 8, s));
        ALTO_TRY(cpd_alloc(t, R, c));
        ALTO_TRY(cpd_upload(t, c, Uh, s, dev_seed));
        ALTO_TRY(fill_value<double>(c.lambda.template as<double>(), R, 1.0, s));
        ALTO_TRY(fill_value<MT>(c.M.template as<MT>(), t->dims[0] * R, (MT)0, s));  // the first MTTKRP accumulates
        ALTO_TRY(normsq_device(t->vals.as<double>(), t->nnz, c.nxpart.template as<double>(), c.nx.template as<double>(), s));
        HIP_TRY(hipMemcpyAsync(&normX, c.nx.p, 8, hipMemcpyDeviceToHost, s));
        ALTO_TRY(fence(s));
        HIP_TRY(hipStreamSynchronize(s));
        return ALTO_OK;
    }
    int nmodes() const override { return c.N; }
    void mode_io(int j, const void** fptr, void** M) override
    {
        for (int m = 0; m < c.N; ++m) fptr[m] = c.st(m);
        *M = c.M.p;
    }
    int mode_rest(int j, hipStream_t st) override
    {
        PhaseTimer off;
        ALTO_TRY(off.init(false, st));
        ALTO_TRY(cpd_mode_rest(t, c, j, it, off, st));
        return fence(st);
    }
    int fit_launch(hipStream_t st, double* fit_dst) override
    {
        ALTO_TRY(cpd_fit_step(t, c, normX, st));
        HIP_TRY(hipMemcpyAsync(fit_dst, c.fitbuf.p, sizeof(double), hipMemcpyDeviceToHost, st));
        return fence(st);
    }
    // the arguments cpd_mode_rest and cpd_fit_step pass to their launches
    void mode_job(int j, DenseModeJob* x) override
    {
        static_assert(std::is_same<MT, double>::value, "the row solves read an FP64 M");
        const int N = c.N;
        x->grams = c.grams.template as<double>();
        x->nmode = N;
        x->mode = j;
        x->P = c.P.template as<double>();
        x->wsA = c.wsA.template as<double>();
        x->wsQ = c.wsQ.template as<double>();
        x->status = c.status.template as<int>();
        x->M = c.M.template as<double>();
        x->I = t->dims[j];
        x->U = c.u(j);
        x->psum = c.psum.template as<double>();
        x->pmax = c.pmax.template as<double>();
        const bool last = j == N - 1;
        x->pin = last && !c.rounds ? c.pin.template as<double>() : nullptr;
        x->Mgram = last && c.rounds ? c.M.template as<double>() : nullptr;
        x->pin_gram = last && c.rounds ? c.pin_gram.template as<double>() : nullptr;
        x->nb = c.nb_solve;
        x->which = it == 0 ? 0 : 1;
        x->lambda = c.lambda.template as<double>();
        x->S = std::is_same<UT, ST>::value ? nullptr : (void*)c.st(j);
        x->G = c.grams.template as<double>() + (size_t)j * R * R;
        x->gpart = c.gpart.template as<double>();
        x->zero = c.M.p;
        x->zero_bytes = t->dims[(j + 1) % N] * R * sizeof(MT);
    }
    void fit_job(double* gather, DenseFitJob* f) override
    {
        f->grams = c.grams.template as<double>();
        f->lambda = c.lambda.template as<double>();
        f->pin = c.rounds ? c.pin_gram.template as<double>() : c.pin.template as<double>();
        f->out = c.fitbuf.template as<double>();
        f->gather = gather;
        f->normX = normX;
        f->nmode = c.N;
        f->nb = c.rounds ? gram_blocks(t->dims[c.N - 1]) : c.nb_solve;
    }
    int iterate(int max_iters, double eps, double* fit_trace, int* ran, int verbose) override
    {
        if (graphs_enabled() && !verbose) return iterate_graph(max_iters, eps, fit_trace, ran);
        ALTO_TRY(pt.init(timed || verbose, s));
        int done = 0;
        const double t0 = now_s();
        for (int k = 0; k < max_iters && !converged; ++k, ++it) {
            const double ti = now_s();
            const double mt_before = acc[PH_MTTKRP];
            for (int j = 0; j < c.N; ++j) ALTO_TRY(cpd_mode_step(t, c, j, it, pt, s));
            // the fit starts at the last mode's last mark
            ALTO_TRY(cpd_fit_step(t, c, normX, s));
            ALTO_TRY(pt.mark(PH_FIT));
            double fb[3];
            HIP_TRY(hipMemcpyAsync(fb, c.fitbuf.p, 24, hipMemcpyDeviceToHost, s));
            ALTO_TRY(fence(s));
            HIP_TRY(hipStreamSynchronize(s));
            ALTO_TRY(pt.collect(acc));
            fit = fb[0];
            if (fit_trace) fit_trace[k] = fit;
            ++done;
            if (verbose)
                printf("it: %d\t fit: %g\t fit-delta: %g\ttime(for MTTKRP): %.4f s (%.4f s)\n", it, fit,
                       fabs(prev_fit - fit), now_s() - ti, acc[PH_MTTKRP] - mt_before);
            if (it > 0 && fabs(prev_fit - fit) < eps) converged = true;
            prev_fit = fit;
        }
        total += now_s() - t0;
        if (ran) *ran = done;
        return ALTO_OK;
    }
    int iterate_graph(int max_iters, double eps, double* fit_trace, int* ran)
    {
        int done = 0;
        const double t0 = now_s();
        for (int k = 0; k < max_iters && !converged; ++k, ++it) {
            const int kind = it == 0 ? 0 : 1;
            if (!exec[kind]) ALTO_TRY(capture(kind));
            HIP_TRY(hipGraphLaunch(exec[kind], s));
            ALTO_TRY(fence(s));
            HIP_TRY(hipStreamSynchronize(s));
            fit = fit_host[0];
            if (fit_trace) fit_trace[k] = fit;
            ++done;
            if (it > 0 && fabs(prev_fit - fit) < eps) converged = true;
            prev_fit = fit;
        }
        total += now_s() - t0;
        if (ran) *ran = done;
        return ALTO_OK;
    }
    int get(double* const* Uh, double* lambda_h) override
    {
        if (Uh) return cpd_download(t, c, Uh, lambda_h, s);
        if (lambda_h) {
            HIP_TRY(hipMemcpyAsync(lambda_h, c.lambda.p, R * 8, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        return ALTO_OK;
    }
};

static int session_new(alto_tensor* t, int R, int prec, double* const* Uh, hipStream_t s, CpdSessionBase** out,
                       uint64_t dev_seed = 0)
{
    if (R < 1 || R > 128) {
        set_error("cpd: rank must be 1..128");
        return ALTO_ERR_ARG;
    }
    CpdSessionBase* b = nullptr;
    int st = ALTO_OK;
    switch (prec) {
    case ALTO_PREC_F64: {
        auto* x = new CpdSession<double, double, double>();
        st = x->init(t, R, prec, Uh, s, dev_seed);
        b = x;
        break;
    }
    case ALTO_PREC_F32: {
        auto* x = new CpdSession<float, float, double>();
        st = x->init(t, R, prec, Uh, s, dev_seed);
        b = x;
        break;
    }
    case ALTO_PREC_BF16: {
        auto* x = new CpdSession<float, bf16_t, double>();
        st = x->init(t, R, prec, Uh, s, dev_seed);
        b = x;
        break;
    }
    default:
        set_error("cpd: unknown precision %d", prec);
        return ALTO_ERR_ARG;
    }
    if (st != ALTO_OK) {
        delete b;
        return st;
    }
    *out = b;
    return ALTO_OK;
}

static void print_totals(const CpdSessionBase* b, int iters)
{
    const double* acc = b->acc;
    printf("Total time (for MTTKRP):\t %.4f s (%.4f s)\n", b->total, acc[PH_MTTKRP]);
    printf("Total     MTTKRP    PseudoInv MemCopy   Normalize Update    Fit\n");
    printf("%07.4f   %07.4f   %07.4f   %07.4f   %07.4f   %07.4f   %07.4f\n", b->total, acc[PH_MTTKRP], acc[PH_PINV],
           acc[PH_COPY], acc[PH_NORM], acc[PH_UPDATE], acc[PH_FIT]);
    if (iters < 1) iters = 1;
    printf("Per iteration\n%07.4f   %07.4f   %07.4f   %07.4f   %07.4f   %07.4f   %07.4f\n", b->total / iters,
           acc[PH_MTTKRP] / iters, acc[PH_PINV] / iters, acc[PH_COPY] / iters, acc[PH_NORM] / iters,
           acc[PH_UPDATE] / iters, acc[PH_FIT] / iters);
}

static void fill_timers(const CpdSessionBase* b, alto_cpd_timers_t* tm)
{
    tm->total = b->total;
    tm->mttkrp = b->acc[PH_MTTKRP];
    tm->pseudoinv = b->acc[PH_PINV];
    tm->memcopy = b->acc[PH_COPY];
    tm->normalize = b->acc[PH_NORM];
    tm->update = b->acc[PH_UPDATE];
    tm->fit = b->acc[PH_FIT];
}

int cpd_als_impl(alto_tensor* t, int R, int max_iters, double eps, int prec, double* const* Uh, double* lambda_h,
                 double* fit_trace, int* iters, alto_cpd_timers_t* tm, int verbose, hipStream_t s)
{
    if (max_iters < 1) {
        set_error("cpd_als: max_iters must be >= 1");
        return ALTO_ERR_ARG;
    }
    CpdSessionBase* b = nullptr;
    ALTO_TRY(session_new(t, R, prec, Uh, s, &b));
    std::unique_ptr<CpdSessionBase> guard(b);
    if (verbose) printf("Running ALTO CP-ALS with %d max iterations and %.2e epsilon\n", max_iters, eps);
    int ran = 0;
    ALTO_TRY(b->iterate(max_iters, eps, fit_trace, &ran, verbose));
    ALTO_TRY(b->get(Uh, lambda_h));
    if (iters) *iters = ran;
    if (tm) fill_timers(b, tm);
    if (verbose) print_totals(b, ran);
    return ALTO_OK;
}

// cpd_fit_alto for a model in host memory: Grams, last-mode MTTKRP, fit
int cpd_fit_impl(alto_tensor* t, int R, double* const* Uh, const double* lambda_h, double* fit)
{
    hipStream_t s = nullptr;
    CpdState<double, double, double> c;
    c.prec = ALTO_PREC_F64;
    ALTO_TRY(prepare_mode_streams(t, 1, 0, s));
    ALTO_TRY(cpd_alloc(t, R, c));
    ALTO_TRY(cpd_upload(t, c, Uh, s));
    HIP_TRY(hipMemcpy(c.lambda.p, lambda_h, R * 8, hipMemcpyHostToDevice));
    const void* fptr[ALTO_MAXN];
    for (int m = 0; m < c.N; ++m) fptr[m] = c.st(m);
    ALTO_TRY(c.scratch.alloc(t->dims[c.N - 1] * R * 8));
    ALTO_TRY(mttkrp_device(t, c.N - 1, R, fptr, c.scratch.p, ALTO_PREC_F64, 0, s, 0));
    ALTO_TRY(normsq_device(t->vals.as<double>(), t->nnz, c.nxpart.template as<double>(), c.nx.template as<double>(), s));
    double normX = 0;
    HIP_TRY(hipMemcpy(&normX, c.nx.p, 8, hipMemcpyDeviceToHost));
    // the model is given, so the inner product comes from its own partials
    ALTO_TRY((inner_partial_device<double, double>(c.u(c.N - 1), c.scratch.as<double>(), t->dims[c.N - 1], R,
                                                   c.ipart.as<double>(), s)));
    ALTO_TRY(fit_device(c.grams.as<double>(), c.N, R, c.lambda.as<double>(), c.ipart.as<double>(), normX,
                        c.fitbuf.as<double>(), s));
    double fb[3];
    HIP_TRY(hipMemcpy(fb, c.fitbuf.p, 24, hipMemcpyDeviceToHost));
    *fit = fb[0];
    return ALTO_OK;
}

}  // namespace alto

int alto::cpd_session_set_timing(alto_cpd_session_t* sess, int on)
{
    if (!sess) return ALTO_ERR_ARG;
    reinterpret_cast<alto::CpdSessionBase*>(sess)->timed = on != 0;
    return ALTO_OK;
}

int alto::cpd_session_set_stream(alto_cpd_session_t* sess, hipStream_t s)
{
    if (!sess) return ALTO_ERR_ARG;
    auto* b = reinterpret_cast<alto::CpdSessionBase*>(sess);
    HIP_TRY(hipStreamSynchronize(b->s));
    b->s = s;
    return ALTO_OK;
}

int alto::cpd_session_state(const alto_cpd_session_t* sess, int* iterations, int* converged)
{
    if (!sess) return ALTO_ERR_ARG;
    const auto* b = reinterpret_cast<const alto::CpdSessionBase*>(sess);
    if (iterations) *iterations = b->it;
    if (converged) *converged = b->converged ? 1 : 0;
    return ALTO_OK;
}

using namespace alto;

extern "C" {

int alto_cpd_als(alto_tensor_t* t, uint64_t rank, int max_iters, double epsilon, int precision, double* const* U,
                 double* lambda, double* fit_trace, int* iters, alto_cpd_timers_t* timers, int verbose,
                 alto_stream_t stream)
{
    if (!t || !U || !lambda) {
        set_error("alto_cpd_als: null argument");
        return ALTO_ERR_ARG;
    }
    return cpd_als_impl(t, (int)rank, max_iters, epsilon, precision, U, lambda, fit_trace, iters, timers, verbose,
                        (hipStream_t)stream);
}

int alto_cpd_session_create(alto_tensor_t* t, uint64_t rank, int precision, double* const* U, alto_stream_t stream,
                            alto_cpd_session_t** out)
{
    if (!t || !U || !out) {
        set_error("alto_cpd_session_create: null argument");
        return ALTO_ERR_ARG;
    }
    CpdSessionBase* b = nullptr;
    ALTO_TRY(session_new(t, (int)rank, precision, U, (hipStream_t)stream, &b));
    *out = reinterpret_cast<alto_cpd_session_t*>(b);
    return ALTO_OK;
}

int alto_cpd_session_create_seeded(alto_tensor_t* t, uint64_t rank, int precision, uint64_t seed,
                                   alto_stream_t stream, alto_cpd_session_t** out)
{
    if (!t || !out) {
        set_error("alto_cpd_session_create_seeded: null argument");
        return ALTO_ERR_ARG;
    }
    CpdSessionBase* b = nullptr;
    ALTO_TRY(session_new(t, (int)rank, precision, nullptr, (hipStream_t)stream, &b, seed));
    *out = reinterpret_cast<alto_cpd_session_t*>(b);
    return ALTO_OK;
}

int alto_cpd_session_iterate(alto_cpd_session_t* sess, int max_iters, double epsilon, double* fit_trace, int* iters,
                             int* converged, int verbose)
{
    if (!sess || max_iters < 0) {
        set_error("alto_cpd_session_iterate: invalid argument");
        return ALTO_ERR_ARG;
    }
    CpdSessionBase* b = reinterpret_cast<CpdSessionBase*>(sess);
    ALTO_TRY(b->iterate(max_iters, epsilon, fit_trace, iters, verbose));
    if (converged) *converged = b->converged ? 1 : 0;
    return ALTO_OK;
}

int alto_cpd_sessions_iterate(int n, alto_cpd_session_t* const* sessions, int max_iters, double epsilon,
                              alto_stream_t stream, int* iters)
{
    if (n < 0 || max_iters < 0 || (n > 0 && !sessions)) {
        set_error("alto_cpd_sessions_iterate: invalid argument");
        return ALTO_ERR_ARG;
    }
    std::vector<CpdSessionBase*> b(n);
    for (int i = 0; i < n; ++i) {
        b[i] = reinterpret_cast<CpdSessionBase*>(sessions[i]);
        if (!b[i]) {
            set_error("alto_cpd_sessions_iterate: null session %d", i);
            return ALTO_ERR_ARG;
        }
        if (b[i]->R != b[0]->R || b[i]->prec != b[0]->prec) {
            set_error("alto_cpd_sessions_iterate: session %d differs in rank or precision from session 0", i);
            return ALTO_ERR_ARG;
        }
        for (int k = 0; k < i; ++k)
            if (b[k] == b[i]) {
                set_error("alto_cpd_sessions_iterate: session %d is listed twice", i);
                return ALTO_ERR_ARG;
            }
        if (iters) iters[i] = 0;
    }
    if (n == 0) return ALTO_OK;
    hipStream_t s = (hipStream_t)stream;
    for (auto* x : b) HIP_TRY(hipStreamSynchronize(x->s));  // earlier work of each session
    // pinned fits, one buffer per thread, grown as needed and kept until the
    // process ends (callers that step their sessions one iteration per call
    // make many calls)
    struct Pinned {
        double* p = nullptr;
        int cap = 0;
    };
    static thread_local Pinned pinned;
    if (pinned.cap < n) {
        if (pinned.p) HIP_TRY(hipHostFree(pinned.p));
        pinned.p = nullptr;
        pinned.cap = 0;
        HIP_TRY(hipHostMalloc((void**)&pinned.p, sizeof(double) * (size_t)std::max(n, 64), hipHostMallocDefault));
        pinned.cap = std::max(n, 64);
    }
    double* fits = pinned.p;
    // the sessions' solve chains are independent: they run on NAUX private
    // streams, forked from `s` after each batched MTTKRP and joined back
    // before the next one (the next MTTKRP accumulates into the M that each
    // chain's last launch clears).  The streams and events are created once
    // per thread and device and kept until the process ends.
    constexpr int NAUX = 4;
    struct Aux {
        int device = -1;
        hipStream_t st[NAUX] = {};
        hipEvent_t fork = nullptr, join[NAUX] = {};
    };
    static thread_local Aux aux_cache[8];
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    Aux& aux = aux_cache[dev & 7];
    const int naux = n > 1 ? NAUX : 0;
    if (naux && aux.device != dev) {
        for (int k = 0; k < NAUX; ++k) {
            HIP_TRY(hipStreamCreateWithFlags(&aux.st[k], hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&aux.join[k], hipEventDisableTiming));
        }
        HIP_TRY(hipEventCreateWithFlags(&aux.fork, hipEventDisableTiming));
        aux.device = dev;
    }
    auto fork = [&]() -> int {
        if (!naux) return ALTO_OK;
        HIP_TRY(hipEventRecord(aux.fork, s));
        for (int k = 0; k < naux; ++k) HIP_TRY(hipStreamWaitEvent(aux.st[k], aux.fork, 0));
        return ALTO_OK;
    };
    auto join = [&]() -> int {
        for (int k = 0; k < naux; ++k) {
            HIP_TRY(hipEventRecord(aux.join[k], aux.st[k]));
            HIP_TRY(hipStreamWaitEvent(s, aux.join[k], 0));
        }
        return ALTO_OK;
    };
    auto lane = [&](int q) { return naux ? aux.st[q % naux] : s; };
    // rank <= 64: each mode's updates and the fits of all sessions in one
    // launch per kernel (in groups of DENSE_BATCH_MAX), on `s`
    const int R = (int)b[0]->R, prec = b[0]->prec;
    const bool dense_batched = lockstep_dense_batch_enabled() && R <= 64;
    DevBuf gathered;
    if (dense_batched) ALTO_TRY(gathered.alloc((size_t)n * sizeof(double)));
    std::vector<DenseModeJob> djobs;
    std::vector<DenseFitJob> fjobs;
    auto dense_groups = [&](hipStream_t st) -> int {
        for (size_t k0 = 0; k0 < djobs.size(); k0 += DENSE_BATCH_MAX) {
            const int cnt = (int)std::min<size_t>(DENSE_BATCH_MAX, djobs.size() - k0);
            if (prec == ALTO_PREC_F64) ALTO_TRY((dense_mode_batch<double, double>(cnt, &djobs[k0], R, true, st)));
            else if (prec == ALTO_PREC_F32)
                ALTO_TRY((dense_mode_batch<float, float>(cnt, &djobs[k0], R, gram_f64(prec), st)));
            else ALTO_TRY((dense_mode_batch<float, bf16_t>(cnt, &djobs[k0], R, gram_f64(prec), st)));
        }
        return ALTO_OK;
    };
    std::vector<alto_tensor_t*> ts;
    std::vector<int> modes;
    std::vector<std::vector<const void*>> fp(n, std::vector<const void*>(ALTO_MAXN));
    std::vector<const void* const*> fpp;
    std::vector<void*> outs;
    std::vector<int> act;
    int maxN = 0;
    // the batched MTTKRP of mode index j over the active sessions, on st
    auto mttkrp_step = [&](int j, hipStream_t st) -> int {
        ts.clear();
        modes.clear();
        fpp.clear();
        outs.clear();
        for (int i : act)
            if (j < b[i]->nmodes()) {
                void* M = nullptr;
                b[i]->mode_io(j, fp[i].data(), &M);
                ts.push_back(b[i]->t);
                modes.push_back(j);
                fpp.push_back(fp[i].data());
                outs.push_back(M);
            }
        // M is zero here (cleared at set-up and by each mode's last launch)
        return mttkrp_batched_impl((int)ts.size(), ts.data(), modes.data(), (uint64_t)R, fpp.data(), outs.data(),
                                   prec, 1, st);
    };
    // one iteration of the active sessions with batched dense launches, on
    // st; the fits land in fits[i]
    auto batched_iteration = [&](hipStream_t st) -> int {
        for (int j = 0; j < maxN; ++j) {
            ALTO_TRY(mttkrp_step(j, st));
            djobs.clear();
            for (int i : act)
                if (j < b[i]->nmodes()) {
                    djobs.emplace_back();
                    b[i]->mode_job(j, &djobs.back());
                }
            ALTO_TRY(dense_groups(st));
        }
        fjobs.clear();
        for (int i : act) {
            fjobs.emplace_back();
            b[i]->fit_job(gathered.as<double>() + i, &fjobs.back());
        }
        for (size_t k0 = 0; k0 < fjobs.size(); k0 += DENSE_BATCH_MAX)
            ALTO_TRY(fit_batch((int)std::min<size_t>(DENSE_BATCH_MAX, fjobs.size() - k0), &fjobs[k0], R, st));
        HIP_TRY(hipMemcpyAsync(fits, gathered.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
        return ALTO_OK;
    };
    // With ALTO_TUNE_LOCKSTEP_GRAPHS an iteration whose active sessions and
    // first-iteration flags repeat the previous iteration's is captured as
    // a HIP graph (on a private stream: `s` may be the null stream) and
    // replayed while they stay the same.  The first iteration of a set runs
    // eagerly, so buffers that a first use allocates exist before a capture.
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        std::vector<int> sig;
        ~StepGraph()
        {
            if (exec) (void)hipGraphExecDestroy(exec);
        }
    } graph;
    static thread_local hipStream_t capture_stream[8] = {};
    const bool use_graph = dense_batched && lockstep_graphs_enabled();
    if (use_graph && !capture_stream[dev & 7])
        HIP_TRY(hipStreamCreateWithFlags(&capture_stream[dev & 7], hipStreamNonBlocking));
    std::vector<int> sig, prev_sig;
    auto capture = [&]() -> int {
        hipStream_t cs = capture_stream[dev & 7];
        if (graph.exec) {
            (void)hipGraphExecDestroy(graph.exec);
            graph.exec = nullptr;
        }
        HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        const int st = batched_iteration(cs);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(cs, &g);
        if (st != ALTO_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (ec != hipSuccess) {
            set_error_hip(ec, "hipStreamEndCapture (lockstep iteration)");
            return ALTO_ERR_HIP;
        }
        const hipError_t ei = hipGraphInstantiate(&graph.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            graph.exec = nullptr;
            set_error_hip(ei, "hipGraphInstantiate (lockstep iteration)");
            return ALTO_ERR_HIP;
        }
        graph.sig = sig;
        return ALTO_OK;
    };
    for (int k = 0; k < max_iters; ++k) {
        act.clear();
        sig.clear();
        maxN = 0;
        for (int i = 0; i < n; ++i)
            if (!b[i]->converged) {
                act.push_back(i);
                sig.push_back(2 * i + (b[i]->it == 0 ? 1 : 0));
                maxN = std::max(maxN, b[i]->nmodes());
            }
        if (act.empty()) break;
        const double t0 = now_s();
        if (use_graph && sig == prev_sig) {
            if (!graph.exec || graph.sig != sig) ALTO_TRY(capture());
            HIP_TRY(hipGraphLaunch(graph.exec, s));
        } else if (dense_batched) {
            ALTO_TRY(batched_iteration(s));
        } else {
            for (int j = 0; j < maxN; ++j) {
                ALTO_TRY(mttkrp_step(j, s));
                ALTO_TRY(fork());
                int q = 0;
                for (int i : act)
                    if (j < b[i]->nmodes()) ALTO_TRY(b[i]->mode_rest(j, lane(q++)));
                ALTO_TRY(join());
            }
            ALTO_TRY(fork());
            int q = 0;
            for (int i : act) ALTO_TRY(b[i]->fit_launch(lane(q++), fits + i));
            ALTO_TRY(join());
        }
        prev_sig = sig;
        if (dense_batched)
            for (int i : act) ALTO_TRY(b[i]->fence(s));
        HIP_TRY(hipStreamSynchronize(s));
        const double dt = now_s() - t0;
        for (int i : act) {
            CpdSessionBase* x = b[i];
            x->fit = fits[i];
            if (x->it > 0 && fabs(x->prev_fit - x->fit) < epsilon) x->converged = true;
            x->prev_fit = x->fit;
            ++x->it;
            x->total += dt;
            if (iters) ++iters[i];
        }
    }
    return ALTO_OK;
}

int alto_cpd_session_get(alto_cpd_session_t* sess, double* const* U, double* lambda, double* fit,
                         alto_cpd_timers_t* timers)
{
    if (!sess) {
        set_error("alto_cpd_session_get: null session");
        return ALTO_ERR_ARG;
    }
    CpdSessionBase* b = reinterpret_cast<CpdSessionBase*>(sess);
    ALTO_TRY(b->get(U, lambda));
    if (fit) *fit = b->fit;
    if (timers) fill_timers(b, timers);
    return ALTO_OK;
}

int alto_cpd_session_destroy(alto_cpd_session_t* sess)
{
    delete reinterpret_cast<CpdSessionBase*>(sess);
    return ALTO_OK;
}

int alto_cpd_fit(alto_tensor_t* t, uint64_t rank, double* const* U, const double* lambda, double* fit)
{
    if (!t || !U || !lambda || !fit) {
        set_error("alto_cpd_fit: null argument");
        return ALTO_ERR_ARG;
    }
    return cpd_fit_impl(t, (int)rank, U, lambda, fit);
}

int alto_cpd_fit_parts(int nmodes, const uint64_t* dims, uint64_t rank, const double* const* grams,
                       const double* lambda, const double* U_last, const double* U_mttkrp, double normX, int ptr_kind,
                       double* fit)
{
    if (nmodes < 1 || nmodes > ALTO_MAXN || !dims || !grams || !lambda || !U_last || !U_mttkrp || !fit || rank < 1 ||
        rank > 128) {
        set_error("alto_cpd_fit_parts: invalid argument");
        return ALTO_ERR_ARG;
    }
    hipStream_t s = nullptr;
    const int R = (int)rank, N = nmodes;
    const uint64_t I = dims[N - 1], n = I * R;
    const hipMemcpyKind k = ptr_kind == ALTO_PTR_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    DevBuf g, lam, u, m, ipart, out;
    ALTO_TRY(g.alloc((size_t)N * R * R * 8));
    ALTO_TRY(lam.alloc((size_t)R * 8));
    ALTO_TRY(u.alloc(n * 8));
    ALTO_TRY(m.alloc(n * 8));
    ALTO_TRY(ipart.alloc((size_t)64 * R * 8));
    ALTO_TRY(out.alloc(32));
    for (int q = 0; q < N; ++q)
        HIP_TRY(hipMemcpyAsync(g.as<double>() + (size_t)q * R * R, grams[q], (size_t)R * R * 8, k, s));
    HIP_TRY(hipMemcpyAsync(lam.p, lambda, (size_t)R * 8, k, s));
    HIP_TRY(hipMemcpyAsync(u.p, U_last, n * 8, k, s));
    HIP_TRY(hipMemcpyAsync(m.p, U_mttkrp, n * 8, k, s));
    ALTO_TRY((inner_partial_device<double, double>(u.as<double>(), m.as<double>(), I, R, ipart.as<double>(), s)));
    ALTO_TRY(fit_device(g.as<double>(), N, R, lam.as<double>(), ipart.as<double>(), normX, out.as<double>(), s));
    double fb[3];
    HIP_TRY(hipMemcpyAsync(fb, out.p, 24, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *fit = fb[0];
    return ALTO_OK;
}

int alto_pseudo_inverse(const double* grams, int nmodes, int mode, uint64_t rank, double* U, uint64_t dim,
                        int* used_fallback, alto_stream_t stream)
{
    if (!grams || !U || rank < 1 || rank > 128 || mode < 0 || mode >= nmodes || nmodes > ALTO_MAXN) {
        set_error("alto_pseudo_inverse: invalid argument");
        return ALTO_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    const int R = (int)rank;
    DevBuf L, P, A, Q, st, psum, pmax;
    const int nb = solve_blocks(dim, R);
    ALTO_TRY(L.alloc((size_t)R * R * 8));
    ALTO_TRY(P.alloc((size_t)R * R * 8));
    ALTO_TRY(A.alloc((size_t)R * R * 8));
    ALTO_TRY(Q.alloc((size_t)R * R * 8));
    ALTO_TRY(st.alloc(8));
    ALTO_TRY(psum.alloc((size_t)nb * R * 8));
    ALTO_TRY(pmax.alloc((size_t)nb * R * 8));
    ALTO_TRY(chol_device(grams, nmodes, mode, R, L.as<double>(), P.as<double>(), A.as<double>(), Q.as<double>(),
                         st.as<int>(), s));
    ALTO_TRY((solve_device<double, double>(U, dim, R, L.as<double>(), P.as<double>(), st.as<int>(), U,
                                           psum.as<double>(), pmax.as<double>(), nb, s)));
    int h = 0;
    HIP_TRY(hipMemcpyAsync(&h, st.p, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (used_fallback) *used_fallback = h;
    return ALTO_OK;
}

int alto_update_gram(const void* U, uint64_t dim, uint64_t rank, int precision, double* gram, alto_stream_t stream)
{
    if (!U || !gram || rank < 1 || rank > 128) {
        set_error("alto_update_gram: invalid argument");
        return ALTO_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    DevBuf part;
    ALTO_TRY(part.alloc((size_t)gram_blocks(dim) * rank * rank * 8));
    int st;
    if (precision == ALTO_PREC_F64)
        st = gram_device<double>((const double*)U, dim, (int)rank, gram, part.as<double>(), s, true);
    else if (precision == ALTO_PREC_F32)
        st = gram_device<float>((const float*)U, dim, (int)rank, gram, part.as<double>(), s, false);
    else
        st = gram_device<bf16_t>((const bf16_t*)U, dim, (int)rank, gram, part.as<double>(), s, false);
    if (st != ALTO_OK) return st;
    HIP_TRY(hipStreamSynchronize(s));
    return ALTO_OK;
}

int alto_update_gram_noalloc(const void* U, uint64_t dim, uint64_t rank, int precision, double* gram,
                             alto_stream_t stream)
{
    if (rank > 128) {
        set_error("alto_update_gram_noalloc: rank %llu above 128", (unsigned long long)rank);
        return ALTO_ERR_ARG;
    }
    return gram_noalloc(U, dim, (int)rank, precision, gram, (hipStream_t)stream);
}

}  // extern "C"
