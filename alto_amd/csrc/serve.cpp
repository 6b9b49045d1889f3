// Continuous-batching CP-ALS server (no reference counterpart: the reference
// runs one decomposition per process).  Jobs (a COO tensor, rank,
// iteration limit, tolerance, seed, precision) are submitted at any time;
// an admission loop admits them in arrival order as soon as a worker slot is
// free and their estimated HBM footprint fits the budget (a later, smaller
// job may pass a head job that does not fit yet).  Each worker owns a
// non-blocking HIP stream and runs create_alto + CP-ALS for its job, so the
// work of all admitted jobs overlaps on the device; finished jobs release
// their footprint and the next arrivals are admitted immediately.
//
// Iteration-level batching (ALTO_TUNE_SERVE_LOCKSTEP_NNZ > 0): a job of at
// most that many nonzeros and rank <= 64 is created by a worker as usual,
// then handed to the stepper thread, which keeps such sessions in groups of
// one rank, precision and tolerance and runs one lockstep iteration of each
// group at a time (alto_cpd_sessions_iterate: one batched MTTKRP per mode
// index, the dense steps of all members in one launch per kernel).  Sessions
// join a group between its iterations and leave it when they converge or
// reach their iteration limit, so the group changes every few iterations
// while the worker that created the session takes the next job.  At most
// ALTO_TUNE_SERVE_LOCKSTEP_SESSIONS such jobs are admitted at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "alto_hip.h"
#include "alto_internal.hpp"

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Job {
    uint64_t id = 0;
    int nmodes = 0;
    std::vector<uint64_t> dims;
    uint64_t nnz = 0;
    std::vector<std::vector<uint64_t>> cidx;
    std::vector<double> vals;
    std::vector<const uint64_t*> dev_cidx;  // device inputs, used in place
    const double* dev_vals = nullptr;
    uint64_t rank = 0;
    int max_iters = 0;
    double epsilon = 0;
    unsigned seed = 0;
    int precision = 0;
    int want_factors = 0;
    uint64_t footprint = 0;
    bool lockstep = false;  // iterated by the stepper
    bool stepping = false;  // handed to the stepper (counted as running)
    // results
    int status = 0;  // 0 queued, 1 running, 2 done, < 0 error code
    int iters = 0;
    double fit = 0;
    double t_submit = 0, t_start = 0, t_end = 0;
    std::vector<std::vector<double>> U;
    std::vector<double> lambda;
    std::string error;
};

// bytes the job holds on the device: ALTO keys + values, one stream per mode
// (packed coordinates + value in the job's precision), factors and the
// per-mode / per-iteration work buffers
uint64_t footprint_of(const Job& j)
{
    int bits = 0;
    uint64_t sum_dims = 0, max_dim = 0;
    for (uint64_t d : j.dims) {
        bits += d > 1 ? 64 - __builtin_clzll(d - 1) : 0;
        sum_dims += d;
        max_dim = std::max(max_dim, d);
    }
    const uint64_t W = bits <= 64 ? 8 : 16;
    const uint64_t sv = j.precision == ALTO_PREC_F64 ? 8 : 4;
    uint64_t b = j.nnz * (W + 8);                           // ALTO keys + values
    b += (uint64_t)j.nmodes * j.nnz * (W + sv);             // per-mode streams
    b += j.nnz * (W + 8 + 8);                               // construction temporaries (sort buffers)
    b += sum_dims * j.rank * 8 * 2 + 2 * max_dim * j.rank * 8;  // factors, copies, M, scratch
    b += 64ull << 20;                                        // LUTs, Grams, partials
    return b;
}

}  // namespace

struct alto_server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_step;
    std::deque<Job*> queue;                  // waiting for admission, arrival order
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    uint64_t budget = 0, in_use = 0, peak = 0, next_id = 1;
    int slots = 1, running = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    int device = 0;
    // iteration-level batching of small jobs
    struct Member {
        Job* j;
        alto_tensor_t* t;
        alto_cpd_session_t* sess;
    };
    uint64_t lock_max_nnz = 0;  // 0: off
    int lock_max = 0, lock_active = 0;
    int stepping = 0;  // jobs handed to the stepper, not yet finished
    std::vector<Member> inbox;  // sessions created, waiting to join a group
    std::thread stepper;
    hipStream_t step_stream = nullptr;  // the groups' iterations
    bool stop_stepper = false;

    bool small(const Job* j) const { return lock_max_nnz > 0 && j->nnz <= lock_max_nnz && j->rank <= 64; }

    Job* admit_locked()
    {
        if (running >= slots) return nullptr;
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            Job* j = *it;
            const bool sm = small(j);
            if (sm && lock_active >= lock_max) continue;
            if (in_use + j->footprint <= budget || (running == 0 && in_use == 0)) {
                queue.erase(it);
                in_use += j->footprint;
                peak = std::max(peak, in_use);
                ++running;
                j->status = 1;
                j->lockstep = sm;
                if (sm) ++lock_active;
                return j;
            }
        }
        return nullptr;
    }

    // create_alto and a device-seeded CP-ALS session for the job, on s
    int create_job(Job* j, hipStream_t s, alto_tensor_t** t, alto_cpd_session_t** sess)
    {
        j->t_start = now_s();
        const bool dev = j->dev_vals != nullptr;
        std::vector<const uint64_t*> rows(j->nmodes);
        for (int m = 0; m < j->nmodes; ++m) rows[m] = dev ? j->dev_cidx[m] : j->cidx[m].data();
        int st = alto_tensor_create(j->nmodes, j->dims.data(), j->nnz, rows.data(), dev ? j->dev_vals : j->vals.data(),
                                    dev ? ALTO_PTR_DEVICE : ALTO_PTR_HOST, ALTO_VAL_F64, 0, 1, (alto_stream_t)s, t);
        if (st == ALTO_OK) st = alto_cpd_session_create_seeded(*t, j->rank, j->precision, j->seed, (alto_stream_t)s, sess);
        if (st == ALTO_OK) st = alto::cpd_session_set_timing(*sess, 0);  // phase times are not reported
        return st;
    }

    // runs the job; returns its final status, which the caller publishes
    // under the server's lock after every result field is written
    int run_job(Job* j, hipStream_t s)
    {
        alto_tensor_t* t = nullptr;
        alto_cpd_session_t* sess = nullptr;
        int st = create_job(j, s, &t, &sess);
        int ran = 0, conv = 0;
        if (st == ALTO_OK) st = alto_cpd_session_iterate(sess, j->max_iters, j->epsilon, nullptr, &ran, &conv, 0);
        return finish_job(j, t, sess, st, ran);
    }

    // results of a job whose iterations ended (st: their status), then its
    // session and tensor are released
    int finish_job(Job* j, alto_tensor_t* t, alto_cpd_session_t* sess, int st, int ran)
    {
        if (st == ALTO_OK && j->want_factors) {
            j->U.assign(j->nmodes, {});
            std::vector<double*> up(j->nmodes);
            for (int m = 0; m < j->nmodes; ++m) {
                j->U[m].resize(j->dims[m] * j->rank);
                up[m] = j->U[m].data();
            }
            j->lambda.assign(j->rank, 0.0);
            st = alto_cpd_session_get(sess, up.data(), j->lambda.data(), &j->fit, nullptr);
        } else if (st == ALTO_OK) {
            st = alto_cpd_session_get(sess, nullptr, nullptr, &j->fit, nullptr);
        }
        if (st != ALTO_OK) j->error = alto_last_error();
        if (sess) alto_cpd_session_destroy(sess);
        if (t) alto_tensor_destroy(t);
        j->iters = ran;
        // the inputs are no longer needed
        std::vector<std::vector<uint64_t>>().swap(j->cidx);
        std::vector<double>().swap(j->vals);
        j->t_end = now_s();
        return st == ALTO_OK ? 2 : st;
    }

    // publishes a job's final status and returns its footprint
    void publish(Job* j, int status)
    {
        {
            std::lock_guard<std::mutex> g(mu);
            j->status = status;
            in_use -= j->footprint;
            if (j->lockstep) --lock_active;
            if (j->stepping) --stepping;
        }
        cv_work.notify_all();
        cv_done.notify_all();
    }

    void worker_main()
    {
        (void)hipSetDevice(device);
        hipStream_t s = nullptr;
        (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        for (;;) {
            Job* j = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || (j = admit_locked()) != nullptr; });
                if (!j) break;
            }
            if (j->lockstep) {
                // create here, iterate in the stepper's groups
                alto_tensor_t* t = nullptr;
                alto_cpd_session_t* sess = nullptr;
                int st = create_job(j, s, &t, &sess);
                // the set-up has drained (session creation ends with a
                // synchronization): later work goes to the stepper's stream
                if (st == ALTO_OK) st = alto::cpd_session_set_stream(sess, step_stream);
                if (st == ALTO_OK) {
                    {
                        std::lock_guard<std::mutex> g(mu);
                        inbox.push_back(Member{j, t, sess});
                        j->stepping = true;
                        ++stepping;
                        --running;
                    }
                    cv_step.notify_all();
                    cv_work.notify_all();
                    continue;
                }
                const int status = finish_job(j, t, sess, st, 0);
                {
                    std::lock_guard<std::mutex> g(mu);
                    --running;
                }
                publish(j, status);
                continue;
            }
            const int status = run_job(j, s);
            {
                std::lock_guard<std::mutex> g(mu);
                --running;
            }
            publish(j, status);
        }
        if (s) (void)hipStreamDestroy(s);
    }

    // Steps the lockstep groups: each round takes the sessions the workers
    // created since the last round into their groups, runs one iteration
    // of every group and retires the sessions that converged or reached
    // their iteration limit.
    void stepper_main()
    {
        (void)hipSetDevice(device);
        hipStream_t s = step_stream;
        struct Key {
            uint64_t rank;
            int precision;
            double epsilon;
            bool operator<(const Key& o) const
            {
                if (rank != o.rank) return rank < o.rank;
                if (precision != o.precision) return precision < o.precision;
                return epsilon < o.epsilon;
            }
        };
        std::map<Key, std::vector<Member>> groups;
        size_t active = 0;
        std::vector<alto_cpd_session_t*> sv;
        std::vector<int> iters;
        for (;;) {
            std::vector<Member> fresh;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_step.wait(lk, [&] { return stop_stepper || !inbox.empty() || active > 0; });
                if (stop_stepper && inbox.empty() && active == 0) break;
                fresh.swap(inbox);
            }
            for (auto& m : fresh) {
                groups[Key{m.j->rank, m.j->precision, m.j->epsilon}].push_back(m);
                ++active;
            }
            for (auto& kv : groups) {
                std::vector<Member>& g = kv.second;
                if (g.empty()) continue;
                sv.resize(g.size());
                iters.assign(g.size(), 0);
                for (size_t k = 0; k < g.size(); ++k) sv[k] = g[k].sess;
                const int st = alto_cpd_sessions_iterate((int)g.size(), sv.data(), 1, kv.first.epsilon,
                                                         (alto_stream_t)s, iters.data());
                std::vector<Member> keep;
                for (auto& m : g) {
                    int it = 0, conv = 0;
                    (void)alto::cpd_session_state(m.sess, &it, &conv);
                    if (st != ALTO_OK || conv || it >= m.j->max_iters) {
                        publish(m.j, finish_job(m.j, m.t, m.sess, st, it));
                        --active;
                    } else {
                        keep.push_back(m);
                    }
                }
                g.swap(keep);
            }
        }
    }
};

extern "C" {

int alto_server_create(uint64_t hbm_budget_bytes, int max_concurrent, alto_server_t** out)
{
    if (!out || max_concurrent < 1 || hbm_budget_bytes == 0) return ALTO_ERR_ARG;
    alto_server* S = new alto_server();
    S->budget = hbm_budget_bytes;
    S->slots = max_concurrent;
    alto::serve_lockstep_settings(&S->lock_max_nnz, &S->lock_max);
    if (S->lock_max < 1) S->lock_max_nnz = 0;
    (void)hipGetDevice(&S->device);
    if (S->lock_max_nnz) {
        if (hipStreamCreateWithFlags(&S->step_stream, hipStreamNonBlocking) != hipSuccess) S->lock_max_nnz = 0;
        else S->stepper = std::thread([S] { S->stepper_main(); });
    }
    for (int k = 0; k < max_concurrent; ++k) S->workers.emplace_back([S] { S->worker_main(); });
    *out = S;
    return ALTO_OK;
}

int alto_server_submit(alto_server_t* S, const alto_job_desc_t* d, uint64_t* job_id)
{
    if (!S || !d || !job_id || d->nmodes < 1 || d->nmodes > ALTO_MAX_MODES || !d->dims || !d->cidx || !d->vals ||
        d->rank < 1 || d->rank > 128 || d->max_iters < 1 || d->nnz < 1)
        return ALTO_ERR_ARG;
    auto j = std::make_unique<Job>();
    j->nmodes = d->nmodes;
    j->dims.assign(d->dims, d->dims + d->nmodes);
    j->nnz = d->nnz;
    if (d->ptr_kind == ALTO_PTR_DEVICE) {
        j->dev_cidx.assign(d->cidx, d->cidx + d->nmodes);
        j->dev_vals = d->vals;
    } else {
        j->cidx.resize(d->nmodes);
        for (int m = 0; m < d->nmodes; ++m) j->cidx[m].assign(d->cidx[m], d->cidx[m] + d->nnz);
        j->vals.assign(d->vals, d->vals + d->nnz);
    }
    j->rank = d->rank;
    j->max_iters = d->max_iters;
    j->epsilon = d->epsilon;
    j->seed = d->seed;
    j->precision = d->precision;
    j->want_factors = d->want_factors;
    j->footprint = footprint_of(*j);
    j->t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(S->mu);
        j->id = S->next_id++;
        *job_id = j->id;
        S->queue.push_back(j.get());
        S->jobs[j->id] = std::move(j);
    }
    S->cv_work.notify_all();
    return ALTO_OK;
}

int alto_server_result(alto_server_t* S, uint64_t job_id, int wait, alto_job_result_t* res, double* const* U,
                       double* lambda)
{
    if (!S || !res) return ALTO_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) return ALTO_ERR_ARG;
    Job* j = it->second.get();
    if (wait) S->cv_done.wait(lk, [&] { return j->status == 2 || j->status < 0; });
    res->status = j->status;
    res->iters = j->iters;
    res->fit = j->fit;
    res->t_submit = j->t_submit;
    res->t_start = j->t_start;
    res->t_end = j->t_end;
    res->footprint = j->footprint;
    if (j->status == 2 && j->want_factors) {
        if (U)
            for (int m = 0; m < j->nmodes; ++m)
                if (U[m]) std::copy(j->U[m].begin(), j->U[m].end(), U[m]);
        if (lambda) std::copy(j->lambda.begin(), j->lambda.end(), lambda);
    }
    return ALTO_OK;
}

int alto_server_release(alto_server_t* S, uint64_t job_id)
{
    if (!S) return ALTO_ERR_ARG;
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) return ALTO_ERR_ARG;
    if (it->second->status == 0 || it->second->status == 1) return ALTO_ERR_BUSY;  // queued or running
    S->jobs.erase(it);
    return ALTO_OK;
}

int alto_server_stats(alto_server_t* S, alto_server_stats_t* st)
{
    if (!S || !st) return ALTO_ERR_ARG;
    std::lock_guard<std::mutex> g(S->mu);
    st->queued = (int)S->queue.size();
    st->running = S->running + S->stepping;  // with the jobs in lockstep groups
    st->hbm_in_use = S->in_use;
    st->hbm_peak = S->peak;
    st->hbm_budget = S->budget;
    return ALTO_OK;
}

int alto_server_destroy(alto_server_t* S)
{
    if (!S) return ALTO_OK;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_done.wait(lk, [&] { return S->queue.empty() && S->running == 0 && S->lock_active == 0; });
        S->stop = true;
        S->stop_stepper = true;
    }
    S->cv_work.notify_all();
    S->cv_step.notify_all();
    for (auto& w : S->workers) w.join();
    if (S->stepper.joinable()) S->stepper.join();
    if (S->step_stream) (void)hipStreamDestroy(S->step_stream);
    delete S;
    return ALTO_OK;
}

}  // extern "C"
