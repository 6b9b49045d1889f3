// CP-APR (multiplicative update) on the device: cp_apr_alto
// (include/apr.hpp:984-1376) with Phi, the KKT reduction, the
// multiplicative update, the kappa fix, lambda redistribution and the L1
// normalisations as kernels; the host only reads the KKT value of each
// inner iteration (for the early break) unless running in bench mode.
//
// Final statistics: tt_logLikelihood<LIT> (include/apr.hpp:850-982) and the
// least-squares fit of the final model (include/apr.hpp:1299-1330).
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "alto_internal.hpp"
#include "dense.hpp"
#include "kernels_common.hpp"

namespace alto {

constexpr int COL_BLOCKS = 128;

// partial column sums of |U| (thread t: column t % R, rows t / R + k * (256/R))
__global__ void k_colabs_partial(const double* __restrict__ U, uint64_t I, int R, double* __restrict__ part)
{
    __shared__ double s[256];
    const int RP = 256 / R, t = threadIdx.x;
    double a = 0;
    if (t < RP * R) {
        const int c = t % R, ri = t / R;
        for (uint64_t i = (uint64_t)blockIdx.x * RP + ri; i < I; i += (uint64_t)gridDim.x * RP) a += fabs(U[i * R + c]);
    }
    s[t] = a;
    __syncthreads();
    for (int c = t; c < R; c += blockDim.x) {
        double v = 0;
        for (int ri = 0; ri < RP; ++ri) v += s[ri * R + c];
        part[(size_t)blockIdx.x * R + c] = v;
    }
}

// norm_r = sum of partials; lambda_r *= norm_r
__global__ void k_colnorm_final(const double* __restrict__ part, int nb, int R, double* __restrict__ norm,
                                double* __restrict__ lambda)
{
    for (int c = threadIdx.x; c < R; c += blockDim.x) {
        double a = 0;
        for (int p = 0; p < nb; ++p) a += part[(size_t)p * R + c];
        norm[c] = a;
        lambda[c] *= a;
    }
}

// U[:, r] *= lambda_r then lambda = 1 (RedistributeLambda, src/kruskal_model.cpp:355-378)
__global__ void k_mul_cols(double* __restrict__ U, uint64_t n, int R, const double* __restrict__ lambda)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        U[k] *= lambda[k % R];
}

// B += kappa where B < kappa_tol and Phi > 1 (include/apr.hpp:1176-1189)
__global__ void k_kappa_fix(double* __restrict__ B, const double* __restrict__ Phi, uint64_t n, double kappa_tol,
                            double kappa, int* __restrict__ any)
{
    int hit = 0;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        if (B[k] < kappa_tol && Phi[k] > 1.0) {
            B[k] += kappa;
            hit = 1;
        }
    if (hit) atomicOr(any, 1);
}

// partial max of |min(B, 1 - Phi)|; with mu != 0 also B *= Phi (bench mode:
// the update does not depend on the KKT value)
__global__ void k_kkt_partial(double* __restrict__ B, const double* __restrict__ Phi, uint64_t n, int mu,
                              double* __restrict__ part)
{
    __shared__ double s[256];
    double m = 0;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const double b = B[k], p = Phi[k];
        m = fmax(m, fabs(fmin(b, 1.0 - p)));
        if (mu) B[k] = b * p;
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// max of the nb per-block partials (one block of 256 threads)
__global__ void __launch_bounds__(256) k_max_final(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    __shared__ double s[256];
    double m = 0;
    for (int p = threadIdx.x; p < nb; p += blockDim.x) m = fmax(m, part[p]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

__global__ void k_mu(double* __restrict__ B, const double* __restrict__ Phi, uint64_t n)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        B[k] *= Phi[k];
}

// log-likelihood partials over one mode's stream: x * log(max(sum_r prod_n U_n[row_n, r], eps))
template <int G, int E, int PW>
__global__ void __launch_bounds__(256) k_loglik(StreamArgs<double, double, double> a, const double* __restrict__ Un,
                                                double eps, double* __restrict__ part)
{
    __shared__ double s[256];
    const int lane = threadIdx.x & 63;
    const int gl = lane & (G - 1);
    const uint64_t gpb = 256 / G;
    const int R = a.rank;
    double ll = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * gpb + threadIdx.x / G; i < a.nnz; i += (uint64_t)gridDim.x * gpb) {
        uint64_t k[PW];
#pragma unroll
        for (int w = 0; w < PW; ++w) k[w] = a.pk[i * PW + w];
        double pr[E];
        const double* u0 = Un + field<PW>(k, a.offn, a.bitsn) * (uint64_t)R;
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int c = gl + q * G;
            pr[q] = c < R ? u0[c] : 0.0;
        }
        for (int o = 0; o < a.nother; ++o) {
            const double* u = a.Uo[o] + field<PW>(k, a.offo[o], a.bitso[o]) * (uint64_t)R;
#pragma unroll
            for (int q = 0; q < E; ++q) {
                const int c = gl + q * G;
                pr[q] *= c < R ? u[c] : 0.0;
            }
        }
        double sum = 0;
#pragma unroll
        for (int q = 0; q < E; ++q) sum += pr[q];
        sum = group_sum<G>(sum);
        if (sum < eps) sum = eps;
        if (gl == 0) ll += a.val[i] * log(sum);
    }
    s[threadIdx.x] = ll;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_sum_final(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double m = 0;
        for (int p = 0; p < nb; ++p) m += part[p];
        *out = m;
    }
}

// sum of all entries of U (the factor sum of tt_logLikelihood)
__global__ void k_sum_partial(const double* __restrict__ U, uint64_t n, double* __restrict__ part)
{
    __shared__ double s[256];
    double a = 0;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) a += U[k];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

static unsigned ew_blocks(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    if (b > 2048) b = 2048;
    return (unsigned)(b ? b : 1);
}

struct AprState {
    int N = 0, R = 0;
    DevBuf U[ALTO_MAXN], Phi[ALTO_MAXN];
    DevBuf lambda, ones, norm, cpart, kpart, kval, any, llpart, llval, grams, gpart, Mlast, ipart, fitbuf, nx, nxpart;
    DevBuf pic;  // products of the other modes' rows per nonzero, for the inner iterations of one mode
};

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// KruskalModelNormalizeMode (src/kruskal_model.cpp:178-222)
static int normalize_mode(AprState& c, uint64_t dim, int n, hipStream_t s)
{
    const int R = c.R;
    hipLaunchKernelGGL(k_colabs_partial, dim3(COL_BLOCKS), dim3(256), 0, s, c.U[n].as<double>(), dim, R, c.cpart.as<double>());
    hipLaunchKernelGGL(k_colnorm_final, dim3(1), dim3(128), 0, s, c.cpart.as<double>(), COL_BLOCKS, R, c.norm.as<double>(),
                       c.lambda.as<double>());
    HIP_TRY(hipGetLastError());
    return scale_device<double, double>(c.U[n].as<double>(), dim, R, c.norm.as<double>(), nullptr, s);
}

static int redistribute(AprState& c, uint64_t dim, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_mul_cols, dim3(ew_blocks(dim * c.R)), dim3(256), 0, s, c.U[n].as<double>(), dim * c.R, c.R,
                       c.lambda.as<double>());
    HIP_TRY(hipGetLastError());
    return fill_value<double>(c.lambda.as<double>(), c.R, 1.0, s);
}

static int loglik(alto_tensor* t, AprState& c, double eps, double* out, hipStream_t s)
{
    // RedistributeLambda(M, 0) then the nonzero sum over mode 0's stream
    ALTO_TRY(redistribute(c, t->dims[0], 0, s));
    StreamArgs<double, double, double> a;
    fill_stream_args(t, 0, c.R, 1, a);
    const double* Uall[ALTO_MAXN];
    for (int m = 0; m < c.N; ++m) Uall[m] = c.U[m].as<double>();
    ALTO_TRY(reserve_rank_copies(t, (uint64_t)c.R, 8, s));
    ALTO_TRY(set_factors(t, 0, Uall, a, s));
    a.val = t->ms[0].val64.as<double>();
    const int nb = 1024;
    RANK_DISPATCH(c.R, t->pack.pw, {
        hipLaunchKernelGGL((k_loglik<G, E, PW>), dim3(nb), dim3(256), 0, s, a, Uall[0], eps, c.llpart.as<double>());
    });
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(1), 0, s, c.llpart.as<double>(), nb, c.llval.as<double>());
    hipLaunchKernelGGL(k_sum_partial, dim3(256), dim3(256), 0, s, c.U[0].as<double>(), t->dims[0] * c.R,
                       c.llpart.as<double>() + nb);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(1), 0, s, c.llpart.as<double>() + nb, 256, c.llval.as<double>() + 1);
    HIP_TRY(hipGetLastError());
    double v[2];
    HIP_TRY(hipMemcpyAsync(v, c.llval.p, 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *out = v[0] - v[1];
    // renormalize U_0 (L1 columns) into lambda
    return normalize_mode(c, t->dims[0], 0, s);
}

// rule: ALTO_APR_STOP_ALTO -- cp_apr_alto's outer stop, |KKT - previous
// KKT| <= eps^2 (include/apr.hpp:1261); ALTO_APR_BENCH -- every inner and
// outer iteration runs (its dobench); ALTO_APR_STOP_COO -- cp_apr_mu_alto's
// (src/apr.cpp:345-448): an outer iteration in which every mode's first
// inner step already met the KKT tolerance ends the run, and that
// iteration is counted.
int cp_apr_impl(alto_tensor* t, int R, int max_iters, int max_inner, double epsilon, double eps_div_zero,
                double kappa_tol, double kappa, int rule, double* const* Uh, double* lambda_h, alto_apr_result_t* res,
                int verbose, hipStream_t s)
{
    if (rule != ALTO_APR_STOP_ALTO && rule != ALTO_APR_BENCH && rule != ALTO_APR_STOP_COO) {
        set_error("cp_apr: unknown stop rule %d", rule);
        return ALTO_ERR_ARG;
    }
    const int dobench = rule == ALTO_APR_BENCH;
    const bool coo = rule == ALTO_APR_STOP_COO;
    if (R < 1 || R > 128 || max_iters < 1 || max_inner < 1) {
        set_error("cp_apr: rank must be 1..128 and iteration counts >= 1");
        return ALTO_ERR_ARG;
    }
    const double tp0 = now_s();
    AprState c;
    c.N = t->nmode;
    c.R = R;
    const int N = c.N;
    ALTO_TRY(prepare_mode_streams(t, 1, 0, s));
    uint64_t maxd = 0;
    for (int m = 0; m < N; ++m) {
        maxd = std::max<uint64_t>(maxd, t->dims[m]);
        ALTO_TRY(c.U[m].alloc(t->dims[m] * R * 8));
        ALTO_TRY(c.Phi[m].alloc(t->dims[m] * R * 8));
        ALTO_TRY(fill_zero(c.Phi[m].as<double>(), t->dims[m] * R, s));
        HIP_TRY(hipMemcpyAsync(c.U[m].p, Uh[m], t->dims[m] * R * 8, hipMemcpyHostToDevice, s));
    }
    ALTO_TRY(c.lambda.alloc(R * 8));
    ALTO_TRY(c.norm.alloc(R * 8));
    ALTO_TRY(c.cpart.alloc((size_t)COL_BLOCKS * R * 8));
    ALTO_TRY(c.kpart.alloc(2048 * 8));
    ALTO_TRY(c.kval.alloc(8));
    ALTO_TRY(c.any.alloc(8));
    ALTO_TRY(c.llpart.alloc((1024 + 256) * 8));
    ALTO_TRY(c.llval.alloc(16));
    ALTO_TRY(c.grams.alloc((size_t)N * R * R * 8));
    ALTO_TRY(c.gpart.alloc((size_t)gram_blocks(maxd) * R * R * 8));
    ALTO_TRY(c.Mlast.alloc(t->dims[N - 1] * R * 8));
    ALTO_TRY(c.ipart.alloc((size_t)64 * R * 8));
    ALTO_TRY(c.fitbuf.alloc(32));
    ALTO_TRY(c.nx.alloc(8));
    ALTO_TRY(c.nxpart.alloc(256 * 8));
    // the model's lambda (KruskalModelRandomInit sets it to 1), then KruskalModelNormalize
    HIP_TRY(hipMemcpyAsync(c.lambda.p, lambda_h, R * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (verbose) {
        printf("Running ALTO CP-APR with: \n\tmax iters:           %d\n\tmax inner iters:     %d\n", max_iters, max_inner);
        printf("\ttolerance:           %.2e\n\teps_div_zero:        %.2e\n\tkappa_tol:           %.2e\n", epsilon,
               eps_div_zero, kappa_tol);
        printf("\tkappa:               %.2e\n", kappa);
    }
    const double t_pre = now_s() - tp0;

    const double ta0 = now_s();
    for (int n = 0; n < N; ++n) ALTO_TRY(normalize_mode(c, t->dims[n], n, s));
    std::vector<double> kkt(max_iters + 1, 0.0);
    std::vector<int> n_inner(max_iters + 1, 0), n_viol(max_iters + 1, 0);
    double kkt_mode[ALTO_MAXN] = {0};
    int inner_per_mode[ALTO_MAXN] = {0};
    double t_phi = 0;
    // one event pair per inner iteration, read once per mode, so that the
    // Phi timer does not stop the stream after every inner iteration
    std::vector<hipEvent_t> ev0(max_inner, nullptr), ev1(max_inner, nullptr);
    struct EventsGuard {
        std::vector<hipEvent_t>* v[2];
        ~EventsGuard()
        {
            for (auto* vv : v)
                for (auto e : *vv)
                    if (e) (void)hipEventDestroy(e);
        }
    } ev_guard{{&ev0, &ev1}};
    for (int k = 0; k < max_inner; ++k) {
        HIP_TRY(hipEventCreate(&ev0[k]));
        HIP_TRY(hipEventCreate(&ev1[k]));
    }
    const double* Uptr[ALTO_MAXN];
    for (int m = 0; m < N; ++m) Uptr[m] = c.U[m].as<double>();
    // the inner iterations of a mode keep the other factors: their row
    // products are stored per nonzero by the first Phi of the mode and read
    // back by the others, when nnz x R doubles fit in half the free memory
    // (at most 64 GiB)
    bool cache = apr_pi_cache_enabled() && !det_enabled() && max_inner > 1;
    if (cache) {
        size_t fr = 0, tot = 0;
        HIP_TRY(hipMemGetInfo(&fr, &tot));
        const uint64_t want = t->nnz * (uint64_t)R * sizeof(double);
        cache = want <= fr / 2 && want <= (64ull << 30) && c.pic.alloc(want) == ALTO_OK;
    }
    int iter;
    for (iter = 0; iter < max_iters; ++iter) {
        bool converged = coo;
        for (int n = 0; n < N; ++n) {
            const uint64_t dim = t->dims[n], ne = dim * R;
            double* B = c.U[n].as<double>();
            double* Ph = c.Phi[n].as<double>();
            if (iter > 0) {
                ALTO_TRY(fill_zero(c.any.as<int>(), 1, s));
                hipLaunchKernelGGL(k_kappa_fix, dim3(ew_blocks(ne)), dim3(256), 0, s, B, Ph, ne, kappa_tol, kappa,
                                   c.any.as<int>());
                HIP_TRY(hipGetLastError());
                int hit = 0;
                HIP_TRY(hipMemcpyAsync(&hit, c.any.p, 4, hipMemcpyDeviceToHost, s));
                HIP_TRY(hipStreamSynchronize(s));
                if (hit) n_viol[iter]++;
            }
            ALTO_TRY(redistribute(c, dim, n, s));
            int timed = 0;
            for (int inner = 0; inner < max_inner; ++inner) {
                n_inner[iter]++;
                inner_per_mode[n]++;
                HIP_TRY(hipEventRecord(ev0[inner], s));
                ALTO_TRY(phi_device_cached(t, n, R, Uptr, Ph, eps_div_zero, s, 0, c.pic.as<double>(),
                                           cache ? (inner == 0 ? 1 : 2) : 0));
                HIP_TRY(hipEventRecord(ev1[inner], s));
                timed = inner + 1;
                const unsigned nb = ew_blocks(ne);
                hipLaunchKernelGGL(k_kkt_partial, dim3(nb), dim3(256), 0, s, B, Ph, ne, dobench ? 1 : 0,
                                   c.kpart.as<double>());
                hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, s, c.kpart.as<double>(), (int)nb, c.kval.as<double>());
                HIP_TRY(hipGetLastError());
                double kv = 0;
                if (!dobench || inner == max_inner - 1) {
                    HIP_TRY(hipMemcpyAsync(&kv, c.kval.p, 8, hipMemcpyDeviceToHost, s));
                    HIP_TRY(hipStreamSynchronize(s));
                    kkt_mode[n] = kv;
                }
                if (dobench) continue;  // B *= Phi already applied
                if (kv < epsilon) break;
                converged = false;
                hipLaunchKernelGGL(k_mu, dim3(nb), dim3(256), 0, s, B, Ph, ne);
                HIP_TRY(hipGetLastError());
            }
            if (timed) HIP_TRY(hipEventSynchronize(ev1[timed - 1]));
            for (int k = 0; k < timed; ++k) {
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, ev0[k], ev1[k]));
                t_phi += ms * 1e-3;
            }
            ALTO_TRY(normalize_mode(c, dim, n, s));
        }
        kkt[iter] = kkt_mode[0];
        for (int n = 1; n < N; ++n) kkt[iter] = fmax(kkt[iter], kkt_mode[n]);
        if (verbose && iter % 10 == 0)
            printf("Iter %d: Inner its = %d KKT violation = %.6e, nViolations = %d\n", iter, n_inner[iter], kkt[iter],
                   n_viol[iter]);
        if (!coo && iter > 0 && fabs(kkt[iter] - kkt[iter - 1]) <= epsilon * epsilon && !dobench) {
            converged = true;
            if (verbose) {
                printf("Iter %d: Inner its = %d KKT violation = %.6e, nViolations = %d\n", iter, n_inner[iter],
                       kkt[iter], n_viol[iter]);
                printf("ErrorNorm delta below threshold. Assume convergence\n");
            }
        }
        if (converged) {
            if (verbose) printf("Exiting since all subproblems reached KKT tol\n");
            break;
        }
        if (verbose && iter == max_iters - 1)
            printf("Iter %d: Inner its = %d KKT violation = %.6e, nViolations = %d\n", iter, n_inner[iter], kkt[iter],
                   n_viol[iter]);
    }
    HIP_TRY(hipStreamSynchronize(s));
    const double t_apr = now_s() - ta0;

    const double tq0 = now_s();
    // cp_apr_mu_alto counts the iteration it stopped in (src/apr.cpp:415-423)
    if (coo && iter == max_iters) iter = max_iters - 1;
    const int counted = coo ? iter + 1 : iter;
    int total_inner = 0;
    for (int i = 0; i < counted; ++i) total_inner += n_inner[i];
    double obj = 0;
    ALTO_TRY(loglik(t, c, eps_div_zero, &obj, s));
    // least-squares fit of the final model
    for (int m = 0; m < N; ++m)
        ALTO_TRY(gram_device<double>(c.U[m].as<double>(), t->dims[m], R, c.grams.as<double>() + (size_t)m * R * R,
                                     c.gpart.as<double>(), s, true));
    const void* fptr[ALTO_MAXN];
    for (int m = 0; m < N; ++m) fptr[m] = c.U[m].p;
    ALTO_TRY(mttkrp_device(t, N - 1, R, fptr, c.Mlast.p, ALTO_PREC_F64, 0, s, 0));
    ALTO_TRY(normsq_device(t->vals.as<double>(), t->nnz, c.nxpart.as<double>(), c.nx.as<double>(), s));
    double normX = 0;
    HIP_TRY(hipMemcpyAsync(&normX, c.nx.p, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    ALTO_TRY((inner_partial_device<double, double>(c.U[N - 1].as<double>(), c.Mlast.as<double>(), t->dims[N - 1], R,
                                                   c.ipart.as<double>(), s)));
    ALTO_TRY(fit_device(c.grams.as<double>(), N, R, c.lambda.as<double>(), c.ipart.as<double>(), normX,
                        c.fitbuf.as<double>(), s));
    double fb[3];
    HIP_TRY(hipMemcpyAsync(fb, c.fitbuf.p, 24, hipMemcpyDeviceToHost, s));
    for (int m = 0; m < N; ++m) HIP_TRY(hipMemcpyAsync(Uh[m], c.U[m].p, t->dims[m] * R * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(lambda_h, c.lambda.p, R * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const double t_post = now_s() - tq0;

    const double kkt_last = coo ? kkt[iter] : iter > 0 ? kkt[iter - 1] : 0.0;
    if (res) {
        res->loglik = obj;
        res->ls_fit = fb[0];
        res->kkt = kkt_last;
        res->outer_iters = iter;
        res->total_inner = total_inner;
        for (int m = 0; m < ALTO_MAXN; ++m) res->inner_per_mode[m] = m < N ? inner_per_mode[m] : 0;
        res->t_pre = t_pre;
        res->t_apr = t_apr;
        res->t_phi = t_phi;
        res->t_post = t_post;
    }
    if (verbose) {
        printf("  Final log-likelihood = %e\n", obj);
        printf("  Final least squares fit = %e\n", fb[0]);
        printf("  Final KKT violation = %7.7e\n", kkt_last);
        printf("  Total outer iterations = %d\n", iter);
        printf("  Total inner iterations = %d\n", total_inner);
        printf("CP-APR pre-processing time  = %f (s)\n", t_pre);
        printf("CP-APR main iteration time  = %f (s)\n", t_apr);
        printf("CP-APR Phi accumulated time = %f (s)\n", t_phi);
        printf("CP-APR post-processing time = %f (s)\n", t_post);
        printf("CP-APR TOTAL time           = %f (s)\n", t_pre + t_apr + t_post);
        printf("Time per iteration          = %f (s)\n\n", (t_pre + t_apr + t_post) / (total_inner ? total_inner : 1));
    }
    return ALTO_OK;
}

// ------------------------------------------------------------------ Kruskal model utilities
// per-block column statistics: which 0 = sum of squares, 1 = maximum (signed)
__global__ void k_colstat_partial(const double* __restrict__ U, uint64_t I, int R, int which, double* __restrict__ part)
{
    __shared__ double s[256];
    const int RP = 256 / R, t = threadIdx.x;
    double a = which ? -INFINITY : 0.0;
    if (t < RP * R) {
        const int c = t % R, ri = t / R;
        for (uint64_t i = (uint64_t)blockIdx.x * RP + ri; i < I; i += (uint64_t)gridDim.x * RP) {
            const double v = U[i * R + c];
            a = which ? fmax(a, v) : a + v * v;
        }
    }
    s[t] = a;
    __syncthreads();
    for (int c = t; c < R; c += blockDim.x) {
        double v = which ? -INFINITY : 0.0;
        for (int ri = 0; ri < RP; ++ri) v = which ? fmax(v, s[ri * R + c]) : v + s[ri * R + c];
        part[(size_t)blockIdx.x * R + c] = v;
    }
}

// Mat2Norm / MatMaxNorm (src/kruskal_model.cpp:224-305): lambda_r = the
// column's 2-norm, or max(1, its largest entry) (signed)
__global__ void k_colstat_final(const double* __restrict__ part, int nb, int R, int which, double* __restrict__ lambda)
{
    for (int c = threadIdx.x; c < R; c += blockDim.x) {
        double a = which ? -INFINITY : 0.0;
        for (int p = 0; p < nb; ++p) a = which ? fmax(a, part[(size_t)p * R + c]) : a + part[(size_t)p * R + c];
        lambda[c] = which ? fmax(a, 1.0) : sqrt(a);
    }
}

// one Kruskal utility on device factors U (dim x R) and lambda (R)
static int kruskal_op(double* U, uint64_t dim, int R, double* lambda, int kind, double* part, double* norm,
                      hipStream_t s)
{
    if (kind == ALTO_KRUSKAL_L1) {  // KruskalModelNormalizeMode (src/kruskal_model.cpp:178-222)
        hipLaunchKernelGGL(k_colabs_partial, dim3(COL_BLOCKS), dim3(256), 0, s, U, dim, R, part);
        hipLaunchKernelGGL(k_colnorm_final, dim3(1), dim3(128), 0, s, part, COL_BLOCKS, R, norm, lambda);
        HIP_TRY(hipGetLastError());
        return scale_device<double, double>(U, dim, R, norm, nullptr, s);
    }
    if (kind == ALTO_KRUSKAL_NORM2 || kind == ALTO_KRUSKAL_NORMMAX) {  // KruskalModelNorm (:307-341)
        const int which = kind == ALTO_KRUSKAL_NORMMAX;
        hipLaunchKernelGGL(k_colstat_partial, dim3(COL_BLOCKS), dim3(256), 0, s, U, dim, R, which, part);
        hipLaunchKernelGGL(k_colstat_final, dim3(1), dim3(128), 0, s, part, COL_BLOCKS, R, which, lambda);
        HIP_TRY(hipGetLastError());
        return scale_device<double, double>(U, dim, R, lambda, nullptr, s);
    }
    // RedistributeLambda (:355-378)
    hipLaunchKernelGGL(k_mul_cols, dim3(ew_blocks(dim * R)), dim3(256), 0, s, U, dim * R, R, lambda);
    HIP_TRY(hipGetLastError());
    return fill_value<double>(lambda, R, 1.0, s);
}

}  // namespace alto

using namespace alto;

extern "C" int alto_kruskal_normalize(int nmodes, const uint64_t* dims, uint64_t rank, double* const* U, double* lambda,
                                      int mode, int kind, int ptr_kind, alto_stream_t stream)
{
    if (nmodes < 1 || nmodes > ALTO_MAXN || !dims || !U || !lambda || rank < 1 || rank > 128 ||
        mode < -1 || mode >= nmodes || (mode < 0 && kind != ALTO_KRUSKAL_L1) ||
        kind < ALTO_KRUSKAL_L1 || kind > ALTO_KRUSKAL_REDISTRIBUTE) {
        set_error("alto_kruskal_normalize: invalid argument");
        return ALTO_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    const int R = (int)rank;
    const bool host = ptr_kind == ALTO_PTR_HOST;
    DevBuf part, norm, lam, Ud[ALTO_MAXN];
    ALTO_TRY(part.alloc((size_t)COL_BLOCKS * R * 8));
    ALTO_TRY(norm.alloc((size_t)R * 8));
    double* ld = lambda;
    if (host) {
        ALTO_TRY(lam.alloc((size_t)R * 8));
        HIP_TRY(hipMemcpyAsync(lam.p, lambda, (size_t)R * 8, hipMemcpyHostToDevice, s));
        ld = lam.as<double>();
    }
    const int m0 = mode < 0 ? 0 : mode, m1 = mode < 0 ? nmodes : mode + 1;
    for (int m = m0; m < m1; ++m) {
        double* u = U[m];
        if (host) {
            ALTO_TRY(Ud[m].alloc(dims[m] * R * 8));
            HIP_TRY(hipMemcpyAsync(Ud[m].p, U[m], dims[m] * R * 8, hipMemcpyHostToDevice, s));
            u = Ud[m].as<double>();
        }
        ALTO_TRY(kruskal_op(u, dims[m], R, ld, kind, part.as<double>(), norm.as<double>(), s));
        if (host) HIP_TRY(hipMemcpyAsync(U[m], u, dims[m] * R * 8, hipMemcpyDeviceToHost, s));
    }
    if (host) HIP_TRY(hipMemcpyAsync(lambda, ld, (size_t)R * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));  // temporaries are released below
    return ALTO_OK;
}

extern "C" int alto_cp_apr(alto_tensor_t* t, uint64_t rank, int max_iters, int max_inner, double epsilon,
                           double eps_div_zero, double kappa_tol, double kappa, int dobench, double* const* U,
                           double* lambda, alto_apr_result_t* res, int verbose, alto_stream_t stream)
{
    if (!t || !U || !lambda) {
        set_error("alto_cp_apr: null argument");
        return ALTO_ERR_ARG;
    }
    return cp_apr_impl(t, (int)rank, max_iters, max_inner, epsilon, eps_div_zero, kappa_tol, kappa,
                       dobench ? ALTO_APR_BENCH : ALTO_APR_STOP_ALTO, U, lambda, res, verbose, (hipStream_t)stream);
}

extern "C" int alto_cp_apr_ex(alto_tensor_t* t, uint64_t rank, int max_iters, int max_inner, double epsilon,
                              double eps_div_zero, double kappa_tol, double kappa, int stop_rule, double* const* U,
                              double* lambda, alto_apr_result_t* res, int verbose, alto_stream_t stream)
{
    if (!t || !U || !lambda) {
        set_error("alto_cp_apr_ex: null argument");
        return ALTO_ERR_ARG;
    }
    return cp_apr_impl(t, (int)rank, max_iters, max_inner, epsilon, eps_div_zero, kappa_tol, kappa, stop_rule, U,
                       lambda, res, verbose, (hipStream_t)stream);
}
