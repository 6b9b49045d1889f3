// cpd64 / cpd128: the reference's command line (src/main.cpp) over the
// MI355X engine.  Same flags, defaults and output lines that
// scripts/run_cpd.sh parses ("ALTO creation time", "ALTO runtime", the
// "it:" lines and the CP-ALS / CP-APR timer tables).  The binary name picks
// the ALTO key width (cpd64 / cpd128) as ALTO_MASK_LENGTH does
// (src/main.cpp:32-39); -l defaults to ALS.
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <chrono>
#include <string>
#include <vector>

#include "alto/alto.hpp"
#include "alto/apr.hpp"
#include "alto/cpd.hpp"
#include "alto/poisson_generator.hpp"
#include "alto/sptensor.hpp"

static const char* short_opt = "hvi:o:b:l:r:m:n:x:d:t:s:e:f:zcp";
static const struct option long_opt[] = {
    {"help", 0, NULL, 'h'},       {"version", 0, NULL, 'v'},     {"input", 1, NULL, 'i'},
    {"output", 1, NULL, 'o'},     {"bin", 1, NULL, 'b'},         {"model", 1, NULL, 'l'},
    {"rank", 1, NULL, 'r'},       {"max-iter", 1, NULL, 'm'},    {"max-inner", 1, NULL, 'n'},
    {"seed", 1, NULL, 'x'},       {"dims", 1, NULL, 'd'},        {"target-mode", 1, NULL, 't'},
    {"sparsity", 1, NULL, 's'},   {"epsilon", 1, NULL, 'e'},     {"file", 1, NULL, 'f'},
    {"zero-based", 0, NULL, 'z'}, {"check", 0, NULL, 'c'},       {"bench", 0, NULL, 'p'},
    {NULL, 0, NULL, 0}};

static double wtime()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void usage(const char* call)
{
    fprintf(stderr, "Usage: %s [OPTIONS]\n", call);
    fprintf(stderr, "Options:\n");
    fprintf(stderr, "\t-h or --help         Display this information\n");
    fprintf(stderr, "\t-v or --version      Display version information\n");
    fprintf(stderr, "\t-i or --input        Input tensor file in text\n");
    fprintf(stderr, "\t-o or --output       Output tensor to file\n");
    fprintf(stderr, "\t-b or --bin          Input tensor file in binary\n");
    fprintf(stderr, "\t-z or --zero-based   Assume input tensor is zero-based (default: one-based)\n");
    fprintf(stderr, "\t-l or --model        CPD model (ALS or APR)\n");
    fprintf(stderr, "\t-f or --file         Save tensor to another format\n");
    fprintf(stderr, "\t-r or --rank         Rank\n");
    fprintf(stderr, "\t-m or --max-iter     Maximum outer iterations\n");
    fprintf(stderr, "\t-n or --max-inner    Maximum inner iterations\n");
    fprintf(stderr, "\t-t or --target-mode  Target mode of tensor\n");
    fprintf(stderr, "\t-e or --epsilon      Convergence criteria\n");
    fprintf(stderr, "\t-x or --seed         Random value seed\n");
    fprintf(stderr, "\t-d or --dims         Dimension lenghts (I,J,K)\n");
    fprintf(stderr, "\t-s or --sparsity     Sparsity of generate tensor\n");
    fprintf(stderr, "\t-c or --check        Run ALTO (par) validation against cpd MTTKRP\n");
    fprintf(stderr, "\t-p or --bench        Run ALTO benchmark mode based on the CPD model with the given CMD line options\n");
}

static void print_info(IType rank, int max_iters, SparseTensor* X)
{
    double tmp = 1;
    for (int i = 0; i < X->nmodes; i++) tmp *= (double)X->dims[i];
    fprintf(stdout, "# Modes         = %u\n", X->nmodes);
    fprintf(stdout, "Rank            = %llu\n", rank);
    fprintf(stdout, "Sparsity        = %f\n", X->nnz / tmp);
    fprintf(stdout, "Max iters       = %d\n", max_iters);
    fprintf(stdout, "Dimensions      = [%llu", X->dims[0]);
    for (int i = 1; i < X->nmodes; i++) fprintf(stdout, " X %llu", X->dims[i]);
    fprintf(stdout, "]\n");
    fprintf(stdout, "NNZ             = %llu\n", X->nnz);
}

template <typename LIT>
static int run(SparseTensor* X, Model model, IType rank, int max_iters, int max_inner, double epsilon, int seed,
               int target_mode, bool do_check, bool do_bench, int nprtn)
{
    const double eps_div_zero = 1.0e-10, kappa_tol = 1.0e-10, kappa = 1.0e-2;
    KruskalModel* M;
    if (do_check) {
        printf("===Check for mode %d===\n", target_mode);
        print_info(rank, 1, X);
        CreateKruskalModel(X->nmodes, X->dims, rank, &M);
        KruskalModelRandomInit(M, (unsigned)seed);
        AltoTensor<LIT>* AT;
        create_alto(X, &AT, nprtn);
        // the truth: the COO MTTKRP (include/cpd.hpp:509-543), straight from
        // the coordinates; then the ALTO MTTKRP from the same factors
        for (int m = 0; m < X->nmodes; ++m) {
            if (target_mode != -1 && m != target_mode) continue;
            std::vector<FType> saved(M->U[m], M->U[m] + X->dims[m] * rank);
            mttkrp(X, M, (IType)m);
            std::vector<FType> truth(M->U[m], M->U[m] + X->dims[m] * rank);
            memcpy(M->U[m], saved.data(), saved.size() * sizeof(FType));
            mttkrp_alto_par(m, M->U, rank, AT, nullptr, nullptr);
            IType cnt = 0;
            for (IType k = 0; k < X->dims[m] * rank; ++k)
                if (fabs(truth[k] - M->U[m][k]) > 1.0e-6) cnt++;
            printf("mode %d: ", m);
            fflush(stdout);
            if (cnt == 0) fprintf(stderr, "Results of mttkrp_alto is correct\n");
            else fprintf(stderr, "Results of mttkrp_alto is incorrect by %llu\n", cnt);
            memcpy(M->U[m], saved.data(), saved.size() * sizeof(FType));
        }
        destroy_alto(AT);
        DestroyKruskalModel(M);
        return 0;
    }
    print_info(rank, max_iters, X);
    CreateKruskalModel(X->nmodes, X->dims, rank, &M);
    KruskalModelRandomInit(M, (unsigned)seed);
    AltoTensor<LIT>* AT;
    double t0 = wtime();
    create_alto(X, &AT, nprtn);
    double tc = wtime() - t0;
    if (do_bench) {
        printf("ALTO creation time:   %f\n", tc);
        if (model == APR) {
            cp_apr_alto(AT, M, max_iters, max_inner, 0.0, 0.0, kappa_tol, kappa, true);
        } else {
            alto_dropin::DeviceFactors d(AT->nmode, AT->dims, rank, M->U);
            auto pass = [&]() {
                for (int m = 0; m < AT->nmode; ++m)
                    if (target_mode == -1 || m == target_mode)
                        alto_dropin::check(alto_mttkrp(AT->dev, m, rank, (const void* const*)d.p.data(), d.p[m],
                                                       ALTO_PREC_F64, 1, nullptr),
                                           "mttkrp");
            };
            pass();  // warm-up
            alto_dropin::check(alto_device_synchronize(), "sync");
            t0 = wtime();
            for (int i = 0; i < max_iters; ++i) pass();
            alto_dropin::check(alto_device_synchronize(), "sync");
            printf("ALTO runtime:   %f\n", wtime() - t0);
        }
    } else if (model == ALS) {
        cpd_alto(AT, M, max_iters, epsilon);
    } else {
        cp_apr_alto(AT, M, max_iters, max_inner, epsilon, eps_div_zero, kappa_tol, kappa, false);
    }
    destroy_alto(AT);
    DestroyKruskalModel(M);
    return 0;
}

int main(int argc, char** argv)
{
    int max_iters = 10, max_inner = 10, target_mode = -1, save = 0, zero_based = 0;
    IType rank = 16;
    double epsilon = 1e-5, sparsity = 0.1;
    int seed = (int)time(NULL);
    bool do_check = false, do_bench = false;
    Model model = ALS;
    std::string text_file, text_out, bin_file, dims_arg;
    int c;
    while ((c = getopt_long(argc, argv, short_opt, long_opt, NULL)) != -1) {
        switch (c) {
        case 'h': usage(argv[0]); return 0;
        case 'v': fprintf(stdout, "%s version 0.1.2 (%s)\n", argv[0], alto_version()); return 0;
        case 'i': text_file = optarg; break;
        case 'o': text_out = optarg; break;  // accepted and unused, as in the reference
        case 'b': bin_file = optarg; break;
        case 'l':
            if (!strcmp(optarg, "ALS") || !strcmp(optarg, "als")) model = ALS;
            else if (!strcmp(optarg, "APR") || !strcmp(optarg, "apr")) model = APR;
            else { fprintf(stderr, "Invalid -model: %s.\n", optarg); return -1; }
            break;
        case 'r': rank = (IType)atoll(optarg); if ((long long)rank < 1) { fprintf(stderr, "Invalid -rank: %s.\n", optarg); return -1; } break;
        case 'm': max_iters = atoi(optarg); if (max_iters <= 0) { fprintf(stderr, "Invalid -max-iter: %s.\n", optarg); return -1; } break;
        case 'n': max_inner = atoi(optarg); if (max_inner <= 0) { fprintf(stderr, "Invalid -max-inner: %s.\n", optarg); return -1; } break;
        case 'x': seed = atoi(optarg); if (seed < 0) { fprintf(stderr, "Invalid -seed: %s.\n", optarg); return -1; } break;
        case 'd': dims_arg = optarg; break;
        case 't': target_mode = atoi(optarg); if (target_mode < -1) { fprintf(stderr, "Invalid -target-mode: %s.\n", optarg); return -1; } break;
        case 's':
            sparsity = atof(optarg);
            if (sparsity <= 0.0) { fprintf(stderr, "Invalid -sparsity: %s.\n", optarg); return -1; }
            break;
        case 'e': epsilon = atof(optarg); if (epsilon <= 0.0) { fprintf(stderr, "Invalid -epsilon: %s.\n", optarg); return -1; } break;
        case 'f': save = atoi(optarg); if (save < 0) { fprintf(stderr, "Invalid -file: %s.\n", optarg); return -1; } break;
        case 'z': zero_based = 1; break;
        case 'c': do_check = true; break;
        case 'p': do_bench = true; break;
        default: usage(argv[0]); return -1;
        }
    }
    (void)text_out;
    SparseTensor* X = nullptr;
    if (!bin_file.empty()) {
        ImportSparseTensor(bin_file.c_str(), BINARY_FORMAT, &X, zero_based);
        if (save) ExportSparseTensor(NULL, TEXT_FORMAT, X);
    } else if (!text_file.empty()) {
        ImportSparseTensor(text_file.c_str(), TEXT_FORMAT, &X, zero_based);
        if (save) ExportSparseTensor(NULL, BINARY_FORMAT, X);
    } else if (dims_arg.empty()) {
        fprintf(stderr, "No dims specified... exiting\n");
        usage(argv[0]);
        return -1;
    } else {
        // MakeSparseTensor (src/main.cpp:653-675): sparsity x prod(dims) Poisson
        // draws from a random rank-`rank` model; the reference splits its
        // uniform draws over its OpenMP threads, so the thread count
        // (OMP_NUM_THREADS, else the hardware threads) selects the tensor
        std::vector<IType> dims;
        size_t pos = 0;
        while (pos != std::string::npos) {
            const size_t e = dims_arg.find(',', pos);
            dims.push_back((IType)atoll(dims_arg.substr(pos, e == std::string::npos ? e : e - pos).c_str()));
            pos = e == std::string::npos ? e : e + 1;
        }
        IType cells = 1;
        for (IType d : dims) {
            if (d < rank) {
                fprintf(stderr, "Invalid -dims: every dimension must be at least the rank (%llu)\n",
                        (unsigned long long)rank);
                return -1;
            }
            cells *= d;
        }
        if (dims.empty() || dims.size() > MAX_NUM_MODES) {
            fprintf(stderr, "Invalid -dims: %s.\n", dims_arg.c_str());
            return -1;
        }
        const IType edges = (IType)(sparsity * (double)cells);
        if (edges < 1) {
            fprintf(stderr, "Invalid -sparsity: no nonzeros for %s.\n", dims_arg.c_str());
            return -1;
        }
        X = alto_gen::poisson_tensor((int)dims.size(), dims.data(), edges, rank, alto_gen::default_threads());
        if (save) {
            ExportSparseTensor(NULL, TEXT_FORMAT, X);
            ExportSparseTensor(NULL, BINARY_FORMAT, X);
        }
    }
    const char* base = strrchr(argv[0], '/');
    base = base ? base + 1 : argv[0];
    const int nprtn = 16;  // reference P = thread count; fixed here for reproducible partitions
    const int rc = strstr(base, "128") ? run<unsigned __int128>(X, model, rank, max_iters, max_inner, epsilon, seed,
                                                                 target_mode, do_check, do_bench, nprtn)
                                       : run<unsigned long long>(X, model, rank, max_iters, max_inner, epsilon, seed,
                                                                 target_mode, do_check, do_bench, nprtn);
    DestroySparseTensor(X);
    return rc;
}
