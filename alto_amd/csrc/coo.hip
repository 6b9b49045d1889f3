// MTTKRP straight from coordinate (COO) lists: the reference's sequential
// and locked-parallel COO MTTKRPs (include/cpd.hpp:460-543) that its -c
// check and COO CP-ALS use.  No ALTO tensor is built: each nonzero is read
// once from its coordinate arrays and its Khatri-Rao row product is added
// into the output row with FP64 atomics.
#include "alto_internal.hpp"
#include "kernels_common.hpp"

namespace alto {

struct CooArgs {
    const uint64_t* cidx[ALTO_MAXN];
    const double* U[ALTO_MAXN];
    const double* vals;
    double* out;
    uint64_t dims[ALTO_MAXN];
    uint64_t nnz;
    int nmodes, mode, R;
};

// L lanes per nonzero (L a power of two up to 64, at least the rank when
// the rank is below 64); lane c handles columns c, c + L, ...
template <int L>
__global__ void __launch_bounds__(256) k_coo_mttkrp(CooArgs a)
{
    const int lane = threadIdx.x % L;
    const uint64_t g0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / L;
    const uint64_t gs = (uint64_t)gridDim.x * blockDim.x / L;
    for (uint64_t i = g0; i < a.nnz; i += gs) {
        uint64_t row[ALTO_MAXN];
        bool in = true;
        for (int m = 0; m < a.nmodes; ++m) {
            row[m] = a.cidx[m][i];
            in = in && row[m] < a.dims[m];
        }
        if (!in) continue;  // a coordinate out of range adds nothing
        const double v = a.vals[i];
        for (int c = lane; c < a.R; c += L) {
            double p = v;
            for (int m = 0; m < a.nmodes; ++m)
                if (m != a.mode) p *= a.U[m][row[m] * a.R + c];
            atomicAdd(&a.out[row[a.mode] * a.R + c], p);
        }
    }
}

template <int L>
static void launch_coo(const CooArgs& a, hipStream_t s)
{
    const uint64_t groups = (a.nnz + 256 / L - 1) / (256 / L);
    const unsigned blocks = (unsigned)(groups < 8192 ? (groups ? groups : 1) : 8192);
    hipLaunchKernelGGL(k_coo_mttkrp<L>, dim3(blocks), dim3(256), 0, s, a);
}

}  // namespace alto

using namespace alto;

extern "C" int alto_coo_mttkrp(int nmodes, const uint64_t* dims, uint64_t nnz, const uint64_t* const* cidx,
                               const double* vals, int mode, uint64_t rank, const double* const* factors, double* out,
                               int accumulate, int ptr_kind, alto_stream_t stream)
{
    if (nmodes < 1 || nmodes > ALTO_MAXN || !dims || !cidx || !factors || !out || (nnz && !vals) || mode < 0 ||
        mode >= nmodes || rank < 1 || rank > (1u << 20) || (ptr_kind != ALTO_PTR_HOST && ptr_kind != ALTO_PTR_DEVICE)) {
        set_error("alto_coo_mttkrp: invalid argument");
        return ALTO_ERR_ARG;
    }
    for (int m = 0; m < nmodes; ++m)
        if (!dims[m] || (m != mode && !factors[m]) || (nnz && !cidx[m])) {
            set_error("alto_coo_mttkrp: invalid argument (mode %d)", m);
            return ALTO_ERR_ARG;
        }
    hipStream_t s = (hipStream_t)stream;
    const int R = (int)rank;
    CooArgs a{};
    a.nnz = nnz;
    a.nmodes = nmodes;
    a.mode = mode;
    a.R = R;
    for (int m = 0; m < nmodes; ++m) a.dims[m] = dims[m];
    const bool host = ptr_kind == ALTO_PTR_HOST;
    DevBuf ci[ALTO_MAXN], Ud[ALTO_MAXN], vd, od;
    if (host) {
        // host coordinates are checked here; on the device a nonzero with a
        // coordinate out of range is skipped
        for (int m = 0; m < nmodes; ++m)
            for (uint64_t i = 0; i < nnz; ++i)
                if (cidx[m][i] >= dims[m]) {
                    set_error("alto_coo_mttkrp: coordinate %llu of nonzero %llu in mode %d out of range",
                              (unsigned long long)cidx[m][i], (unsigned long long)i, m);
                    return ALTO_ERR_ARG;
                }
        for (int m = 0; m < nmodes; ++m) {
            if (m != mode) {
                ALTO_TRY(Ud[m].alloc(dims[m] * rank * 8));
                HIP_TRY(hipMemcpyAsync(Ud[m].p, factors[m], dims[m] * rank * 8, hipMemcpyHostToDevice, s));
                a.U[m] = Ud[m].as<double>();
            }
            if (nnz) {
                ALTO_TRY(ci[m].alloc(nnz * 8));
                HIP_TRY(hipMemcpyAsync(ci[m].p, cidx[m], nnz * 8, hipMemcpyHostToDevice, s));
            }
            a.cidx[m] = ci[m].as<uint64_t>();
        }
        if (nnz) {
            ALTO_TRY(vd.alloc(nnz * 8));
            HIP_TRY(hipMemcpyAsync(vd.p, vals, nnz * 8, hipMemcpyHostToDevice, s));
        }
        a.vals = vd.as<double>();
        ALTO_TRY(od.alloc(dims[mode] * rank * 8));
        if (accumulate) HIP_TRY(hipMemcpyAsync(od.p, out, dims[mode] * rank * 8, hipMemcpyHostToDevice, s));
        a.out = od.as<double>();
    } else {
        for (int m = 0; m < nmodes; ++m) {
            a.U[m] = m != mode ? factors[m] : nullptr;
            a.cidx[m] = cidx[m];
        }
        a.vals = vals;
        a.out = out;
    }
    if (!accumulate) ALTO_TRY(fill_zero<double>(a.out, dims[mode] * rank, s));
    if (nnz) {
        if (R >= 64) launch_coo<64>(a, s);
        else if (R > 16) launch_coo<32>(a, s);
        else if (R > 8) launch_coo<16>(a, s);
        else if (R > 4) launch_coo<8>(a, s);
        else launch_coo<4>(a, s);
        HIP_TRY(hipGetLastError());
    }
    if (host) {
        HIP_TRY(hipMemcpyAsync(out, od.p, dims[mode] * rank * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));  // temporaries are released below
    }
    return ALTO_OK;
}
