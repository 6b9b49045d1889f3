// ALTO format construction on gfx950: linearization (bit deposit) and
// delinearization (bit extract) as byte-LUT bit permutations, a stable LSD
// radix sort written for wave64, line-balanced partition pointers,
// partition intervals, conflict-resolution (CR) bits, and the per-mode
// chunk-ordered nonzero streams consumed by MTTKRP / Phi.
//
// Reference behaviour followed: create_alto (include/alto.hpp:108-294),
// setup_packed_alto (:669-770), sort_alto (:772-803), prtn_alto (:805-1028).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>

#include "alto_internal.hpp"
#include "kernels_common.hpp"

namespace alto {

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

void set_error(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

void set_error_hip(hipError_t e, const char* what)
{
    set_error("%s failed: %s", what, hipGetErrorString(e));
}

const char* last_error() { return g_err.c_str(); }

// Caching device allocator: freed blocks are kept per device in power-of-two
// size classes and reused, so concurrent jobs never call hipFree (which
// synchronizes the device) on their hot path.  alto_device_trim returns the
// cache to the runtime.
namespace {
struct Pool {
    std::mutex mu;
    std::multimap<size_t, void*> free_blocks[64];  // per device: class bytes -> ptr
    size_t cached[64] = {0};
    size_t in_use[64] = {0};
};
Pool& pool()
{
    static Pool* p = new Pool();
    return *p;
}
size_t size_class(size_t n)
{
    if (n <= 256) return 256;
    size_t c = 256;
    while (c < n) c <<= 1;
    return c;
}
}  // namespace

int DevBuf::alloc(size_t n)
{
    release();
    if (n == 0) n = 8;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    const size_t cls = size_class(n);
    Pool& P = pool();
    {
        std::lock_guard<std::mutex> g(P.mu);
        auto& fb = P.free_blocks[dev & 63];
        auto it = fb.find(cls);
        if (it != fb.end()) {
            p = it->second;
            fb.erase(it);
            P.cached[dev & 63] -= cls;
            P.in_use[dev & 63] += cls;
            bytes = n;
            cap = cls;
            device = dev;
            return ALTO_OK;
        }
    }
    hipError_t e = hipMalloc(&p, cls);
    if (e == hipErrorOutOfMemory) {  // give the cache back and retry once
        (void)hipGetLastError();
        alto_device_trim();
        e = hipMalloc(&p, cls);
    }
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc(%zu bytes) failed: %s", cls, hipGetErrorString(e));
        (void)hipGetLastError();
        return e == hipErrorOutOfMemory ? ALTO_ERR_OOM : ALTO_ERR_HIP;
    }
    {
        std::lock_guard<std::mutex> g(P.mu);
        P.in_use[dev & 63] += cls;
    }
    bytes = n;
    cap = cls;
    device = dev;
    return ALTO_OK;
}

void DevBuf::release()
{
    if (p) {
        Pool& P = pool();
        std::lock_guard<std::mutex> g(P.mu);
        P.free_blocks[device & 63].emplace(cap, p);
        P.cached[device & 63] += cap;
        P.in_use[device & 63] -= cap;
    }
    p = nullptr;
    bytes = 0;
    cap = 0;
}

size_t pool_in_use(int dev)
{
    Pool& P = pool();
    std::lock_guard<std::mutex> g(P.mu);
    return P.in_use[dev & 63];
}

}  // namespace alto

int alto_tensor::note_use(hipStream_t s)
{
    std::lock_guard<std::mutex> g(use_mu);
    hipEvent_t e = nullptr;
    for (auto& u : uses)
        if (u.first == s) e = u.second;
    if (!e) {
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        uses.emplace_back(s, e);
    }
    HIP_TRY(hipEventRecord(e, s));
    return ALTO_OK;
}

int alto_tensor::wait_uses()
{
    std::lock_guard<std::mutex> g(use_mu);
    for (auto& u : uses) HIP_TRY(hipEventSynchronize(u.second));
    return ALTO_OK;
}

alto_tensor::~alto_tensor()
{
    for (auto& u : uses) {
        (void)hipEventSynchronize(u.second);
        (void)hipEventDestroy(u.second);
    }
}

namespace alto {

// ------------------------------------------------------------------ masks
// setup_packed_alto(at, LSB_FIRST, SHORT_FIRST): modes ordered by bit count
// (stable), one bit per still-active mode per level, LSB first.
int compute_masks(int nmode, const uint64_t* dims, int W, uint64_t* masks, uint64_t* alto_mask)
{
    int mode[ALTO_MAXN], bits[ALTO_MAXN], total = 0, maxb = 0;
    for (int n = 0; n < nmode; ++n) {
        uint64_t d = dims[n] - 1;
        mode[n] = n;
        bits[n] = d ? 64 - __builtin_clzll(d) : 0;
        total += bits[n];
        maxb = std::max(maxb, bits[n]);
    }
    if (total > 64 * W) return -1;
    std::stable_sort(mode, mode + nmode, [&](int a, int b) { return bits[a] < bits[b]; });
    memset(masks, 0, sizeof(uint64_t) * nmode * W);
    int shift = 0;
    for (int level = 0; level < maxb; ++level)
        for (int k = 0; k < nmode; ++k)
            if (level < bits[mode[k]]) {
                masks[mode[k] * W + (shift >> 6)] |= 1ull << (shift & 63);
                ++shift;
            }
    for (int w = 0; w < W; ++w) {
        alto_mask[w] = 0;
        for (int n = 0; n < nmode; ++n) alto_mask[w] |= masks[n * W + w];
    }
    return total;
}

// Packed layout: mode n at bit offset sum_{m<n} bits[m].  The key->packed
// map is a bit permutation: the j-th set bit of mask_n goes to off[n]+j.
int build_luts(alto_tensor* t, hipStream_t s)
{
    PackLayout& L = t->pack;
    L.nmode = t->nmode;
    int off = 0;
    int keypos_of[128];   // packed bit -> key bit
    for (int n = 0; n < t->nmode; ++n) {
        L.off[n] = off;
        int j = 0;
        for (int q = 0; q < 64 * t->W; ++q)
            if ((t->masks[n * t->W + (q >> 6)] >> (q & 63)) & 1) keypos_of[off + j++] = q;
        L.bits[n] = j;
        off += j;
    }
    L.pw = off > 64 ? 2 : 1;
    int packpos_of[128];
    for (int q = 0; q < 128; ++q) packpos_of[q] = -1;
    for (int p = 0; p < off; ++p) packpos_of[keypos_of[p]] = p;

    const int nbits = off;
    auto fill = [nbits, s](BitPermLut& lut, int in_w, int out_w, const int* map /* in bit -> out bit or -1 */) -> int {
        lut.in_words = in_w;
        lut.out_words = out_w;
        lut.nb = std::min(in_w * 8, (nbits + 7) / 8);
        lut.host.assign((size_t)in_w * 8 * 256 * out_w, 0);
        for (int b = 0; b < in_w * 8; ++b)
            for (int v = 0; v < 256; ++v) {
                uint64_t* e = &lut.host[((size_t)b * 256 + v) * out_w];
                for (int i = 0; i < 8; ++i) {
                    if (!((v >> i) & 1)) continue;
                    int o = map[b * 8 + i];
                    if (o >= 0) e[o >> 6] |= 1ull << (o & 63);
                }
            }
        ALTO_TRY(lut.dev.alloc(lut.host.size() * 8));
        HIP_TRY(hipMemcpyAsync(lut.dev.p, lut.host.data(), lut.host.size() * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        return ALTO_OK;
    };
    int map_delin[128], map_lin[128];
    for (int q = 0; q < 128; ++q) map_delin[q] = q < 64 * t->W ? packpos_of[q] : -1;
    for (int p = 0; p < 128; ++p) map_lin[p] = p < off ? keypos_of[p] : -1;
    ALTO_TRY(fill(t->delin, t->W, L.pw, map_delin));
    ALTO_TRY(fill(t->lin, L.pw, t->W, map_lin));
    return ALTO_OK;
}

// ------------------------------------------------------------------ kernels
// Byte-LUT bit permutation: out = OR_b lut[b][byte_b(in)], b < nb (the bytes
// that can carry index bits).  The table is staged in LDS when it fits.
template <int IW, int OW>
__device__ inline void lut_apply(const uint64_t* __restrict__ lut, int nb, const uint64_t (&in)[IW], uint64_t (&out)[OW])
{
#pragma unroll
    for (int w = 0; w < OW; ++w) out[w] = 0;
#pragma unroll
    for (int b = 0; b < IW * 8; ++b) {
        if (b < nb) {
            unsigned v = (unsigned)(in[b >> 3] >> ((b & 7) * 8)) & 0xffu;
            const uint64_t* e = lut + ((size_t)b * 256 + v) * OW;
#pragma unroll
            for (int w = 0; w < OW; ++w) out[w] |= e[w];
        }
    }
}

struct LutRef {
    const uint64_t* g;  // device table
    int nb;             // input bytes covered
    int lds;            // staged in LDS?
};

template <int OW>
__device__ inline const uint64_t* stage_lut(uint64_t* s_lut, LutRef L)
{
    if (!L.lds) return L.g;
    const int n = L.nb * 256 * OW;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s_lut[i] = L.g[i];
    __syncthreads();
    return s_lut;
}

struct CoordPtrs {
    const uint64_t* c[ALTO_MAXN];
};

struct DimArr {
    uint64_t d[ALTO_MAXN];
};

// linearize (include/alto.hpp:184-204): pack the coordinates, then permute
// the packed bits into ALTO key order; vals cast to ValType.
template <int PW, int W>
__global__ void __launch_bounds__(256) k_linearize(CoordPtrs cp, int nmode, PackLayout L, const double* __restrict__ vin,
                                                   int val_u64, LutRef g_lut, uint64_t n,
                                                   uint64_t* __restrict__ keys, double* __restrict__ vout,
                                                   uint32_t* __restrict__ payload)
{
    extern __shared__ uint64_t s_lut[];
    const uint64_t* lut = stage_lut<W>(s_lut, g_lut);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t pk[PW];
#pragma unroll
        for (int w = 0; w < PW; ++w) pk[w] = 0;
        for (int m = 0; m < nmode; ++m) {
            uint64_t c = cp.c[m][i];
            int o = L.off[m];
            if (PW == 1) {
                pk[0] |= c << o;
            } else {
                if (o < 64) {
                    pk[0] |= c << o;
                    if (o > 0) pk[PW - 1] |= c >> (64 - o);
                } else {
                    pk[PW - 1] |= c << (o - 64);
                }
            }
        }
        uint64_t k[W];
        lut_apply<PW, W>(lut, g_lut.nb, pk, k);
#pragma unroll
        for (int w = 0; w < W; ++w) keys[i * W + w] = k[w];
        if (vout) {
            double v = vin[i];
            vout[i] = val_u64 ? (double)(unsigned long long)v : v;
        }
        if (payload) payload[i] = (uint32_t)i;
    }
}

// coordinates must lie inside dims (the reference asserts it in
// CreateSparseTensor, src/sptensor.cpp:276-281); a violation sets *bad
__global__ void __launch_bounds__(256) k_validate(CoordPtrs cp, int nmode, DimArr dims, uint64_t n, int* __restrict__ bad)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        for (int m = 0; m < nmode; ++m)
            if (cp.c[m][i] >= dims.d[m]) atomicOr(bad, 1);
}

// generic bit permutation of an array of keys (pdep / pext of all modes)
template <int IW, int OW>
__global__ void __launch_bounds__(256) k_bitperm(const uint64_t* __restrict__ in, LutRef g_lut,
                                                 uint64_t n, uint64_t* __restrict__ out)
{
    extern __shared__ uint64_t s_lut[];
    const uint64_t* lut = stage_lut<OW>(s_lut, g_lut);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t a[IW], b[OW];
#pragma unroll
        for (int w = 0; w < IW; ++w) a[w] = in[i * IW + w];
        lut_apply<IW, OW>(lut, g_lut.nb, a, b);
#pragma unroll
        for (int w = 0; w < OW; ++w) out[i * OW + w] = b[w];
    }
}

// delinearize all modes: keys -> coordinates (nmode x n)
template <int W, int PW>
__global__ void __launch_bounds__(256) k_delin_coords(const uint64_t* __restrict__ keys, LutRef g_lut,
                                                      PackLayout L, uint64_t n, uint64_t* __restrict__ coords)
{
    extern __shared__ uint64_t s_lut[];
    const uint64_t* lut = stage_lut<PW>(s_lut, g_lut);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k[W], pk[PW];
#pragma unroll
        for (int w = 0; w < W; ++w) k[w] = keys[i * W + w];
        lut_apply<W, PW>(lut, g_lut.nb, k, pk);
        for (int m = 0; m < L.nmode; ++m) coords[(uint64_t)m * n + i] = field<PW>(pk, L.off[m], L.bits[m]);
    }
}

// packed coords -> coordinates (used by alto_pdep_keys inputs is the reverse)
template <int PW, int W>
__global__ void __launch_bounds__(256) k_pack_lin(const uint64_t* __restrict__ coords, PackLayout L, LutRef g_lut,
                                                  uint64_t n, uint64_t* __restrict__ keys)
{
    extern __shared__ uint64_t s_lut[];
    const uint64_t* lut = stage_lut<W>(s_lut, g_lut);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t pk[PW];
#pragma unroll
        for (int w = 0; w < PW; ++w) pk[w] = 0;
        for (int m = 0; m < L.nmode; ++m) {
            uint64_t c = coords[(uint64_t)m * n + i];
            int o = L.off[m];
            if (PW == 1) pk[0] |= c << o;
            else if (o < 64) {
                pk[0] |= c << o;
                if (o > 0) pk[PW - 1] |= c >> (64 - o);
            } else pk[PW - 1] |= c << (o - 64);
        }
        uint64_t k[W];
        lut_apply<PW, W>(lut, g_lut.nb, pk, k);
#pragma unroll
        for (int w = 0; w < W; ++w) keys[i * W + w] = k[w];
    }
}

static int grid_for(uint64_t n, int threads = 256, int max_blocks = 4096)
{
    uint64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > (uint64_t)max_blocks) b = max_blocks;
    return (int)b;
}

static LutRef lutref(const BitPermLut& L)
{
    LutRef r;
    r.g = L.dev.as<uint64_t>();
    r.nb = L.nb;
    r.lds = (size_t)L.nb * 256 * L.out_words * 8 <= 48 * 1024;
    return r;
}
static size_t lut_smem(const BitPermLut& L)
{
    size_t b = (size_t)L.nb * 256 * L.out_words * 8;
    return b <= 48 * 1024 ? b : 0;
}

// ------------------------------------------------------------------ radix sort
// Stable LSD radix sort, 8-bit digits, tiles of 256 threads x 16 keys with a
// striped layout (key j*256+t of a tile belongs to round j, lane order t):
// per-round wave ranks come from 8 ballots (match on the digit), cross-wave
// offsets from per-wave digit counts in LDS.
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

template <int W>
__device__ inline unsigned digit_of(const uint64_t* keys, uint64_t i, int shift)
{
    uint64_t w = keys[i * W + (shift >> 6)];
    return (unsigned)(w >> (shift & 63)) & 0xffu;
}

template <int W>
__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint64_t* __restrict__ keys, uint64_t n, int shift,
                                                        uint32_t* __restrict__ hist, int nblocks)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int j = 0; j < RS_ITEMS; ++j) {
        uint64_t i = base + (uint64_t)j * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[digit_of<W>(keys, i, shift)], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan, three phases over tiles of 1024 elements
__global__ void __launch_bounds__(1024) k_scan_tiles(uint32_t* __restrict__ a, uint64_t m, uint32_t* __restrict__ sums)
{
    __shared__ uint32_t s[1024];
    uint64_t i = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
    uint32_t v = i < m ? a[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < m) a[i] = s[threadIdx.x] - v;
    if (threadIdx.x == 1023) sums[blockIdx.x] = s[1023];
}

__global__ void __launch_bounds__(1024) k_scan_sums(uint32_t* __restrict__ sums, int nt)
{
    // single block: sequential chunks of 1024
    __shared__ uint32_t s[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nt; base += 1024) {
        int i = base + threadIdx.x;
        uint32_t v = i < nt ? sums[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint32_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nt) sums[i] = s[threadIdx.x] - v + carry;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(1024) k_scan_add(uint32_t* __restrict__ a, uint64_t m, const uint32_t* __restrict__ sums)
{
    uint64_t i = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
    if (i < m) a[i] += sums[blockIdx.x];
}

template <int W>
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ pay,
                                                           uint64_t n, int shift, const uint32_t* __restrict__ offs,
                                                           int nblocks, uint64_t* __restrict__ keys_out,
                                                           uint32_t* __restrict__ pay_out)
{
    __shared__ uint32_t running[256];
    __shared__ uint32_t wcnt[RS_WAVES][256];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    running[t] = offs[(size_t)t * nblocks + blockIdx.x];
    for (int w = 0; w < RS_WAVES; ++w) wcnt[w][t] = 0;
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1;
    uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    for (int j = 0; j < RS_ITEMS; ++j) {
        uint64_t i = base + (uint64_t)j * RS_THREADS + t;
        bool valid = i < n;
        unsigned d = valid ? digit_of<W>(keys, i, shift) : 0;
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bb = __ballot(valid && ((d >> b) & 1));
            match &= ((d >> b) & 1) ? bb : ~bb;
        }
        unsigned rank = (unsigned)__popcll(match & lt);
        if (valid && rank == 0) wcnt[wave][d] = (unsigned)__popcll(match);
        __syncthreads();
        if (valid) {
            uint32_t pos = running[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
#pragma unroll
            for (int q = 0; q < W; ++q) keys_out[(uint64_t)pos * W + q] = keys[i * W + q];
            pay_out[pos] = pay[i];
        }
        __syncthreads();
        uint32_t tot = 0;
        for (int w = 0; w < RS_WAVES; ++w) {
            tot += wcnt[w][t];
            wcnt[w][t] = 0;
        }
        running[t] += tot;
        __syncthreads();
    }
}

static int scan_exclusive(uint32_t* a, uint64_t m, DevBuf& tmp, hipStream_t s)
{
    uint64_t nt = (m + 1023) / 1024;
    if (nt > (1ull << 31)) {
        set_error("scan too large");
        return ALTO_ERR_UNSUPPORTED;
    }
    if (tmp.bytes < nt * 4) ALTO_TRY(tmp.alloc(nt * 4));
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(1024), 0, s, a, m, tmp.as<uint32_t>());
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, tmp.as<uint32_t>(), (int)nt);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nt), dim3(1024), 0, s, a, m, tmp.as<uint32_t>());
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

// Sorts (keys, payload) by the low `bits` bits of the keys; results end in
// the input buffers.
template <int W>
static int radix_sort_impl(uint64_t* keys, uint32_t* pay, uint64_t n, int bits, hipStream_t s)
{
    if (n <= 1 || bits <= 0) return ALTO_OK;
    if (bits > 64 * W) {
        set_error("radix sort: %d key bits exceed the %d-bit key", bits, 64 * W);
        return ALTO_ERR_UNSUPPORTED;
    }
    if (n >= (1ull << 32)) {
        set_error("radix sort supports fewer than 2^32 nonzeros");
        return ALTO_ERR_UNSUPPORTED;
    }
    int nblocks = (int)((n + RS_TILE - 1) / RS_TILE);
    DevBuf kb, pb, hist, tmp;
    ALTO_TRY(kb.alloc(n * W * 8));
    ALTO_TRY(pb.alloc(n * 4));
    ALTO_TRY(hist.alloc((size_t)256 * nblocks * 4));
    uint64_t* ka = keys;
    uint64_t* kz = kb.as<uint64_t>();
    uint32_t* pa = pay;
    uint32_t* pz = pb.as<uint32_t>();
    int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        int shift = p * 8;
        hipLaunchKernelGGL(k_rs_hist<W>, dim3(nblocks), dim3(RS_THREADS), 0, s, ka, n, shift, hist.as<uint32_t>(), nblocks);
        ALTO_TRY(scan_exclusive(hist.as<uint32_t>(), (uint64_t)256 * nblocks, tmp, s));
        hipLaunchKernelGGL(k_rs_scatter<W>, dim3(nblocks), dim3(RS_THREADS), 0, s, ka, pa, n, shift, hist.as<uint32_t>(),
                           nblocks, kz, pz);
        HIP_TRY(hipGetLastError());
        std::swap(ka, kz);
        std::swap(pa, pz);
    }
    if (ka != keys) {
        HIP_TRY(hipMemcpyAsync(keys, ka, n * W * 8, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(pay, pa, n * 4, hipMemcpyDeviceToDevice, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return ALTO_OK;
}

int radix_sort_pairs(uint64_t* keys, int W, uint32_t* payload, uint64_t n, int bits, hipStream_t s)
{
    return W == 1 ? radix_sort_impl<1>(keys, payload, n, bits, s) : radix_sort_impl<2>(keys, payload, n, bits, s);
}

// ------------------------------------------------------------------ post-sort
__global__ void k_gather_f64(const double* __restrict__ in, const uint32_t* __restrict__ perm, uint64_t n,
                             double* __restrict__ out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[perm[i]];
}

// partition intervals (include/alto.hpp:874-899): per (partition, mode) the
// [min, max] row; partition p holds [p*per, min((p+1)*per, nnz)).
template <int W, int PW>
__global__ void __launch_bounds__(256) k_intervals(const uint64_t* __restrict__ keys, LutRef g_lut,
                                                   PackLayout L, uint64_t n, uint64_t per, unsigned long long* __restrict__ iv)
{
    extern __shared__ uint64_t s_lut[];
    __shared__ unsigned long long s_min[ALTO_MAXN], s_max[ALTO_MAXN];
    const uint64_t* lut = stage_lut<PW>(s_lut, g_lut);
    for (uint64_t tile = blockIdx.x; tile * 256 < n; tile += gridDim.x) {
        uint64_t base = tile * 256;
        uint64_t p0 = base / per;
        if (threadIdx.x < ALTO_MAXN) {
            s_min[threadIdx.x] = ~0ull;
            s_max[threadIdx.x] = 0;
        }
        __syncthreads();
        uint64_t i = base + threadIdx.x;
        if (i < n) {
            uint64_t k[W], pk[PW];
#pragma unroll
            for (int w = 0; w < W; ++w) k[w] = keys[i * W + w];
            lut_apply<W, PW>(lut, g_lut.nb, k, pk);
            uint64_t p = i / per;
            for (int m = 0; m < L.nmode; ++m) {
                unsigned long long r = field<PW>(pk, L.off[m], L.bits[m]);
                if (p == p0) {
                    atomicMin(&s_min[m], r);
                    atomicMax(&s_max[m], r);
                } else {
                    atomicMin(&iv[(p * L.nmode + m) * 2], r);
                    atomicMax(&iv[(p * L.nmode + m) * 2 + 1], r);
                }
            }
        }
        __syncthreads();
        if (threadIdx.x < (unsigned)L.nmode) {
            int m = threadIdx.x;
            if (s_max[m] >= s_min[m] || s_min[m] != ~0ull) {
                atomicMin(&iv[(p0 * L.nmode + m) * 2], s_min[m]);
                atomicMax(&iv[(p0 * L.nmode + m) * 2 + 1], s_max[m]);
            }
        }
        __syncthreads();
    }
}

// CR ownership (include/alto.hpp:970-1020): rows reached from more than one
// partition get the mode's CR bit in every one of their keys.
template <int W, int PW>
__global__ void __launch_bounds__(256) k_cr_owner(const uint64_t* __restrict__ keys, LutRef g_lut,
                                                  PackLayout L, int mode, uint64_t n, uint64_t per, int* __restrict__ lo,
                                                  int* __restrict__ hi)
{
    extern __shared__ uint64_t s_lut[];
    const uint64_t* lut = stage_lut<PW>(s_lut, g_lut);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k[W], pk[PW];
#pragma unroll
        for (int w = 0; w < W; ++w) k[w] = keys[i * W + w];
        lut_apply<W, PW>(lut, g_lut.nb, k, pk);
        uint64_t r = field<PW>(pk, L.off[mode], L.bits[mode]);
        const int p = (int)(i / per);
        // most updates change nothing (many nonzeros per row and partition):
        // read before issuing the atomic
        if (lo[r] > p) atomicMin(&lo[r], p);
        if (hi[r] < p) atomicMax(&hi[r], p);
    }
}

template <int W, int PW>
__global__ void __launch_bounds__(256) k_cr_set(uint64_t* __restrict__ keys, LutRef g_lut, PackLayout L,
                                                int mode, uint64_t n, const int* __restrict__ lo, const int* __restrict__ hi,
                                                uint64_t cm0, uint64_t cm1)
{
    extern __shared__ uint64_t s_lut[];
    const uint64_t* lut = stage_lut<PW>(s_lut, g_lut);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k[W], pk[PW];
#pragma unroll
        for (int w = 0; w < W; ++w) k[w] = keys[i * W + w];
        lut_apply<W, PW>(lut, g_lut.nb, k, pk);
        uint64_t r = field<PW>(pk, L.off[mode], L.bits[mode]);
        if (lo[r] != hi[r]) {
            keys[i * W] = k[0] | cm0;
            if (W == 2) keys[i * W + W - 1] = k[W - 1] | cm1;
        }
    }
}

__global__ void k_fill_i32(int* __restrict__ a, uint64_t n, int v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) a[i] = v;
}

__global__ void k_fill_u64(unsigned long long* __restrict__ a, uint64_t n, unsigned long long v0, unsigned long long v1)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = (i & 1) ? v1 : v0;
}

// per-mode streams: composite key (chunk << bits_mode) | row
template <int W, int PW>
__global__ void __launch_bounds__(256) k_mode_key(const uint64_t* __restrict__ keys, LutRef g_lut,
                                                  PackLayout L, int mode, int chunk_log2, uint64_t n,
                                                  uint64_t* __restrict__ ck, uint32_t* __restrict__ pay)
{
    extern __shared__ uint64_t s_lut[];
    const uint64_t* lut = stage_lut<PW>(s_lut, g_lut);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k[W], pk[PW];
#pragma unroll
        for (int w = 0; w < W; ++w) k[w] = keys[i * W + w];
        lut_apply<W, PW>(lut, g_lut.nb, k, pk);
        uint64_t r = field<PW>(pk, L.off[mode], L.bits[mode]);
        ck[i] = ((i >> chunk_log2) << L.bits[mode]) | r;
        pay[i] = (uint32_t)i;
    }
}

template <int W, int PW>
__global__ void __launch_bounds__(256) k_mode_gather(const uint64_t* __restrict__ keys, const double* __restrict__ vals,
                                                     LutRef g_lut, const uint32_t* __restrict__ perm,
                                                     uint64_t n, uint64_t* __restrict__ pk_out, double* __restrict__ v64,
                                                     float* __restrict__ v32)
{
    extern __shared__ uint64_t s_lut[];
    const uint64_t* lut = stage_lut<PW>(s_lut, g_lut);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t src = perm[i];
        uint64_t k[W], pk[PW];
#pragma unroll
        for (int w = 0; w < W; ++w) k[w] = keys[src * W + w];
        lut_apply<W, PW>(lut, g_lut.nb, k, pk);
#pragma unroll
        for (int w = 0; w < PW; ++w) pk_out[i * PW + w] = pk[w];
        double v = vals[src];
        if (v64) v64[i] = v;
        if (v32) v32[i] = (float)v;
    }
}

// ------------------------------------------------------------------ fiber layout
// Sort key of a fiber-ordered stream: (chunk, row of `mode`, row of `amode`)
// of the nonzero at ALTO position idx[j] (j itself when idx is null); the
// chunk is that of the ALTO position.
template <int W, int PW>
__global__ void __launch_bounds__(256) k_fiber_key(const uint64_t* __restrict__ keys, LutRef g_lut, PackLayout L,
                                                   int mode, int amode, int chunk_log2, uint64_t n,
                                                   const uint32_t* __restrict__ idx, uint64_t* __restrict__ ck,
                                                   uint32_t* __restrict__ pay)
{
    extern __shared__ uint64_t s_lut[];
    const uint64_t* lut = stage_lut<PW>(s_lut, g_lut);
    const int ba = L.bits[amode], bn = L.bits[mode];
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = idx ? idx[j] : j;
        uint64_t k[W], pk[PW];
#pragma unroll
        for (int w = 0; w < W; ++w) k[w] = keys[i * W + w];
        lut_apply<W, PW>(lut, g_lut.nb, k, pk);
        const uint64_t r = field<PW>(pk, L.off[mode], bn), a = field<PW>(pk, L.off[amode], ba);
        ck[j] = (((i >> chunk_log2) << bn | r) << ba) | a;
        pay[j] = (uint32_t)i;
    }
}

// h[i] = 1 where a fiber starts in the sorted keys
__global__ void k_fiber_heads(const uint64_t* __restrict__ ck, uint64_t n, uint32_t* __restrict__ h)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        h[i] = (i == 0 || ck[i] != ck[i - 1]) ? 1u : 0u;
}

// id: in, the exclusive scan of the heads; out, the fiber of each nonzero.
// fstart[f] = first position of fiber f, fstart[nfibers] = n.
__global__ void k_fiber_starts(const uint64_t* __restrict__ ck, uint64_t n, uint32_t* __restrict__ id,
                               uint32_t* __restrict__ fstart)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t head = (i == 0 || ck[i] != ck[i - 1]) ? 1u : 0u;
        const uint32_t f = id[i] + head - 1;
        id[i] = f;
        if (head) fstart[f] = (uint32_t)i;
        if (i == n - 1) fstart[f + 1] = (uint32_t)n;
    }
}

__device__ inline uint32_t fiber_long(const uint32_t* id, const uint32_t* fstart, uint64_t i, uint32_t min_len)
{
    const uint32_t f = id[i];
    return fstart[f + 1] - fstart[f] >= min_len ? 1u : 0u;
}

__global__ void k_fiber_class(const uint32_t* __restrict__ id, const uint32_t* __restrict__ fstart, uint64_t n,
                              uint32_t min_len, uint32_t* __restrict__ cls)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        cls[i] = fiber_long(id, fstart, i, min_len);
}

// stable split: long-fiber nonzeros to [0, nlong), the rest after them;
// pos = exclusive scan of the long flags
__global__ void k_fiber_scatter(const uint32_t* __restrict__ id, const uint32_t* __restrict__ fstart,
                                const uint32_t* __restrict__ pos, const uint32_t* __restrict__ pay, uint64_t n,
                                uint32_t min_len, uint32_t* __restrict__ perm, uint64_t* __restrict__ nlong_out)
{
    const uint64_t nlong = (uint64_t)pos[n - 1] + fiber_long(id, fstart, n - 1, min_len);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t p = fiber_long(id, fstart, i, min_len) ? pos[i] : nlong + (i - pos[i]);
        perm[p] = pay[i];
        if (i == 0) *nlong_out = nlong;
    }
}

#define DISPATCH_WPW(W_, PW_, ...)                                              \
    do {                                                                        \
        if ((W_) == 1 && (PW_) == 1) { constexpr int W = 1, PW = 1; __VA_ARGS__; } \
        else if ((W_) == 2 && (PW_) == 1) { constexpr int W = 2, PW = 1; __VA_ARGS__; } \
        else { constexpr int W = 2, PW = 2; __VA_ARGS__; }                      \
    } while (0)

int delinearize_all(const alto_tensor* t, uint64_t* coords, hipStream_t s)
{
    DISPATCH_WPW(t->W, t->pack.pw, {
        hipLaunchKernelGGL((k_delin_coords<W, PW>), dim3(grid_for(t->nnz)), dim3(256), lut_smem(t->delin), s,
                           t->key.as<uint64_t>(), lutref(t->delin), t->pack, t->nnz, coords);
    });
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

// ------------------------------------------------------- row frequency ranks
// Nonzeros per row of mode m, from stream m (inside each chunk the rows are
// contiguous): each run adds (end - start) through two wrapping additions.
template <int PW>
__global__ void __launch_bounds__(256) k_row_counts(const uint64_t* __restrict__ pk, uint64_t n, int off, int bits,
                                                    int chunk_log2, uint32_t* __restrict__ cnt)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k[PW], kp[PW], kn[PW];
#pragma unroll
        for (int w = 0; w < PW; ++w) k[w] = pk[i * PW + w];
        const uint64_t r = field<PW>(k, off, bits);
        bool start = i == 0, end = i + 1 == n;
        if (!start) {
#pragma unroll
            for (int w = 0; w < PW; ++w) kp[w] = pk[(i - 1) * PW + w];
            start = field<PW>(kp, off, bits) != r || (i >> chunk_log2) != ((i - 1) >> chunk_log2);
        }
        if (!end) {
#pragma unroll
            for (int w = 0; w < PW; ++w) kn[w] = pk[(i + 1) * PW + w];
            end = field<PW>(kn, off, bits) != r || (i >> chunk_log2) != ((i + 1) >> chunk_log2);
        }
        if (start) atomicAdd(&cnt[r], 0u - (uint32_t)i);
        if (end) atomicAdd(&cnt[r], (uint32_t)(i + 1));
    }
}

__global__ void k_rank_keys(const uint32_t* __restrict__ cnt, uint64_t d, int bits, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ pay)
{
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < d; r += (uint64_t)gridDim.x * blockDim.x) {
        keys[r] = ((uint64_t)(0xffffffffu - cnt[r]) << bits) | r;
        pay[r] = (uint32_t)r;
    }
}

__global__ void k_rank_scatter(const uint32_t* __restrict__ pay, uint64_t d, uint32_t* __restrict__ rank_row,
                               uint32_t* __restrict__ row_rank)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < d; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = pay[k];
        rank_row[k] = r;
        row_rank[r] = (uint32_t)k;
    }
}

struct RankTabs {
    const uint32_t* row_rank[ALTO_MAXN];
};

// rewrite the other modes' fields of one stream from rows to ranks
template <int PW>
__global__ void __launch_bounds__(256) k_stream_rank(uint64_t* __restrict__ pk, uint64_t n, PackLayout L, int mode,
                                                     RankTabs T)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k[PW];
#pragma unroll
        for (int w = 0; w < PW; ++w) k[w] = pk[i * PW + w];
        for (int o = 0; o < L.nmode; ++o) {
            if (o == mode || L.bits[o] == 0) continue;
            set_field<PW>(k, L.off[o], L.bits[o], T.row_rank[o][field<PW>(k, L.off[o], L.bits[o])]);
        }
#pragma unroll
        for (int w = 0; w < PW; ++w) pk[i * PW + w] = k[w];
    }
}

// rank tables from the (row-ordered) streams, then every unranked stream
// rewritten; streams must all exist
static int rank_streams(alto_tensor* t, hipStream_t s)
{
    const uint64_t n = t->nnz;
    for (int m = 0; m < t->nmode; ++m) {
        if (t->rank_row[m].p) continue;
        const uint64_t d = t->dims[m];
        DevBuf cnt, keys, pay;
        ALTO_TRY(cnt.alloc(d * 4));
        ALTO_TRY(keys.alloc(d * 8));
        ALTO_TRY(pay.alloc(d * 4));
        ALTO_TRY(t->rank_row[m].alloc(d * 4));
        ALTO_TRY(t->row_rank[m].alloc(d * 4));
        ALTO_TRY(fill_zero(cnt.as<uint32_t>(), d, s));
        RANK_DISPATCH_PW(t->pack.pw, {
            hipLaunchKernelGGL((k_row_counts<PW>), dim3(grid_for(n)), dim3(256), 0, s, t->ms[m].pk.as<uint64_t>(), n,
                               t->pack.off[m], t->pack.bits[m], t->ms[m].chunk_log2, cnt.as<uint32_t>());
        });
        hipLaunchKernelGGL(k_rank_keys, dim3(grid_for(d)), dim3(256), 0, s, cnt.as<uint32_t>(), d, t->pack.bits[m],
                           keys.as<uint64_t>(), pay.as<uint32_t>());
        HIP_TRY(hipGetLastError());
        ALTO_TRY(radix_sort_pairs(keys.as<uint64_t>(), 1, pay.as<uint32_t>(), d, 32 + t->pack.bits[m], s));
        hipLaunchKernelGGL(k_rank_scatter, dim3(grid_for(d)), dim3(256), 0, s, pay.as<uint32_t>(), d,
                           t->rank_row[m].as<uint32_t>(), t->row_rank[m].as<uint32_t>());
        HIP_TRY(hipGetLastError());
        // the temporaries go back to the shared cache: drain first
        HIP_TRY(hipStreamSynchronize(s));
    }
    RankTabs T;
    for (int m = 0; m < ALTO_MAXN; ++m) T.row_rank[m] = m < t->nmode ? t->row_rank[m].as<uint32_t>() : nullptr;
    for (int m = 0; m < t->nmode; ++m) {
        ModeStream& M = t->ms[m];
        if (M.ranked) continue;
        RANK_DISPATCH_PW(t->pack.pw, {
            hipLaunchKernelGGL((k_stream_rank<PW>), dim3(grid_for(n)), dim3(256), 0, s, M.pk.as<uint64_t>(), n,
                               t->pack, m, T);
        });
        HIP_TRY(hipGetLastError());
        M.ranked = 1;
    }
    return ALTO_OK;
}

int reserve_rank_copies(alto_tensor* t, uint64_t rank, size_t elem_bytes, hipStream_t s)
{
    bool ranked = false;
    for (int m = 0; m < t->nmode; ++m) ranked = ranked || t->ms[m].ranked;
    if (!ranked) return ALTO_OK;
    bool drained = false;
    for (int m = 0; m < t->nmode; ++m) {
        const size_t need = t->dims[m] * rank * elem_bytes + 64;
        if (t->pcopy[m].bytes >= need) continue;
        if (!drained) {  // the old copy goes back to the shared cache
            HIP_TRY(hipStreamSynchronize(s));
            ALTO_TRY(t->wait_uses());
            drained = true;
        }
        t->pcopy[m].release();
        ALTO_TRY(t->pcopy[m].alloc(need));
    }
    return ALTO_OK;
}

// ------------------------------------------------------- hot-pattern split
struct HotK {
    int k[ALTO_MAXN - 1];
};

// sort key of the split: (pattern, chunk, row); the stream is ranked, so the
// other modes' fields are ranks
template <int PW>
__global__ void __launch_bounds__(256) k_split_key(const uint64_t* __restrict__ pk, uint64_t n, PackLayout L, int mode,
                                                   HotK hk, int chunk_log2, int cbits, uint64_t* __restrict__ ck,
                                                   uint32_t* __restrict__ pay, unsigned int* __restrict__ pat_count)
{
    __shared__ unsigned int s_cnt[16];
    if (threadIdx.x < 16) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k[PW];
#pragma unroll
        for (int w = 0; w < PW; ++w) k[w] = pk[i * PW + w];
        uint64_t pat = 0;
        int o = 0;
        for (int m = 0; m < L.nmode; ++m) {
            if (m == mode) continue;
            if (field<PW>(k, L.off[m], L.bits[m]) < (uint64_t)hk.k[o]) pat |= 1ull << o;
            ++o;
        }
        const uint64_t r = field<PW>(k, L.off[mode], L.bits[mode]);
        ck[i] = (((pat << cbits) | (i >> chunk_log2)) << L.bits[mode]) | r;
        pay[i] = (uint32_t)i;
        atomicAdd(&s_cnt[pat], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 16 && s_cnt[threadIdx.x]) atomicAdd(&pat_count[threadIdx.x], s_cnt[threadIdx.x]);
}

template <int PW>
__global__ void __launch_bounds__(256) k_split_gather(const uint64_t* __restrict__ pk, const double* __restrict__ v64,
                                                      const float* __restrict__ v32, const uint32_t* __restrict__ perm,
                                                      uint64_t n, uint64_t* __restrict__ pk_out,
                                                      double* __restrict__ v64_out, float* __restrict__ v32_out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t src = perm[i];
#pragma unroll
        for (int w = 0; w < PW; ++w) pk_out[i * PW + w] = pk[src * PW + w];
        if (v64) v64_out[i] = v64[src];
        if (v32) v32_out[i] = v32[src];
    }
}

// Groups a ranked stream by hot pattern (see ModeStream); hot_k[o] = rows of
// the o-th other mode counted as hot.  Synchronous on `s`.
int split_stream(alto_tensor* t, int mode, const int* hot_k, hipStream_t s)
{
    ModeStream& M = t->ms[mode];
    if (!M.ranked) {
        set_error("split_stream: stream %d is not ranked", mode);
        return ALTO_ERR_UNSUPPORTED;
    }
    const int no = t->nmode - 1;
    const uint64_t n = t->nnz;
    const int npat = 1 << no;  // at most 16 (five modes)
    uint64_t nchunks = (n + (1ull << M.chunk_log2) - 1) >> M.chunk_log2;
    int cbits = 0;
    while ((1ull << cbits) < nchunks) ++cbits;
    const int bits = no + cbits + t->pack.bits[mode];
    if (bits > 64) {
        set_error("split_stream: sort key of %d bits", bits);
        return ALTO_ERR_UNSUPPORTED;
    }
    HotK hk;
    for (int o = 0; o < ALTO_MAXN - 1; ++o) hk.k[o] = o < no ? hot_k[o] : 0;
    DevBuf ck, pay, cnt, pk2, v64, v32;
    ALTO_TRY(ck.alloc(n * 8));
    ALTO_TRY(pay.alloc(n * 4));
    ALTO_TRY(cnt.alloc(64 * 4));
    ALTO_TRY(fill_zero(cnt.as<uint32_t>(), 64, s));
    ALTO_TRY(pk2.alloc(n * t->pack.pw * 8));
    if (M.val64.p) ALTO_TRY(v64.alloc(n * 8));
    if (M.val32.p) ALTO_TRY(v32.alloc(n * 4));
    RANK_DISPATCH_PW(t->pack.pw, {
        hipLaunchKernelGGL((k_split_key<PW>), dim3(grid_for(n)), dim3(256), 0, s, M.pk.as<uint64_t>(), n, t->pack, mode,
                           hk, M.chunk_log2, cbits, ck.as<uint64_t>(), pay.as<uint32_t>(), cnt.as<unsigned int>());
    });
    HIP_TRY(hipGetLastError());
    ALTO_TRY(radix_sort_pairs(ck.as<uint64_t>(), 1, pay.as<uint32_t>(), n, bits, s));
    RANK_DISPATCH_PW(t->pack.pw, {
        hipLaunchKernelGGL((k_split_gather<PW>), dim3(grid_for(n)), dim3(256), 0, s, M.pk.as<uint64_t>(),
                           M.val64.as<double>(), M.val32.as<float>(), pay.as<uint32_t>(), n, pk2.as<uint64_t>(),
                           v64.as<double>(), v32.as<float>());
    });
    HIP_TRY(hipGetLastError());
    uint32_t hc[64];
    HIP_TRY(hipMemcpyAsync(hc, cnt.p, sizeof(hc), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    M.pk.swap(pk2);
    if (M.val64.p) M.val64.swap(v64);
    if (M.val32.p) M.val32.swap(v32);
    M.pat_start.assign(npat + 1, 0);
    for (int p = 0; p < npat; ++p) M.pat_start[p + 1] = M.pat_start[p] + hc[p];
    for (int o = 0; o < ALTO_MAXN - 1; ++o) M.hot_k[o] = hk.k[o];
    M.split = 1;
    return ALTO_OK;
}

// Ranked streams need every mode's rows below 2^32 (32-bit ranks) and a sort
// key of 32 count bits + row bits.
static bool rank_applicable(const alto_tensor* t)
{
    if (!hot_rows_enabled() || t->nmode < 2) return false;
    for (int m = 0; m < t->nmode; ++m)
        if (t->pack.bits[m] > 32) return false;
    return true;
}

// the order a mode's stream should have with the current settings
struct StreamLayout {
    int chunk_log2, fib_mode, fib_mode2, fib_min;
};
static StreamLayout wanted_layout(const alto_tensor* t, int m, int ranked)
{
    StreamLayout L{t->chunk_log2, -1, -1, 0};
    if (det_enabled()) {  // one chunk: each row's nonzeros contiguous
        L.chunk_log2 = 63;
        return L;
    }
    int fmin = 0, fchunk = 0;
    uint64_t fnnz = 0;
    const int tiers = fiber_settings(&fmin, &fchunk, &fnnz);
    if (ranked || !tiers || t->nmode < 3 || t->nnz < fnnz) return L;
    const uint64_t nch = (t->nnz + (1ull << fchunk) - 1) >> fchunk;
    int cb = 0;
    while ((1ull << cb) < nch) ++cb;
    // the fiber modes: the other modes with the fewest (rule 1: most) rows,
    // in that order, whose sort keys fit 64 bits
    const int most = fiber_rule();
    int pick[2] = {-1, -1};
    for (int k = 0; k < tiers; ++k) {
        int a = -1;
        for (int q = 0; q < t->nmode; ++q) {
            if (q == m || q == pick[0] || cb + t->pack.bits[m] + t->pack.bits[q] > 64) continue;
            if (a < 0 || (most ? t->dims[q] > t->dims[a] : t->dims[q] < t->dims[a])) a = q;
        }
        pick[k] = a;
        if (a < 0) break;
    }
    if (pick[0] < 0) return L;
    L.chunk_log2 = fchunk;
    L.fib_mode = pick[0];
    L.fib_mode2 = pick[1];
    L.fib_min = fmin;
    return L;
}

// One fiber split: the nonzeros at ALTO positions idx[0, n) (0..n-1 when
// idx is null) ordered by (chunk, row_m, row_a) and split stably into the
// nonzeros of fibers of at least `min_len` and the others; the ALTO
// positions in that order go to perm[0, n), the length of the first part
// to *nl_dev.
static int fiber_split(alto_tensor* t, int m, int amode, int chunk_log2, int cb, uint32_t min_len,
                       const uint32_t* idx, uint64_t n, uint64_t* ck, uint32_t* pay, uint32_t* id, uint32_t* fstart,
                       uint32_t* cls, uint32_t* perm, DevBuf& tmp, uint64_t* nl_dev, hipStream_t s)
{
    DISPATCH_WPW(t->W, t->pack.pw, {
        hipLaunchKernelGGL((k_fiber_key<W, PW>), dim3(grid_for(n)), dim3(256), lut_smem(t->delin), s,
                           t->key.as<uint64_t>(), lutref(t->delin), t->pack, m, amode, chunk_log2, n, idx, ck, pay);
    });
    HIP_TRY(hipGetLastError());
    ALTO_TRY(radix_sort_pairs(ck, 1, pay, n, cb + t->pack.bits[m] + t->pack.bits[amode], s));
    hipLaunchKernelGGL(k_fiber_heads, dim3(grid_for(n)), dim3(256), 0, s, ck, n, id);
    ALTO_TRY(scan_exclusive(id, n, tmp, s));
    hipLaunchKernelGGL(k_fiber_starts, dim3(grid_for(n)), dim3(256), 0, s, ck, n, id, fstart);
    hipLaunchKernelGGL(k_fiber_class, dim3(grid_for(n)), dim3(256), 0, s, id, fstart, n, min_len, cls);
    ALTO_TRY(scan_exclusive(cls, n, tmp, s));
    hipLaunchKernelGGL(k_fiber_scatter, dim3(grid_for(n)), dim3(256), 0, s, id, fstart, cls, pay, n, min_len, perm,
                       nl_dev);
    HIP_TRY(hipGetLastError());
    return ALTO_OK;
}

// Orders the nonzeros of stream m by (chunk, row_m, row_a) and splits them
// into long- and short-fiber parts; with a second fiber mode b the short
// part is ordered by (chunk, row_m, row_b) and split again.  The
// permutation ends in `pay`, the lengths of the fiber parts in *nlong and
// *nlong2.  Synchronous on `s`.
static int fiber_order(alto_tensor* t, int m, const StreamLayout& L, DevBuf& ck, DevBuf& pay, hipStream_t s,
                       uint64_t* nlong, uint64_t* nlong2)
{
    const uint64_t n = t->nnz;
    const uint64_t nch = (n + (1ull << L.chunk_log2) - 1) >> L.chunk_log2;
    int cb = 0;
    while ((1ull << cb) < nch) ++cb;
    DevBuf id, fstart, cls, perm, tmp, nl;
    ALTO_TRY(id.alloc(n * 4));
    ALTO_TRY(fstart.alloc((n + 1) * 4));
    ALTO_TRY(cls.alloc(n * 4));
    ALTO_TRY(perm.alloc(n * 4));
    ALTO_TRY(nl.alloc(16));
    uint64_t* ckp = ck.as<uint64_t>();
    uint32_t *payp = pay.as<uint32_t>(), *permp = perm.as<uint32_t>();
    ALTO_TRY(fiber_split(t, m, L.fib_mode, L.chunk_log2, cb, (uint32_t)L.fib_min, nullptr, n, ckp, payp,
                         id.as<uint32_t>(), fstart.as<uint32_t>(), cls.as<uint32_t>(), permp, tmp, nl.as<uint64_t>(),
                         s));
    HIP_TRY(hipMemcpyAsync(nlong, nl.p, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *nlong2 = 0;
    const uint64_t rest = n - *nlong;
    if (L.fib_mode2 >= 0 && rest > 0) {
        // the short part [nlong, n) of perm holds the ALTO positions it orders
        const uint64_t o = *nlong;
        ALTO_TRY(fiber_split(t, m, L.fib_mode2, L.chunk_log2, cb, (uint32_t)L.fib_min, permp + o, rest, ckp + o,
                             payp + o, id.as<uint32_t>(), fstart.as<uint32_t>(), cls.as<uint32_t>(), permp + o, tmp,
                             nl.as<uint64_t>() + 1, s));
        HIP_TRY(hipMemcpyAsync(nlong2, nl.as<uint64_t>() + 1, 8, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipMemcpyAsync(pay.p, perm.p, n * 4, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ALTO_OK;
}

int prepare_mode_streams(alto_tensor* t, int need64, int need32, hipStream_t s)
{
    uint64_t n = t->nnz;
    const int want_chunk = tuned_chunk_log2();
    if (want_chunk > 0 && want_chunk != t->chunk_log2) t->chunk_log2 = want_chunk;
    const int ranked = det_enabled() ? 0 : (int)rank_applicable(t);
    StreamLayout want[ALTO_MAXN];
    bool any = false, reorder = false;
    for (int m = 0; m < t->nmode; ++m) {
        const ModeStream& M = t->ms[m];
        want[m] = wanted_layout(t, m, ranked);
        if (M.chunk_log2 != want[m].chunk_log2 || M.ranked != ranked || M.fib_mode != want[m].fib_mode ||
            M.fib_mode2 != want[m].fib_mode2 || M.fib_min != want[m].fib_min)
            reorder = true;
        if ((need64 && !M.val64.p) || (need32 && !M.val32.p) || !M.pk.p) any = true;
    }
    if (!any && !reorder) return ALTO_OK;
    // buffers go back to the shared cache: drain work that may still read them
    HIP_TRY(hipStreamSynchronize(s));
    ALTO_TRY(t->wait_uses());
    for (int m = 0; m < t->nmode; ++m) {
        ModeStream& M = t->ms[m];
        if (M.chunk_log2 != want[m].chunk_log2 || M.ranked != ranked || M.fib_mode != want[m].fib_mode ||
            M.fib_mode2 != want[m].fib_mode2 || M.fib_min != want[m].fib_min) {
            M.pk.release();
            M.val64.release();
            M.val32.release();
            M.ranked = 0;
            M.split = 0;
            M.pat_start.clear();
            M.fib_mode = M.fib_mode2 = -1;
            M.fib_min = 0;
            M.nlong = M.nlong2 = 0;
        }
    }
    DevBuf ck, pay;
    ALTO_TRY(ck.alloc(n * 8));
    ALTO_TRY(pay.alloc(n * 4));
    for (int m = 0; m < t->nmode; ++m) {
        ModeStream& M = t->ms[m];
        if (M.pk.p && (!need64 || M.val64.p) && (!need32 || M.val32.p)) continue;
        M.chunk_log2 = want[m].chunk_log2;
        M.ranked = 0;
        M.split = 0;
        M.pat_start.clear();
        // the stream is rewritten in its base order: every value array it
        // has is rewritten with it
        const bool w64 = need64 || M.val64.p, w32 = need32 || M.val32.p;
        if (!M.pk.p) ALTO_TRY(M.pk.alloc(n * t->pack.pw * 8));
        if (w64 && !M.val64.p) ALTO_TRY(M.val64.alloc(n * 8));
        if (w32 && !M.val32.p) ALTO_TRY(M.val32.alloc(n * 4));
        if (want[m].fib_mode >= 0) {
            uint64_t nlong = 0, nlong2 = 0;
            ALTO_TRY(fiber_order(t, m, want[m], ck, pay, s, &nlong, &nlong2));
            M.fib_mode = want[m].fib_mode;
            M.fib_mode2 = want[m].fib_mode2;
            M.fib_min = want[m].fib_min;
            M.nlong = nlong;
            M.nlong2 = nlong2;
        } else {
            uint64_t nchunks = (n + (1ull << M.chunk_log2) - 1) >> M.chunk_log2;
            int cbits = 0;
            while ((1ull << cbits) < nchunks) ++cbits;
            DISPATCH_WPW(t->W, t->pack.pw, {
                hipLaunchKernelGGL((k_mode_key<W, PW>), dim3(grid_for(n)), dim3(256), lut_smem(t->delin), s,
                                   t->key.as<uint64_t>(), lutref(t->delin), t->pack, m, M.chunk_log2, n,
                                   ck.as<uint64_t>(), pay.as<uint32_t>());
            });
            HIP_TRY(hipGetLastError());
            ALTO_TRY(radix_sort_pairs(ck.as<uint64_t>(), 1, pay.as<uint32_t>(), n, t->pack.bits[m] + cbits, s));
            M.fib_mode = M.fib_mode2 = -1;
            M.fib_min = 0;
            M.nlong = M.nlong2 = 0;
        }
        DISPATCH_WPW(t->W, t->pack.pw, {
            hipLaunchKernelGGL((k_mode_gather<W, PW>), dim3(grid_for(n)), dim3(256), lut_smem(t->delin), s,
                               t->key.as<uint64_t>(), t->vals.as<double>(), lutref(t->delin),
                               pay.as<uint32_t>(), n, M.pk.as<uint64_t>(), w64 ? M.val64.as<double>() : nullptr,
                               w32 ? M.val32.as<float>() : nullptr);
        });
        HIP_TRY(hipGetLastError());
    }
    if (rank_applicable(t)) ALTO_TRY(rank_streams(t, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ALTO_OK;
}

static int create_impl(int nmode, const uint64_t* dims, uint64_t nnz, const uint64_t* const* cidx,
                       const double* vals, int ptr_kind, int val_kind, int lit_bits, int nprtn, hipStream_t s,
                       alto_tensor** out)
{
    if (nmode < 1 || nmode > ALTO_MAXN || !dims || !cidx || !vals || !out || nprtn < 1 || nnz < 1) {
        set_error("alto_tensor_create: invalid argument");
        return ALTO_ERR_ARG;
    }
    if (nnz >= (1ull << 32)) {
        set_error("alto_tensor_create: nnz must be below 2^32");
        return ALTO_ERR_UNSUPPORTED;
    }
    for (int m = 0; m < nmode; ++m)
        if (dims[m] < 1) {
            set_error("alto_tensor_create: dims must be >= 1");
            return ALTO_ERR_ARG;
        }
    alto_tensor* t = new alto_tensor();
    std::unique_ptr<alto_tensor> guard(t);
    t->nmode = nmode;
    t->nnz = nnz;
    t->nprtn = nprtn;
    t->val_kind = val_kind;
    HIP_TRY(hipGetDevice(&t->device));
    for (int m = 0; m < nmode; ++m) t->dims[m] = dims[m];
    int total = 0;
    for (int m = 0; m < nmode; ++m) total += dims[m] > 1 ? 64 - __builtin_clzll(dims[m] - 1) : 0;
    if (lit_bits == 0) lit_bits = total <= 64 ? 64 : 128;
    if (lit_bits != 64 && lit_bits != 128) {
        set_error("lit_bits must be 64 or 128");
        return ALTO_ERR_ARG;
    }
    t->W = lit_bits / 64;
    t->alto_bits = compute_masks(nmode, dims, t->W, t->masks, t->alto_mask);
    if (t->alto_bits < 0) {
        set_error("alto_tensor_create: %d index bits exceed the %d-bit ALTO mask", total, lit_bits);
        return ALTO_ERR_ARG;
    }
    ALTO_TRY(build_luts(t, s));

    // inputs to device
    DevBuf dcoord[ALTO_MAXN], dvals;
    CoordPtrs cp = {};
    const double* dv = vals;
    if (ptr_kind == ALTO_PTR_HOST) {
        for (int m = 0; m < nmode; ++m) {
            ALTO_TRY(dcoord[m].alloc(nnz * 8));
            HIP_TRY(hipMemcpyAsync(dcoord[m].p, cidx[m], nnz * 8, hipMemcpyHostToDevice, s));
            cp.c[m] = dcoord[m].as<uint64_t>();
        }
        ALTO_TRY(dvals.alloc(nnz * 8));
        HIP_TRY(hipMemcpyAsync(dvals.p, vals, nnz * 8, hipMemcpyHostToDevice, s));
        dv = dvals.as<double>();
    } else {
        for (int m = 0; m < nmode; ++m) cp.c[m] = cidx[m];
    }

    {
        DevBuf bad;
        ALTO_TRY(bad.alloc(4));
        ALTO_TRY(fill_zero(bad.as<int>(), 1, s));
        DimArr da = {};
        for (int m = 0; m < nmode; ++m) da.d[m] = dims[m];
        hipLaunchKernelGGL(k_validate, dim3(grid_for(nnz)), dim3(256), 0, s, cp, nmode, da, nnz, bad.as<int>());
        HIP_TRY(hipGetLastError());
        int hbad = 0;
        HIP_TRY(hipMemcpyAsync(&hbad, bad.p, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (hbad) {
            set_error("alto_tensor_create: a coordinate is outside dims");
            return ALTO_ERR_ARG;
        }
    }
    ALTO_TRY(t->key.alloc(nnz * t->W * 8));
    ALTO_TRY(t->vals.alloc(nnz * 8));
    DevBuf vtmp, pay;
    ALTO_TRY(vtmp.alloc(nnz * 8));
    ALTO_TRY(pay.alloc(nnz * 4));
    DISPATCH_WPW(t->W, t->pack.pw, {
        hipLaunchKernelGGL((k_linearize<PW, W>), dim3(grid_for(nnz)), dim3(256), lut_smem(t->lin), s, cp, nmode, t->pack, dv,
                           val_kind == ALTO_VAL_U64, lutref(t->lin), nnz, t->key.as<uint64_t>(),
                           vtmp.as<double>(), pay.as<uint32_t>());
    });
    HIP_TRY(hipGetLastError());
    // sort_alto: by the ALTO key (unique after duplicate merging; stable otherwise)
    ALTO_TRY(radix_sort_pairs(t->key.as<uint64_t>(), t->W, pay.as<uint32_t>(), nnz, t->alto_bits, s));
    hipLaunchKernelGGL(k_gather_f64, dim3(grid_for(nnz)), dim3(256), 0, s, vtmp.as<double>(), pay.as<uint32_t>(), nnz,
                       t->vals.as<double>());
    HIP_TRY(hipGetLastError());
    // the staging buffers are released only after the stream has drained
    // (end of this function): the allocator hands freed blocks to other streams

    // oidx flag (include/alto.hpp:250-258)
    for (int m = 0; m < nmode; ++m)
        if (nnz / dims[m] <= 4) t->is_oidx = 1;

    // prtn_alto: pointers (closed form) and intervals
    uint64_t per = (nnz + nprtn - 1) / nprtn;
    t->prtn_ptr.assign(nprtn + 1, 0);
    for (int p = 0; p < nprtn; ++p) t->prtn_ptr[p + 1] = std::min<uint64_t>((uint64_t)(p + 1) * per, nnz);
    {
        DevBuf iv;
        uint64_t niv = (uint64_t)nprtn * nmode * 2;
        ALTO_TRY(iv.alloc(niv * 8));
        // start = dims[n] (empty partition convention), stop = 0: fill then fix starts
        std::vector<unsigned long long> init(niv);
        for (int p = 0; p < nprtn; ++p)
            for (int m = 0; m < nmode; ++m) {
                init[(p * nmode + m) * 2] = dims[m];
                init[(p * nmode + m) * 2 + 1] = 0;
            }
        HIP_TRY(hipMemcpyAsync(iv.p, init.data(), niv * 8, hipMemcpyHostToDevice, s));
        DISPATCH_WPW(t->W, t->pack.pw, {
            hipLaunchKernelGGL((k_intervals<W, PW>), dim3(grid_for(nnz, 256, 8192)), dim3(256), lut_smem(t->delin), s,
                               t->key.as<uint64_t>(), lutref(t->delin), t->pack, nnz, per,
                               iv.as<unsigned long long>());
        });
        HIP_TRY(hipGetLastError());
        t->intervals.assign(niv, 0);
        HIP_TRY(hipMemcpyAsync(t->intervals.data(), iv.p, niv * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }

    // CR bits (include/alto.hpp:930-1027)
    int free_bits = 64 * t->W - t->alto_bits;
    for (int m = 0; m < nmode; ++m) {
        if (nnz / dims[m] <= 4 && free_bits) {
            int pos = 64 * t->W - free_bits;
            t->cr_masks[m * t->W + (pos >> 6)] = 1ull << (pos & 63);
            t->alto_cr_mask[pos >> 6] |= 1ull << (pos & 63);
            --free_bits;
        }
    }
    for (int m = 0; m < nmode; ++m) {
        uint64_t cm0 = t->cr_masks[m * t->W], cm1 = t->W == 2 ? t->cr_masks[m * t->W + 1] : 0;
        // with one partition no row is shared, so no key gets the bit
        if (!(cm0 | cm1) || nprtn == 1) continue;
        DevBuf lo, hi;
        ALTO_TRY(lo.alloc(dims[m] * 4));
        ALTO_TRY(hi.alloc(dims[m] * 4));
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(dims[m])), dim3(256), 0, s, lo.as<int>(), dims[m], nprtn);
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(dims[m])), dim3(256), 0, s, hi.as<int>(), dims[m], -1);
        DISPATCH_WPW(t->W, t->pack.pw, {
            hipLaunchKernelGGL((k_cr_owner<W, PW>), dim3(grid_for(nnz)), dim3(256), lut_smem(t->delin), s, t->key.as<uint64_t>(),
                               lutref(t->delin), t->pack, m, nnz, per, lo.as<int>(), hi.as<int>());
            hipLaunchKernelGGL((k_cr_set<W, PW>), dim3(grid_for(nnz)), dim3(256), lut_smem(t->delin), s, t->key.as<uint64_t>(),
                               lutref(t->delin), t->pack, m, nnz, lo.as<int>(), hi.as<int>(), cm0, cm1);
        });
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    *out = guard.release();
    return ALTO_OK;
}

}  // namespace alto

using namespace alto;

extern "C" {

const char* alto_last_error(void) { return alto::last_error(); }
const char* alto_version(void) { return "alto-mi355x 0.1.2"; }

int alto_tensor_create(int nmodes, const uint64_t* dims, uint64_t nnz, const uint64_t* const* cidx,
                       const double* vals, int ptr_kind, int val_kind, int lit_bits, int nprtn,
                       alto_stream_t stream, alto_tensor_t** out)
{
    return create_impl(nmodes, dims, nnz, cidx, vals, ptr_kind, val_kind, lit_bits, nprtn, (hipStream_t)stream, out);
}

int alto_tensor_destroy(alto_tensor_t* t)
{
    if (!t) return ALTO_OK;
    const int st = t->wait_uses();  // kernels of asynchronous calls may still read the buffers
    delete t;
    return st;
}

int alto_tensor_info(const alto_tensor_t* t, alto_tensor_info_t* info)
{
    if (!t || !info) {
        set_error("alto_tensor_info: null argument");
        return ALTO_ERR_ARG;
    }
    memset(info, 0, sizeof(*info));
    info->nmodes = t->nmode;
    info->lit_words = t->W;
    info->alto_bits = t->alto_bits;
    info->nprtn = t->nprtn;
    info->is_oidx = t->is_oidx;
    info->val_kind = t->val_kind;
    info->nnz = t->nnz;
    for (int m = 0; m < t->nmode; ++m) info->dims[m] = t->dims[m];
    memcpy(info->mode_masks, t->masks, sizeof(t->masks));
    memcpy(info->alto_mask, t->alto_mask, sizeof(t->alto_mask));
    memcpy(info->cr_masks, t->cr_masks, sizeof(t->cr_masks));
    memcpy(info->alto_cr_mask, t->alto_cr_mask, sizeof(t->alto_cr_mask));
    uint64_t b = t->key.bytes + t->vals.bytes + t->lin.dev.bytes + t->delin.dev.bytes;
    for (int m = 0; m < t->nmode; ++m) b += t->ms[m].pk.bytes + t->ms[m].val64.bytes + t->ms[m].val32.bytes;
    info->device_bytes = b;
    return ALTO_OK;
}

int alto_tensor_export(const alto_tensor_t* t, uint64_t* idx, double* vals, uint64_t* prtn_ptr, uint64_t* intervals)
{
    if (!t) {
        set_error("alto_tensor_export: null tensor");
        return ALTO_ERR_ARG;
    }
    if (idx) HIP_TRY(hipMemcpy(idx, t->key.p, t->nnz * t->W * 8, hipMemcpyDeviceToHost));
    if (vals) HIP_TRY(hipMemcpy(vals, t->vals.p, t->nnz * 8, hipMemcpyDeviceToHost));
    if (prtn_ptr) memcpy(prtn_ptr, t->prtn_ptr.data(), (t->nprtn + 1) * 8);
    if (intervals) memcpy(intervals, t->intervals.data(), t->intervals.size() * 8);
    return ALTO_OK;
}

int alto_tensor_delinearize(const alto_tensor_t* t, uint64_t* coords)
{
    if (!t || !coords) {
        set_error("alto_tensor_delinearize: null argument");
        return ALTO_ERR_ARG;
    }
    DevBuf c;
    ALTO_TRY(c.alloc(t->nnz * t->nmode * 8));
    ALTO_TRY(delinearize_all(t, c.as<uint64_t>(), 0));
    HIP_TRY(hipMemcpy(coords, c.p, t->nnz * t->nmode * 8, hipMemcpyDeviceToHost));
    return ALTO_OK;
}

int alto_tensor_export_oidx(const alto_tensor_t* t, int mode, uint64_t* oidx)
{
    if (!t || !oidx || mode < 0 || mode >= t->nmode) {
        set_error("alto_tensor_export_oidx: invalid argument");
        return ALTO_ERR_ARG;
    }
    uint64_t n = t->nnz;
    DevBuf ck, pay;
    ALTO_TRY(ck.alloc(n * 8));
    ALTO_TRY(pay.alloc(n * 4));
    // rows with chunk_log2 = 63 -> chunk id 0 for every nonzero: a plain stable sort by row
    DISPATCH_WPW(t->W, t->pack.pw, {
        hipLaunchKernelGGL((k_mode_key<W, PW>), dim3(grid_for(n)), dim3(256), lut_smem(t->delin), 0, t->key.as<uint64_t>(),
                           lutref(t->delin), t->pack, mode, 63, n, ck.as<uint64_t>(), pay.as<uint32_t>());
    });
    HIP_TRY(hipGetLastError());
    ALTO_TRY(radix_sort_pairs(ck.as<uint64_t>(), 1, pay.as<uint32_t>(), n, t->pack.bits[mode], 0));
    std::vector<uint32_t> h(n);
    HIP_TRY(hipMemcpy(h.data(), pay.p, n * 4, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < n; ++i) oidx[i] = h[i];
    return ALTO_OK;
}

int alto_pdep_keys(const alto_tensor_t* t, const uint64_t* coords, uint64_t n, uint64_t* keys, alto_stream_t stream)
{
    if (!t || !coords || !keys) {
        set_error("alto_pdep_keys: null argument");
        return ALTO_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    DISPATCH_WPW(t->W, t->pack.pw, {
        hipLaunchKernelGGL((k_pack_lin<PW, W>), dim3(grid_for(n)), dim3(256), lut_smem(t->lin), s, coords, t->pack,
                           lutref(t->lin), n, keys);
    });
    HIP_TRY(hipGetLastError());
    return const_cast<alto_tensor*>(t)->note_use(s);
}

int alto_pext_keys(const alto_tensor_t* t, const uint64_t* keys, uint64_t n, uint64_t* coords, alto_stream_t stream)
{
    if (!t || !coords || !keys) {
        set_error("alto_pext_keys: null argument");
        return ALTO_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    DISPATCH_WPW(t->W, t->pack.pw, {
        hipLaunchKernelGGL((k_delin_coords<W, PW>), dim3(grid_for(n)), dim3(256), lut_smem(t->delin), s, keys,
                           lutref(t->delin), t->pack, n, coords);
    });
    HIP_TRY(hipGetLastError());
    return const_cast<alto_tensor*>(t)->note_use(s);
}

int alto_device_malloc(void** p, uint64_t bytes)
{
    HIP_TRY(hipMalloc(p, bytes ? bytes : 8));
    return ALTO_OK;
}
int alto_device_free(void* p)
{
    HIP_TRY(hipFree(p));
    return ALTO_OK;
}
int alto_memcpy_h2d(void* dst, const void* src, uint64_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return ALTO_OK;
}
int alto_memcpy_d2h(void* dst, const void* src, uint64_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return ALTO_OK;
}
int alto_memcpy_d2d(void* dst, const void* src, uint64_t bytes, alto_stream_t stream)
{
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return ALTO_OK;
}
int alto_device_trim(void)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    Pool& P = pool();
    std::lock_guard<std::mutex> g(P.mu);
    for (auto& kv : P.free_blocks[dev & 63]) (void)hipFree(kv.second);
    P.free_blocks[dev & 63].clear();
    P.cached[dev & 63] = 0;
    return ALTO_OK;
}

int alto_device_synchronize(void)
{
    HIP_TRY(hipDeviceSynchronize());
    return ALTO_OK;
}

}  // extern "C"
