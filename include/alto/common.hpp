// Common types and helpers of the reference ALTO library (include/common.hpp,
// src/common.cpp), header-only: the index / value types, the cache-line
// aligned allocators the reference frees its structures with, ParMemcpy /
// ParMemset, and the TSC timers its command line prints through.
#ifndef ALTO_DROPIN_COMMON_HPP_
#define ALTO_DROPIN_COMMON_HPP_

#include <inttypes.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <time.h>

#include <algorithm>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef MAX_NUM_MODES
#define MAX_NUM_MODES 5
#endif

// include/common.hpp:38-42: the library's device engine takes 64-bit
// coordinates, so IDX_TYPE32 is not offered
#ifdef IDX_TYPE32
#error "IDX_TYPE32 is not supported: coordinates are 64-bit"
#endif
typedef unsigned long long IType;

// include/common.hpp:45-58: factor matrices are double on the host side of
// this API (the device engine also runs FP32 / BF16 through alto_hip.h)
#ifdef FP_TYPE32
#error "FP_TYPE32 is not supported by the drop-in headers; use ALTO_PREC_F32 through alto_hip.h"
#endif
typedef double FType;

// include/common.hpp:62-69
#ifdef VAL_TYPE_INT32
typedef uint32_t ValType;
#elif defined(VAL_TYPE_INT64)
typedef unsigned long long ValType;
#else
typedef FType ValType;
#endif

#define CACHELINE 64
#define ROW 1

#ifndef TIME
#define TIME 0
#endif

typedef enum Model_ { ALS, APR } Model;

// include/common.hpp:77-111: the x86 time-stamp counter (a nanosecond clock
// elsewhere)
inline uint64_t ReadTSC(void)
{
#if defined(__x86_64__) || defined(__i386__)
    uint32_t lo, hi;
    __asm__ __volatile__("rdtsc" : "=a"(lo), "=d"(hi));
    return ((uint64_t)hi << 32) | lo;
#else
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
#endif
}

// include/common.hpp:114-131: 64-byte aligned blocks released with free()
inline void* AlignedMalloc(size_t size)
{
    void* p = NULL;
    if (posix_memalign(&p, CACHELINE, size ? size : 1) != 0) p = NULL;
    return p;
}

inline void AlignedFree(void* addr)
{
    free(addr);
}

// include/common.hpp:133-165: huge-page backed buffers (2 MiB pages); a
// failed mapping ends the program as there
inline void* HPMalloc(size_t nbytes)
{
    const size_t page = (size_t)2 << 20;
    nbytes = (nbytes + page - 1) / page * page;
    void* p = mmap(NULL, nbytes, PROT_READ | PROT_WRITE, MAP_ANONYMOUS | MAP_PRIVATE | MAP_HUGETLB, -1, 0);
    if (p == MAP_FAILED) {
        fprintf(stderr, "mmap call failed\n");
        exit(1);
    }
    return p;
}

inline void HPFree(void* ptr, size_t nbytes)
{
    const size_t page = (size_t)2 << 20;
    munmap(ptr, (nbytes + page - 1) / page * page);
}

// include/common.hpp:167-197: the copy / fill split evenly over the OpenMP
// threads (one thread without OpenMP)
static inline void ParMemcpy(void* const dst, void* const src, size_t const bytes)
{
#ifdef _OPENMP
#pragma omp parallel
    {
        const size_t nt = (size_t)omp_get_num_threads(), t = (size_t)omp_get_thread_num();
        const size_t per = (bytes + nt - 1) / nt, b = std::min(per * t, bytes), e = std::min(b + per, bytes);
        memcpy((char*)dst + b, (char*)src + b, e - b);
    }
#else
    memcpy(dst, src, bytes);
#endif
}

static inline void ParMemset(void* dst, int val, size_t bytes)
{
#ifdef _OPENMP
#pragma omp parallel
    {
        const size_t nt = (size_t)omp_get_num_threads(), t = (size_t)omp_get_thread_num();
        const size_t per = (bytes + nt - 1) / nt, b = std::min(per * t, bytes), e = std::min(b + per, bytes);
        memset((char*)dst + b, val, e - b);
    }
#else
    memset(dst, val, bytes);
#endif
}

namespace alto_dropin {
// ticks per second of ReadTSC, set by InitTSC
inline double& ticks_per_second()
{
    static double t = 0.0;
    return t;
}
}  // namespace alto_dropin

// src/common.cpp:12-19: the tick rate, calibrated here over 50 ms of the
// monotonic clock instead of a one-second sleep
inline void InitTSC(void)
{
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    const uint64_t t0 = ReadTSC();
    struct timespec nap = {0, 50 * 1000 * 1000};
    nanosleep(&nap, NULL);
    const uint64_t t1 = ReadTSC();
    clock_gettime(CLOCK_MONOTONIC, &b);
    const double s = (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
    alto_dropin::ticks_per_second() = (double)(t1 - t0) / s;
}

// src/common.cpp:22-31
inline double ElapsedTime(uint64_t ticks)
{
    if (alto_dropin::ticks_per_second() == 0.0) {
        fprintf(stderr, "TSC timer has not been initialized.\n");
        return 0.0;
    }
    return (double)ticks / alto_dropin::ticks_per_second();
}

// include/common.hpp:200-216 and src/common.cpp:33-57: the timers are live
// when TIME is defined non-zero, as there
static inline void BEGIN_TIMER(uint64_t* ticks)
{
#if TIME
    *ticks = ReadTSC();
#else
    (void)ticks;
#endif
}

static inline void END_TIMER(uint64_t* ticks)
{
#if TIME
    *ticks = ReadTSC();
#else
    (void)ticks;
#endif
}

inline void ELAPSED_TIME(uint64_t start, uint64_t end, double* t_elapsed)
{
#if TIME
    if (alto_dropin::ticks_per_second() == 0.0)
        fprintf(stderr, "TSC timer has not been initialized.\n");
    else
        *t_elapsed = (double)(end - start) / alto_dropin::ticks_per_second();
#else
    (void)start;
    (void)end;
    (void)t_elapsed;
#endif
}

inline void PRINT_TIMER(const char* message, double t)
{
#if TIME
    printf("%s: %f s\n", message, t);
#else
    (void)message;
    (void)t;
#endif
}

// src/common.cpp:5-9
__attribute__((noinline)) inline void ASMTrace(char* str)
{
    fprintf(stderr, "HERE %s.\n", str);
}

// src/common.cpp:59-83: row-major m x n matrices on stderr
inline void PrintFPMatrix(char* name, FType* a, size_t m, size_t n)
{
    fprintf(stderr, "%s:\n", name);
    for (size_t i = 0; i < m; i++) {
        for (size_t j = 0; j < n; j++) fprintf(stderr, "%.6g ", a[i * n + j]);
        fprintf(stderr, "\n");
    }
}

inline void PrintIntMatrix(char* name, size_t* a, size_t m, size_t n)
{
    fprintf(stderr, "%s:\n", name);
    for (size_t i = 0; i < m; i++) {
        for (size_t j = 0; j < n; j++) fprintf(stderr, "%zu ", a[i * n + j]);
        fprintf(stderr, "\n");
    }
}

#endif  // ALTO_DROPIN_COMMON_HPP_
