// Helpers shared by the drop-in headers: error checks and device buffers
// over the C ABI of libalto_hip.  Errors: the reference asserts or exits; a
// failed call here prints alto_last_error() and aborts, so a misuse stops
// the program as it does there.
#ifndef ALTO_DROPIN_DEVICE_HPP_
#define ALTO_DROPIN_DEVICE_HPP_

#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "../alto_hip.h"
#include "common.hpp"

namespace alto_dropin {

inline void check(int status, const char* what)
{
    if (status != ALTO_OK) {
        fprintf(stderr, "%s failed: %s\n", what, alto_last_error());
        abort();
    }
}

// one device allocation, freed with the object
struct DeviceBuffer {
    void* p = nullptr;
    explicit DeviceBuffer(size_t bytes) { check(alto_device_malloc(&p, bytes ? bytes : 8), "alto_device_malloc"); }
    DeviceBuffer(const void* host, size_t bytes) : DeviceBuffer(bytes)
    {
        if (bytes) check(alto_memcpy_h2d(p, host, bytes), "alto_memcpy_h2d");
    }
    ~DeviceBuffer() { alto_device_free(p); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    template <typename T> T* as() const { return (T*)p; }
};

// device copies of host factor matrices (dims[m] x rank, row-major) for one call
struct DeviceFactors {
    std::vector<void*> p;
    DeviceFactors(int nmodes, const IType* dims, IType rank, FType* const* U)
    {
        p.resize(nmodes);
        for (int m = 0; m < nmodes; ++m) {
            check(alto_device_malloc(&p[m], dims[m] * rank * sizeof(FType)), "alto_device_malloc");
            check(alto_memcpy_h2d(p[m], U[m], dims[m] * rank * sizeof(FType)), "alto_memcpy_h2d");
        }
    }
    ~DeviceFactors()
    {
        for (void* q : p) alto_device_free(q);
    }
    DeviceFactors(const DeviceFactors&) = delete;
    DeviceFactors& operator=(const DeviceFactors&) = delete;
};

}  // namespace alto_dropin

#endif  // ALTO_DROPIN_DEVICE_HPP_
