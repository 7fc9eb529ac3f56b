// Internals shared by the C ABI translation units (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/nrelax.h"
#include "device.hpp"
#include "network.hpp"

struct nr_network {
    nr::Network net;
};

namespace nr {

// Records msg as the calling thread's last error and returns code.
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
void to_report(const NetOut &o, nr_report &r);

// Runs body() and turns an exception escaping it into a status, so that no
// exception crosses the C ABI: NR_E_ALLOC for host memory, `code` otherwise.
template <class F>
int guarded(int code, F &&body) noexcept
{
    try {
        return body();
    } catch (const std::bad_alloc &) {
        return fail(NR_E_ALLOC, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(code, e.what());
    } catch (...) {
        return fail(code, "unexpected exception");
    }
}

// Owning device array.
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t count)
    {
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
        n = count;
        return hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(count, 1));
    }
    hipError_t upload(const std::vector<T> &h)
    {
        hipError_t e = alloc(h.size());
        if (e != hipSuccess || h.empty()) return e;
        return hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
    }
};

}  // namespace nr

#define HIP_TRY(expr)                                         \
    do {                                                      \
        hipError_t e_ = (expr);                               \
        if (e_ != hipSuccess) return nr::hip_fail(e_, #expr); \
    } while (0)
