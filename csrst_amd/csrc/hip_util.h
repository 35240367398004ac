// hip_util.h -- error handling, device buffers and reductions shared by the
// CSRST kernel library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

namespace hu {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HU_CHECK(expr)                                                                               \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            throw hu::Error(-3, std::string(#expr) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + \
                                    ":" + std::to_string(__LINE__) + ")");                           \
    } while (0)

#define HU_LAUNCH_CHECK() HU_CHECK(hipGetLastError())

// Last error of a call that had no handle to store it in (one per thread,
// shared by every translation unit of the library).
inline std::string &global_error()
{
    static thread_local std::string e;
    return e;
}

inline std::string fmt(const char *f, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof(buf), f, ap);
    va_end(ap);
    return buf;
}

// Owning device buffer.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count)
    {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            throw Error(-2, fmt("hipMalloc of %zu bytes failed", count * sizeof(T)));
        }
        n = count;
    }
    void zero(hipStream_t s) { if (n) HU_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), s)); }
    size_t bytes() const { return n * sizeof(T); }
};

// Monotone mapping of doubles onto unsigned 64-bit integers so that min/max
// reductions can use integer atomics (order-independent, hence deterministic).
__device__ __forceinline__ unsigned long long ord(double v)
{
    unsigned long long u = __double_as_longlong(v);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double unord(unsigned long long u)
{
    u = (u & 0x8000000000000000ull) ? (u & 0x7fffffffffffffffull) : ~u;
    double d;
    memcpy(&d, &u, sizeof(d));
    return d;
}

// Wave-level (64-lane) reductions.
template <class T>
__device__ __forceinline__ T wave_min(T v)
{
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v)
{
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

}  // namespace hu
