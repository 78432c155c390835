// common.hpp - shared definitions of the MI355X gravity FMM library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace vg {

// 21 octal digits fit one 63-bit path code; a tree deeper than that is refused.
constexpr int MAXD = 21;
// Expansion order limit: (p+1)^2 <= 169 real coefficients per cell.
constexpr int MAXP = 12;
constexpr int MAXK = (MAXP + 1) * (MAXP + 1);

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define VG_HIP(expr)                                                                                              \
    do {                                                                                                          \
        hipError_t _e = (expr);                                                                                   \
        if (_e != hipSuccess)                                                                                     \
            throw ::vg::Error(4, std::string("HIP error ") + hipGetErrorString(_e) + " at " + __FILE__ + ":" +  \
                                     std::to_string(__LINE__) + " (" #expr ")");                                 \
    } while (0)

#define VG_LAUNCH_CHECK() VG_HIP(hipGetLastError())

inline unsigned ceil_div(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

// Device buffer that can grow; grow() keeps the old contents when asked.
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    void ensure(size_t n, bool keep = false, hipStream_t st = nullptr) {
        if (n <= cap) return;
        size_t nc = cap ? cap : 1024;
        while (nc < n) nc *= 2;
        T *q = nullptr;
        VG_HIP(hipMalloc(&q, nc * sizeof(T)));
        if (keep && p && cap) VG_HIP(hipMemcpyAsync(q, p, cap * sizeof(T), hipMemcpyDeviceToDevice, st));
        if (p) {
            if (keep) VG_HIP(hipStreamSynchronize(st));
            (void)hipFree(p);
        }
        p = q;
        cap = nc;
    }
    void swap(DevBuf &o) noexcept {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
    }
    operator T *() const { return p; }
};

// Real-packed index of a degree-n, order-m (m >= 0) coefficient: degree n owns
// entries [n*n, (n+1)*(n+1)): Re(m=0), then (Re, Im) for m = 1..n.
__host__ __device__ inline int kidx_re(int n, int m) { return m == 0 ? n * n : n * n + 2 * m - 1; }
__host__ __device__ inline int kidx_im(int n, int m) { return n * n + 2 * m; }

}  // namespace vg
