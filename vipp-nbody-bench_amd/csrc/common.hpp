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
// Operators built per evaluation: 8 M2M, 8 L2L and the 189 M2L offsets of
// octant 0's walk (m2l_walk below; the other octants reuse them mirrored).
constexpr int NOPS_USED = 16 + 189;

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

// Block-wide minimum and maximum of each coordinate and a non-finite flag
// over a block of 256 threads, stored as the 7 doubles of one partial box
// (lo xyz, hi xyz, flag) at part.  Every thread of the block must call it.
__device__ inline void aabb_block_store(const double *lo, const double *hi, double bad, double *part) {
    __shared__ double s[7][256];
    for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = lo[a];
        s[3 + a][threadIdx.x] = hi[a];
    }
    s[6][threadIdx.x] = bad;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < static_cast<unsigned>(w))
            for (int a = 0; a < 7; ++a)
                s[a][threadIdx.x] = a < 3 ? fmin(s[a][threadIdx.x], s[a][threadIdx.x + w]) : fmax(s[a][threadIdx.x], s[a][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 7) part[threadIdx.x] = s[threadIdx.x][0];
}

// M2L by mirror symmetry.  The 189 well-separated (parent neighbour d, source
// octant q) pairs of a target of octant 0 are walked in one fixed order; for a
// target of octant qt the geometry is octant 0's mirrored in the axes where qt
// has a bit set (d -> mirror(d), q -> q ^ qt), so the cell offset, and with it
// the operator T_e, is octant 0's.  m2l_walk fills, per step of that order,
// the operator index o (offset + (3,3,3) in base 7), the parent-neighbour slot
// dn (0..26) of the mirrored neighbour and the mirrored source octant qs.
__host__ __device__ inline void m2l_walk(int qt, int16_t *o, int16_t *dn, int16_t *qs) {
    int st = 0;
    for (int d = 0; d < 27; ++d) {
        const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
        for (int q = 0; q < 8; ++q) {
            const int ex = 2 * dx - (q & 1), ey = 2 * dy - ((q >> 1) & 1), ez = 2 * dz - ((q >> 2) & 1);
            if (ex >= -1 && ex <= 1 && ey >= -1 && ey <= 1 && ez >= -1 && ez <= 1) continue;  // adjacent: near field
            const int mx = (qt & 1) ? -dx : dx, my = (qt & 2) ? -dy : dy, mz = (qt & 4) ? -dz : dz;
            o[st] = static_cast<int16_t>((ex + 3) + 7 * (ey + 3) + 49 * (ez + 3));
            dn[st] = static_cast<int16_t>((mx + 1) + 3 * (my + 1) + 9 * (mz + 1));
            qs[st] = static_cast<int16_t>(q ^ qt);
            ++st;
        }
    }
}

// The mirrored operator is D_qt T_e D_qt with D_qt diagonal +-1; its entry for
// real-packed coefficient k (x -> -x: (-1)^m on cosine, -(-1)^m on sine
// coefficients; y -> -y: -1 on sine; z -> -z: (-1)^(n+m)).
__host__ __device__ inline int m2l_mirror_sign(int qt, int k) {
    int nn = 0;  // k = n^2 + r: degree n, r = 0 (m = 0), 2m - 1 (cosine) or 2m (sine)
    while ((nn + 1) * (nn + 1) <= k) ++nn;
    const int r = k - nn * nn, m = (r + 1) >> 1;
    const bool im = r > 0 && (r & 1) == 0;
    int sg = 1;
    if (qt & 1) sg *= ((m & 1) ? -1 : 1) * (im ? -1 : 1);
    if ((qt & 2) && im) sg = -sg;
    if ((qt & 4) && ((nn + m) & 1)) sg = -sg;
    return sg;
}

}  // namespace vg
