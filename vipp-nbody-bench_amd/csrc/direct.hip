// direct.hip - FP64 direct-sum accelerations of a subset of particles against
// all N, the force oracle of solver_ref_naive.hpp:103-119:
//   a_i = sum_{j != i} G m_j (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^{3/2}
// (and, on request, the potential sum_{j != i} m_j / sqrt(|x_j - x_i|^2 + eps^2)
// without G).  Used to measure the FMM's force error on large samples.
//
// Layout: one thread per target, the sources streamed through LDS in tiles of
// 256 (x, y, z, m) FP64 records.  The source range is cut into S slices (grid
// y) so that a few thousand targets still fill the chip; the S partial sums of
// a target are added in slice order afterwards, so the result does not depend
// on scheduling.
#include <algorithm>
#include <atomic>
#include <climits>
#include <mutex>

#include "gravity.hpp"

namespace vg {
namespace {

constexpr int DT = 256;  // threads per block and sources per LDS tile

// bad (may be null): set to 1 when a target index lies outside [0, n); such
// a target is skipped (the host-side entry point checks indices beforehand)
__global__ __launch_bounds__(DT) void k_direct(const double *__restrict__ x, const double *__restrict__ masses, double mass, int64_t n,
                                               const int64_t *__restrict__ idx, int64_t nidx, int64_t slice, double G, double eps2, int potential,
                                               double *__restrict__ part, double *__restrict__ bad) {
    __shared__ double sx[DT], sy[DT], sz[DT], sm[DT];
    const int64_t k = blockIdx.x * static_cast<int64_t>(DT) + threadIdx.x;
    const int64_t i0 = k < nidx ? idx[k] : -1;
    const bool live = k < nidx && i0 >= 0 && i0 < n;
    if (k < nidx && !live && bad) *bad = 1.0;
    const int64_t i = live ? i0 : -1;
    const double xi = live ? x[3 * i] : 0, yi = live ? x[3 * i + 1] : 0, zi = live ? x[3 * i + 2] : 0;
    const int64_t j0 = blockIdx.y * slice;
    const int64_t j1 = j0 + slice < n ? j0 + slice : n;
    double ax = 0, ay = 0, az = 0, pt = 0;
    for (int64_t b = j0; b < j1; b += DT) {
        __syncthreads();
        const int64_t j = b + threadIdx.x;
        if (j < j1) {
            sx[threadIdx.x] = x[3 * j];
            sy[threadIdx.x] = x[3 * j + 1];
            sz[threadIdx.x] = x[3 * j + 2];
            sm[threadIdx.x] = masses ? masses[j] : mass;
        }
        __syncthreads();
        const int nt = j1 - b < DT ? static_cast<int>(j1 - b) : DT;
        for (int u = 0; u < nt; ++u) {
            if (b + u == i) continue;  // solver_ref_naive.hpp:111-113
            const double dx = sx[u] - xi, dy = sy[u] - yi, dz = sz[u] - zi;
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            const double r = sqrt(r2);
            const double gm = G * sm[u] / (r2 * r);
            ax += gm * dx;
            ay += gm * dy;
            az += gm * dz;
            if (potential) pt += sm[u] / r;
        }
    }
    if (k >= nidx) return;
    double *o = part + (static_cast<int64_t>(blockIdx.y) * nidx + k) * 4;
    o[0] = live ? ax : 0.0;
    o[1] = live ? ay : 0.0;
    o[2] = live ? az : 0.0;
    o[3] = live ? pt : 0.0;
}

__global__ __launch_bounds__(256) void k_direct_sum(const double *__restrict__ part, int64_t nidx, int S, double *__restrict__ a,
                                                    double *__restrict__ pot) {
    const int64_t k = blockIdx.x * 256ll + threadIdx.x;
    if (k >= nidx) return;
    double s[4] = {0, 0, 0, 0};
    for (int y = 0; y < S; ++y)
        for (int c = 0; c < 4; ++c) s[c] += part[(static_cast<int64_t>(y) * nidx + k) * 4 + c];
    a[3 * k] = s[0];
    a[3 * k + 1] = s[1];
    a[3 * k + 2] = s[2];
    if (pot) pot[k] = s[3];
}

__global__ void k_clear(double *__restrict__ w) { *w = 0.0; }

// Scratch of the allocation-free entry point: partial sums of up to
// NOALLOC_RECORDS (target, slice) records of 4 doubles, then one flag word.
// Each device holds its own copy; host code reaches it only through the
// address hipGetSymbolAddress gives for that device (noalloc_scratch()).
constexpr int64_t NOALLOC_RECORDS = int64_t(1) << 19;  // 16 MiB of partial sums
static __device__ __align__(256) double scratch[4 * NOALLOC_RECORDS + 32];

// slices of the source range: enough blocks to fill the chip, each slice at
// least 4096 sources, at most max_records / nidx of them
int64_t direct_slices(int64_t n, int64_t nidx, int64_t max_records) {
    const int64_t tblocks = ceil_div(nidx, DT);
    int64_t S = std::max<int64_t>(1, std::min<int64_t>(ceil_div(4096, tblocks), ceil_div(n, 4096)));
    S = std::min<int64_t>(S, 65535);
    S = std::max<int64_t>(1, std::min<int64_t>(S, max_records / nidx));
    return S;
}

void launch_direct(Ctx &c, const int64_t *idx, int64_t nidx, double *part, int64_t S, double *a, double *pot, double *bad) {
    const int64_t slice = ceil_div(c.n, S);
    S = ceil_div(c.n, slice);
    const dim3 grid(ceil_div(nidx, DT), static_cast<unsigned>(S));
    k_direct<<<grid, DT, 0, c.stream>>>(c.x, c.has_masses ? c.m.p : nullptr, c.mass, c.n, idx, nidx, slice, c.G, c.eps * c.eps, pot ? 1 : 0,
                                        part, bad);
    VG_LAUNCH_CHECK();
    k_direct_sum<<<ceil_div(nidx, 256), 256, 0, c.stream>>>(part, nidx, static_cast<int>(S), a, pot);
    VG_LAUNCH_CHECK();
}

// the device address of `scratch` on each device ordinal, set once per device
// by noalloc_init (the array's host-side name is not a device address)
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
std::atomic<double *> g_scratch_ptr[MAX_DEVICES];

}  // namespace

void noalloc_init(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw Error(1, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device].load()) return;
    int cur = -1;
    VG_HIP(hipGetDevice(&cur));
    if (cur != device) throw Error(4, "noalloc_init: the device must be current");
    void *q = nullptr;
    VG_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device].store(static_cast<double *>(q));
}

void direct_subset_noalloc(Ctx &c, const int64_t *idx, int64_t nidx, double *a, double *pot) {
    if (nidx <= 0) return;
    if (nidx > NOALLOC_RECORDS) throw Error(1, "direct (no allocation): more than 524288 targets in one call");
    double *scr = c.device >= 0 && c.device < MAX_DEVICES ? g_scratch_ptr[c.device].load() : nullptr;
    if (!scr) throw Error(5, "direct (no allocation): scratch of this device not set up");
    double *bad = scr + 4 * NOALLOC_RECORDS;
    k_clear<<<1, 1, 0, c.stream>>>(bad);
    VG_LAUNCH_CHECK();
    launch_direct(c, idx, nidx, scr, direct_slices(c.n, nidx, NOALLOC_RECORDS), a, pot, bad);
    double *hflag = reinterpret_cast<double *>(static_cast<char *>(c.pinned) + Ctx::PINNED_BYTES - 64);  // a pinned host word
    VG_HIP(hipMemcpyAsync(hflag, bad, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    if (*hflag != 0.0) throw Error(1, "direct: particle index out of range");
}

void direct_subset(Ctx &c, const int64_t *idx_host, int64_t nidx, double *a_host, double *pot_host) {
    if (nidx <= 0) return;
    for (int64_t k = 0; k < nidx; ++k)
        if (idx_host[k] < 0 || idx_host[k] >= c.n) throw Error(1, "direct: particle index out of range");
    const int64_t S = direct_slices(c.n, nidx, INT64_MAX);
    DevBuf<int64_t> idx;
    DevBuf<double> part, a, pot;
    idx.ensure(nidx);
    part.ensure(static_cast<size_t>(S * nidx * 4));
    a.ensure(3 * nidx);
    if (pot_host) pot.ensure(nidx);
    VG_HIP(hipMemcpyAsync(idx.p, idx_host, nidx * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
    launch_direct(c, idx, nidx, part, S, a, pot_host ? pot.p : nullptr, nullptr);
    VG_HIP(hipMemcpyAsync(a_host, a.p, 3 * nidx * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (pot_host) VG_HIP(hipMemcpyAsync(pot_host, pot.p, nidx * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));  // the scratch buffers are released on return
}

}  // namespace vg
