// tree.hip - octree build and interaction lists on the GPU.
//
// The reference inserts points one by one into an adaptive octree
// (octree.hpp:79-191).  Its result does not depend on the insertion order:
//  * a point descends by octant o = [x<mid] + 2[y<mid] + 4[z<mid] with
//    mid = (max+min)/2 in FP64 and ties on the high side (first-fit over the
//    children 0..7 with closed bounds, octree.hpp:136,153-158,169-176);
//  * a cell has children iff it holds more than floor(sqrt(8) s) points
//    (octree.hpp:83,143).
// So every point gets a 63-bit path code (3 bits per level, 21 levels) by a
// per-point FP64 midpoint descent, the codes are radix sorted (stable, so equal
// codes keep the caller's order), and each level of cells is cut out of the
// sorted codes with binary searches.  Leaves are then sorted by caller index,
// which is the order the reference's leaf vectors have.
#include "gravity.hpp"

#include <algorithm>
#include <cmath>

namespace vg {
namespace {

constexpr double PADDING = 1e-5;  // Problem::OCTREE_WIDTH_PADDING (problem.hpp:167), (double)1E-5L == 1e-5

__global__ __launch_bounds__(256) void k_aabb(const double *__restrict__ x, int64_t n, double *__restrict__ part) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double bad = 0;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
        for (int a = 0; a < 3; ++a) {
            const double v = x[3 * i + a];
            if (!isfinite(v)) bad = 1;
            lo[a] = fmin(lo[a], v);
            hi[a] = fmax(hi[a], v);
        }
    }
    aabb_block_store(lo, hi, bad, part + 7ll * blockIdx.x);
}

// final reduction; box = padded root bounds (fmm.hpp:403-405), box[6] = non-finite flag
__global__ __launch_bounds__(256) void k_aabb_final(const double *__restrict__ part, int nb, double *__restrict__ box) {
    __shared__ double s[7][256];
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0};
    for (int b = threadIdx.x; b < nb; b += 256) {
        for (int a = 0; a < 3; ++a) {
            v[a] = fmin(v[a], part[7 * b + a]);
            v[3 + a] = fmax(v[3 + a], part[7 * b + 3 + a]);
        }
        v[6] = fmax(v[6], part[7 * b + 6]);
    }
    for (int a = 0; a < 7; ++a) s[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < static_cast<unsigned>(w))
            for (int a = 0; a < 7; ++a)
                s[a][threadIdx.x] = a < 3 ? fmin(s[a][threadIdx.x], s[a][threadIdx.x + w]) : fmax(s[a][threadIdx.x], s[a][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        box[threadIdx.x] = s[threadIdx.x][0] - PADDING;          // min and max are exact: the order of the reduction does not matter
        box[3 + threadIdx.x] = s[3 + threadIdx.x][0] + PADDING;
    }
    if (threadIdx.x == 0) box[6] = s[6][0];
}

// 63-bit path code: level-1 octant in bits 62..60, level-21 octant in bits 2..0;
// the octants of levels 1..levels only (the bits of deeper levels stay 0: the
// tree reads no digit below the levels it sorts).
__global__ __launch_bounds__(256) void k_codes(const double *__restrict__ x, int64_t n, const double *__restrict__ box, int levels,
                                               uint64_t *__restrict__ key, uint32_t *__restrict__ val) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= n) return;
    double lo[3] = {box[0], box[1], box[2]}, hi[3] = {box[3], box[4], box[5]};
    const double p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    uint64_t code = 0;
    for (int l = 0; l < levels; ++l) {
        unsigned o = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double mid = (hi[a] + lo[a]) / 2;  // octree.hpp:169
            const bool low = p[a] < mid;
            o |= static_cast<unsigned>(low) << a;
            if (low)
                hi[a] = mid;
            else
                lo[a] = mid;
        }
        code = (code << 3) | o;
    }
    key[i] = code << (3 * (MAXD - levels));
    val[i] = static_cast<uint32_t>(i);
}

__global__ void k_root(const double *__restrict__ box, int64_t n, uint32_t thr, uint32_t *start, uint32_t *count, int32_t *child, int32_t *parent,
                       uint8_t *octant, int8_t *level, double *bmin, double *bmax, int32_t *grid, int32_t *anc, int32_t *int_ids, uint32_t *nint) {
    if (threadIdx.x != 0) return;
    start[0] = 0;
    count[0] = static_cast<uint32_t>(n);
    child[0] = -1;
    parent[0] = -1;
    octant[0] = 0;
    level[0] = 0;
    for (int a = 0; a < 3; ++a) {
        // an empty input keeps the zero box of the reference (fmm.hpp:404-405)
        bmin[a] = n ? box[a] : 0.0;
        bmax[a] = n ? box[3 + a] : 0.0;
        grid[a] = 0;
    }
    for (int l = 0; l <= MAXD; ++l) anc[l] = l == 0 ? 0 : -1;
    const bool internal = static_cast<uint64_t>(n) > thr;
    if (internal) int_ids[0] = 0;
    *nint = internal ? 1u : 0u;
}

__device__ inline uint32_t digit_lower_bound(const uint64_t *__restrict__ key, uint32_t lo, uint32_t hi, int shift, unsigned q) {
    // first position in [lo, hi) whose octal digit at `shift` is >= q (digits are sorted there)
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (((key[mid] >> shift) & 7u) < q)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// A child's row from its parent's: everything of the parent is read before
// any store, so the loads issue together (the same arrays are written for the
// new cells in the same kernel, so the compiler keeps loads and stores in
// program order).
__device__ __forceinline__ void read_parent(int64_t c, int32_t P, int d, const double *bmin, const double *bmax, const int32_t *grid,
                                            const int32_t *anc, double *pmn, double *pmx, int32_t *pg, int32_t *pa) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        pmn[a] = bmin[3ll * P + a];
        pmx[a] = bmax[3ll * P + a];
        pg[a] = grid[3ll * P + a];
    }
#pragma unroll
    for (int l = 0; l <= MAXD; ++l) pa[l] = l <= d ? anc[static_cast<int64_t>(P) * (MAXD + 1) + l] : (l == d + 1 ? static_cast<int32_t>(c) : -1);
}
__device__ __forceinline__ void write_child(int64_t c, int32_t P, unsigned q, int d, uint32_t lo, uint32_t hi, uint32_t *start, uint32_t *count,
                                            int32_t *child, int32_t *parent, uint8_t *octant, int8_t *level, double *bmin, double *bmax,
                                            int32_t *grid, int32_t *anc, const double *pmn, const double *pmx, const int32_t *pg,
                                            const int32_t *pa) {
    start[c] = lo;
    count[c] = hi - lo;
    child[c] = -1;
    parent[c] = P;
    octant[c] = static_cast<uint8_t>(q);
    level[c] = static_cast<int8_t>(d + 1);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double mid = (pmx[a] + pmn[a]) / 2;
        const bool low = (q >> a) & 1u;
        bmin[3 * c + a] = low ? pmn[a] : mid;
        bmax[3 * c + a] = low ? mid : pmx[a];
        grid[3 * c + a] = 2 * pg[a] + (low ? 0 : 1);
    }
#pragma unroll
    for (int l = 0; l <= MAXD; ++l) anc[c * (MAXD + 1) + l] = pa[l];
}

// Children of the internal cells of level d (octree.hpp:166-191).
__global__ __launch_bounds__(256) void k_expand(int d, const int32_t *__restrict__ ids, int nint, int64_t child_off, const uint64_t *__restrict__ key,
                                                uint32_t thr, uint32_t *start, uint32_t *count, int32_t *child, int32_t *parent, uint8_t *octant,
                                                int8_t *level, double *bmin, double *bmax, int32_t *grid, int32_t *anc, uint8_t *flag) {
    const int64_t t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= 8ll * nint) return;
    const int k = static_cast<int>(t >> 3);
    const unsigned q = static_cast<unsigned>(t & 7);
    const int32_t P = ids[k];
    const uint32_t s = start[P], e = s + count[P];
    const int64_t c = child_off + 8ll * k + q;
    double pmn[3], pmx[3];
    int32_t pg[3], pa[MAXD + 1];
    read_parent(c, P, d, bmin, bmax, grid, anc, pmn, pmx, pg, pa);
    const int shift = 3 * (MAXD - 1 - d);
    // child q ends where child q + 1 starts: the 8 children of a parent are 8
    // consecutive lanes (8-aligned), so one search per lane and a shuffle
    const uint32_t lo = digit_lower_bound(key, s, e, shift, q);
    const uint32_t nx = __shfl_down(lo, 1, 8);
    const uint32_t hi = q == 7 ? e : nx;
    write_child(c, P, q, d, lo, hi, start, count, child, parent, octant, level, bmin, bmax, grid, anc, pmn, pmx, pg, pa);
    flag[8ll * k + q] = (hi - lo) > thr ? 1 : 0;
    if (q == 0) child[P] = static_cast<int32_t>(child_off + 8ll * k);
}

// Ordered compaction in one block: out[j] = base + i for the j-th set flag.
__global__ __launch_bounds__(1024) void k_compact(const uint8_t *__restrict__ flag, int64_t nflag, int64_t base, int32_t *__restrict__ out,
                                                  uint32_t *__restrict__ nout) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const unsigned lane = threadIdx.x & 63;
    for (int64_t off = 0; off < nflag; off += 1024) {
        const int64_t i = off + threadIdx.x;
        const uint32_t f = (i < nflag && flag[i]) ? 1u : 0u;
        const uint64_t bal = __ballot(f);
        const uint32_t before = __popcll(bal & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
        if (lane == 0) wsum[threadIdx.x >> 6] = __popcll(bal);
        __syncthreads();
        uint32_t wb = 0;
        for (unsigned w = 0; w < (threadIdx.x >> 6); ++w) wb += wsum[w];
        const uint32_t c0 = carry;
        if (f) out[c0 + wb + before] = static_cast<int32_t>(base + i);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (int w = 0; w < 16; ++w) tot += wsum[w];
            carry = c0 + tot;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *nout = carry;
}

// Multi-block ordered compaction: per 1024-flag block counts, an exclusive
// scan of the counts in one block, then each block writes its entries.
__global__ __launch_bounds__(1024) void k_flag_counts(const uint8_t *__restrict__ flag, int64_t nflag, uint32_t *__restrict__ bcnt) {
    __shared__ uint32_t wsum[16];
    const int64_t i = blockIdx.x * 1024ll + threadIdx.x;
    const uint64_t bal = __ballot(i < nflag && flag[i]);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < 16; ++w) t += wsum[w];
        bcnt[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(1024) void k_count_scan(uint32_t *__restrict__ bcnt, int nb, uint32_t *__restrict__ nout) {
    __shared__ uint32_t carry;
    __shared__ uint32_t wsum[16];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const unsigned lane = threadIdx.x & 63;
    for (int off = 0; off < nb; off += 1024) {
        const int i = off + threadIdx.x;
        const uint32_t v = i < nb ? bcnt[i] : 0u;
        uint32_t x = v;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(x, d, 64);
            if (lane >= static_cast<unsigned>(d)) x += o;
        }
        if (lane == 63) wsum[threadIdx.x >> 6] = x;
        __syncthreads();
        uint32_t wb = 0;
        for (unsigned w = 0; w < (threadIdx.x >> 6); ++w) wb += wsum[w];
        const uint32_t c0 = carry;
        if (i < nb) bcnt[i] = c0 + wb + x - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c0 + wb + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) *nout = carry;
}

__global__ __launch_bounds__(1024) void k_flag_write(const uint8_t *__restrict__ flag, int64_t nflag, int64_t base, const uint32_t *__restrict__ boff,
                                                     int32_t *__restrict__ out) {
    __shared__ uint32_t wsum[16];
    const int64_t i = blockIdx.x * 1024ll + threadIdx.x;
    const unsigned lane = threadIdx.x & 63;
    const bool f = i < nflag && flag[i];
    const uint64_t bal = __ballot(f);
    if (lane == 0) wsum[threadIdx.x >> 6] = __popcll(bal);
    __syncthreads();
    uint32_t wb = boff[blockIdx.x];
    for (unsigned w = 0; w < (threadIdx.x >> 6); ++w) wb += wsum[w];
    if (f) out[wb + __popcll(bal & ((1ull << lane) - 1))] = static_cast<int32_t>(base + i);
}

__global__ __launch_bounds__(256) void k_leaf_flags(const int32_t *__restrict__ child, const uint32_t *__restrict__ count, int64_t nc,
                                                    uint8_t *__restrict__ flag) {
    const int64_t c = blockIdx.x * 256ll + threadIdx.x;
    if (c < nc) flag[c] = (child[c] < 0 && count[c] > 0) ? 1 : 0;
}

// Sort each leaf's caller indices ascending (a leaf holds at most thr points).
__global__ __launch_bounds__(256) void k_leaf_sort(const int32_t *__restrict__ leaves, const uint32_t *__restrict__ start, const uint32_t *__restrict__ count,
                                                   const uint32_t *__restrict__ in, uint32_t *__restrict__ out) {
    extern __shared__ uint32_t sv[];
    const int32_t c = leaves[blockIdx.x];
    const uint32_t s = start[c], n = count[c];
    for (uint32_t i = threadIdx.x; i < n; i += 256) sv[i] = in[s + i];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
        const uint32_t v = sv[i];
        uint32_t r = 0;
        for (uint32_t j = 0; j < n; ++j) r += sv[j] < v;
        out[s + r] = v;
    }
}

// ---- device-driven level loop: the level sizes stay on the device -------
// lv[d]: first cell of level d; lv[LV_I + d]: first entry of level d's
// internal cells in int_ids; LV_OVF: a capacity or grid bound was exceeded
// (bit 0) or the tree is deeper than the sorted digits (bit 1), and the host
// then builds again level by level; LV_NLEAF: non-empty leaves.
constexpr int LV_I = MAXD + 3, LV_OVF = 2 * (MAXD + 3), LV_NLEAF = LV_OVF + 1, LV_WORDS = LV_OVF + 2;

__global__ void k_lv_init(const uint32_t *__restrict__ nint0, uint32_t *__restrict__ lv) {
    for (int i = threadIdx.x; i < LV_WORDS; i += blockDim.x) lv[i] = 0;
    if (threadIdx.x == 0) {
        lv[1] = 1;
        lv[LV_I + 1] = *nint0;
    }
}

// k_expand with the internal cells of level d and the first new cell read from lv
__global__ __launch_bounds__(256) void k_expand_lv(int d, uint32_t *lv, const int32_t *__restrict__ int_ids, int64_t cap_cells,
                                                   const uint64_t *__restrict__ key, uint32_t thr, uint32_t *start, uint32_t *count, int32_t *child,
                                                   int32_t *parent, uint8_t *octant, int8_t *level, double *bmin, double *bmax, int32_t *grid,
                                                   int32_t *anc, uint8_t *flag) {
    const uint32_t i0 = lv[LV_I + d], nint = lv[LV_I + d + 1] - i0;
    const int64_t child_off = lv[d + 1];
    if (lv[LV_OVF]) return;
    if (child_off + 8ll * nint > cap_cells || 8ll * nint > static_cast<int64_t>(gridDim.x) * 256) {
        if (blockIdx.x == 0 && threadIdx.x == 0) lv[LV_OVF] = 1;  // every block sees the same sizes and returns
        return;
    }
    const int64_t t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= 8ll * nint) return;
    const int k = static_cast<int>(t >> 3);
    const unsigned q = static_cast<unsigned>(t & 7);
    const int32_t P = int_ids[i0 + k];
    const uint32_t s = start[P], e = s + count[P];
    const int64_t c = child_off + 8ll * k + q;
    double pmn[3], pmx[3];
    int32_t pg[3], pa[MAXD + 1];
    read_parent(c, P, d, bmin, bmax, grid, anc, pmn, pmx, pg, pa);
    const int shift = 3 * (MAXD - 1 - d);
    // child q ends where child q + 1 starts: the 8 children of a parent are 8
    // consecutive lanes (8-aligned), so one search per lane and a shuffle
    const uint32_t lo = digit_lower_bound(key, s, e, shift, q);
    const uint32_t nx = __shfl_down(lo, 1, 8);
    const uint32_t hi = q == 7 ? e : nx;
    write_child(c, P, q, d, lo, hi, start, count, child, parent, octant, level, bmin, bmax, grid, anc, pmn, pmx, pg, pa);
    flag[8ll * k + q] = (hi - lo) > thr ? 1 : 0;
    if (q == 0) child[P] = static_cast<int32_t>(child_off + 8ll * k);
}

// What a device-driven compaction works on: the children of level d's
// internal cells (d >= 0: flags at cell lv[d + 1] on, entries appended to
// int_ids at lv[LV_I + d + 1]), or all cells below lv[-d - 1] (d < 0: the
// leaves, entries from 0).
struct LvCmp {
    int64_t nflag, base, out0;
};
__device__ inline LvCmp lv_cmp(const uint32_t *lv, int d) {
    if (lv[LV_OVF]) return {0, 0, 0};
    if (d >= 0) return {8ll * (lv[LV_I + d + 1] - lv[LV_I + d]), lv[d + 1], lv[LV_I + d + 1]};
    return {lv[-d - 1], 0, 0};
}

__global__ __launch_bounds__(1024) void k_flag_counts_lv(const uint8_t *__restrict__ flag, const uint32_t *__restrict__ lv, int d,
                                                         uint32_t *__restrict__ bcnt) {
    __shared__ uint32_t wsum[16];
    const LvCmp a = lv_cmp(lv, d);
    const int64_t i = blockIdx.x * 1024ll + threadIdx.x;
    const uint64_t bal = __ballot(i < a.nflag && flag[i]);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < 16; ++w) t += wsum[w];
        bcnt[blockIdx.x] = t;
    }
}

// exclusive scan of the block counts in one block; the total goes into lv
__global__ __launch_bounds__(1024) void k_count_scan_lv(uint32_t *__restrict__ bcnt, int nb, uint32_t *lv, int d, int64_t cap_out, bool last) {
    __shared__ uint32_t carry;
    __shared__ uint32_t wsum[16];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const unsigned lane = threadIdx.x & 63;
    for (int off = 0; off < nb; off += 1024) {
        const int i = off + threadIdx.x;
        const uint32_t v = i < nb ? bcnt[i] : 0u;
        uint32_t x = v;
        for (int s = 1; s < 64; s <<= 1) {
            const uint32_t o = __shfl_up(x, s, 64);
            if (lane >= static_cast<unsigned>(s)) x += o;
        }
        if (lane == 63) wsum[threadIdx.x >> 6] = x;
        __syncthreads();
        uint32_t wb = 0;
        for (unsigned w = 0; w < (threadIdx.x >> 6); ++w) wb += wsum[w];
        const uint32_t c0 = carry;
        if (i < nb) bcnt[i] = c0 + wb + x - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c0 + wb + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint32_t tot = carry;
        if (d >= 0) {
            const uint32_t nint = lv[LV_I + d + 1] - lv[LV_I + d];
            lv[LV_I + d + 2] = lv[LV_I + d + 1] + tot;
            lv[d + 2] = lv[d + 1] + 8u * nint;
            if (lv[LV_I + d + 2] > cap_out) lv[LV_OVF] |= 1u;
            if (last && tot > 0) lv[LV_OVF] |= 2u;  // cells of the last sorted level split further
        } else {
            lv[LV_NLEAF] = tot;
        }
    }
}

__global__ __launch_bounds__(1024) void k_flag_write_lv(const uint8_t *__restrict__ flag, const uint32_t *__restrict__ lv, int d,
                                                        const uint32_t *__restrict__ boff, int32_t *__restrict__ out, int64_t cap_out) {
    __shared__ uint32_t wsum[16];
    const LvCmp a = lv_cmp(lv, d);
    const int64_t i = blockIdx.x * 1024ll + threadIdx.x;
    const unsigned lane = threadIdx.x & 63;
    const bool f = i < a.nflag && flag[i];
    const uint64_t bal = __ballot(f);
    if (lane == 0) wsum[threadIdx.x >> 6] = __popcll(bal);
    __syncthreads();
    uint32_t wb = boff[blockIdx.x];
    for (unsigned w = 0; w < (threadIdx.x >> 6); ++w) wb += wsum[w];
    const int64_t pos = a.out0 + wb + __popcll(bal & ((1ull << lane) - 1));
    if (f && pos < cap_out) out[pos] = static_cast<int32_t>(a.base + i);
}

// Set flag bytes (each 0 or 1) among the first rem (>= 1) bytes of w.
__device__ inline uint32_t flag_count16(const uint4 w, int64_t rem) {
    const uint32_t v[4] = {w.x, w.y, w.z, w.w};
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int64_t r = rem - 4 * j;
        const uint32_t m = r >= 4 ? 0xffffffffu : (r <= 0 ? 0u : (1u << (8 * r)) - 1u);
        c += __builtin_popcount(v[j] & m);
    }
    return c;
}

// k_flag_counts_lv, k_count_scan_lv and k_flag_write_lv in one block, for up
// to WIDE_FLAGS flags: thread t counts the flags of the contiguous run
// [t per, (t + 1) per) 16 bytes at a time, the block scans the counts once, and
// each thread writes its run's entries in order; then the level table is
// updated as k_count_scan_lv does.  The writes of one wavefront's lanes are
// strided, so this suits the internal cells of a level (a small fraction of
// the flags), not the leaves.
constexpr int64_t WIDE_FLAGS = 1 << 21;
__global__ __launch_bounds__(1024) void k_compact_wide_lv(const uint8_t *__restrict__ flag, uint32_t *lv, int d, int32_t *__restrict__ out,
                                                          int64_t cap_out, bool last) {
    __shared__ uint32_t wsum[16];
    const LvCmp a = lv_cmp(lv, d);  // every thread reads lv before the barrier; the last thread writes it after
    const int64_t per = ((a.nflag + 1023) / 1024 + 15) / 16 * 16;
    const int64_t b0 = threadIdx.x * per, b1 = min(a.nflag, b0 + per);
    uint32_t cnt = 0;
    for (int64_t i = b0; i < b1; i += 16) cnt += flag_count16(*reinterpret_cast<const uint4 *>(flag + i), b1 - i);
    const unsigned lane = threadIdx.x & 63;
    uint32_t x = cnt;
    for (int s = 1; s < 64; s <<= 1) {
        const uint32_t o = __shfl_up(x, s, 64);
        if (lane >= static_cast<unsigned>(s)) x += o;
    }
    if (lane == 63) wsum[threadIdx.x >> 6] = x;
    __syncthreads();
    uint32_t wb = 0;
    for (unsigned w = 0; w < (threadIdx.x >> 6); ++w) wb += wsum[w];
    int64_t pos = a.out0 + wb + x - cnt;
    for (int64_t i = b0; i < b1; i += 16) {
        const uint4 w = *reinterpret_cast<const uint4 *>(flag + i);
        const uint32_t v[4] = {w.x, w.y, w.z, w.w};
        for (int k = 0; k < 16 && i + k < b1; ++k)
            if ((v[k >> 2] >> (8 * (k & 3))) & 0xffu) {
                if (pos < cap_out) out[pos] = static_cast<int32_t>(a.base + i + k);
                ++pos;
            }
    }
    if (threadIdx.x == 1023) {
        const uint32_t tot = wb + x;
        if (d >= 0) {
            const uint32_t nint = lv[LV_I + d + 1] - lv[LV_I + d];
            lv[LV_I + d + 2] = lv[LV_I + d + 1] + tot;
            lv[d + 2] = lv[d + 1] + 8u * nint;
            if (lv[LV_I + d + 2] > cap_out) lv[LV_OVF] |= 1u;
            if (last && tot > 0) lv[LV_OVF] |= 2u;  // cells of the last sorted level split further
        } else {
            lv[LV_NLEAF] = tot;
        }
    }
}

__global__ __launch_bounds__(256) void k_leaf_flags_lv(const int32_t *__restrict__ child, const uint32_t *__restrict__ count, const uint32_t *lv,
                                                       int last, int64_t cap, uint8_t *__restrict__ flag) {
    const int64_t c = blockIdx.x * 256ll + threadIdx.x;
    if (c >= cap) return;
    const int64_t nc = lv[last];
    flag[c] = (c < nc && child[c] < 0 && count[c] > 0) ? 1 : 0;
}

// k_leaf_sort with a wavefront per leaf (LS_WAVES leaves per block, LDS of
// LS_WAVES thr words), fused with k_gather: the leaves hold about s points, so
// a block per leaf would leave most of its threads idle.  nl leaves, or with
// lv the lv[LV_NLEAF] leaves of a device-driven build (none when it did not
// complete).  A lane's first LS_REG points are gathered as soon as their
// caller indices are read, so the random loads of their coordinates overlap
// the ranking; the points of larger leaves beyond 64 LS_REG are gathered after
// it.  Each point is written to its sorted slot with the arithmetic of
// k_gather.
constexpr int LS_WAVES = 4;
constexpr int LS_REG = 3;
constexpr uint32_t LS_MAXTHR = 4096;  // LDS bound of the wavefront form
__global__ __launch_bounds__(64 * LS_WAVES) void k_leaf_sort_w(const int32_t *__restrict__ leaves, uint32_t nl, const uint32_t *lv,
                                                               const uint32_t *__restrict__ start, const uint32_t *__restrict__ count, uint32_t thr,
                                                               const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                               const double *__restrict__ x, const double *__restrict__ m, double mass,
                                                               const double *__restrict__ box, double *__restrict__ xs64, float *__restrict__ xs32) {
    extern __shared__ uint32_t sv[];
    if (lv) {
        if (lv[LV_OVF]) return;
        nl = lv[LV_NLEAF];
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t *sw = sv + w * thr;
    const double cx = (box[0] + box[3]) / 2, cy = (box[1] + box[4]) / 2, cz = (box[2] + box[5]) / 2;
    auto put = [&](uint32_t k, uint32_t i, double px, double py, double pz, double mi) {
        out[k] = i;
        xs64[4ll * k] = px;
        xs64[4ll * k + 1] = py;
        xs64[4ll * k + 2] = pz;
        xs64[4ll * k + 3] = mi;
        xs32[4ll * k] = static_cast<float>(px - cx);
        xs32[4ll * k + 1] = static_cast<float>(py - cy);
        xs32[4ll * k + 2] = static_cast<float>(pz - cz);
        xs32[4ll * k + 3] = static_cast<float>(mi);
    };
    for (uint32_t j = blockIdx.x * LS_WAVES + w; j < nl; j += gridDim.x * LS_WAVES) {
        const int32_t c = leaves[j];
        const uint32_t s = start[c], n = count[c];
        if (n > thr) continue;  // not a leaf of a complete build (same n for the whole wavefront)
        uint32_t vi[LS_REG];
        double px[LS_REG], py[LS_REG], pz[LS_REG], pm[LS_REG];
#pragma unroll
        for (int u = 0; u < LS_REG; ++u) {
            const uint32_t i = lane + 64u * u;
            vi[u] = i < n ? in[s + i] : 0u;
        }
#pragma unroll
        for (int u = 0; u < LS_REG; ++u) {
            const uint32_t i = lane + 64u * u;
            if (i < n) {
                px[u] = x[3ll * vi[u]];
                py[u] = x[3ll * vi[u] + 1];
                pz[u] = x[3ll * vi[u] + 2];
                pm[u] = m ? m[vi[u]] : mass;
                sw[i] = vi[u];
            }
        }
        for (uint32_t i = lane + 64u * LS_REG; i < n; i += 64) sw[i] = in[s + i];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int u = 0; u < LS_REG; ++u) {
            const uint32_t i = lane + 64u * u;
            if (i < n) {
                uint32_t r = 0;
                for (uint32_t k = 0; k < n; ++k) r += sw[k] < vi[u];
                put(s + r, vi[u], px[u], py[u], pz[u], pm[u]);
            }
        }
        for (uint32_t i = lane + 64u * LS_REG; i < n; i += 64) {
            const uint32_t v = sw[i];
            uint32_t r = 0;
            for (uint32_t k = 0; k < n; ++k) r += sw[k] < v;
            put(s + r, v, x[3ll * v], x[3ll * v + 1], x[3ll * v + 2], m ? m[v] : mass);
        }
        __builtin_amdgcn_wave_barrier();  // every lane has read sw before the next leaf overwrites it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// lv (optional): skipped when a device-driven build did not complete
__global__ __launch_bounds__(256) void k_gather(const uint32_t *__restrict__ perm, int64_t n, const double *__restrict__ x, const double *__restrict__ m,
                                                double mass, const double *__restrict__ box, double *__restrict__ xs64, float *__restrict__ xs32,
                                                const uint32_t *lv) {
    const int64_t k = blockIdx.x * 256ll + threadIdx.x;
    if (k >= n || (lv && lv[LV_OVF])) return;
    const uint32_t i = perm[k];
    const double px = x[3ll * i], py = x[3ll * i + 1], pz = x[3ll * i + 2];
    const double mi = m ? m[i] : mass;
    xs64[4 * k] = px;
    xs64[4 * k + 1] = py;
    xs64[4 * k + 2] = pz;
    xs64[4 * k + 3] = mi;
    const double cx = (box[0] + box[3]) / 2, cy = (box[1] + box[4]) / 2, cz = (box[2] + box[5]) / 2;
    xs32[4 * k] = static_cast<float>(px - cx);
    xs32[4 * k + 1] = static_cast<float>(py - cy);
    xs32[4 * k + 2] = static_cast<float>(pz - cz);
    xs32[4 * k + 3] = static_cast<float>(mi);
}

// ---------------------------------------------------------------------------
// Interaction lists, level by level (calc_link, solver_ref.hpp:115-162)
// ---------------------------------------------------------------------------
__global__ void k_lists_root(const int32_t *__restrict__ child, int32_t *nb, int32_t *ni, int32_t *nb_cnt, int32_t *ni_cnt, int32_t *ap_cnt, int32_t *nbr27) {
    if (threadIdx.x != 0) return;
    for (int q = 0; q < 27; ++q) nbr27[q] = -1;
    if (child[0] < 0) {  // solver_ref.hpp:149-150
        ni[0] = 0;
        ni_cnt[0] = 1;
        nb_cnt[0] = 0;
    } else {  // solver_ref.hpp:151-152
        nb[0] = 0;
        nb_cnt[0] = 1;
        ni_cnt[0] = 0;
        nbr27[13] = 0;
    }
    ap_cnt[0] = 0;
}

// Lists of the 8 children of one internal cell per wavefront (4 per block).
// The siblings share their candidates: the children of the parent's NB
// cells, taken in the order of solver_ref.hpp:129-147 (NB entry e, child
// offset q).  Lane l tests candidate 64 r + l of round r against all 8
// siblings; each sibling's NB / NI / AP entries are appended in candidate
// order through ballots and prefix popcounts, so the lists come out in the
// reference's order.
__global__ __launch_bounds__(256) void k_lists_level(const int32_t *__restrict__ pids, int np, const int32_t *__restrict__ child,
                                                     const int32_t *__restrict__ grid, int32_t *__restrict__ nb, int32_t *__restrict__ ni,
                                                     int32_t *__restrict__ ap, int32_t *__restrict__ nb_cnt, int32_t *__restrict__ ni_cnt,
                                                     int32_t *__restrict__ ap_cnt, int32_t *__restrict__ nbr27, uint32_t *__restrict__ overflow) {
    const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wv >= np) return;  // whole wavefront
    const unsigned lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1;
    const int32_t P = pids[wv];
    const int32_t cb = child[P];
    const int npn = nb_cnt[P];
    for (int i = lane; i < 8 * 27; i += 64) nbr27[27 * static_cast<int64_t>(cb) + i] = -1;
    int gx[8], gy[8], gz[8];
    bool leaf[8];
    int nbn[8], nin[8], apn[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        gx[t] = grid[3 * (cb + t)];
        gy[t] = grid[3 * (cb + t) + 1];
        gz[t] = grid[3 * (cb + t) + 2];
        leaf[t] = child[cb + t] < 0;
        nbn[t] = nin[t] = apn[t] = 0;
    }
    const int ncand = 8 * npn;
    for (int r0 = 0; r0 < ncand; r0 += 64) {
        const int cand = r0 + static_cast<int>(lane);
        const bool valid = cand < ncand;
        int32_t cc = 0;
        int cx = 0, cy = 0, cz = 0;
        bool cint = false;
        if (valid) {
            cc = child[nb[27 * static_cast<int64_t>(P) + cand / 8]] + cand % 8;
            cx = grid[3 * cc];
            cy = grid[3 * cc + 1];
            cz = grid[3 * cc + 2];
            cint = child[cc] >= 0;
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int32_t ct = cb + t;
            const int dx = cx - gx[t], dy = cy - gy[t], dz = cz - gz[t];
            const bool adj = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;  // fmm.hpp:475-493
            const bool isnb = valid && adj && !leaf[t] && cint;
            const bool isni = valid && adj && !isnb;
            const bool isap = valid && !adj;
            const uint64_t mnb = __ballot(isnb), mni = __ballot(isni), map = __ballot(isap);
            if (isnb) {
                const int k = nbn[t] + __popcll(mnb & below);
                if (k < 27) nb[27 * static_cast<int64_t>(ct) + k] = cc;
                nbr27[27 * static_cast<int64_t>(ct) + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)] = cc;
            }
            if (isni) {
                const int k = nin[t] + __popcll(mni & below);
                if (k < 27) ni[27 * static_cast<int64_t>(ct) + k] = cc;
            }
            if (isap) {
                const int k = apn[t] + __popcll(map & below);
                if (k < 189) ap[189 * static_cast<int64_t>(ct) + k] = cc;
            }
            nbn[t] += __popcll(mnb);
            nin[t] += __popcll(mni);
            apn[t] += __popcll(map);
        }
    }
    if (lane < 8) {
        int a = 0, b = 0, d = 0;
#pragma unroll
        for (int t = 0; t < 8; ++t)
            if (lane == static_cast<unsigned>(t)) a = nbn[t], b = nin[t], d = apn[t];
        if (a > 27 || b > 27 || d > 189) atomicOr(overflow, 1u);
        nb_cnt[cb + lane] = min(a, 27);
        ni_cnt[cb + lane] = min(b, 27);
        ap_cnt[cb + lane] = min(d, 189);
    }
}

// ---------------------------------------------------------------------------
// Reference numbering (for export): children of the internal cell that
// subdivided r-th are 1 + 8r .. 8 + 8r (octree.hpp:170-177).  A cell
// subdivides when its (thr+1)-th point in caller order arrives; cells that
// subdivide on the same arrival do so parent first.  So internal cells are
// ranked by (t = (thr+1)-th smallest caller index of the cell, depth).
// S(cell) = its thr+1 smallest caller indices, merged bottom-up.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_smallest(const int32_t *__restrict__ ids, int n_ids, const int32_t *__restrict__ child, const uint32_t *__restrict__ start,
                                                  const uint32_t *__restrict__ count, const uint32_t *__restrict__ perm, const int32_t *__restrict__ int_index,
                                                  uint32_t *__restrict__ slist, uint32_t keep, uint64_t *__restrict__ rkey, uint32_t *__restrict__ rval,
                                                  const int8_t *__restrict__ level, int pow2) {
    extern __shared__ uint32_t buf[];
    const int32_t P = ids[blockIdx.x];
    const int32_t base = child[P];
    // gather candidates
    __shared__ uint32_t off[9];
    if (threadIdx.x == 0) {
        uint32_t o = 0;
        for (int q = 0; q < 8; ++q) {
            off[q] = o;
            const int32_t c = base + q;
            o += child[c] < 0 ? count[c] : keep;
        }
        off[8] = o;
    }
    __syncthreads();
    for (int q = 0; q < 8; ++q) {
        const int32_t c = base + q;
        const uint32_t n = off[q + 1] - off[q];
        const uint32_t *src = child[c] < 0 ? perm + start[c] : slist + static_cast<size_t>(int_index[c]) * keep;
        for (uint32_t i = threadIdx.x; i < n; i += 256) buf[off[q] + i] = src[i];
    }
    for (uint32_t i = off[8] + threadIdx.x; i < static_cast<uint32_t>(pow2); i += 256) buf[i] = 0xFFFFFFFFu;
    __syncthreads();
    // bitonic sort of pow2 entries
    for (int k = 2; k <= pow2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < pow2; i += 256) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t a = buf[i], b = buf[l];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        buf[i] = b;
                        buf[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    const int32_t me = int_index[P];
    for (uint32_t i = threadIdx.x; i < keep; i += 256) slist[static_cast<size_t>(me) * keep + i] = buf[i];
    if (threadIdx.x == 0) {
        rkey[me] = (static_cast<uint64_t>(buf[keep - 1]) << 5) | static_cast<uint64_t>(level[P]);
        rval[me] = static_cast<uint32_t>(me);
    }
}

__global__ __launch_bounds__(256) void k_int_index(const int32_t *__restrict__ int_ids, int64_t nint, int32_t *__restrict__ int_index) {
    const int64_t k = blockIdx.x * 256ll + threadIdx.x;
    if (k < nint) int_index[int_ids[k]] = static_cast<int32_t>(k);
}

__global__ __launch_bounds__(256) void k_rank_of(const uint32_t *__restrict__ sorted_val, int64_t nint, int32_t *__restrict__ rank) {
    const int64_t r = blockIdx.x * 256ll + threadIdx.x;
    if (r < nint) rank[sorted_val[r]] = static_cast<int32_t>(r);
}

__global__ __launch_bounds__(256) void k_ref_ids(int64_t nc, const int32_t *__restrict__ parent, const uint8_t *__restrict__ octant,
                                                 const int32_t *__restrict__ int_index, const int32_t *__restrict__ rank, int32_t *__restrict__ ref_id) {
    const int64_t c = blockIdx.x * 256ll + threadIdx.x;
    if (c >= nc) return;
    const int32_t P = parent[c];
    ref_id[c] = P < 0 ? 0 : 1 + 8 * rank[int_index[P]] + octant[c];
}

}  // namespace

// out[0..k) = base + i for the flagged i, in order; k in *nout (device)
static void compact_flags(Ctx &c, const uint8_t *flag, int64_t n, int64_t base, int32_t *out, uint32_t *nout) {
    if (n <= 8192) {
        k_compact<<<1, 1024, 0, c.stream>>>(flag, n, base, out, nout);
    } else {
        const int nb = static_cast<int>(ceil_div(n, 1024));
        c.cmp_cnt.ensure(nb);
        k_flag_counts<<<nb, 1024, 0, c.stream>>>(flag, n, c.cmp_cnt);
        k_count_scan<<<1, 1024, 0, c.stream>>>(c.cmp_cnt, nb, nout);
        k_flag_write<<<nb, 1024, 0, c.stream>>>(flag, n, base, c.cmp_cnt, out);
    }
    VG_LAUNCH_CHECK();
}

static void ensure_cells(TreeDev &t, int64_t ncell, bool keep, hipStream_t st) {
    t.start.ensure(ncell, keep, st);
    t.count.ensure(ncell, keep, st);
    t.child.ensure(ncell, keep, st);
    t.parent.ensure(ncell, keep, st);
    t.octant.ensure(ncell, keep, st);
    t.level.ensure(ncell, keep, st);
    t.bmin.ensure(3 * ncell, keep, st);
    t.bmax.ensure(3 * ncell, keep, st);
    t.grid.ensure(3 * ncell, keep, st);
    t.anc.ensure((MAXD + 1) * ncell, keep, st);
}

// Steps 2-4 of tree_build with the level sizes kept on the device: every
// level of the sorted digits is launched at once, on grids sized by bounds
// (level d holds at most min(8^d, n / (thr + 1)) internal cells, whose points
// are disjoint sets of more than thr each), and the host reads the level
// table once. Returns 1 when built, 0 when the cell capacity (twice the
// previous tree's cells) was too small, -1 when the tree is deeper than the
// sorted digits; the caller then builds level by level.
static int tree_build_lv(Ctx &c, int SH, uint64_t *&k0, uint32_t *&v0, uint64_t *&k1, uint32_t *&v1) {
    hipStream_t st = c.stream;
    TreeDev &t = c.t;
    const int64_t n = c.n;
    k_codes<<<ceil_div(n, 256), 256, 0, st>>>(c.x, n, c.box, SH, k0, v0);
    VG_LAUNCH_CHECK();
    radix_sort_pairs(st, k0, v0, k1, v1, n, 3 * (MAXD - SH), 3 * MAXD, c.rs_hist);
    const int64_t cap = std::max<int64_t>(1024, 2 * t.ncells);
    ensure_cells(t, cap, false, st);
    t.int_ids.ensure(cap, false, st);
    t.leaf_ids.ensure(cap, false, st);
    const int64_t per = n / (static_cast<int64_t>(c.thr) + 1);
    std::vector<int64_t> nflag(SH);
    int64_t fmax = cap;
    for (int d = 0; d < SH; ++d) {
        const int64_t b = d < 7 ? std::min<int64_t>(per, 1ll << (3 * d)) : per;
        nflag[d] = std::max<int64_t>(1, 8 * b);
        fmax = std::max(fmax, nflag[d]);
    }
    t.flags.ensure(ceil_div(fmax, 1024) * 1024, false, st);
    c.cmp_cnt.ensure(ceil_div(fmax, 1024));
    t.lv.ensure(LV_WORDS);
    k_root<<<1, 64, 0, st>>>(c.box, n, c.thr, t.start, t.count, t.child, t.parent, t.octant, t.level, t.bmin, t.bmax, t.grid, t.anc, t.int_ids,
                            c.scal.p);
    k_lv_init<<<1, 64, 0, st>>>(c.scal.p, t.lv);
    for (int d = 0; d < SH; ++d) {
        const unsigned gc = static_cast<unsigned>(ceil_div(nflag[d], 1024));
        k_expand_lv<<<ceil_div(nflag[d], 256), 256, 0, st>>>(d, t.lv, t.int_ids, cap, k0, c.thr, t.start, t.count, t.child, t.parent, t.octant,
                                                              t.level, t.bmin, t.bmax, t.grid, t.anc, t.flags);
        if (nflag[d] <= WIDE_FLAGS) {  // one launch instead of three
            k_compact_wide_lv<<<1, 1024, 0, st>>>(t.flags, t.lv, d, t.int_ids, cap, d == SH - 1);
        } else {
            k_flag_counts_lv<<<gc, 1024, 0, st>>>(t.flags, t.lv, d, c.cmp_cnt);
            k_count_scan_lv<<<1, 1024, 0, st>>>(c.cmp_cnt, static_cast<int>(gc), t.lv, d, cap, d == SH - 1);
            k_flag_write_lv<<<gc, 1024, 0, st>>>(t.flags, t.lv, d, c.cmp_cnt, t.int_ids, cap);
        }
    }
    VG_LAUNCH_CHECK();
    // non-empty leaves among the lv[SH + 1] cells
    const unsigned gl = static_cast<unsigned>(ceil_div(cap, 1024));
    // (most cells are leaves: the multi-block compaction writes them coalesced)
    k_leaf_flags_lv<<<ceil_div(cap, 256), 256, 0, st>>>(t.child, t.count, t.lv, SH + 1, cap, t.flags);
    k_flag_counts_lv<<<gl, 1024, 0, st>>>(t.flags, t.lv, -(SH + 2), c.cmp_cnt);
    k_count_scan_lv<<<1, 1024, 0, st>>>(c.cmp_cnt, static_cast<int>(gl), t.lv, -(SH + 2), cap, false);
    k_flag_write_lv<<<gl, 1024, 0, st>>>(t.flags, t.lv, -(SH + 2), c.cmp_cnt, t.leaf_ids, cap);
    c.perm.ensure(std::max<int64_t>(1, n));
    // a wavefront per leaf of the previous tree and a quarter more (the leaf
    // count changes little; the blocks loop over any excess)
    const int64_t gls = ceil_div(std::min<int64_t>(cap, std::max<int64_t>(1024, t.nleaves + t.nleaves / 4)), LS_WAVES);
    c.xs64.ensure(4 * std::max<int64_t>(1, n));
    c.xs32.ensure(4 * std::max<int64_t>(1, n));
    k_leaf_sort_w<<<static_cast<unsigned>(gls), 64 * LS_WAVES, LS_WAVES * c.thr * sizeof(uint32_t), st>>>(
        t.leaf_ids, 0, t.lv, t.start, t.count, c.thr, v0, c.perm, c.x, c.has_masses ? c.m.p : nullptr, c.mass, c.box, c.xs64, c.xs32);
    VG_LAUNCH_CHECK();
    double *hbox = static_cast<double *>(c.pinned);                                         // 8 doubles
    uint32_t *hlv = reinterpret_cast<uint32_t *>(static_cast<char *>(c.pinned) + 512);  // LV_WORDS pinned host words
    static_assert(512 + 4 * LV_WORDS <= Ctx::PINNED_BYTES, "pinned words");
    VG_HIP(hipMemcpyAsync(hbox, c.box, 7 * sizeof(double), hipMemcpyDeviceToHost, st));
    VG_HIP(hipMemcpyAsync(hlv, t.lv, LV_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    VG_HIP(hipStreamSynchronize(st));
    if (hbox[6] != 0) throw Error(2, "non-finite particle coordinates");
    if (hlv[LV_OVF] & 1u) return 0;
    if (hlv[LV_OVF]) return -1;
    std::vector<int64_t> lvl_off{0, 1}, int_off{0, hlv[LV_I + 1]};
    int64_t nint = hlv[LV_I + 1];
    for (int d = 0; nint > 0; ++d) {
        if (d >= SH) return -1;  // level d + 1 digits were not sorted
        lvl_off.push_back(hlv[d + 2]);
        int_off.push_back(hlv[LV_I + d + 2]);
        nint = static_cast<int64_t>(hlv[LV_I + d + 2]) - hlv[LV_I + d + 1];
    }
    for (int a = 0; a < 6; ++a) t.box[a] = hbox[a];
    t.lvl_off.swap(lvl_off);
    t.int_off.swap(int_off);
    t.depth = static_cast<int>(t.lvl_off.size()) - 2;
    t.ncells = t.lvl_off.back();
    t.nleaves = hlv[LV_NLEAF];
    t.ref_valid = false;
    return 1;
}

void tree_build(Ctx &c) {
    hipStream_t st = c.stream;
    TreeDev &t = c.t;
    const int64_t n = c.n;
    // 1. root box
    //    (the RK4 stage that wrote x may have left its partial boxes)
    c.box.ensure(8);
    c.scal.ensure(16);
    if (c.aabb_parts > 0) {
        k_aabb_final<<<1, 256, 0, st>>>(c.aabb_part, c.aabb_parts, c.box);
    } else {
        const int nb = static_cast<int>(std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256)));
        c.aabb_part.ensure(7 * static_cast<size_t>(nb));
        k_aabb<<<nb, 256, 0, st>>>(c.x, n, c.aabb_part);
        k_aabb_final<<<1, 256, 0, st>>>(c.aabb_part, nb, c.box);
    }
    c.aabb_parts = 0;
    VG_LAUNCH_CHECK();
    // 2. path codes and a stable sort of the digits of levels 1..SHALLOW
    //    (bits below the deepest level do not matter: leaves are re-sorted by
    //    caller index); a tree that needs a deeper level is rebuilt with all
    //    21 digits sorted.
    // Levels sorted: two beyond the previous tree's depth (consecutive
    // evaluations differ little), 16 for the first tree.
    const int SHALLOW = t.depth > 0 ? std::min(MAXD, t.depth + 2) : 16;
    int64_t ncells = 1;
    uint32_t *hbuf = reinterpret_cast<uint32_t *>(static_cast<char *>(c.pinned) + 64);  // pinned host words
    for (int b = 0; b < 2; ++b) {
        c.key[b].ensure(std::max<int64_t>(1, n));
        c.val[b].ensure(std::max<int64_t>(1, n));
    }
    uint64_t *k0 = c.key[0].p, *k1 = c.key[1].p;
    uint32_t *v0 = c.val[0].p, *v1 = c.val[1].p;
    int first = 0;
    if (t.depth > 0 && n > 0 && c.thr <= LS_MAXTHR) {  // a previous tree sizes the capacity and the digits sorted
        const int r = tree_build_lv(c, SHALLOW, k0, v0, k1, v1);
        if (r > 0) return;
        if (r < 0) first = 1;
    }
    for (int attempt = first; attempt < 2; ++attempt) {
    const int low_bit = attempt == 0 ? 3 * (MAXD - SHALLOW) : 0;
    if (n) {
        k_codes<<<ceil_div(n, 256), 256, 0, st>>>(c.x, n, c.box, attempt == 0 ? SHALLOW : MAXD, k0, v0);
        VG_LAUNCH_CHECK();
        radix_sort_pairs(st, k0, v0, k1, v1, n, low_bit, 3 * MAXD, c.rs_hist);
    }
    // 3. cells, level by level
    t.lvl_off.assign(1, 0);
    t.int_off.assign(1, 0);
    int64_t cap = 1024;
    ensure_cells(t, cap, false, st);
    t.int_ids.ensure(1024, false, st);
    k_root<<<1, 64, 0, st>>>(c.box, n, c.thr, t.start, t.count, t.child, t.parent, t.octant, t.level, t.bmin, t.bmax, t.grid, t.anc, t.int_ids,
                            c.scal.p);
    VG_LAUNCH_CHECK();
    double *hbox = static_cast<double *>(c.pinned);  // 8 doubles
    VG_HIP(hipMemcpyAsync(hbox, c.box, 7 * sizeof(double), hipMemcpyDeviceToHost, st));
    VG_HIP(hipMemcpyAsync(hbuf, c.scal.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    VG_HIP(hipStreamSynchronize(st));
    if (hbox[6] != 0) throw Error(2, "non-finite particle coordinates");
    for (int a = 0; a < 6; ++a) t.box[a] = hbox[a];
    int64_t nint = hbuf[0];
    ncells = 1;
    t.lvl_off.push_back(1);
    t.int_off.push_back(nint);
    int d = 0;
    bool deeper = false;
    while (nint > 0) {
        if (d >= MAXD)
            throw Error(3, "octree deeper than 21 levels: more than floor(sqrt(8) s) points closer than the root box / 2^21");
        if (low_bit > 0 && d >= SHALLOW) {  // level d + 1 digits were not sorted
            deeper = true;
            break;
        }
        const int64_t nnew = 8 * nint;
        if (ncells + nnew > cap) {
            while (cap < ncells + nnew) cap *= 2;
            ensure_cells(t, cap, true, st);
        }
        c.t.flags.ensure(nnew, false, st);
        t.int_ids.ensure(t.int_off.back() + nnew, true, st);
        k_expand<<<ceil_div(nnew, 256), 256, 0, st>>>(d, t.int_ids.p + t.int_off[d], static_cast<int>(nint), ncells, k0, c.thr, t.start, t.count,
                                                       t.child, t.parent, t.octant, t.level, t.bmin, t.bmax, t.grid, t.anc, t.flags);
        compact_flags(c, t.flags, nnew, ncells, t.int_ids.p + t.int_off[d + 1], c.scal.p);
        VG_HIP(hipMemcpyAsync(hbuf, c.scal.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        VG_HIP(hipStreamSynchronize(st));
        ncells += nnew;
        t.lvl_off.push_back(ncells);
        nint = hbuf[0];
        t.int_off.push_back(t.int_off.back() + nint);
        ++d;
    }
    if (!deeper) break;
    }
    t.depth = static_cast<int>(t.lvl_off.size()) - 2;
    t.ncells = ncells;
    // 4. non-empty leaves, leaf order by caller index, particles in sorted order
    t.flags.ensure(ncells, false, st);
    t.leaf_ids.ensure(ncells, false, st);
    k_leaf_flags<<<ceil_div(ncells, 256), 256, 0, st>>>(t.child, t.count, ncells, t.flags);
    compact_flags(c, t.flags, ncells, 0, t.leaf_ids, c.scal.p);
    VG_LAUNCH_CHECK();
    VG_HIP(hipMemcpyAsync(hbuf, c.scal.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    VG_HIP(hipStreamSynchronize(st));
    t.nleaves = hbuf[0];
    c.perm.ensure(std::max<int64_t>(1, n));
    c.xs64.ensure(4 * std::max<int64_t>(1, n));
    c.xs32.ensure(4 * std::max<int64_t>(1, n));
    if (t.nleaves && c.thr <= LS_MAXTHR) {  // sort and gather in one pass
        k_leaf_sort_w<<<static_cast<unsigned>(ceil_div(t.nleaves, LS_WAVES)), 64 * LS_WAVES, LS_WAVES * c.thr * sizeof(uint32_t), st>>>(
            t.leaf_ids, static_cast<uint32_t>(t.nleaves), nullptr, t.start, t.count, c.thr, v0, c.perm, c.x, c.has_masses ? c.m.p : nullptr,
            c.mass, c.box, c.xs64, c.xs32);
        VG_LAUNCH_CHECK();
    } else if (t.nleaves) {
        k_leaf_sort<<<static_cast<unsigned>(t.nleaves), 256, c.thr * sizeof(uint32_t), st>>>(t.leaf_ids, t.start, t.count, v0, c.perm);
        k_gather<<<ceil_div(n, 256), 256, 0, st>>>(c.perm, n, c.x, c.has_masses ? c.m.p : nullptr, c.mass, c.box, c.xs64, c.xs32, nullptr);
        VG_LAUNCH_CHECK();
    }
    t.ref_valid = false;
}

void lists_build(Ctx &c) {
    hipStream_t st = c.stream;
    TreeDev &t = c.t;
    const int64_t nc = t.ncells;
    t.nb.ensure(27 * nc);
    t.ni.ensure(27 * nc);
    t.ap.ensure(189 * nc);
    t.nb_cnt.ensure(nc);
    t.ni_cnt.ensure(nc);
    t.ap_cnt.ensure(nc);
    t.nbr27.ensure(27 * nc);
    VG_HIP(hipMemsetAsync(c.scal.p + 4, 0, sizeof(uint32_t), st));
    k_lists_root<<<1, 64, 0, st>>>(t.child, t.nb, t.ni, t.nb_cnt, t.ni_cnt, t.ap_cnt, t.nbr27);
    for (int d = 1; d <= t.depth; ++d) {  // children at level d of the internal cells of level d - 1
        const int64_t a = t.int_off[d - 1], b = t.int_off[d];
        if (b > a)
            k_lists_level<<<ceil_div(b - a, 4), 256, 0, st>>>(t.int_ids.p + a, static_cast<int>(b - a), t.child, t.grid, t.nb, t.ni, t.ap, t.nb_cnt,
                                                              t.ni_cnt, t.ap_cnt, t.nbr27, c.scal.p + 4);
    }
    VG_LAUNCH_CHECK();
}

void ref_numbering(Ctx &c) {
    hipStream_t st = c.stream;
    TreeDev &t = c.t;
    const int64_t nc = t.ncells;
    const int64_t nint = t.int_off.back();
    t.ref_id.ensure(std::max<int64_t>(1, nc));
    if (nint == 0) {
        k_ref_ids<<<ceil_div(nc, 256), 256, 0, st>>>(nc, t.parent, t.octant, nullptr, nullptr, t.ref_id);
        VG_LAUNCH_CHECK();
        t.ref_valid = true;
        return;
    }
    const uint32_t keep = c.thr + 1;
    int pow2 = 1;
    while (static_cast<uint32_t>(pow2) < 8 * keep) pow2 <<= 1;
    if (pow2 * sizeof(uint32_t) > 64 * 1024) throw Error(1, "export of the reference numbering supports floor(sqrt(8) s) < 2048");
    DevBuf<int32_t> int_index, rank;
    int_index.ensure(nc);
    rank.ensure(nint);
    c.slist.ensure(static_cast<size_t>(nint) * keep);
    for (int b = 0; b < 2; ++b) {
        c.rkey[b].ensure(nint);
        c.rval[b].ensure(nint);
    }
    k_int_index<<<ceil_div(nint, 256), 256, 0, st>>>(t.int_ids, nint, int_index);
    for (int d = t.depth; d >= 0; --d) {
        const int64_t a = t.int_off[d], b = t.int_off[d + 1];
        if (b > a)
            k_smallest<<<static_cast<unsigned>(b - a), 256, pow2 * sizeof(uint32_t), st>>>(t.int_ids.p + a, static_cast<int>(b - a), t.child, t.start,
                                                                                            t.count, c.perm, int_index, c.slist, keep, c.rkey[0], c.rval[0],
                                                                                            t.level, pow2);
    }
    VG_LAUNCH_CHECK();
    uint64_t *k0 = c.rkey[0].p, *k1 = c.rkey[1].p;
    uint32_t *v0 = c.rval[0].p, *v1 = c.rval[1].p;
    radix_sort_pairs(st, k0, v0, k1, v1, nint, 0, 64, c.rs_hist);
    k_rank_of<<<ceil_div(nint, 256), 256, 0, st>>>(v0, nint, rank);
    k_ref_ids<<<ceil_div(nc, 256), 256, 0, st>>>(nc, t.parent, t.octant, int_index, rank, t.ref_id);
    VG_LAUNCH_CHECK();
    VG_HIP(hipStreamSynchronize(st));
    t.ref_valid = true;
}

}  // namespace vg
