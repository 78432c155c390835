// p2p.hip - near field (P2P) of every non-empty leaf.
//
// The reference sums, for each point i of a leaf, over every cell of the
// leaf's NI list and over all points of that cell's subtree, skipping i
// itself (fmm.hpp:377-385,513-540, solver_ref.hpp:209-210).  The NI list of a
// cell is its parent's NI list followed by its own new entries
// (solver_ref.hpp:126), so a leaf's list is the concatenation of the local
// entries of its ancestors and itself.  Every subtree is one contiguous range
// of the sorted particles, so each entry is a (start, count) range.
//
// Work items: a plan kernel counts each non-empty leaf's sources; every leaf
// is cut into chunks of at most P2P_CHUNK sources (fewer when the leaves are
// too few to fill the device) and each (leaf, chunk) pair is one 256-thread
// block, so leaves with long inherited lists spread over many blocks.  A
// block flattens its ranges in LDS, streams the chunk's sources through LDS
// tiles and gives each lane T = 1..6 targets, in groups of 8 to 64 lanes, by
// leaf-size class (the fewest idle lanes).
// A leaf with one chunk writes its result; split leaves write partial sums,
// and the work item that finishes last adds them in chunk order
// (deterministic).  With a nonzero
// softening the i == j term is exactly zero (dx = dy = dz = 0), so it needs no
// branch; without softening a select removes it.
#include "gravity.hpp"

namespace vg {
namespace {

constexpr int P2P_THREADS = 256;
constexpr int MAXE = 27 * (MAXD + 1);
constexpr uint32_t P2P_CHUNK = 16384;  // sources per work item (at most; 8192 and 32768 measured slower on the 4M Plummer input)

// Lane layout of a leaf with nt targets and T targets per lane: groups of
// TPW lanes, the smallest TPW in {8, 16, 32, 64} with TPW T >= nt (0: none).
__host__ __device__ inline int tpw_for(uint32_t nt, int T) {
    for (int tp = 8; tp <= 64; tp *= 2)
        if (static_cast<uint32_t>(tp * T) >= nt) return tp;
    return 0;
}
constexpr int TMAX = 6;  // targets per lane of the largest regular class

template <class R>
struct Vec4;
template <>
struct Vec4<float> {
    using T = float4;
};
template <>
struct Vec4<double> {
    using T = double4;
};

__device__ inline float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
// 1/sqrt(x) in FP64: the hardware estimate refined by two Newton steps
// y <- y + y (1 - x y^2) / 2, each squaring the relative error, so the result
// is within a few ulps of the correctly rounded value (x > 0; x = 0 gives NaN,
// which the self-pair select replaces)
__device__ inline double rsq(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double e = fma(-x, y * y, 1.0);
        y = fma(y * e, 0.5, y);
    }
    return y;
}

// Sources [src0, src1) of the leaf's flattened list against targets pt0 ..
// pt0+nt-1.  out == nullptr: write the final G-scaled result to acc; else
// write the raw partial sums to out[t].  UM (every particle has mass mu): the
// pair terms leave the mass out and the sums are scaled by mu once.
template <class R, bool SELF_GUARD, bool POT, bool UM, int T, class V4>
__device__ __forceinline__ void p2p_leaf(V4 *tile, V4 *red, const uint32_t *ent_start, const uint32_t *ent_off, int E, uint32_t src0, uint32_t src1,
                                         uint32_t nt, uint32_t pt0, const V4 *__restrict__ xs, R eps2, R G, R mu, V4 *__restrict__ acc,
                                         V4 *__restrict__ out) {
    constexpr int CH = sizeof(R) == 4 ? 1024 : 512;
    const int tid = threadIdx.x;
    const int TPW = tpw_for(nt, T);  // nt <= 64 T here
    const int lane = tid & 63, wave = tid >> 6;
    const int groups = 64 / TPW;
    const int nsl = 4 * groups;
    const int sl = wave * groups + lane / TPW;
    const int ti = lane % TPW;
    V4 xi[T];
    R ax[T], ay[T], az[T], ph[T];
#pragma unroll
    for (int q = 0; q < T; ++q) {
        const uint32_t t = ti + q * TPW;
        xi[q] = t < nt ? xs[pt0 + t] : V4{R(1e30), R(1e30), R(1e30), R(0)};
        ax[q] = ay[q] = az[q] = ph[q] = 0;
    }
    for (uint32_t c0 = src0; c0 < src1; c0 += CH) {
        const uint32_t len = min(static_cast<uint32_t>(CH), src1 - c0);
#pragma unroll
        for (int u = 0; u < CH / P2P_THREADS; ++u) {
            const uint32_t f = tid + u * P2P_THREADS;
            const uint32_t g = c0 + min(f, len - 1);
            int lo = 0;  // last e with ent_off[e] <= g (fixed trip count: E <= 1023)
#pragma unroll
            for (int step = 512; step > 0; step >>= 1)
                if (lo + step <= E && ent_off[lo + step] <= g) lo += step;
            if (f < len) tile[f] = xs[ent_start[lo] + (g - ent_off[lo])];
        }
        __syncthreads();
        // slice sl takes the contiguous range [b0, b1) of this tile
        const uint32_t per = (len + nsl - 1) / nsl;
        const uint32_t b0 = min(len, sl * per), b1 = min(len, b0 + per);
        uint32_t j = b0;
        for (; j + 4 <= b1; j += 4) {
            V4 sj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) sj[u] = tile[j + u];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int q = 0; q < T; ++q) {
                    const R dx = sj[u].x - xi[q].x, dy = sj[u].y - xi[q].y, dz = sj[u].z - xi[q].z;
                    const R r2 = eps2 + dx * dx + dy * dy + dz * dz;
                    R rinv = rsq(r2);
                    if (SELF_GUARD) rinv = r2 > R(0) ? rinv : R(0);
                    const R mr = UM ? rinv : sj[u].w * rinv;
                    const R mr3 = mr * rinv * rinv;
                    ax[q] += mr3 * dx;
                    ay[q] += mr3 * dy;
                    az[q] += mr3 * dz;
                    if (POT) ph[q] += mr;
                }
        }
        for (; j < b1; ++j) {
            const V4 sj = tile[j];
#pragma unroll
            for (int q = 0; q < T; ++q) {
                const R dx = sj.x - xi[q].x, dy = sj.y - xi[q].y, dz = sj.z - xi[q].z;
                const R r2 = eps2 + dx * dx + dy * dy + dz * dz;
                R rinv = rsq(r2);
                if (SELF_GUARD) rinv = r2 > R(0) ? rinv : R(0);
                const R mr = UM ? rinv : sj.w * rinv;
                const R mr3 = mr * rinv * rinv;
                ax[q] += mr3 * dx;
                ay[q] += mr3 * dy;
                az[q] += mr3 * dz;
                if (POT) ph[q] += mr;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < T; ++q) red[q * P2P_THREADS + tid] = V4{ax[q], ay[q], az[q], ph[q]};
    __syncthreads();
    for (uint32_t t = tid; t < nt; t += P2P_THREADS) {
        // target t lives in lane-slot (q, ti) of every slice
        const int q = static_cast<int>(t / TPW), tl = static_cast<int>(t % TPW);
        R sx = 0, sy = 0, sz = 0, sp = 0;
        for (int s2 = 0; s2 < nsl; ++s2) {
            const int owner = (s2 / groups) * 64 + (s2 % groups) * TPW + tl;
            const V4 v = red[q * P2P_THREADS + owner];
            sx += v.x;
            sy += v.y;
            sz += v.z;
            sp += v.w;
        }
        if (UM) {
            sx *= mu;
            sy *= mu;
            sz *= mu;
            sp *= mu;
        }
        if (out) {
            out[t] = V4{sx, sy, sz, sp};
        } else {
            if (POT && !SELF_GUARD) sp -= xs[pt0 + t].w * rsq(eps2);  // drop the m_i / eps self term
            acc[pt0 + t] = V4{G * sx, G * sy, G * sz, sp};
        }
    }
}

// Lane layout: a wavefront holds 64/TPW groups of TPW lanes; the lanes of a
// group take targets ti, ti+TPW, ... (up to TMAX per lane) and read the same
// source at each step (an LDS broadcast).  The block's 4 x (64/TPW) groups
// split every staged tile of sources.
template <class R, bool SELF_GUARD, bool POT, bool UM, int T>
__global__ __launch_bounds__(P2P_THREADS) void k_p2p(const int2 *__restrict__ items, const int32_t *__restrict__ leaves, const uint32_t *__restrict__ nchunk,
                                                     const uint32_t *__restrict__ item_off, const int8_t *__restrict__ level,
                                                     const uint32_t *__restrict__ start, const uint32_t *__restrict__ count, const int32_t *__restrict__ anc,
                                                     const int32_t *__restrict__ ni, const int32_t *__restrict__ ni_cnt,
                                                     const typename Vec4<R>::T *__restrict__ xs, R eps2, R G, R mu, typename Vec4<R>::T *__restrict__ acc,
                                                     typename Vec4<R>::T *__restrict__ partial, uint32_t pstride, uint32_t *__restrict__ ticket,
                                                     uint32_t chunk) {
    using V4 = typename Vec4<R>::T;
    constexpr int CH = sizeof(R) == 4 ? 1024 : 512;
    constexpr int TR = T <= TMAX ? T : 4;  // T = TMAX + 1: leaves in batches of 256 targets, 4 per lane
    // source tiles while streaming, per-lane partial sums after (a barrier apart)
    __shared__ V4 buf[CH > P2P_THREADS * TR ? CH : P2P_THREADS * TR];
    V4 *tile = buf;
    __shared__ uint32_t ent_start[MAXE];
    __shared__ uint32_t ent_off[MAXE + 1];
    __shared__ int32_t an[MAXD + 1];
    __shared__ int32_t an_off[MAXD + 2];
    V4 *red = buf;

    const int2 item = items[blockIdx.x];  // (leaf index, chunk)
    const int32_t c = leaves[item.x];
    const int lev = level[c];
    const uint32_t nt = count[c], pt0 = start[c];
    const int tid = threadIdx.x;
    if (tid <= lev) an[tid] = anc[static_cast<int64_t>(c) * (MAXD + 1) + tid];
    __syncthreads();
    if (tid < 64) {  // entry counts per ancestor, scanned by one wavefront
        const int cntl = tid <= lev ? ni_cnt[an[tid]] : 0;
        int incl = cntl;
        for (int d = 1; d < 32; d <<= 1) {
            const int o = __shfl_up(incl, d, 64);
            if ((tid & 63) >= d) incl += o;
        }
        if (tid <= lev) an_off[tid] = incl - cntl;
        if (tid == lev) an_off[lev + 1] = incl;
    }
    __syncthreads();
    const int E = an_off[lev + 1];
    for (int e = tid; e < E; e += P2P_THREADS) {
        int l = 0;
        while (an_off[l + 1] <= e) ++l;
        const int32_t cell = ni[27 * static_cast<int64_t>(an[l]) + (e - an_off[l])];
        ent_start[e] = start[cell];
        ent_off[e + 1] = count[cell];
    }
    __syncthreads();
    if (tid < 64) {  // prefix sum of the entry sizes
        uint32_t carry = 0;
        for (int b = 0; b < E; b += 64) {
            const uint32_t v = (b + tid < E) ? ent_off[b + tid + 1] : 0u;
            uint32_t incl = v;
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t o = __shfl_up(incl, d, 64);
                if (tid >= d) incl += o;
            }
            if (b + tid < E) ent_off[b + tid + 1] = carry + incl;
            carry += __shfl(incl, 63, 64);
        }
        if (tid == 0) ent_off[0] = 0;
    }
    __syncthreads();
    const uint32_t total = ent_off[E];
    const uint32_t src0 = static_cast<uint32_t>(item.y) * chunk, src1 = min(total, src0 + chunk);
    V4 *out = nchunk[item.x] > 1 ? partial + static_cast<size_t>(item_off[item.x] + item.y) * pstride : nullptr;
    if constexpr (T <= TMAX) {
        p2p_leaf<R, SELF_GUARD, POT, UM, T>(tile, red, ent_start, ent_off, E, src0, src1, nt, pt0, xs, eps2, G, mu, acc, out);
    } else {  // leaves above 64 TMAX points: batches of 256 targets, each streaming the sources again
        for (uint32_t tb = 0; tb < nt; tb += 256) {
            p2p_leaf<R, SELF_GUARD, POT, UM, 4>(tile, red, ent_start, ent_off, E, src0, src1, min(nt - tb, 256u), pt0 + tb, xs, eps2, G, mu, acc,
                                                out ? out + tb : nullptr);
            __syncthreads();
        }
    }
    if (out) {
        // A leaf split over several work items: the block that finishes last
        // adds the partial sums in chunk order (the same order every run).
        __shared__ uint32_t s_last;
        const uint32_t nch = nchunk[item.x];
        __threadfence();  // this block's partial sums are visible device-wide
        __syncthreads();
        if (tid == 0) s_last = atomicAdd(&ticket[item.x], 1u) == nch - 1 ? 1u : 0u;
        __syncthreads();
        if (s_last) {
            __threadfence();  // and the other blocks' ones to this block
            const V4 *pk = partial + static_cast<size_t>(item_off[item.x]) * pstride;
            for (uint32_t t = tid; t < nt; t += P2P_THREADS) {
                R sx = 0, sy = 0, sz = 0, sp = 0;
                for (uint32_t q = 0; q < nch; ++q) {
                    const V4 v = pk[static_cast<size_t>(q) * pstride + t];
                    sx += v.x;
                    sy += v.y;
                    sz += v.z;
                    sp += v.w;
                }
                if (POT && !SELF_GUARD) sp -= xs[pt0 + t].w * rsq(eps2);
                acc[pt0 + t] = V4{G * sx, G * sy, G * sz, sp};
            }
        }
    }
}

// Per leaf: number of sources in its NI chain, split into chunks of at most
// chunk sources (bounded work per block), and its class by targets per lane.
// Sources of each cell's own NI entries (the leaf's full list is its
// ancestors' entries and its own, solver_ref.hpp:126).
__global__ __launch_bounds__(256) void k_ni_local(int64_t ncells, const uint32_t *__restrict__ count, const int32_t *__restrict__ ni,
                                                  const int32_t *__restrict__ ni_cnt, uint32_t *__restrict__ local) {
    const int64_t c = blockIdx.x * 256ll + threadIdx.x;
    if (c >= ncells) return;
    const int ne = ni_cnt[c];
    uint32_t s = 0;
    for (int e = 0; e < ne; ++e) s += count[ni[27 * c + e]];
    local[c] = s;
}

// Leaf classes: class T - 1 (T = 1 .. TMAX) takes T targets per lane in groups
// of tpw_for(nt, T) lanes, T chosen for the fewest lane slots TPW T >= nt
// (ties: fewer targets per lane); class TMAX: more than 64 TMAX targets, in
// batches of 256.
constexpr int NCLS = TMAX + 1;
constexpr int NSC = NCLS + 2;  // scanned per leaf: chunks, chunks of each class, split flag
__device__ inline int leaf_class(uint32_t nt) {
    if (nt > 64u * TMAX) return TMAX;
    int best = 0;
    uint32_t slots = ~0u;
    for (int T = 1; T <= TMAX; ++T) {
        const int tp = tpw_for(nt, T);
        if (tp && static_cast<uint32_t>(tp * T) < slots) {
            slots = tp * T;
            best = T - 1;
        }
    }
    return best;
}

// Per leaf: source count -> chunks and class; block-local exclusive scans of
// the NSC components (one leaf per thread, 256 per block) and block totals.
__global__ __launch_bounds__(256) void k_p2p_plan(const int32_t *__restrict__ leaves, int nleaves, const int8_t *__restrict__ level,
                                                  const uint32_t *__restrict__ count, const int32_t *__restrict__ anc, const uint32_t *__restrict__ local,
                                                  uint32_t *__restrict__ nchunk, uint32_t *__restrict__ pre, uint32_t *__restrict__ btot,
                                                  uint32_t chunk) {
    __shared__ uint32_t wtot[4][NSC];
    const int k = blockIdx.x * 256 + threadIdx.x;
    const unsigned lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t v[NSC] = {};
    if (k < nleaves) {
        const int32_t c = leaves[k];
        uint32_t src = 0;
        for (int l = 0; l <= level[c]; ++l) src += local[anc[static_cast<int64_t>(c) * (MAXD + 1) + l]];
        const uint32_t nch = max(1u, (src + chunk - 1) / chunk);
        nchunk[k] = nch;
        v[0] = nch;
        v[1 + leaf_class(count[c])] = nch;
        v[NSC - 1] = nch > 1 ? 1u : 0u;
    }
    uint32_t inc[NSC];
#pragma unroll
    for (int q = 0; q < NSC; ++q) {
        uint32_t x = v[q];
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(x, d, 64);
            if (lane >= static_cast<unsigned>(d)) x += o;
        }
        inc[q] = x;
        if (lane == 63) wtot[w][q] = x;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NSC; ++q) {
        uint32_t base = 0;
        for (unsigned u = 0; u < w; ++u) base += wtot[u][q];
        if (k < nleaves) pre[static_cast<int64_t>(k) * NSC + q] = base + inc[q] - v[q];
        if (threadIdx.x == 255) btot[blockIdx.x * NSC + q] = base + inc[q];
    }
}

// One block: exclusive scan of the block totals; writes the class counts and
// bases, the item total and the split-leaf total into scal:
//   scal[0..NCLS) items per class, scal[NCLS..2 NCLS) class bases,
//   scal[2 NCLS] items, scal[2 NCLS + 1] split leaves.
__global__ __launch_bounds__(256) void k_p2p_plan_blocks(uint32_t *__restrict__ btot, int nblocks, uint32_t *__restrict__ scal) {
    // exclusive scan of each component over the plan blocks, one wavefront per
    // component (wavefront-wide inclusive scans, carried across runs of 64)
    __shared__ uint32_t tot[NSC];
    const unsigned lane = threadIdx.x & 63;
    for (int q = threadIdx.x >> 6; q < NSC; q += 4) {
        uint32_t carry = 0;
        for (int b0 = 0; b0 < nblocks; b0 += 64) {
            const int b = b0 + static_cast<int>(lane);
            const uint32_t v = b < nblocks ? btot[b * NSC + q] : 0u;
            uint32_t x = v;
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t o = __shfl_up(x, d, 64);
                if (lane >= static_cast<unsigned>(d)) x += o;
            }
            if (b < nblocks) btot[b * NSC + q] = carry + x - v;
            carry += __shfl(x, 63, 64);
        }
        if (lane == 0) tot[q] = carry;
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // scal: [class items] [class bases] total items, split leaves
        scal[2 * NCLS] = tot[0];
        scal[2 * NCLS + 1] = tot[NSC - 1];
        uint32_t base = 0;
        for (int j = 0; j < NCLS; ++j) {
            scal[j] = tot[1 + j];
            scal[NCLS + j] = base;
            base += tot[1 + j];
        }
    }
}

// Item offsets, items (leaf, chunk) bucketed by class in leaf order, and the
// list of leaves whose sources span several chunks.
__global__ __launch_bounds__(256) void k_p2p_items(const int32_t *__restrict__ leaves, int nleaves, const uint32_t *__restrict__ count,
                                                   const uint32_t *__restrict__ nchunk, const uint32_t *__restrict__ pre,
                                                   const uint32_t *__restrict__ boff, const uint32_t *__restrict__ scal,
                                                   uint32_t *__restrict__ item_off, int2 *__restrict__ items) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nleaves) return;
    const uint32_t *pk = pre + static_cast<int64_t>(k) * NSC;
    const uint32_t *bk = boff + blockIdx.x * NSC;
    const uint32_t nch = nchunk[k];
    item_off[k] = bk[0] + pk[0];
    const int cls = leaf_class(count[leaves[k]]);
    const uint32_t pos = scal[NCLS + cls] + bk[1 + cls] + pk[1 + cls];
    for (uint32_t q = 0; q < nch; ++q) items[pos + q] = make_int2(k, static_cast<int>(q));
}

// Work items of the P2P: counts per leaf, then the plan scalars (items per
// class) copied back to pinned host words, with ev_plan recorded after the
// copy.  Nothing here waits, so the caller can enqueue other work before
// p2p_run needs the scalars.
void p2p_plan(Ctx &c) {
    TreeDev &t = c.t;
    if (!t.nleaves) return;
    const int nl = static_cast<int>(t.nleaves);
    const int nb = static_cast<int>(ceil_div(nl, 256));
    c.p2p_scal.ensure(16);
    c.p2p_nchunk.ensure(nl);
    c.p2p_item_off.ensure(nl);
    c.p2p_local.ensure(t.ncells);
    c.p2p_pre.ensure(static_cast<size_t>(nl) * NSC);
    c.p2p_btot.ensure(static_cast<size_t>(nb) * NSC);
    c.p2p_ticket.ensure(nl);
    k_ni_local<<<ceil_div(t.ncells, 256), 256, 0, c.stream>>>(t.ncells, t.count, t.ni, t.ni_cnt, c.p2p_local);
    // under 8 leaves per CU (the 16k Plummer sphere has 430): chunks of 8192
    // sources scaled down by the leaf shortfall (at least 1024), so that the
    // work items still give every CU several blocks
    uint32_t chunk = P2P_CHUNK;
    if (c.ncu > 0 && nl < 8 * c.ncu)
        chunk = std::max<uint32_t>(1024, static_cast<uint32_t>(8192ull * nl / (8 * c.ncu)) / 256 * 256);
    c.p2p_chunk = chunk;
    k_p2p_plan<<<nb, 256, 0, c.stream>>>(t.leaf_ids, nl, t.level, t.count, t.anc, c.p2p_local, c.p2p_nchunk, c.p2p_pre, c.p2p_btot, chunk);
    k_p2p_plan_blocks<<<1, 256, 0, c.stream>>>(c.p2p_btot, nb, c.p2p_scal);
    VG_LAUNCH_CHECK();
    static_assert(2 * NCLS + 2 <= 16, "plan scalars");
    uint32_t *hs = reinterpret_cast<uint32_t *>(static_cast<char *>(c.pinned) + 128);  // 16 pinned host words
    VG_HIP(hipMemcpyAsync(hs, c.p2p_scal.p, 16 * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipEventRecord(c.ev_plan, c.stream));
}

// The P2P itself, on the stream p2p_plan used: waits on the host for the plan
// scalars, then launches one kernel per leaf class.
template <class R>
void p2p_run(Ctx &c, bool potential) {
    TreeDev &t = c.t;
    if (!t.nleaves) return;
    using V4 = typename Vec4<R>::T;
    const V4 *xs = reinterpret_cast<const V4 *>(sizeof(R) == 4 ? static_cast<const void *>(c.xs32.p) : static_cast<const void *>(c.xs64.p));
    V4 *acc;
    if (sizeof(R) == 4) {
        c.as32.ensure(4 * std::max<int64_t>(1, c.n));
        acc = reinterpret_cast<V4 *>(c.as32.p);
    } else {
        c.as64.ensure(4 * std::max<int64_t>(1, c.n));
        acc = reinterpret_cast<V4 *>(c.as64.p);
    }
    const R eps2 = static_cast<R>(c.eps * c.eps);
    const bool guard = !(eps2 > 0);
    // one mass for every particle (no per-particle masses uploaded): force
    // kernels take it out of the pair terms (the potential keeps the general path)
    const bool um = !c.has_masses;
    const R mu = static_cast<R>(c.mass);
    const int nl = static_cast<int>(t.nleaves);
    const int nb = static_cast<int>(ceil_div(nl, 256));
    const uint32_t *hs = reinterpret_cast<const uint32_t *>(static_cast<const char *>(c.pinned) + 128);
    VG_HIP(hipEventSynchronize(c.ev_plan));
    const uint32_t nitems = hs[2 * NCLS], nsplit = hs[2 * NCLS + 1];
    const uint32_t *hbase = hs + NCLS;
    c.p2p_items.ensure(std::max<uint32_t>(1, nitems));
    int2 *items = reinterpret_cast<int2 *>(c.p2p_items.p);
    k_p2p_items<<<nb, 256, 0, c.stream>>>(t.leaf_ids, nl, t.count, c.p2p_nchunk, c.p2p_pre, c.p2p_btot, c.p2p_scal, c.p2p_item_off, items);
    VG_LAUNCH_CHECK();
    const bool split = nsplit > 0;
    const uint32_t pstride = c.thr;
    V4 *partial = nullptr;
    if (split) {
        VG_HIP(hipMemsetAsync(c.p2p_ticket.p, 0, static_cast<size_t>(nl) * sizeof(uint32_t), c.stream));
        if (sizeof(R) == 4) {
            c.p2p_part32.ensure(4 * static_cast<size_t>(nitems) * pstride);
            partial = reinterpret_cast<V4 *>(c.p2p_part32.p);
        } else {
            c.p2p_part64.ensure(4 * static_cast<size_t>(nitems) * pstride);
            partial = reinterpret_cast<V4 *>(c.p2p_part64.p);
        }
    }
    for (int cls = 0; cls < NCLS; ++cls) {
        if (!hs[cls]) continue;
        const dim3 grid(hs[cls]);
        const int2 *it = items + hbase[cls];
#define VG_P2P_T(G_, P_, U_, T_)                                                                                                                \
    k_p2p<R, G_, P_, U_, T_><<<grid, P2P_THREADS, 0, c.stream>>>(it, t.leaf_ids, c.p2p_nchunk, c.p2p_item_off, t.level, t.start, t.count, t.anc, \
                                                                 t.ni, t.ni_cnt, xs, eps2, static_cast<R>(c.G), mu, acc, partial, pstride, c.p2p_ticket, \
                                                                 c.p2p_chunk)
#define VG_P2P(G_, P_, U_)           \
    do {                             \
        if (cls == 0)                \
            VG_P2P_T(G_, P_, U_, 1); \
        else if (cls == 1)           \
            VG_P2P_T(G_, P_, U_, 2); \
        else if (cls == 2)           \
            VG_P2P_T(G_, P_, U_, 3); \
        else if (cls == 3)           \
            VG_P2P_T(G_, P_, U_, 4); \
        else if (cls == 4)           \
            VG_P2P_T(G_, P_, U_, 5); \
        else if (cls == 5)           \
            VG_P2P_T(G_, P_, U_, 6); \
        else /* batches */           \
            VG_P2P_T(G_, P_, U_, 7); \
    } while (0)
        if (guard) {
            if (potential)
                VG_P2P(true, true, false);
            else if (um)
                VG_P2P(true, false, true);
            else
                VG_P2P(true, false, false);
        } else {
            if (potential)
                VG_P2P(false, true, false);
            else if (um)
                VG_P2P(false, false, true);
            else
                VG_P2P(false, false, false);
        }
#undef VG_P2P
#undef VG_P2P_T
    }
    VG_LAUNCH_CHECK();
}

}  // namespace

void fmm_p2p_plan(Ctx &c) { p2p_plan(c); }

void fmm_p2p_run(Ctx &c, bool potential) {
    if (c.prec == 0)
        p2p_run<double>(c, potential);
    else
        p2p_run<float>(c, potential);
}

void fmm_p2p(Ctx &c, bool potential) {
    fmm_p2p_plan(c);
    fmm_p2p_run(c, potential);
}

}  // namespace vg
