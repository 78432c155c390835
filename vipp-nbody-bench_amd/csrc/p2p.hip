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
// One 256-thread block per target leaf: the ranges are flattened in LDS and
// streamed in tiles of CH source particles; each target is shared by
// 256/ntargets threads that split the tile, then reduce.  With a nonzero
// softening the i == j term is exactly zero (dx = dy = dz = 0), so it needs no
// branch; without softening a select removes it.
#include "gravity.hpp"

namespace vg {
namespace {

constexpr int P2P_THREADS = 256;
constexpr int MAXE = 27 * (MAXD + 1);

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
__device__ inline double rsq(double x) { return 1.0 / sqrt(x); }

template <class R, bool SELF_GUARD, bool POT>
__global__ __launch_bounds__(P2P_THREADS) void k_p2p(const int32_t *__restrict__ leaves, const int8_t *__restrict__ level, const uint32_t *__restrict__ start,
                                                     const uint32_t *__restrict__ count, const int32_t *__restrict__ anc, const int32_t *__restrict__ ni,
                                                     const int32_t *__restrict__ ni_cnt, const typename Vec4<R>::T *__restrict__ xs, R eps2, R G,
                                                     typename Vec4<R>::T *__restrict__ acc) {
    using V4 = typename Vec4<R>::T;
    constexpr int CH = sizeof(R) == 4 ? 1024 : 512;
    __shared__ V4 tile[CH];
    __shared__ uint32_t ent_start[MAXE];
    __shared__ uint32_t ent_off[MAXE + 1];
    __shared__ int32_t an[MAXD + 1];
    __shared__ int32_t an_off[MAXD + 2];
    __shared__ V4 red[P2P_THREADS];

    const int32_t c = leaves[blockIdx.x];
    const int lev = level[c];
    const uint32_t nt = count[c], pt0 = start[c];
    const int tid = threadIdx.x;
    if (tid <= lev) an[tid] = anc[static_cast<int64_t>(c) * (MAXD + 1) + tid];
    __syncthreads();
    if (tid == 0) {
        int o = 0;
        for (int l = 0; l <= lev; ++l) {
            an_off[l] = o;
            o += ni_cnt[an[l]];
        }
        an_off[lev + 1] = o;
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
    if (tid == 0) {
        uint32_t o = 0;
        ent_off[0] = 0;
        for (int e = 0; e < E; ++e) {
            o += ent_off[e + 1];
            ent_off[e + 1] = o;
        }
    }
    __syncthreads();
    const uint32_t total = ent_off[E];
    const int S = nt ? max(1, min(16, P2P_THREADS / static_cast<int>(nt))) : 1;
    const int ti = tid % max(1u, nt), sl = tid / max(1u, nt);
    const bool active = nt && sl < S;
    V4 xi{};
    if (active) xi = xs[pt0 + ti];
    R ax = 0, ay = 0, az = 0, ph = 0;
    for (uint32_t c0 = 0; c0 < total; c0 += CH) {
        const uint32_t len = min(static_cast<uint32_t>(CH), total - c0);
        for (uint32_t f = tid; f < len; f += P2P_THREADS) {
            const uint32_t g = c0 + f;
            int lo = 0, hi = E;  // last e with ent_off[e] <= g
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (ent_off[mid] <= g)
                    lo = mid;
                else
                    hi = mid;
            }
            tile[f] = xs[ent_start[lo] + (g - ent_off[lo])];
        }
        __syncthreads();
        if (active) {
            for (uint32_t j = sl; j < len; j += S) {
                const V4 sj = tile[j];
                const R dx = sj.x - xi.x, dy = sj.y - xi.y, dz = sj.z - xi.z;
                const R r2 = dx * dx + dy * dy + dz * dz + eps2;
                R rinv = rsq(r2);
                if (SELF_GUARD) rinv = r2 > R(0) ? rinv : R(0);
                const R mr = sj.w * rinv;
                const R mr3 = mr * rinv * rinv;
                ax += mr3 * dx;
                ay += mr3 * dy;
                az += mr3 * dz;
                if (POT) ph += mr;
            }
        }
        __syncthreads();
    }
    red[tid] = V4{ax, ay, az, ph};
    __syncthreads();
    if (active && sl == 0) {
        R sx = 0, sy = 0, sz = 0, sp = 0;
        for (int q = 0; q < S; ++q) {
            const V4 v = red[ti + q * nt];
            sx += v.x;
            sy += v.y;
            sz += v.z;
            sp += v.w;
        }
        // the self term of the potential (m_i / eps) is removed here
        if (POT && !SELF_GUARD) sp -= xi.w * rsq(eps2);
        acc[pt0 + ti] = V4{G * sx, G * sy, G * sz, sp};
    }
}

template <class R>
void p2p_t(Ctx &c, bool potential) {
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
    const dim3 grid(static_cast<unsigned>(t.nleaves));
#define VG_P2P(G_, P_) \
    k_p2p<R, G_, P_><<<grid, P2P_THREADS, 0, c.stream>>>(t.leaf_ids, t.level, t.start, t.count, t.anc, t.ni, t.ni_cnt, xs, eps2, static_cast<R>(c.G), acc)
    if (guard) {
        if (potential)
            VG_P2P(true, true);
        else
            VG_P2P(true, false);
    } else {
        if (potential)
            VG_P2P(false, true);
        else
            VG_P2P(false, false);
    }
#undef VG_P2P
    VG_LAUNCH_CHECK();
}

}  // namespace

void fmm_p2p(Ctx &c, bool potential) {
    if (c.prec == 0)
        p2p_t<double>(c, potential);
    else
        p2p_t<float>(c, potential);
}

}  // namespace vg
