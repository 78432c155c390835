// closure.hip - the learned closure inside the force evaluation.
//
// A low-order FMM (order p, normally 4) evaluates every leaf's local
// expansion L.  The network maps the leaf's first 25 scaled coefficients
// (orders <= 4), normalised to unit length, plus three geometry features, to
// the residual of those coefficients against an order-p_hi evaluation
// (normally 12), in the same normalisation:
//   f = [L~_25 / |L~_25|, level / 16, 2^-level, |c - c_root| / W_root]
//   L~_25 += |L~_25| * s_out * MLP((f - shift) * s_in)
// before L2P (shift, s_in, s_out: the standardisation the network was
// trained with).  |s_out * MLP(.)|, the predicted relative correction, is the
// per-leaf error estimate: leaves where it exceeds the bound take the
// order-p_hi far field instead.  That fallback runs a complete order-p_hi
// upward pass (any cell can be a source) and M2L / L2L only along the
// ancestor chains of the flagged leaves, then their L2P at order p_hi; the
// other leaves keep the corrected low-order L2P.
#include "gravity.hpp"

#include <cmath>

namespace vg {
namespace {

constexpr int NC = 25;  // coefficients the closure sees and corrects (orders <= 4)

__global__ __launch_bounds__(256) void k_closure_inputs(const int32_t *__restrict__ leaves, int64_t n, const int8_t *__restrict__ level,
                                                        const double *__restrict__ bmin, const double *__restrict__ bmax, double cx, double cy,
                                                        double cz, double iw, const float *__restrict__ L, int ldk, const float *__restrict__ io,
                                                        float *__restrict__ x, float *__restrict__ norm, float *__restrict__ l25, int32_t *__restrict__ lev) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= n) return;
    const int32_t c = leaves[i];
    const int lv = level[c];
    float v[NC];
    float s = 0.0f;
    const bool far = lv >= 2;  // cells that take L2P (solver_ref.hpp:220-221)
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        v[k] = far ? L[static_cast<int64_t>(c) * ldk + k] : 0.0f;
        s += v[k] * v[k];
    }
    const float nrm = sqrtf(s);
    const float inv = nrm > 0.0f ? 1.0f / nrm : 0.0f;
    const double mx = (bmin[3 * c] + bmax[3 * c]) / 2 - cx, my = (bmin[3 * c + 1] + bmax[3 * c + 1]) / 2 - cy,
                 mz = (bmin[3 * c + 2] + bmax[3 * c + 2]) / 2 - cz;
    float f[28];
#pragma unroll
    for (int k = 0; k < NC; ++k) f[k] = v[k] * inv;
    f[25] = static_cast<float>(lv) / 16.0f;
    f[26] = static_cast<float>(ldexp(1.0, -lv));
    f[27] = static_cast<float>(sqrt(mx * mx + my * my + mz * mz) * iw);
    // standardised for the network (io: shift 28, scale 28), or raw for export
#pragma unroll
    for (int k = 0; k < 28; ++k) x[i * 28 + k] = io ? (f[k] - io[k]) * io[28 + k] : f[k];
    norm[i] = nrm;
    if (l25)
#pragma unroll
        for (int k = 0; k < NC; ++k) l25[i * NC + k] = v[k];
    if (lev) lev[i] = lv;
}

__global__ __launch_bounds__(256) void k_closure_apply(const int32_t *__restrict__ leaves, int64_t n, const float *__restrict__ norm,
                                                       const float *__restrict__ y, const float *__restrict__ oscale, float *__restrict__ L, int ldk,
                                                       float bound, uint8_t *__restrict__ flag, uint32_t *__restrict__ nflag) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= n) return;
    const float nrm = norm[i];
    uint8_t f = 0;
    if (nrm > 0.0f) {
        const int32_t c = leaves[i];
        float e = 0.0f;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const float r = y[i * NC + k] * oscale[k];
            e += r * r;
            L[static_cast<int64_t>(c) * ldk + k] += nrm * r;
        }
        f = sqrtf(e) > bound ? 1 : 0;
    }
    flag[i] = f;
    if (f) atomicAdd(nflag, 1u);
}

// mark every cell on the ancestor chain (self included) of a flagged leaf
__global__ __launch_bounds__(256) void k_mark_chains(const int32_t *__restrict__ leaves, const uint8_t *__restrict__ flag, int64_t n,
                                                     const int8_t *__restrict__ level, const int32_t *__restrict__ anc, uint8_t *__restrict__ need) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const int32_t c = leaves[i];
    for (int l = 0; l <= level[c]; ++l) need[anc[static_cast<int64_t>(c) * (MAXD + 1) + l]] = 1;
}

// ordered compaction of ids[i] with keep(ids[i]) (one block): out, and pos[i] =
// number kept before i (pos[n] = total).  keep = need[ids[i]] or flag[i].
__global__ __launch_bounds__(1024) void k_compact_keep(const int32_t *__restrict__ ids, int64_t n, const uint8_t *__restrict__ need,
                                                       const uint8_t *__restrict__ flag, int32_t *__restrict__ out, uint32_t *__restrict__ pos) {
    __shared__ uint32_t s[1024];
    uint32_t base = 0;
    for (int64_t b = 0; b < n; b += 1024) {
        const int64_t i = b + threadIdx.x;
        const uint32_t k = i < n ? (need ? need[ids[i]] : flag[i]) : 0;
        s[threadIdx.x] = k;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
            const uint32_t v = threadIdx.x >= static_cast<unsigned>(off) ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += v;
            __syncthreads();
        }
        const uint32_t excl = s[threadIdx.x] - k;
        if (i < n) {
            if (pos) pos[i] = base + excl;
            if (k) out[base + excl] = ids[i];
        }
        const uint32_t tot = s[1023];
        __syncthreads();
        base += tot;
    }
    if (threadIdx.x == 0 && pos) pos[n] = base;
}

}  // namespace

void closure_inputs(Ctx &c, bool with_l25) {
    TreeDev &t = c.t;
    const int64_t n = t.nleaves;
    Closure &cl = c.cl;
    cl.x.ensure(std::max<int64_t>(1, n) * 28);
    cl.norm.ensure(std::max<int64_t>(1, n));
    if (with_l25) {
        cl.l25.ensure(std::max<int64_t>(1, n) * NC);
        cl.lev.ensure(std::max<int64_t>(1, n));
    }
    if (n == 0) return;
    const double cx = (t.box[0] + t.box[3]) / 2, cy = (t.box[1] + t.box[4]) / 2, cz = (t.box[2] + t.box[5]) / 2;
    const double iw = 1.0 / (t.box[3] - t.box[0]);
    k_closure_inputs<<<ceil_div(n, 256), 256, 0, c.stream>>>(t.leaf_ids, n, t.level, t.bmin, t.bmax, cx, cy, cz, iw, c.L32, c.ldk,
                                                             with_l25 ? nullptr : cl.io.p, cl.x, cl.norm, with_l25 ? cl.l25.p : nullptr,
                                                             with_l25 ? cl.lev.p : nullptr);
    VG_LAUNCH_CHECK();
}

void closure_apply(Ctx &c) {
    TreeDev &t = c.t;
    Closure &cl = c.cl;
    const int64_t n = t.nleaves;
    cl.flag.ensure(std::max<int64_t>(1, n));
    cl.scal.ensure(1);
    VG_HIP(hipMemsetAsync(cl.scal.p, 0, sizeof(uint32_t), c.stream));
    cl.nflag = 0;
    if (n == 0) return;
    closure_inputs(c, false);
    cl.y.ensure(n * NC);
    cl.net.forward(c.stream, cl.x, n, cl.y);
    k_closure_apply<<<ceil_div(n, 256), 256, 0, c.stream>>>(t.leaf_ids, n, cl.norm, cl.y, cl.io.p + 56, c.L32, c.ldk, static_cast<float>(cl.bound),
                                                            cl.flag, cl.scal);
    VG_LAUNCH_CHECK();
}

bool closure_fallback_prepare(Ctx &c) {
    Closure &cl = c.cl;
    TreeDev &t = c.t;
    if (!(cl.bound < INFINITY) || t.nleaves == 0) return false;
    uint32_t *hnf = reinterpret_cast<uint32_t *>(static_cast<char *>(c.pinned) + 256);  // a pinned host word
    VG_HIP(hipMemcpyAsync(hnf, cl.scal.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    const uint32_t nf = *hnf;
    cl.nflag = nf;
    if (nf == 0) return false;
    const int64_t n = t.nleaves;
    // flagged leaves, in leaf order
    cl.fb_leaves.ensure(n);
    k_compact_keep<<<1, 1024, 0, c.stream>>>(t.leaf_ids, n, nullptr, cl.flag, cl.fb_leaves, nullptr);
    VG_LAUNCH_CHECK();
    // internal cells on their ancestor chains, level by level
    cl.need.ensure(t.ncells);
    VG_HIP(hipMemsetAsync(cl.need.p, 0, t.ncells, c.stream));
    k_mark_chains<<<ceil_div(n, 256), 256, 0, c.stream>>>(t.leaf_ids, cl.flag, n, t.level, t.anc, cl.need);
    VG_LAUNCH_CHECK();
    const int64_t nint = t.int_off.back();
    cl.fb_int.ensure(std::max<int64_t>(1, nint));
    cl.fb_pos.ensure(nint + 1);
    k_compact_keep<<<1, 1024, 0, c.stream>>>(t.int_ids, nint, cl.need, nullptr, cl.fb_int, cl.fb_pos);
    VG_LAUNCH_CHECK();
    std::vector<uint32_t> pos(static_cast<size_t>(nint + 1));
    VG_HIP(hipMemcpyAsync(pos.data(), cl.fb_pos.p, (nint + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    cl.fb_off.resize(t.int_off.size());
    for (size_t d = 0; d < t.int_off.size(); ++d) cl.fb_off[d] = pos[static_cast<size_t>(t.int_off[d])];
    // order-p_hi pass restricted to the chains (the upward pass is complete: any cell can be a source)
    ExpSet &hi = cl.hi;
    hi.p = cl.p_hi;
    hi.K = (cl.p_hi + 1) * (cl.p_hi + 1);
    hi.ldk = c.prec == 2 ? (hi.K + 15) / 16 * 16 : hi.K;
    swap_expansions(c, hi);
    try {
        operators_build(c);
        fmm_upward(c);
        if (t.depth >= 1) {
            const int64_t a = cl.fb_off[1], b = cl.fb_off.back();
            m2l_list(c, cl.fb_int.p + a, static_cast<int>(b - a));
        }
        for (int d = 2; d < t.depth; ++d) l2l_list(c, cl.fb_int.p + cl.fb_off[d], cl.fb_off[d + 1] - cl.fb_off[d]);
    } catch (...) {
        swap_expansions(c, hi);
        throw;
    }
    swap_expansions(c, hi);
    return true;
}

void closure_fallback_l2p(Ctx &c) {
    Closure &cl = c.cl;
    swap_expansions(c, cl.hi);
    try {
        l2p_list(c, cl.fb_leaves, cl.nflag, nullptr);
    } catch (...) {
        swap_expansions(c, cl.hi);
        throw;
    }
    swap_expansions(c, cl.hi);
}

}  // namespace vg
