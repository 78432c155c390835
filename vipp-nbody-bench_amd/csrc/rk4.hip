// rk4.hip - FP64 classic RK4 on the caller-ordered state, the scatter of the
// sorted forces back to caller order, energies and work counts.
//
// The stage arithmetic is the reference's (solver_ref.hpp:36-104):
//   x_out += (dt/6) v ; v_out += (dt/6) a      after the stage-0 force
//   x = x0 + v (dt/2) ; v = v0 + a (dt/2)      (euler, solver_ref.hpp:109-112)
//   ... (dt/3, dt/2), (dt/3, dt), (dt/6, copy out)
// evaluated with the same roundings: contraction into FMA is off here, so the
// FP64 state matches a non-FMA x86 build of the reference bit for bit when the
// forces agree.
#include "gravity.hpp"

#pragma clang fp contract(off)

namespace vg {
namespace {

__global__ __launch_bounds__(256) void k_rk4_init(int64_t n3, const double *__restrict__ x, const double *__restrict__ v, double *__restrict__ x0,
                                                  double *__restrict__ v0, double *__restrict__ xo, double *__restrict__ vo) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= n3) return;
    x0[i] = x[i];
    xo[i] = x[i];
    v0[i] = v[i];
    vo[i] = v[i];
}

// accumulate with weight cw, then take the next Euler state with step h
// (last stage: copy the accumulated state out instead).  Grid-stride over the
// 3n scalars; each block also leaves the partial bounding box of the positions
// it wrote (7 doubles at part, the layout of k_aabb), so the next tree build
// needs no pass of its own over x.
template <class A>
__global__ __launch_bounds__(256) void k_rk4_stage(int64_t n3, double cw, double h, int last, const A *__restrict__ a, double *__restrict__ x,
                                                   double *__restrict__ v, const double *__restrict__ x0, const double *__restrict__ v0,
                                                   double *__restrict__ xo, double *__restrict__ vo, double *__restrict__ part) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double bad = 0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
    const int s3 = static_cast<int>(stride % 3);
    int64_t i = blockIdx.x * 256ll + threadIdx.x;
    int comp = static_cast<int>(i % 3);
    for (; i < n3; i += stride) {
        const double ai = static_cast<double>(a[i]);
        const double vi = v[i];
        const double xon = xo[i] + cw * vi;
        const double von = vo[i] + cw * ai;
        double xn;
        if (last) {
            xn = xon;
            v[i] = von;
        } else {
            xo[i] = xon;
            vo[i] = von;
            xn = x0[i] + vi * h;
            v[i] = v0[i] + ai * h;
        }
        x[i] = xn;
        if (!isfinite(xn)) bad = 1;
        lo[0] = comp == 0 ? fmin(lo[0], xn) : lo[0];
        lo[1] = comp == 1 ? fmin(lo[1], xn) : lo[1];
        lo[2] = comp == 2 ? fmin(lo[2], xn) : lo[2];
        hi[0] = comp == 0 ? fmax(hi[0], xn) : hi[0];
        hi[1] = comp == 1 ? fmax(hi[1], xn) : hi[1];
        hi[2] = comp == 2 ? fmax(hi[2], xn) : hi[2];
        comp += s3;
        if (comp >= 3) comp -= 3;
    }
    aabb_block_store(lo, hi, bad, part + 7ll * blockIdx.x);
}

template <class R, class A>
__global__ __launch_bounds__(256) void k_scatter(int64_t n, const uint32_t *__restrict__ perm, const R *__restrict__ acc4, A *__restrict__ a,
                                                 double *__restrict__ pot) {
    const int64_t k = blockIdx.x * 256ll + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = perm[k];
    a[3ll * i] = static_cast<A>(acc4[4 * k]);
    a[3ll * i + 1] = static_cast<A>(acc4[4 * k + 1]);
    a[3ll * i + 2] = static_cast<A>(acc4[4 * k + 2]);
    if (pot) pot[i] = static_cast<double>(acc4[4 * k + 3]);
}

// FP64 block reduction of per-particle terms: kind 0 = m |v|^2, kind 1 = m * pot
__global__ __launch_bounds__(256) void k_energy(int64_t n, int kind, const double *__restrict__ v, const double *__restrict__ m, double mass,
                                                const double *__restrict__ pot, double *__restrict__ part) {
    __shared__ double s[256];
    double acc = 0;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
        const double mi = m ? m[i] : mass;
        if (kind == 0)
            acc += mi * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
        else
            acc += mi * pot[i];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < static_cast<unsigned>(w)) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// work counts: [0] P2P pairs (i != j), [1] M2L pairs of targets at level >= 2,
// [2] M2L pairs whose target and source both hold particles, [3] list
// entries stored (AP, NB and each cell's own NI entries)
__global__ __launch_bounds__(256) void k_count(int64_t nc, const int8_t *__restrict__ level, const int32_t *__restrict__ child, const uint32_t *__restrict__ count,
                                               const int32_t *__restrict__ anc, const int32_t *__restrict__ ni, const int32_t *__restrict__ ni_cnt,
                                               const int32_t *__restrict__ ap, const int32_t *__restrict__ ap_cnt, const int32_t *__restrict__ nb_cnt,
                                               unsigned long long *__restrict__ out) {
    const int64_t c = blockIdx.x * 256ll + threadIdx.x;
    unsigned long long pairs = 0, m2l = 0, live = 0, entries = 0;
    if (c < nc) {
        entries = static_cast<unsigned long long>(ap_cnt[c] + nb_cnt[c] + ni_cnt[c]);
        if (level[c] >= 2) {
            m2l = static_cast<unsigned long long>(ap_cnt[c]);
            if (count[c] > 0)
                for (int e = 0; e < ap_cnt[c]; ++e) live += count[ap[189 * c + e]] > 0;
        }
        if (child[c] < 0 && count[c] > 0) {
            unsigned long long src = 0;
            for (int l = 0; l <= level[c]; ++l) {
                const int32_t a = anc[c * (MAXD + 1) + l];
                for (int e = 0; e < ni_cnt[a]; ++e) src += count[ni[27 * static_cast<int64_t>(a) + e]];
            }
            pairs = static_cast<unsigned long long>(count[c]) * (src - 1);
        }
    }
    if (pairs) atomicAdd(&out[0], pairs);
    if (m2l) atomicAdd(&out[1], m2l);
    if (live) atomicAdd(&out[2], live);
    if (entries) atomicAdd(&out[3], entries);
}

}  // namespace

void rk4_init(Ctx &c) {
    const int64_t n3 = 3 * c.n;
    k_rk4_init<<<ceil_div(n3, 256), 256, 0, c.stream>>>(n3, c.x, c.v, c.x0, c.v0, c.xo, c.vo);
    VG_LAUNCH_CHECK();
}

void rk4_stage(Ctx &c, int stage, double dt) {
    const int64_t n3 = 3 * c.n;
    // weights and Euler steps of solver_ref.hpp:58-99
    const double cw = (stage == 0 || stage == 3) ? dt / 6 : dt / 3;
    const double h = stage == 2 ? dt : dt / 2;
    const int last = stage == 3;
    const int nb = static_cast<int>(std::min<int64_t>(1024, std::max<int64_t>(1, ceil_div(n3, 256))));
    c.aabb_part.ensure(7 * static_cast<size_t>(nb));
    c.aabb_parts = 0;
    if (c.prec == 0)
        k_rk4_stage<double><<<nb, 256, 0, c.stream>>>(n3, cw, h, last, c.acc64, c.x, c.v, c.x0, c.v0, c.xo, c.vo, c.aabb_part);
    else
        k_rk4_stage<float><<<nb, 256, 0, c.stream>>>(n3, cw, h, last, c.acc32, c.x, c.v, c.x0, c.v0, c.xo, c.vo, c.aabb_part);
    VG_LAUNCH_CHECK();
    c.aabb_parts = nb;
}

void scatter_accel(Ctx &c, bool potential) {
    const int64_t n = c.n;
    if (potential) c.pot64.ensure(std::max<int64_t>(1, n));
    double *pot = potential ? c.pot64.p : nullptr;
    if (c.prec == 0) {
        c.acc64.ensure(3 * std::max<int64_t>(1, n));
        k_scatter<double, double><<<ceil_div(n, 256), 256, 0, c.stream>>>(n, c.perm, c.as64, c.acc64.p, pot);
    } else {
        c.acc32.ensure(3 * std::max<int64_t>(1, n));
        k_scatter<float, float><<<ceil_div(n, 256), 256, 0, c.stream>>>(n, c.perm, c.as32, c.acc32.p, pot);
    }
    VG_LAUNCH_CHECK();
}

static double reduce_energy(Ctx &c, int kind) {
    const int nb = 256;
    DevBuf<double> part;
    part.ensure(nb);
    k_energy<<<nb, 256, 0, c.stream>>>(c.n, kind, c.v, c.has_masses ? c.m.p : nullptr, c.mass, c.pot64.p, part);
    VG_LAUNCH_CHECK();
    std::vector<double> h(nb);
    VG_HIP(hipMemcpyAsync(h.data(), part.p, nb * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    double s = 0;
    for (double q : h) s += q;  // fixed order: deterministic
    return s;
}

double kinetic_energy(Ctx &c) { return 0.5 * reduce_energy(c, 0); }

double potential_energy(Ctx &c) { return -0.5 * c.G * reduce_energy(c, 1); }

void count_work(Ctx &c) {
    TreeDev &t = c.t;
    DevBuf<unsigned long long> out;
    out.ensure(4);
    VG_HIP(hipMemsetAsync(out.p, 0, 4 * sizeof(unsigned long long), c.stream));
    k_count<<<ceil_div(t.ncells, 256), 256, 0, c.stream>>>(t.ncells, t.level, t.child, t.count, t.anc, t.ni, t.ni_cnt, t.ap, t.ap_cnt, t.nb_cnt, out);
    VG_LAUNCH_CHECK();
    unsigned long long h[4];
    VG_HIP(hipMemcpyAsync(h, out.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    c.counts[0] = static_cast<double>(h[0]);
    c.counts[1] = static_cast<double>(h[1]);
    c.counts[2] = static_cast<double>(t.ncells);
    c.counts[3] = static_cast<double>(t.nleaves);
    c.counts[4] = static_cast<double>(t.depth);
    c.counts[5] = static_cast<double>(h[2]);
    c.counts[7] = static_cast<double>(h[3]);
}

}  // namespace vg
