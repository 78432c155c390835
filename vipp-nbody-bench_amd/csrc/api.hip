// api.hip - the C ABI of include/vipp_gravity.h over the device phases.
#include "../../include/vipp_gravity.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <random>

#include "gravity.hpp"

struct vipp_grav_ctx {
    vg::Ctx c;
};

namespace vg {

void fmm_m2l_plain(Ctx &c);
void fmm_m2l_split(Ctx &c);
void fmm_m2l_plain_list(Ctx &c, const int32_t *parents, int np);
void fmm_m2l_split_list(Ctx &c, const int32_t *parents, int np);

void m2l_list(Ctx &c, const int32_t *parents, int np) {
    if (c.prec == 2)
        fmm_m2l_split_list(c, parents, np);
    else
        fmm_m2l_plain_list(c, parents, np);
}

void swap_expansions(Ctx &c, ExpSet &e) {
    std::swap(c.p, e.p);
    std::swap(c.K, e.K);
    std::swap(c.ldk, e.ldk);
    c.harm.swap(e.harm);
    c.op64.swap(e.op64);
    c.op32.swap(e.op32);
    c.M32.swap(e.M32);
    c.L32.swap(e.L32);
    c.opsplit.swap(e.opsplit);
    c.opsplit_ml.swap(e.opsplit_ml);
    c.msplit.swap(e.msplit);
}
void fmm_constants_init();

void phase_begin(Ctx &c, Phase p) {
    if (c.timing) VG_HIP(hipEventRecord(c.ev[2 * p], c.stream));
}

void phase_end(Ctx &c, Phase p) {
    if (c.timing) VG_HIP(hipEventRecord(c.ev[2 * p + 1], c.stream));
}

static void collect_times(Ctx &c, const bool *used) {
    if (!c.timing) return;
    VG_HIP(hipStreamSynchronize(c.stream));
    for (int p = 0; p < PH_COUNT; ++p) {
        if (!used[p]) continue;
        float ms = 0;
        VG_HIP(hipEventElapsedTime(&ms, c.ev[2 * p], c.ev[2 * p + 1]));
        c.phase_ms[p] += ms;
    }
}

void fmm_m2l(Ctx &c) {
    if (c.prec == 2)
        fmm_m2l_split(c);
    else
        fmm_m2l_plain(c);
}

void join_p2p(Ctx &c) {
    if (!c.p2p_pending) return;
    VG_HIP(hipStreamWaitEvent(c.stream, c.ev_join, 0));
    c.p2p_pending = false;
}

// Untimed split-mode evaluations run the lists and P2P (VALU-bound, about
// half the evaluation) on a second stream once the tree exists, so that the
// operators and the upward pass (which need only the tree), and M2L
// (MFMA-bound, after the lists), run beside them; L2P waits for P2P. The P2P
// plan is enqueued with the lists and its item counts are read back while the
// host enqueues the operators and the upward pass. The FP32 and FP64 modes,
// whose M2L is VALU-bound too, and timed evaluations keep one stream (the
// events of a phase then bracket its own kernels only).
static void evaluate_overlapped(Ctx &c, bool potential) {
    VG_HIP(hipEventRecord(c.ev_fork, c.stream));
    VG_HIP(hipStreamWaitEvent(c.stream2, c.ev_fork, 0));
    struct Swap {
        Ctx &c;
        explicit Swap(Ctx &c_) : c(c_) { std::swap(c.stream, c.stream2); }
        ~Swap() { std::swap(c.stream, c.stream2); }
    };
    try {
        {
            Swap on_stream2(c);
            lists_build(c);
        }
        VG_HIP(hipEventRecord(c.ev_lists, c.stream2));
        {
            Swap on_stream2(c);
            fmm_p2p_plan(c);
        }
        operators_build(c);
        fmm_upward(c);  // enqueued while the P2P plan's item counts come back
        {
            Swap on_stream2(c);
            fmm_p2p_run(c, potential);
        }
        VG_HIP(hipEventRecord(c.ev_join, c.stream2));
    } catch (...) {
        (void)hipStreamSynchronize(c.stream2);  // nothing of this evaluation stays in flight
        throw;
    }
    c.p2p_pending = true;
    VG_HIP(hipStreamWaitEvent(c.stream, c.ev_lists, 0));  // M2L reads the parents' neighbour tables
    fmm_m2l(c);
    fmm_downward(c, potential);
    join_p2p(c);
    scatter_accel(c, potential);
}

void evaluate(Ctx &c, bool potential) {
    join_p2p(c);  // a P2P left pending by an evaluation that threw must end before the tree changes
    // the tree, lists and expansions are rebuilt in place below: until this
    // evaluation has finished, nothing of the previous one may be exported
    c.evaluated = false;
    c.t.ref_valid = false;
    bool used[PH_COUNT] = {};
    if (!c.timing && c.stream2 && c.prec == 2) {
        tree_build(c);
        evaluate_overlapped(c, potential);
        c.evaluated = true;
        return;
    }
    phase_begin(c, PH_TREE);
    tree_build(c);
    phase_end(c, PH_TREE);
    used[PH_TREE] = true;
    phase_begin(c, PH_LISTS);
    lists_build(c);
    phase_end(c, PH_LISTS);
    used[PH_LISTS] = true;
    phase_begin(c, PH_OPS);
    operators_build(c);
    phase_end(c, PH_OPS);
    used[PH_OPS] = true;
    phase_begin(c, PH_UP);
    fmm_upward(c);
    phase_end(c, PH_UP);
    used[PH_UP] = true;
    phase_begin(c, PH_P2P);
    fmm_p2p(c, potential);
    phase_end(c, PH_P2P);
    used[PH_P2P] = true;
    phase_begin(c, PH_M2L);
    fmm_m2l(c);
    phase_end(c, PH_M2L);
    used[PH_M2L] = true;
    phase_begin(c, PH_DOWN);
    fmm_downward(c, potential);
    scatter_accel(c, potential);
    phase_end(c, PH_DOWN);
    used[PH_DOWN] = true;
    collect_times(c, used);
    c.evaluated = true;
}

}  // namespace vg

namespace {

std::mutex g_err_mu;
std::string g_err;

void set_global_error(const std::string &m) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = m;
}

template <class F>
int guarded(vipp_grav_ctx *ctx, F &&f) {
    try {
        f();
        return VIPP_GRAV_OK;
    } catch (const vg::Error &e) {
        if (ctx) ctx->c.err = e.what();
        set_global_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        if (ctx) ctx->c.err = e.what();
        set_global_error(e.what());
        return VIPP_GRAV_E_INVALID;
    }
}

void need(bool ok, int code, const char *msg) {
    if (!ok) throw vg::Error(code, msg);
}

}  // namespace

namespace {
void ensure_tree(vg::Ctx &c) {
    need(c.uploaded, VIPP_GRAV_E_STATE, "no particles uploaded");
    if (!c.evaluated) vg::evaluate(c, false);
    if (!c.t.ref_valid) vg::ref_numbering(c);
}

template <class T>
std::vector<T> fetch(vg::Ctx &c, const T *d, size_t n) {
    std::vector<T> h(n);
    if (n) VG_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, c.stream));
    return h;
}
struct HostLists {
    std::vector<int32_t> rid, par, lev, nb, ni, ap, nbc, nic, apc;
};
HostLists host_lists(vg::Ctx &c) {
    const vg::TreeDev &t = c.t;
    const size_t nc = static_cast<size_t>(t.ncells);
    HostLists h;
    h.rid = fetch(c, t.ref_id.p, nc);
    h.par = fetch(c, t.parent.p, nc);
    std::vector<int8_t> lv = fetch(c, t.level.p, nc);
    h.lev.assign(lv.begin(), lv.end());
    h.nb = fetch(c, t.nb.p, 27 * nc);
    h.ni = fetch(c, t.ni.p, 27 * nc);
    h.ap = fetch(c, t.ap.p, 189 * nc);
    h.nbc = fetch(c, t.nb_cnt.p, nc);
    h.nic = fetch(c, t.ni_cnt.p, nc);
    h.apc = fetch(c, t.ap_cnt.p, nc);
    VG_HIP(hipStreamSynchronize(c.stream));
    uint32_t overflow = 0;
    VG_HIP(hipMemcpy(&overflow, c.scal.p + 4, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (overflow) throw vg::Error(VIPP_GRAV_E_DEVICE, "interaction list overflow");
    return h;
}
// full NI list of a cell: local entries of its ancestors, root first (solver_ref.hpp:126)
int64_t ni_full_len(const HostLists &h, int32_t g) {
    int64_t s = 0;
    for (int32_t a = g; a >= 0; a = h.par[static_cast<size_t>(a)]) s += h.nic[static_cast<size_t>(a)];
    return s;
}
}  // namespace

namespace {
struct TmpStream {
    hipStream_t s = nullptr;
    explicit TmpStream(int device) {
        VG_HIP(hipSetDevice(device));
        VG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    }
    ~TmpStream() {
        if (s) (void)hipStreamDestroy(s);
    }
};
}  // namespace

extern "C" {

int vipp_mlp_forward(int32_t device, const float *params, const float *x, int64_t n, float *y) {
    return guarded(nullptr, [&] {
        need(params && (n == 0 || (x && y)) && n >= 0, VIPP_GRAV_E_INVALID, "null argument");
        TmpStream st(device);
        vg::mlp::Net net;
        net.par.ensure(vg::mlp::NPAR);
        VG_HIP(hipMemcpyAsync(net.par.p, params, vg::mlp::NPAR * sizeof(float), hipMemcpyHostToDevice, st.s));
        net.set_params(st.s, net.par);
        if (n > 0) {
            vg::DevBuf<float> xd, yd;
            xd.ensure(n * vg::mlp::IN);
            yd.ensure(n * vg::mlp::OUT);
            VG_HIP(hipMemcpyAsync(xd.p, x, n * vg::mlp::IN * sizeof(float), hipMemcpyHostToDevice, st.s));
            net.forward(st.s, xd, n, yd);
            VG_HIP(hipMemcpyAsync(y, yd.p, n * vg::mlp::OUT * sizeof(float), hipMemcpyDeviceToHost, st.s));
            VG_HIP(hipStreamSynchronize(st.s));
        }
        VG_HIP(hipStreamSynchronize(st.s));
    });
}

int vipp_mlp_loss_grad(int32_t device, const float *params, const float *x, const float *t, const float *w, int64_t n, float *loss,
                       float *grad) {
    return guarded(nullptr, [&] {
        need(params && x && t && loss && grad && n > 0, VIPP_GRAV_E_INVALID, "null argument or n <= 0");
        TmpStream st(device);
        vg::mlp::Net net;
        net.par.ensure(vg::mlp::NPAR);
        VG_HIP(hipMemcpyAsync(net.par.p, params, vg::mlp::NPAR * sizeof(float), hipMemcpyHostToDevice, st.s));
        net.set_params(st.s, net.par);
        vg::DevBuf<float> xd, td, wd, gd;
        xd.ensure(n * vg::mlp::IN);
        td.ensure(n * vg::mlp::OUT);
        gd.ensure(vg::mlp::NPAR);
        VG_HIP(hipMemcpyAsync(xd.p, x, n * vg::mlp::IN * sizeof(float), hipMemcpyHostToDevice, st.s));
        VG_HIP(hipMemcpyAsync(td.p, t, n * vg::mlp::OUT * sizeof(float), hipMemcpyHostToDevice, st.s));
        double wsum = 0;
        if (w) {
            for (int64_t i = 0; i < n; ++i) {
                need(std::isfinite(w[i]) && w[i] >= 0, VIPP_GRAV_E_INVALID, "sample weights must be finite and >= 0");
                wsum += w[i];
            }
            need(wsum > 0, VIPP_GRAV_E_INVALID, "sample weights sum to 0");
            wd.ensure(n);
            VG_HIP(hipMemcpyAsync(wd.p, w, n * sizeof(float), hipMemcpyHostToDevice, st.s));
        }
        *loss = net.loss_grad(st.s, xd, td, w ? wd.p : nullptr, wsum, n, gd);
        VG_HIP(hipMemcpyAsync(grad, gd.p, vg::mlp::NPAR * sizeof(float), hipMemcpyDeviceToHost, st.s));
        VG_HIP(hipStreamSynchronize(st.s));
    });
}

int vipp_mlp_train(int32_t device, float *params, const float *x, const float *t, const float *w, int64_t n, uint32_t epochs, uint32_t batch,
                   float lr, uint64_t seed, float *epoch_loss) {
    return guarded(nullptr, [&] {
        need(params && x && t && n > 0 && batch > 0 && epochs > 0, VIPP_GRAV_E_INVALID, "null argument or empty training set");
        need(std::isfinite(lr) && lr > 0, VIPP_GRAV_E_INVALID, "learning rate must be positive");
        TmpStream st(device);
        using namespace vg::mlp;
        Net net;
        net.par.ensure(NPAR);
        VG_HIP(hipMemcpyAsync(net.par.p, params, NPAR * sizeof(float), hipMemcpyHostToDevice, st.s));
        net.set_params(st.s, net.par);
        vg::DevBuf<float> xd, td, wd, bx, bt, bw, gd;
        vg::DevBuf<uint32_t> idx;
        xd.ensure(n * IN), td.ensure(n * OUT), idx.ensure(n), gd.ensure(NPAR);
        const int64_t B = std::min<int64_t>(batch, n);
        bx.ensure(B * IN), bt.ensure(B * OUT);
        VG_HIP(hipMemcpyAsync(xd.p, x, n * IN * sizeof(float), hipMemcpyHostToDevice, st.s));
        VG_HIP(hipMemcpyAsync(td.p, t, n * OUT * sizeof(float), hipMemcpyHostToDevice, st.s));
        if (w) {
            for (int64_t i = 0; i < n; ++i) need(std::isfinite(w[i]) && w[i] >= 0, VIPP_GRAV_E_INVALID, "sample weights must be finite and >= 0");
            wd.ensure(n), bw.ensure(B);
            VG_HIP(hipMemcpyAsync(wd.p, w, n * sizeof(float), hipMemcpyHostToDevice, st.s));
        }
        std::vector<uint32_t> perm(static_cast<size_t>(n));
        int step = 0;
        for (uint32_t e = 0; e < epochs; ++e) {
            std::iota(perm.begin(), perm.end(), 0u);
            std::mt19937_64 rng(seed + e);
            for (int64_t i = n - 1; i > 0; --i) {  // Fisher-Yates with a fixed generator: reproducible order
                const int64_t j = static_cast<int64_t>(rng() % static_cast<uint64_t>(i + 1));
                std::swap(perm[static_cast<size_t>(i)], perm[static_cast<size_t>(j)]);
            }
            VG_HIP(hipMemcpyAsync(idx.p, perm.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st.s));
            const float lr_e = lr * (0.05f + 0.95f * 0.5f * (1.0f + std::cos(3.14159265358979f * static_cast<float>(e) / static_cast<float>(epochs))));
            double sum = 0, wall = 0;
            for (int64_t off = 0; off < n; off += B) {
                const int64_t nb = std::min<int64_t>(B, n - off);
                gather_rows(st.s, xd, IN, idx.p + off, nb, bx);
                gather_rows(st.s, td, OUT, idx.p + off, nb, bt);
                double wsum = static_cast<double>(nb);
                if (w) {
                    gather_rows(st.s, wd, 1, idx.p + off, nb, bw);
                    wsum = 0;
                    for (int64_t i = 0; i < nb; ++i) wsum += w[perm[static_cast<size_t>(off + i)]];
                    if (!(wsum > 0)) continue;  // a batch of zero-weight rows carries no gradient
                }
                sum += static_cast<double>(net.loss_grad(st.s, bx, bt, w ? bw.p : nullptr, wsum, nb, gd)) * wsum;
                wall += wsum;
                net.adam_step(st.s, gd, lr_e, ++step);
            }
            if (epoch_loss) epoch_loss[e] = static_cast<float>(wall > 0 ? sum / wall : 0.0);
        }
        VG_HIP(hipMemcpyAsync(params, net.par.p, NPAR * sizeof(float), hipMemcpyDeviceToHost, st.s));
        VG_HIP(hipStreamSynchronize(st.s));
    });
}

int vipp_grav_create(const vipp_grav_config *cfg, vipp_grav_ctx **out) {
    return guarded(nullptr, [&] {
        need(cfg && out, VIPP_GRAV_E_INVALID, "null argument");
        need(cfg->n != 0, VIPP_GRAV_E_INVALID, "N must not be 0");  // problem.hpp:179-181
        need(cfg->s != 0, VIPP_GRAV_E_INVALID, "s must not be 0");  // problem.hpp:182-184
        need(cfg->p >= 1 && cfg->p <= static_cast<uint32_t>(vg::MAXP), VIPP_GRAV_E_INVALID, "p must be in [1, 12]");
        need(cfg->precision <= 2, VIPP_GRAV_E_INVALID, "unknown precision mode");
        need(cfg->n < (1ull << 31), VIPP_GRAV_E_INVALID, "N must be below 2^31");
        const double thr = std::floor(std::sqrt(8) * static_cast<double>(cfg->s));  // octree.hpp:83
        need(thr >= 1 && thr <= 16384, VIPP_GRAV_E_INVALID, "floor(sqrt(8) s) must be in [1, 16384]");
        need(std::isfinite(cfg->epsilon) && std::isfinite(cfg->G) && std::isfinite(cfg->mass), VIPP_GRAV_E_INVALID, "non-finite parameter");
        auto *ctx = new vipp_grav_ctx;
        vg::Ctx &c = ctx->c;
        try {
            c.n = static_cast<int64_t>(cfg->n);
            c.s = static_cast<int64_t>(cfg->s);
            c.thr = static_cast<uint32_t>(thr);
            c.p = static_cast<int>(cfg->p);
            c.K = (c.p + 1) * (c.p + 1);
            c.prec = static_cast<int>(cfg->precision);
            c.ldk = c.prec == 2 ? (c.K + 15) / 16 * 16 : c.K;  // the MFMA path reads whole 16-coefficient k-steps
            c.eps = cfg->epsilon;
            c.G = cfg->G;
            c.mass = cfg->mass;
            c.flags = cfg->flags;
            c.device = cfg->device;
            VG_HIP(hipSetDevice(c.device));
            VG_HIP(hipDeviceGetAttribute(&c.ncu, hipDeviceAttributeMultiprocessorCount, c.device));
            VG_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
            VG_HIP(hipStreamCreateWithFlags(&c.stream2, hipStreamNonBlocking));
            VG_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
            VG_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
            VG_HIP(hipEventCreateWithFlags(&c.ev_lists, hipEventDisableTiming));
            VG_HIP(hipEventCreateWithFlags(&c.ev_plan, hipEventDisableTiming));
            VG_HIP(hipHostMalloc(&c.pinned, vg::Ctx::PINNED_BYTES, hipHostMallocDefault));
            for (auto &e : c.ev) VG_HIP(hipEventCreate(&e));
            vg::fmm_constants_init();
            vg::noalloc_init(c.device);
            const size_t n3 = 3 * static_cast<size_t>(c.n);
            c.x.ensure(n3);
            c.v.ensure(n3);
            c.x0.ensure(n3);
            c.v0.ensure(n3);
            c.xo.ensure(n3);
            c.vo.ensure(n3);
        } catch (...) {
            delete ctx;
            throw;
        }
        *out = ctx;
    });
}

int vipp_grav_destroy(vipp_grav_ctx *ctx) {
    if (!ctx) return VIPP_GRAV_OK;
    int rc = guarded(ctx, [&] {
        VG_HIP(hipSetDevice(ctx->c.device));
        if (ctx->c.stream) VG_HIP(hipStreamSynchronize(ctx->c.stream));
        if (ctx->c.stream2) VG_HIP(hipStreamSynchronize(ctx->c.stream2));
    });
    for (auto &e : ctx->c.ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->c.ev_fork) (void)hipEventDestroy(ctx->c.ev_fork);
    if (ctx->c.ev_join) (void)hipEventDestroy(ctx->c.ev_join);
    if (ctx->c.ev_lists) (void)hipEventDestroy(ctx->c.ev_lists);
    if (ctx->c.ev_plan) (void)hipEventDestroy(ctx->c.ev_plan);
    if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
    if (ctx->c.stream2) (void)hipStreamDestroy(ctx->c.stream2);
    if (ctx->c.pinned) (void)hipHostFree(ctx->c.pinned);
    delete ctx;
    return rc;
}

const char *vipp_grav_last_error(const vipp_grav_ctx *ctx) {
    if (ctx) return ctx->c.err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_err.c_str();
}

int vipp_grav_upload(vipp_grav_ctx *ctx, const double *x, const double *v, const double *masses) {
    return guarded(ctx, [&] {
        need(ctx && x && v, VIPP_GRAV_E_INVALID, "null argument");
        vg::Ctx &c = ctx->c;
        VG_HIP(hipSetDevice(c.device));
        const size_t n3 = 3 * static_cast<size_t>(c.n);
        VG_HIP(hipMemcpyAsync(c.x.p, x, n3 * sizeof(double), hipMemcpyHostToDevice, c.stream));
        VG_HIP(hipMemcpyAsync(c.v.p, v, n3 * sizeof(double), hipMemcpyHostToDevice, c.stream));
        c.has_masses = masses != nullptr;
        if (masses) {
            c.m.ensure(static_cast<size_t>(c.n));
            VG_HIP(hipMemcpyAsync(c.m.p, masses, c.n * sizeof(double), hipMemcpyHostToDevice, c.stream));
        }
        VG_HIP(hipStreamSynchronize(c.stream));
        c.uploaded = true;
        c.evaluated = false;
    });
}

int vipp_grav_download(vipp_grav_ctx *ctx, double *x, double *v) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        vg::Ctx &c = ctx->c;
        need(c.uploaded, VIPP_GRAV_E_STATE, "download before upload");
        VG_HIP(hipSetDevice(c.device));
        const size_t n3 = 3 * static_cast<size_t>(c.n);
        if (x) VG_HIP(hipMemcpyAsync(x, c.x.p, n3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        if (v) VG_HIP(hipMemcpyAsync(v, c.v.p, n3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        VG_HIP(hipStreamSynchronize(c.stream));
    });
}

int vipp_grav_step(vipp_grav_ctx *ctx, double dt, uint32_t nsteps) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        vg::Ctx &c = ctx->c;
        need(c.uploaded, VIPP_GRAV_E_STATE, "step before upload");
        need(std::isfinite(dt), VIPP_GRAV_E_INVALID, "non-finite dt");
        VG_HIP(hipSetDevice(c.device));
        // partial boxes of the RK4 stages serve the next tree build inside this
        // call only: outside it the caller may have written x (device_state)
        c.aabb_parts = 0;
        for (uint32_t s = 0; s < nsteps; ++s) {
            vg::rk4_init(c);
            try {
                for (int stage = 0; stage < 4; ++stage) {
                    vg::evaluate(c, false);
                    vg::phase_begin(c, vg::PH_RK4);
                    vg::rk4_stage(c, stage, dt);
                    vg::phase_end(c, vg::PH_RK4);
                    if (c.timing) {
                        VG_HIP(hipStreamSynchronize(c.stream));
                        float ms = 0;
                        VG_HIP(hipEventElapsedTime(&ms, c.ev[2 * vg::PH_RK4], c.ev[2 * vg::PH_RK4 + 1]));
                        c.phase_ms[vg::PH_RK4] += ms;
                    }
                }
            } catch (...) {
                // a stage state the tree cannot hold (E_INPUT, E_DEPTH): put back
                // the state this step started from, so x and v hold the last
                // completed step and the next query rebuilds from it
                const size_t n3 = 3 * static_cast<size_t>(c.n);
                (void)hipMemcpyAsync(c.x.p, c.x0.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, c.stream);
                (void)hipMemcpyAsync(c.v.p, c.v0.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, c.stream);
                (void)hipStreamSynchronize(c.stream);
                c.evaluated = false;
                c.t.ref_valid = false;
                c.aabb_parts = 0;
                throw;
            }
        }
        c.aabb_parts = 0;
        VG_HIP(hipStreamSynchronize(c.stream));
        // the last evaluation saw the stage-3 positions, not the stepped ones:
        // tree and list queries after a step rebuild at the new positions
        if (nsteps) {
            c.evaluated = false;
            c.t.ref_valid = false;
        }
    });
}

int vipp_grav_accelerations(vipp_grav_ctx *ctx, double *a) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        vg::Ctx &c = ctx->c;
        need(c.uploaded, VIPP_GRAV_E_STATE, "accelerations before upload");
        VG_HIP(hipSetDevice(c.device));
        vg::evaluate(c, false);
        if (a) {
            const size_t n3 = 3 * static_cast<size_t>(c.n);
            if (c.prec == 0) {
                VG_HIP(hipMemcpyAsync(a, c.acc64.p, n3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
            } else {
                std::vector<float> tmp(n3);
                VG_HIP(hipMemcpyAsync(tmp.data(), c.acc32.p, n3 * sizeof(float), hipMemcpyDeviceToHost, c.stream));
                VG_HIP(hipStreamSynchronize(c.stream));
                for (size_t i = 0; i < n3; ++i) a[i] = tmp[i];
            }
        }
        VG_HIP(hipStreamSynchronize(c.stream));
    });
}

int vipp_grav_energy(vipp_grav_ctx *ctx, double *kinetic, double *potential) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        vg::Ctx &c = ctx->c;
        need(c.uploaded, VIPP_GRAV_E_STATE, "energy before upload");
        VG_HIP(hipSetDevice(c.device));
        vg::evaluate(c, true);
        const double pe = vg::potential_energy(c);
        const double ke = vg::kinetic_energy(c);
        if (kinetic) *kinetic = ke;
        if (potential) *potential = pe;
    });
}


int vipp_grav_direct(vipp_grav_ctx *ctx, const int64_t *idx, int64_t nidx, double *a, double *pot) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        need(nidx >= 0 && (nidx == 0 || (idx && a)), VIPP_GRAV_E_INVALID, "null argument");
        vg::Ctx &c = ctx->c;
        need(c.uploaded, VIPP_GRAV_E_STATE, "direct before upload");
        VG_HIP(hipSetDevice(c.device));
        vg::direct_subset(c, idx, nidx, a, pot);
    });
}

int vipp_grav_direct_noalloc(vipp_grav_ctx *ctx, const int64_t *idx, int64_t nidx, double *a, double *pot) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        need(nidx >= 0 && (nidx == 0 || (idx && a)), VIPP_GRAV_E_INVALID, "null argument");
        vg::Ctx &c = ctx->c;
        need(c.uploaded, VIPP_GRAV_E_STATE, "direct before upload");
        VG_HIP(hipSetDevice(c.device));
        vg::direct_subset_noalloc(c, idx, nidx, a, pot);
    });
}

int vipp_grav_leaf_inputs(vipp_grav_ctx *ctx, int64_t *n, float *x, float *norm, float *l25, int32_t *level) {
    return guarded(ctx, [&] {
        need(ctx != nullptr && n != nullptr, VIPP_GRAV_E_INVALID, "null argument");
        vg::Ctx &c = ctx->c;
        need(c.prec != VIPP_GRAV_F64, VIPP_GRAV_E_INVALID, "leaf inputs need FP32 expansions (precision f32 or split)");
        need(c.p >= 4, VIPP_GRAV_E_INVALID, "leaf inputs need p >= 4 (25 coefficients)");
        need(c.uploaded, VIPP_GRAV_E_STATE, "no particles uploaded");
        VG_HIP(hipSetDevice(c.device));
        if (!c.evaluated) vg::evaluate(c, false);
        const int64_t nl = c.t.nleaves;
        *n = nl;
        if (!x && !norm && !l25 && !level) return;
        vg::closure_inputs(c, true);
        if (nl) {
            if (x) VG_HIP(hipMemcpyAsync(x, c.cl.x.p, nl * 28 * sizeof(float), hipMemcpyDeviceToHost, c.stream));
            if (norm) VG_HIP(hipMemcpyAsync(norm, c.cl.norm.p, nl * sizeof(float), hipMemcpyDeviceToHost, c.stream));
            if (l25) VG_HIP(hipMemcpyAsync(l25, c.cl.l25.p, nl * 25 * sizeof(float), hipMemcpyDeviceToHost, c.stream));
            if (level) VG_HIP(hipMemcpyAsync(level, c.cl.lev.p, nl * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
        }
        VG_HIP(hipStreamSynchronize(c.stream));
    });
}

int vipp_grav_set_closure(vipp_grav_ctx *ctx, const float *params, const float *io, double bound, uint32_t p_fallback) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        vg::Ctx &c = ctx->c;
        VG_HIP(hipSetDevice(c.device));
        if (!params) {
            c.cl.on = false;
            c.evaluated = false;
            return;
        }
        need(c.prec != VIPP_GRAV_F64, VIPP_GRAV_E_INVALID, "the closure needs FP32 expansions (precision f32 or split)");
        need(c.p >= 4, VIPP_GRAV_E_INVALID, "the closure needs p >= 4 (it corrects 25 coefficients)");
        need(!(bound < 0), VIPP_GRAV_E_INVALID, "bound must be >= 0 (use INFINITY for no fallback)");
        need(p_fallback > static_cast<uint32_t>(c.p) && p_fallback <= static_cast<uint32_t>(vg::MAXP), VIPP_GRAV_E_INVALID,
             "fallback order must be in (p, 12]");
        c.cl.net.par.ensure(vg::mlp::NPAR);
        VG_HIP(hipMemcpyAsync(c.cl.net.par.p, params, vg::mlp::NPAR * sizeof(float), hipMemcpyHostToDevice, c.stream));
        c.cl.net.set_params(c.stream, c.cl.net.par);
        std::vector<float> iov(81);
        for (int k = 0; k < 81; ++k) iov[static_cast<size_t>(k)] = io ? io[k] : (k < 28 ? 0.0f : 1.0f);
        c.cl.io.ensure(81);
        VG_HIP(hipMemcpyAsync(c.cl.io.p, iov.data(), 81 * sizeof(float), hipMemcpyHostToDevice, c.stream));
        VG_HIP(hipStreamSynchronize(c.stream));
        c.cl.on = true;
        c.cl.bound = bound;
        c.cl.p_hi = static_cast<int>(p_fallback);
        c.evaluated = false;
    });
}

int vipp_grav_tree_info(vipp_grav_ctx *ctx, int64_t *ncells, int64_t *nleaves, int32_t *depth) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        vg::Ctx &c = ctx->c;
        VG_HIP(hipSetDevice(c.device));
        need(c.uploaded, VIPP_GRAV_E_STATE, "no particles uploaded");
        if (!c.evaluated) vg::evaluate(c, false);
        if (ncells) *ncells = c.t.ncells;
        if (nleaves) *nleaves = c.t.nleaves;
        if (depth) *depth = c.t.depth;
    });
}


int vipp_grav_export_tree(vipp_grav_ctx *ctx, int64_t *parent, int32_t *level, double *bmin, double *bmax, int64_t *child_base, int64_t *pt_off,
                          int64_t *pts) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        vg::Ctx &c = ctx->c;
        VG_HIP(hipSetDevice(c.device));
        ensure_tree(c);
        const vg::TreeDev &t = c.t;
        const size_t nc = static_cast<size_t>(t.ncells);
        auto rid = fetch(c, t.ref_id.p, nc);
        auto par = fetch(c, t.parent.p, nc);
        auto lev = fetch(c, t.level.p, nc);
        auto bmn = fetch(c, t.bmin.p, 3 * nc);
        auto bmx = fetch(c, t.bmax.p, 3 * nc);
        auto chd = fetch(c, t.child.p, nc);
        auto st = fetch(c, t.start.p, nc);
        auto cnt = fetch(c, t.count.p, nc);
        auto perm = fetch(c, c.perm.p, static_cast<size_t>(c.n));
        VG_HIP(hipStreamSynchronize(c.stream));
        std::vector<int64_t> cnt_ref(nc, 0);
        for (size_t g = 0; g < nc; ++g) {
            const size_t r = static_cast<size_t>(rid[g]);
            if (parent) parent[r] = par[g] < 0 ? -1 : rid[static_cast<size_t>(par[g])];
            if (level) level[r] = lev[g];
            for (int a = 0; a < 3; ++a) {
                if (bmin) bmin[3 * r + a] = bmn[3 * g + a];
                if (bmax) bmax[3 * r + a] = bmx[3 * g + a];
            }
            if (child_base) child_base[r] = chd[g] < 0 ? -1 : rid[static_cast<size_t>(chd[g])];
            cnt_ref[r] = chd[g] < 0 ? cnt[g] : 0;
        }
        if (pt_off) {
            int64_t o = 0;
            for (size_t r = 0; r < nc; ++r) {
                pt_off[r] = o;
                o += cnt_ref[r];
            }
            pt_off[nc] = o;
            if (pts)
                for (size_t g = 0; g < nc; ++g)
                    if (chd[g] < 0)
                        for (uint32_t i = 0; i < cnt[g]; ++i) pts[pt_off[rid[g]] + i] = perm[st[g] + i];
        }
    });
}


int vipp_grav_export_lists_sizes(vipp_grav_ctx *ctx, int64_t *ap_total, int64_t *ni_total, int64_t *nb_total) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        vg::Ctx &c = ctx->c;
        VG_HIP(hipSetDevice(c.device));
        ensure_tree(c);
        HostLists h = host_lists(c);
        int64_t a = 0, b = 0, d = 0;
        for (size_t g = 0; g < h.rid.size(); ++g) {
            a += h.apc[g];
            d += h.nbc[g];
            b += ni_full_len(h, static_cast<int32_t>(g));
        }
        if (ap_total) *ap_total = a;
        if (ni_total) *ni_total = b;
        if (nb_total) *nb_total = d;
    });
}

int vipp_grav_export_lists(vipp_grav_ctx *ctx, int64_t *ap_off, int64_t *ap, int64_t *ni_off, int64_t *ni, int64_t *nb_off, int64_t *nb) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        vg::Ctx &c = ctx->c;
        VG_HIP(hipSetDevice(c.device));
        ensure_tree(c);
        HostLists h = host_lists(c);
        const size_t nc = h.rid.size();
        std::vector<int32_t> g_of(nc);
        for (size_t g = 0; g < nc; ++g) g_of[static_cast<size_t>(h.rid[g])] = static_cast<int32_t>(g);
        int64_t a = 0, b = 0, d = 0;
        std::vector<int32_t> chain;
        for (size_t r = 0; r < nc; ++r) {
            const size_t g = static_cast<size_t>(g_of[r]);
            ap_off[r] = a;
            ni_off[r] = b;
            nb_off[r] = d;
            for (int e = 0; e < h.apc[g]; ++e) ap[a++] = h.rid[static_cast<size_t>(h.ap[189 * g + e])];
            for (int e = 0; e < h.nbc[g]; ++e) nb[d++] = h.rid[static_cast<size_t>(h.nb[27 * g + e])];
            chain.clear();
            for (int32_t q = static_cast<int32_t>(g); q >= 0; q = h.par[static_cast<size_t>(q)]) chain.push_back(q);
            for (auto it = chain.rbegin(); it != chain.rend(); ++it)
                for (int e = 0; e < h.nic[static_cast<size_t>(*it)]; ++e) ni[b++] = h.rid[static_cast<size_t>(h.ni[27 * static_cast<size_t>(*it) + e])];
        }
        ap_off[nc] = a;
        ni_off[nc] = b;
        nb_off[nc] = d;
    });
}

int vipp_grav_export_expansions(vipp_grav_ctx *ctx, double *M, double *L) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        vg::Ctx &c = ctx->c;
        VG_HIP(hipSetDevice(c.device));
        ensure_tree(c);
        const vg::TreeDev &t = c.t;
        const size_t nc = static_cast<size_t>(t.ncells), K = static_cast<size_t>(c.ldk);
        std::vector<double> Mh(nc * K), Lh(nc * K);
        if (c.prec == 0) {
            Mh = fetch(c, c.M64.p, nc * K);
            Lh = fetch(c, c.L64.p, nc * K);
        } else {
            auto m32 = fetch(c, c.M32.p, nc * K);
            auto l32 = fetch(c, c.L32.p, nc * K);
            VG_HIP(hipStreamSynchronize(c.stream));
            for (size_t i = 0; i < nc * K; ++i) {
                Mh[i] = m32[i];
                Lh[i] = l32[i];
            }
        }
        auto rid = fetch(c, t.ref_id.p, nc);
        auto lev = fetch(c, t.level.p, nc);
        auto cnt = fetch(c, t.count.p, nc);
        VG_HIP(hipStreamSynchronize(c.stream));
        const int p = c.p;
        const size_t stride = static_cast<size_t>((p + 1) * (2 * p + 1));
        const double wx0 = t.box[3] - t.box[0];
        for (size_t g = 0; g < nc; ++g) {
            const size_t r = static_cast<size_t>(rid[g]);
            const bool has = lev[g] >= 2 && cnt[g] > 0;
            const double w = std::ldexp(wx0, -lev[g]);
            for (int n = 0; n <= p; ++n) {
                const double ms = std::pow(w, n), ls = 1.0 / std::pow(w, n + 1);  // undo the level scaling
                for (int m = 0; m <= n; ++m) {
                    const double mre = has ? Mh[g * K + vg::kidx_re(n, m)] * ms : 0.0, mim = has && m ? Mh[g * K + vg::kidx_im(n, m)] * ms : 0.0;
                    const double lre = has ? Lh[g * K + vg::kidx_re(n, m)] * ls : 0.0, lim = has && m ? Lh[g * K + vg::kidx_im(n, m)] * ls : 0.0;
                    const size_t o_p = r * stride + static_cast<size_t>(n * (2 * p + 1) + p + m);
                    const size_t o_m = r * stride + static_cast<size_t>(n * (2 * p + 1) + p - m);
                    if (M) {
                        M[2 * o_p] = mre;
                        M[2 * o_p + 1] = mim;
                        M[2 * o_m] = mre;
                        M[2 * o_m + 1] = -mim;
                    }
                    if (L) {
                        L[2 * o_p] = lre;
                        L[2 * o_p + 1] = lim;
                        L[2 * o_m] = lre;
                        L[2 * o_m + 1] = -lim;
                    }
                }
            }
        }
    });
}

int vipp_grav_counts(vipp_grav_ctx *ctx, double *out, int32_t n) {
    return guarded(ctx, [&] {
        need(ctx != nullptr && out != nullptr, VIPP_GRAV_E_INVALID, "null argument");
        vg::Ctx &c = ctx->c;
        VG_HIP(hipSetDevice(c.device));
        need(c.uploaded, VIPP_GRAV_E_STATE, "no particles uploaded");
        if (!c.evaluated) vg::evaluate(c, false);
        vg::count_work(c);
        c.counts[6] = static_cast<double>(c.cl.on ? c.cl.nflag : 0);
        for (int i = 0; i < n && i < 8; ++i) out[i] = c.counts[i];
    });
}

int vipp_grav_phase_times(vipp_grav_ctx *ctx, double *ms, int32_t n) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        vg::Ctx &c = ctx->c;
        if (!ms) {
            for (auto &v : c.phase_ms) v = 0;
            return;
        }
        for (int i = 0; i < n && i < vg::PH_COUNT; ++i) ms[i] = c.phase_ms[i];
    });
}

int vipp_grav_set_timing(vipp_grav_ctx *ctx, int32_t enabled) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        ctx->c.timing = enabled != 0;
    });
}

int vipp_grav_device_state(vipp_grav_ctx *ctx, void **x, void **v) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, VIPP_GRAV_E_INVALID, "null context");
        if (x) *x = ctx->c.x.p;
        if (v) *v = ctx->c.v.p;
    });
}

}  // extern "C"
