// solver_gpu.hpp - gravity::GpuSolver, a header-only drop-in for the
// reference's solver plug-in point over the C ABI of vipp_gravity.h.
//
// The reference selects its solver with
//     template <class M, class T, class X, class V, class A, class G>
//     using Solver = SolverRef<M, T, X, V, A, G>;          (solver.hpp:7-8)
// and drives it through preprocess / solve / postprocess (solver_ref.hpp:11-27,
// called by benchmark.hpp:118-127).  GpuSolver keeps those three signatures:
//   preprocess  uploads x and v (the FP64 state lives on the GPU from here on),
//   solve       runs one RK4 step of problem.delta_t() / t on the GPU and
//               returns after the device finished (so a caller's timer around
//               solve covers the GPU work, as benchmark.hpp:124-126 expects),
//   postprocess downloads x and v into the caller's vectors.
// Unlike SolverRef, x and v are only updated by postprocess.
//
// The Problem and Vector templates are the caller's (problem.hpp:32-154,
// 165-292); this header only declares them, and uses N(), s(), p(),
// epsilon(), get_G(), m() and delta_t() of Problem and x(), y(), z() of Vector.
// Errors of the library are thrown as std::runtime_error with its message,
// like the reference's own checks (problem.hpp:179-184, octree.hpp:87-89).
#ifndef VIPP_GRAVITY_SOLVER_GPU_HPP
#define VIPP_GRAVITY_SOLVER_GPU_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../vipp_gravity.h"

namespace gravity {

template <class M, class T, class X, class V, class A, class G>
struct Problem;
template <class T>
struct Vector;

enum class GpuPrecision : uint32_t {
    f64 = VIPP_GRAV_F64,          // FP64 everywhere
    f32 = VIPP_GRAV_F32,          // FP32 P2P and M2L on the vector units, FP64 state
    split = VIPP_GRAV_F32_SPLIT,  // FP32 P2P, BF16 hi/lo M2L on MFMA, FP64 state
};

struct GpuSolverOptions {
    int device = 0;
    // default: f64 when the acceleration type is double, f32 otherwise
    bool has_precision = false;
    GpuPrecision precision = GpuPrecision::f32;
    // false: G multiplies the near field only, as the reference's FMM does
    // (fmm.hpp:539 versus l2p, fmm.hpp:332-374); true: G on the far field too
    bool far_field_G = false;
};

template <class M, class T, class X, class V, class A, class G>
class GpuSolver final {
   public:
    using problem_type = Problem<M, T, X, V, A, G>;

    explicit GpuSolver(GpuSolverOptions opt = {}) : opt_(opt) {}
    ~GpuSolver() { release(); }
    GpuSolver(const GpuSolver &) = delete;
    GpuSolver &operator=(const GpuSolver &) = delete;
    GpuSolver(GpuSolver &&o) noexcept : opt_(o.opt_), ctx_(std::exchange(o.ctx_, nullptr)), key_(o.key_) {}
    GpuSolver &operator=(GpuSolver &&o) noexcept {
        if (this != &o) {
            release();
            opt_ = o.opt_;
            ctx_ = std::exchange(o.ctx_, nullptr);
            key_ = o.key_;
        }
        return *this;
    }

    /// Upload the state; (re)creates the device context when N, s, p, the
    /// softening, G or m changed since the last call.
    void preprocess(const problem_type &problem, const std::vector<Vector<X>> &x, const std::vector<Vector<V>> &v) {
        const std::size_t n = problem.N();
        if (x.size() != n || v.size() != n) throw std::runtime_error("x and v must hold N particles");
        ensure_context(problem);
        stage_.resize((flat<X>() ? 0 : 3 * n) + (flat<V>() ? 0 : 3 * n));
        double *spare = stage_.data();
        const double *xs = pack(x, spare);
        const double *vs = pack(v, spare);
        check(vipp_grav_upload(ctx_, xs, vs, nullptr));
    }

    /// One RK4 step of problem.delta_t() / t on the device (solver_ref.hpp:21-26).
    void solve(const problem_type &problem, const T t, std::vector<Vector<X>> &, std::vector<Vector<V>> &) {
        if (!ctx_) throw std::runtime_error("GpuSolver::solve called before preprocess");
        const T dt = problem.delta_t() / t;
        check(vipp_grav_step(ctx_, static_cast<double>(dt), 1));
    }

    /// Download the state into x and v.
    void postprocess(const problem_type &problem, std::vector<Vector<X>> &x, std::vector<Vector<V>> &v) {
        if (!ctx_) throw std::runtime_error("GpuSolver::postprocess called before preprocess");
        const std::size_t n = problem.N();
        x.resize(n, Vector<X>{});
        v.resize(n, Vector<V>{});
        stage_.resize((flat<X>() ? 0 : 3 * n) + (flat<V>() ? 0 : 3 * n));
        double *spare = stage_.data();
        double *xs = target(x, spare);
        double *vs = target(v, spare);
        check(vipp_grav_download(ctx_, xs, vs));
        unpack(xs, x);
        unpack(vs, v);
    }

    /// The library context (for the extra queries of vipp_gravity.h), or null before preprocess.
    vipp_grav_ctx *context() const noexcept { return ctx_; }

   private:
    // Vector<double> is three contiguous doubles (problem.hpp:143-146), so a
    // std::vector of them is the ABI's double[3N] as it is and crosses it
    // without a staging copy; other coordinate types are converted through
    // stage_.
    template <class U>
    static constexpr bool flat() {
        return std::is_same<U, double>::value && std::is_standard_layout<Vector<U>>::value && sizeof(Vector<U>) == 3 * sizeof(double);
    }
    template <class U>
    static const double *pack(const std::vector<Vector<U>> &a, double *&spare) {
        if constexpr (flat<U>()) {
            return reinterpret_cast<const double *>(a.data());
        } else {
            double *out = spare;
            for (std::size_t i = 0; i < a.size(); ++i) {
                out[3 * i] = static_cast<double>(a[i].x());
                out[3 * i + 1] = static_cast<double>(a[i].y());
                out[3 * i + 2] = static_cast<double>(a[i].z());
            }
            spare += 3 * a.size();
            return out;
        }
    }
    template <class U>
    static double *target(std::vector<Vector<U>> &a, double *&spare) {
        if constexpr (flat<U>()) {
            return reinterpret_cast<double *>(a.data());
        } else {
            double *out = spare;
            spare += 3 * a.size();
            return out;
        }
    }
    template <class U>
    static void unpack(const double *in, std::vector<Vector<U>> &a) {
        if constexpr (!flat<U>())
            for (std::size_t i = 0; i < a.size(); ++i)
                a[i] = Vector<U>{static_cast<U>(in[3 * i]), static_cast<U>(in[3 * i + 1]), static_cast<U>(in[3 * i + 2])};
    }

    struct Key {
        std::size_t n = 0, s = 0;
        unsigned p = 0;
        double eps = 0, grav = 0, m = 0;
        bool operator==(const Key &o) const { return n == o.n && s == o.s && p == o.p && eps == o.eps && grav == o.grav && m == o.m; }
    };

    void ensure_context(const problem_type &problem) {
        Key k;
        k.n = problem.N();
        k.s = problem.s();
        k.p = problem.p();
        k.eps = static_cast<double>(problem.epsilon());
        k.grav = static_cast<double>(problem.get_G());
        k.m = static_cast<double>(problem.m());
        if (ctx_ && k == key_) return;
        release();
        vipp_grav_config cfg{};
        cfg.n = k.n;
        cfg.s = k.s;
        cfg.p = k.p;
        cfg.precision = static_cast<uint32_t>(opt_.has_precision ? opt_.precision
                                              : std::is_same<A, double>::value ? GpuPrecision::f64
                                                                               : GpuPrecision::f32);
        cfg.epsilon = k.eps;
        cfg.G = k.grav;
        cfg.mass = k.m;
        cfg.device = opt_.device;
        cfg.flags = opt_.far_field_G ? 0u : VIPP_GRAV_FLAG_REF_FAR_G;
        vipp_grav_ctx *c = nullptr;
        const int rc = vipp_grav_create(&cfg, &c);
        if (rc != VIPP_GRAV_OK) throw std::runtime_error(std::string("vipp_grav_create: ") + vipp_grav_last_error(nullptr));
        ctx_ = c;
        key_ = k;
    }

    void check(int rc) const {
        if (rc != VIPP_GRAV_OK) throw std::runtime_error(vipp_grav_last_error(ctx_));
    }

    void release() noexcept {
        if (ctx_) vipp_grav_destroy(ctx_);
        ctx_ = nullptr;
    }

    GpuSolverOptions opt_;
    vipp_grav_ctx *ctx_ = nullptr;
    Key key_;
    std::vector<double> stage_;
};

}  // namespace gravity

#endif  // VIPP_GRAVITY_SOLVER_GPU_HPP
