// solver_gpu_test.cpp - gravity::GpuSolver through its three calls, against
// the C ABI driven directly, with minimal Problem / Vector types of this test
// that offer the members the wrapper uses (the reference's own types are
// exercised by oracle/dropin_main.cpp).  Prints "ok" and exits 0 on success.
#include <cmath>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <vector>

#include "gravity/solver_gpu.hpp"

namespace gravity {
template <class T>
struct Vector {
    Vector() : x_{}, y_{}, z_{} {}
    Vector(T x, T y, T z) : x_(x), y_(y), z_(z) {}
    T x() const { return x_; }
    T y() const { return y_; }
    T z() const { return z_; }

   private:
    T x_, y_, z_;
};
template <class M, class T, class X, class V, class A, class G>
struct Problem {
    std::size_t n, s_;
    unsigned p_;
    X eps;
    std::size_t N() const { return n; }
    std::size_t s() const { return s_; }
    unsigned p() const { return p_; }
    X epsilon() const { return eps; }
    G get_G() const { return G{1}; }
    M m() const { return M{1} / static_cast<M>(n); }
    T delta_t() const { return static_cast<T>(eps / std::sqrt(2.0)); }
};
}  // namespace gravity

using D = double;
using Prob = gravity::Problem<D, D, D, D, D, D>;
using Vec = gravity::Vector<D>;

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                                   \
        }                                                               \
    } while (0)

int main() {
    const std::size_t n = 3000;
    std::mt19937_64 rng(3);
    std::uniform_real_distribution<double> u(-1, 1);
    std::vector<Vec> x, v;
    std::vector<double> xr(3 * n), vr(3 * n);
    for (std::size_t i = 0; i < n; ++i) {
        x.emplace_back(u(rng), u(rng), u(rng));
        v.emplace_back(0.1 * u(rng), 0.1 * u(rng), 0.1 * u(rng));
        xr[3 * i] = x[i].x(), xr[3 * i + 1] = x[i].y(), xr[3 * i + 2] = x[i].z();
        vr[3 * i] = v[i].x(), vr[3 * i + 1] = v[i].y(), vr[3 * i + 2] = v[i].z();
    }
    const Prob prob{n, 8, 5, 1e-3};

    // the wrapper: one RK4 step of delta_t / 2
    gravity::GpuSolver<D, D, D, D, D, D> solver;
    solver.preprocess(prob, x, v);
    vipp_grav_ctx *first = solver.context();
    solver.solve(prob, 2.0, x, v);
    solver.postprocess(prob, x, v);

    // the same through the C ABI
    vipp_grav_config cfg{};
    cfg.n = n, cfg.s = 8, cfg.p = 5, cfg.precision = VIPP_GRAV_F64, cfg.epsilon = 1e-3, cfg.G = 1, cfg.mass = 1.0 / n;
    cfg.flags = VIPP_GRAV_FLAG_REF_FAR_G;
    vipp_grav_ctx *c = nullptr;
    CHECK(vipp_grav_create(&cfg, &c) == VIPP_GRAV_OK);
    CHECK(vipp_grav_upload(c, xr.data(), vr.data(), nullptr) == VIPP_GRAV_OK);
    CHECK(vipp_grav_step(c, prob.delta_t() / 2.0, 1) == VIPP_GRAV_OK);
    CHECK(vipp_grav_download(c, xr.data(), vr.data()) == VIPP_GRAV_OK);
    vipp_grav_destroy(c);
    for (std::size_t i = 0; i < n; ++i) {
        CHECK(x[i].x() == xr[3 * i] && x[i].y() == xr[3 * i + 1] && x[i].z() == xr[3 * i + 2]);
        CHECK(v[i].x() == vr[3 * i] && v[i].y() == vr[3 * i + 1] && v[i].z() == vr[3 * i + 2]);
    }

    // same problem again: the context is kept
    solver.preprocess(prob, x, v);
    CHECK(solver.context() == first);

    // the float type set (main.cpp:9-16): coordinates are converted to and
    // from the ABI's doubles, the device computes in FP32
    using F = float;
    std::vector<gravity::Vector<F>> xf, vf;
    for (std::size_t i = 0; i < n; ++i) {
        xf.emplace_back(static_cast<F>(x[i].x()), static_cast<F>(x[i].y()), static_cast<F>(x[i].z()));
        vf.emplace_back(static_cast<F>(v[i].x()), static_cast<F>(v[i].y()), static_cast<F>(v[i].z()));
        xr[3 * i] = xf[i].x(), xr[3 * i + 1] = xf[i].y(), xr[3 * i + 2] = xf[i].z();
        vr[3 * i] = vf[i].x(), vr[3 * i + 1] = vf[i].y(), vr[3 * i + 2] = vf[i].z();
    }
    const gravity::Problem<F, F, F, F, F, F> probf{n, 8, 5, 1e-3f};
    gravity::GpuSolver<F, F, F, F, F, F> solverf;
    solverf.preprocess(probf, xf, vf);
    solverf.solve(probf, 2.0f, xf, vf);
    solverf.postprocess(probf, xf, vf);
    cfg.precision = VIPP_GRAV_F32, cfg.epsilon = static_cast<double>(1e-3f), cfg.mass = static_cast<double>(F{1} / static_cast<F>(n));
    CHECK(vipp_grav_create(&cfg, &c) == VIPP_GRAV_OK);
    CHECK(vipp_grav_upload(c, xr.data(), vr.data(), nullptr) == VIPP_GRAV_OK);
    CHECK(vipp_grav_step(c, static_cast<double>(probf.delta_t() / 2.0f), 1) == VIPP_GRAV_OK);
    CHECK(vipp_grav_download(c, xr.data(), vr.data()) == VIPP_GRAV_OK);
    vipp_grav_destroy(c);
    for (std::size_t i = 0; i < n; ++i) {
        CHECK(xf[i].x() == static_cast<F>(xr[3 * i]) && xf[i].y() == static_cast<F>(xr[3 * i + 1]) && xf[i].z() == static_cast<F>(xr[3 * i + 2]));
        CHECK(vf[i].x() == static_cast<F>(vr[3 * i]) && vf[i].y() == static_cast<F>(vr[3 * i + 1]) && vf[i].z() == static_cast<F>(vr[3 * i + 2]));
    }

    // errors surface as std::runtime_error
    bool threw = false;
    try {
        std::vector<Vec> short_x(n - 1);
        solver.preprocess(prob, short_x, v);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        gravity::GpuSolver<D, D, D, D, D, D> bad;
        const Prob p0{n, 0, 5, 1e-3};  // s = 0 (problem.hpp:182-184)
        bad.preprocess(p0, x, v);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        gravity::GpuSolver<D, D, D, D, D, D> unprepared;
        unprepared.solve(prob, 1.0, x, v);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    CHECK(threw);
    std::printf("ok\n");
    return 0;
}
