// problem.cpp - the reference's initial conditions, as a workload generator of
// the library (host code only: the stream is one sequential mt19937_64 with
// rejection sampling, nothing a device could run in parallel).
//
// Problem::initialize (gravity/include/problem.hpp:188-211) draws, per
// particle and in this order from one std::mt19937_64 seeded with the
// problem's seed: the mass fraction u on [0, 1 - epsilon), Aarseth's
// rejection pairs (x1, x2) until P_MAX x1 < x2^2 (1 - x2^2)^(7/2), two
// uniforms for the position's direction and two for the velocity's.  The
// radius is the Plummer inverse mass profile r(u) = u^(1/3) / sqrt(1 - u^(2/3))
// (problem.hpp:250-252), the escape speed sqrt(2) (r^2 + 1)^(-1/4)
// (problem.hpp:254-256), and a norm is split into components by
// x = (1 - 2 X) norm, (y, z) = sqrt(norm^2 - x^2) (cos, sin)(2 pi X')
// (problem.hpp:277-284).  The default softening, also the truncation of u, is
// 0.63 N^-0.22 (problem.hpp:290-292).
//
// The stream must be the reference's bit for bit (tests compare against
// vectors the reference's own code wrote), so every operation is evaluated in
// the reference's order with the same library calls (libstdc++'s
// uniform_real_distribution, glibc's pow/sqrt/sincos) and without FMA
// contraction: this file is built by the host compiler the reference is built
// with (g++, -ffp-contract=off), which pairs cos and sin of one angle into one
// sincos call as it does in the reference; another compiler's separate cos and
// sin calls differ from it in the last bit now and then.
#include <cmath>
#include <cstdint>
#include <random>

#include "../../include/vipp_gravity.h"

namespace {

struct PlummerStream {
    std::mt19937_64 gen;
    double eps;
    std::uniform_real_distribution<double> mass_fraction, unit;
    static constexpr double kTwoPi = 3.141592653589793 * 2;  // Constant<double>::PI (problem.hpp:22) times 2

    PlummerStream(uint64_t seed, double epsilon) : gen(seed), eps(epsilon), mass_fraction(0, 1.0 - epsilon), unit(0, 1) {}

    static double radius(double u) { return std::pow(u, 1.0 / 3) / std::pow(1.0 - std::pow(u, 2.0 / 3), 1.0 / 2); }
    static double escape_speed(double r) { return std::sqrt(2.0) * std::pow(r * r + 1, -1.0 / 4); }
    static double density(double q) { return q * q * std::pow(1.0 - q * q, 7.0 / 2); }

    // fraction of the escape speed, by rejection under the bound sqrt(7) 686 / 19683
    double speed_fraction() {
        const double bound = std::sqrt(7.0) * 686 / 19683;
        for (;;) {
            const double a = unit(gen);
            const double b = unit(gen);
            if (bound * a < density(b)) return b;
        }
    }

    static void split(double norm, double along, double angle, double *out) {
        const double x = (1.0 - along * 2) * norm;
        const double perp = std::sqrt(norm * norm - x * x);
        out[0] = x;
        out[1] = perp * std::cos(kTwoPi * angle);
        out[2] = perp * std::sin(kTwoPi * angle);
    }

    void particle(double *x, double *v) {
        const double r = radius(mass_fraction(gen));
        const double speed = speed_fraction() * escape_speed(r);
        const double px = unit(gen);
        const double pa = unit(gen);
        split(r, px, pa, x);
        const double vx = unit(gen);
        const double va = unit(gen);
        split(speed, vx, va, v);
    }
};

}  // namespace

extern "C" int vipp_grav_problem_initialize(uint64_t n, uint64_t seed, double epsilon, double *x, double *v, double *epsilon_used) {
    if (n == 0 || !x || !v) return VIPP_GRAV_E_INVALID;
    const double eps = epsilon < 0 ? 63.0 / 100 * std::pow(static_cast<double>(n), -22.0 / 100) : epsilon;
    if (!(eps < 1.0) || !std::isfinite(eps)) return VIPP_GRAV_E_INVALID;
    PlummerStream st(seed, eps);
    for (uint64_t i = 0; i < n; ++i) st.particle(x + 3 * i, v + 3 * i);
    if (epsilon_used) *epsilon_used = eps;
    return VIPP_GRAV_OK;
}
