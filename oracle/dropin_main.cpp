// dropin_main.cpp - the reference's own validation hook driving GpuSolver.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) from the
// reference's headers where they lie (-I) and include/gravity/solver_gpu.hpp:
//   _ref/dropin      as is;
//   _ref/dropin_out  with the reference's ENABLE_PRINT and GRAVITY_CHECK_TIME
//                    output switches (benchmark.hpp:13-21,119-134,172-174);
//   _ref/dropin_sdt  with GRAVITY_SEARCH_DELTA_T and
//                    GRAVITY_SUBDIVIDE_REF_TIMESTEP=2, the build of
//                    experiment/search_timestep.sh (benchmark.hpp:105-106,
//                    135-146,154-156; problem.hpp:235-238).
// It does what the reference's main.cpp does (main.cpp:33-40) with
// gravity::Solver replaced by gravity::GpuSolver.
//
//   dropin N s bench <types> <precision> [epsilon]
//     Benchmark::benchmark (benchmark.hpp:90-175): one GPU RK4 step of
//     delta_t / t per t against the reference's FP64 SolverRef on the CPU,
//     E(T) = alpha T^c fit and the reference's pass/fail verdict.
//   dropin N s precision <types> <reference>
//     the ENABLE_DIFFERENT_PRECISION build (main.cpp:33-36,
//     experiment/different_precision.sh): Benchmark::benchmark_different_precision
//     (benchmark.hpp:31-87) of the GPU solver against a float reference solver.
//
// <types> is one of main.cpp's type sets (main.cpp:9-32): double (the default
// build), float (GRAVITY_FLOAT) or time_double (GRAVITY_TIME_DOUBLE: float
// with a double time type).  <precision> is the GPU arithmetic: auto (FP64
// when the acceleration type is double, else FP32), f64, f32 or split.
// <reference> is the float reference solver of the different-precision run:
// f32 or split for GpuSolver<float, ...> (what the alias swap of solver.hpp
// gives main.cpp), or cpu for the reference's own SolverRef<float, ...> on the
// CPU.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <iostream>
#include <iterator>
#include <limits>
#include <optional>
#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <variant>
#include <vector>

#define ENABLE_DIFFERENT_PRECISION  // declares Benchmark::benchmark_different_precision (benchmark.hpp:31)
#include <benchmark.hpp>

#include "gravity/solver_gpu.hpp"

namespace {

gravity::GpuSolverOptions options(const std::string &precision) {
    gravity::GpuSolverOptions opt;
    if (precision == "auto") return opt;
    opt.has_precision = true;
    if (precision == "f64")
        opt.precision = gravity::GpuPrecision::f64;
    else if (precision == "f32")
        opt.precision = gravity::GpuPrecision::f32;
    else if (precision == "split")
        opt.precision = gravity::GpuPrecision::split;
    else
        throw std::invalid_argument("unknown precision " + precision);
    return opt;
}

template <class M, class T, class X, class V, class A, class G>
void bench(const std::size_t n, const std::size_t s, const char *eps, const gravity::GpuSolverOptions &opt) {
    gravity::GpuSolver<M, T, X, V, A, G> solver(opt);
    const std::optional<X> epsilon = eps ? std::optional<X>(static_cast<X>(std::stod(eps))) : std::nullopt;  // main.cpp:38
    gravity::Benchmark<M, T, X, V, A, G>().benchmark(solver, n, s, epsilon);
}

template <class M, class T, class X, class V, class A, class G>
void precision(const std::size_t n, const std::size_t s, const std::string &ref) {
    gravity::GpuSolver<M, T, X, V, A, G> solver;
    if (ref == "cpu") {
        gravity::SolverRef<float, float, float, float, float, float> reference;
        gravity::Benchmark<M, T, X, V, A, G>().benchmark_different_precision(solver, reference, n, s);
    } else {
        gravity::GpuSolver<float, float, float, float, float, float> reference(options(ref));
        gravity::Benchmark<M, T, X, V, A, G>().benchmark_different_precision(solver, reference, n, s);
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 6) {
        std::cerr << "usage: dropin N s bench double|float|time_double auto|f64|f32|split [epsilon]\n"
                     "       dropin N s precision double|float|time_double f32|split|cpu\n";
        return 2;
    }
    const std::size_t n = std::stoul(argv[1]);
    const std::size_t s = std::stoul(argv[2]);
    const std::string mode = argv[3], types = argv[4], arg = argv[5];
    const char *eps = argc > 6 ? argv[6] : nullptr;
    try {
        if (mode == "bench") {
            const auto opt = options(arg);
            if (types == "double")
                bench<double, double, double, double, double, double>(n, s, eps, opt);
            else if (types == "float")
                bench<float, float, float, float, float, float>(n, s, eps, opt);
            else if (types == "time_double")
                bench<float, double, float, float, float, float>(n, s, eps, opt);
            else
                throw std::invalid_argument("unknown type set " + types);
        } else if (mode == "precision") {
            if (arg != "f32" && arg != "split" && arg != "cpu") throw std::invalid_argument("unknown reference " + arg);
            if (types == "double")
                precision<double, double, double, double, double, double>(n, s, arg);
            else if (types == "float")
                precision<float, float, float, float, float, float>(n, s, arg);
            else if (types == "time_double")
                precision<float, double, float, float, float, float>(n, s, arg);
            else
                throw std::invalid_argument("unknown type set " + types);
        } else {
            throw std::invalid_argument("unknown mode " + mode);
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
