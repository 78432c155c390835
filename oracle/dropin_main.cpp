// dropin_main.cpp - the reference's own validation hook driving GpuSolver.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) into
// oracle/_ref/dropin from the reference's headers where they lie (-I) and
// include/gravity/solver_gpu.hpp.  It does what the reference's main.cpp does
// (main.cpp:33-40) with gravity::Solver replaced by gravity::GpuSolver: the
// reference's Benchmark::benchmark (benchmark.hpp:90-175) runs the E(T) test,
// comparing one GPU RK4 step of delta_t / t, t = 1, 2, 4, 8, 16, with the
// reference's FP64 SolverRef on the CPU, fits E(T) = alpha T^c and prints its
// pass/fail verdict.
//
//   dropin N s f64|f32|split [epsilon]
//
// With "precision" it does what main.cpp:33-36 does when built with
// ENABLE_DIFFERENT_PRECISION (experiment/different_precision.sh), again with
// gravity::Solver replaced by gravity::GpuSolver: the reference's
// Benchmark::benchmark_different_precision (benchmark.hpp:31-87) runs the
// solver of the chosen type set (main.cpp:9-32: double, float, or float with
// a double time type) against GpuSolver<float, ...> as its reference.
//
//   dropin N s precision double|float|time_double [f32|split]
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

template <class F>
static void run(const std::size_t n, const std::size_t s, const std::optional<F> eps, gravity::GpuSolverOptions opt) {
    gravity::GpuSolver<F, F, F, F, F, F> solver(opt);
    gravity::Benchmark<F, F, F, F, F, F>().benchmark(solver, n, s, eps);
}

// main.cpp:33-36 with gravity::Solver = gravity::GpuSolver; opt_ref sets the
// precision of the FP32 reference solver (f32 or split)
template <class M, class T, class X, class V, class A, class G>
static void run_precision(const std::size_t n, const std::size_t s, gravity::GpuSolverOptions opt_ref) {
    gravity::GpuSolver<M, T, X, V, A, G> solver;
    gravity::GpuSolver<float, float, float, float, float, float> reference(opt_ref);
    gravity::Benchmark<M, T, X, V, A, G>().benchmark_different_precision(solver, reference, n, s);
}

int main(int argc, char **argv) {
    if (argc < 4) {
        std::cerr << "usage: dropin N s f64|f32|split [epsilon]\n";
        return 2;
    }
    const std::size_t n = std::stoul(argv[1]);
    const std::size_t s = std::stoul(argv[2]);
    const std::string mode = argv[3];
    gravity::GpuSolverOptions opt;
    try {
        if (mode == "precision") {
            const std::string types = argc > 4 ? argv[4] : "double";
            const std::string ref = argc > 5 ? argv[5] : "f32";
            gravity::GpuSolverOptions opt_ref;
            opt_ref.has_precision = true;
            opt_ref.precision = ref == "split" ? gravity::GpuPrecision::split : gravity::GpuPrecision::f32;
            if (types == "double")
                run_precision<double, double, double, double, double, double>(n, s, opt_ref);
            else if (types == "float")
                run_precision<float, float, float, float, float, float>(n, s, opt_ref);
            else if (types == "time_double")
                run_precision<float, double, float, float, float, float>(n, s, opt_ref);
            else {
                std::cerr << "unknown type set " << types << "\n";
                return 2;
            }
        } else if (mode == "f64") {
            run<double>(n, s, argc > 4 ? std::optional<double>(std::stod(argv[4])) : std::nullopt, opt);
        } else if (mode == "f32" || mode == "split") {
            opt.has_precision = true;
            opt.precision = mode == "f32" ? gravity::GpuPrecision::f32 : gravity::GpuPrecision::split;
            run<float>(n, s, argc > 4 ? std::optional<float>(std::stof(argv[4])) : std::nullopt, opt);
        } else {
            std::cerr << "unknown mode " << mode << "\n";
            return 2;
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
