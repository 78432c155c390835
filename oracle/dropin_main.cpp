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

#include <benchmark.hpp>

#include "gravity/solver_gpu.hpp"

template <class F>
static void run(const std::size_t n, const std::size_t s, const std::optional<F> eps, gravity::GpuSolverOptions opt) {
    gravity::GpuSolver<F, F, F, F, F, F> solver(opt);
    gravity::Benchmark<F, F, F, F, F, F>().benchmark(solver, n, s, eps);
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
        if (mode == "f64") {
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
