// ref_driver.cpp - C-ABI shim over the reference's own gravity headers.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) into
// oracle/_ref/libvipp_ref.so from the reference sources where they lie
// (gravity/include of the reference repository, given by -I).  Nothing of the
// reference is copied here: this file only instantiates the reference's
// templates in double precision and reads their results through the friend
// hooks the reference declares for its tests under ENABLE_TEST
// (octree.hpp:123-125, fmm.hpp:495-497, solver_ref.hpp:29-32,
// solver_ref_naive.hpp:27-29, problem.hpp:245-247, benchmark.hpp:177-179).
//
// The reference headers rely on transitive includes of libc++; the standard
// headers they use are included up front so that they compile with libstdc++.
#define ENABLE_TEST
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
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
#include <solver_ref.hpp>
#include <solver_ref_naive.hpp>

namespace gravity {
using D = double;
using CellD = Cell<D, D, D, D, D, D>;
using TreeD = Octree<CellD, Vector<D>>;
using SolverD = SolverRef<D, D, D, D, D, D>;
using NaiveD = SolverRefNaive<D, D, D, D, D, D>;
using ProblemD = Problem<D, D, D, D, D, D>;
using BenchD = Benchmark<D, D, D, D, D, D>;
using Lists = std::vector<std::vector<std::size_t>>;

struct FMMTest {
    static void link(const TreeD &tree, Lists &ap, Lists &ni, Lists &nb) {
        const auto &cells = tree.cells();
        ap.assign(cells.size(), {});
        ni.assign(cells.size(), {});
        nb.assign(cells.size(), {});
        SolverD::calc_link(cells.front(), cells, ap, ni, nb, std::size_t{0});
    }
    static void fmm(const std::vector<Vector<D>> &x, D G, D m, D eps, std::vector<Vector<D>> &a, unsigned p, TreeD &tree, const Lists &ap,
                    const Lists &ni) {
        SolverD::fmm(x, G, m, eps, a, p, tree, ap, ni);
    }
    static std::complex<D> M(const CellD &c, unsigned p, unsigned k, int m) { return c._multipole.empty() ? std::complex<D>{} : c.multipole(p, k, m); }
    static std::complex<D> L(const CellD &c, unsigned p, unsigned k, int m) { return c._local.empty() ? std::complex<D>{} : c.local(p, k, m); }
    static std::vector<CellD> &cells(TreeD &t) { return t._cells; }
    static bool has_local(const CellD &c) { return !c._local.empty(); }
    static void presize(CellD &c, unsigned p) {
        c._multipole.assign(p + 1, std::vector<std::complex<D>>(2 * p + 1));
        c._local.assign(p + 1, std::vector<std::complex<D>>(2 * p + 1));
    }
    static void divide(TreeD &t, std::size_t idx, std::size_t thr) { t.subdivide(idx, thr); }
    // the pieces of SolverRef::fmm (solver_ref.hpp:180-232), split for timing
    static void p2m(CellD &c, unsigned p, const std::vector<Vector<D>> &x, D m) { gravity::p2m(c, p, x, m); }
    static void m2m(CellD &c, unsigned p, const std::vector<CellD> &cells) { gravity::m2m(c, p, cells); }
    static void m2l(CellD &c, unsigned p, const std::vector<CellD> &cells, const std::vector<std::size_t> &ap) { gravity::m2l(c, p, cells, ap); }
    static void l2l(CellD &c, unsigned p, const std::vector<CellD> &cells) { gravity::l2l(c, p, cells); }
    static void l2p(const CellD &c, unsigned p, const std::vector<Vector<D>> &x, std::vector<Vector<D>> &a) { gravity::l2p(c, p, x, a); }
    static void p2p(const CellD &c, const std::vector<Vector<D>> &x, std::vector<Vector<D>> &a, const std::vector<CellD> &cells,
                    const std::vector<std::size_t> &ni, D G, D m, D eps) {
        gravity::p2p(c, x, a, cells, ni, G, m, eps);
    }
};
struct SolverRefTest {
    static void rk4(std::vector<Vector<D>> &x, std::vector<Vector<D>> &v, D G, D m, D eps, D dt, std::size_t s, unsigned p) {
        SolverD::runge_kutta(x, v, G, m, eps, dt, s, p);
    }
};
struct SolverRefNaiveTest {
    static void rk4(std::vector<Vector<D>> &x, std::vector<Vector<D>> &v, D G, D m, D eps, D dt) { NaiveD::runge_kutta(x, v, G, m, eps, dt); }
    static Vector<D> acc(const std::vector<Vector<D>> &x, D G, D m, D eps, std::size_t i) { return NaiveD::acceleration(x, G, m, eps, i); }
};
struct BenchmarkTest {
    static std::tuple<D, D> fit(const std::vector<D> &e, const std::vector<D> &t) { return BenchD::curve_fitting(e, t); }
};
}  // namespace gravity

namespace {
using namespace gravity;
thread_local std::string g_err;

std::vector<Vector<D>> to_vec(const double *x, int64_t n) {
    std::vector<Vector<D>> v;
    v.reserve(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) v.emplace_back(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    return v;
}
void from_vec(const std::vector<Vector<D>> &v, double *x) {
    for (size_t i = 0; i < v.size(); ++i) {
        x[3 * i] = v[i].x();
        x[3 * i + 1] = v[i].y();
        x[3 * i + 2] = v[i].z();
    }
}
template <class F>
int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
}
}  // namespace

struct ref_tree {
    std::vector<Vector<D>> x;
    TreeD *tree = nullptr;
    Lists ap, ni, nb;
    ~ref_tree() { delete tree; }
};

extern "C" {

const char *ref_last_error(void) { return g_err.c_str(); }

int ref_tree_build(const double *x, int64_t n, int64_t s, ref_tree **out) {
    return guarded([&] {
        auto *t = new ref_tree;
        t->x = to_vec(x, n);
        try {
            t->tree = new TreeD{t->x, static_cast<std::size_t>(s)};
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
    });
}

void ref_tree_free(ref_tree *t) { delete t; }

int64_t ref_tree_ncells(const ref_tree *t) { return static_cast<int64_t>(t->tree->cells().size()); }

void ref_tree_export(const ref_tree *t, int64_t *parent, int32_t *level, double *bmin, double *bmax, int64_t *child_base, int64_t *pt_off,
                     int64_t *pts) {
    const auto &cells = t->tree->cells();
    int64_t off = 0;
    for (size_t c = 0; c < cells.size(); ++c) {
        const auto &cell = cells[c];
        parent[c] = cell.parent() ? static_cast<int64_t>(cell.parent().value()) : -1;
        int32_t lev = 0;
        for (auto q = cell.parent(); q; q = cells[q.value()].parent()) ++lev;
        level[c] = lev;
        bmin[3 * c] = cell.min().x();
        bmin[3 * c + 1] = cell.min().y();
        bmin[3 * c + 2] = cell.min().z();
        bmax[3 * c] = cell.max().x();
        bmax[3 * c + 1] = cell.max().y();
        bmax[3 * c + 2] = cell.max().z();
        pt_off[c] = off;
        if (cell.is_leaf()) {
            child_base[c] = -1;
            for (const auto i : cell.points()) pts[off++] = static_cast<int64_t>(i);
        } else {
            child_base[c] = static_cast<int64_t>(*cell.cells().begin());
        }
    }
    pt_off[cells.size()] = off;
}

int ref_tree_subdivide(ref_tree *t, int64_t idx, int64_t threshold) {
    return guarded([&] {
        FMMTest::divide(*t->tree, static_cast<std::size_t>(idx), static_cast<std::size_t>(threshold));
        t->ap.clear();
    });
}

int ref_tree_lists(ref_tree *t, int64_t *ap_total, int64_t *ni_total, int64_t *nb_total) {
    return guarded([&] {
        FMMTest::link(*t->tree, t->ap, t->ni, t->nb);
        int64_t a = 0, b = 0, c = 0;
        for (size_t k = 0; k < t->ap.size(); ++k) {
            a += static_cast<int64_t>(t->ap[k].size());
            b += static_cast<int64_t>(t->ni[k].size());
            c += static_cast<int64_t>(t->nb[k].size());
        }
        *ap_total = a;
        *ni_total = b;
        *nb_total = c;
    });
}

void ref_tree_lists_export(const ref_tree *t, int64_t *ap_off, int64_t *ap, int64_t *ni_off, int64_t *ni, int64_t *nb_off, int64_t *nb) {
    int64_t a = 0, b = 0, c = 0;
    for (size_t k = 0; k < t->ap.size(); ++k) {
        ap_off[k] = a;
        ni_off[k] = b;
        nb_off[k] = c;
        for (const auto v : t->ap[k]) ap[a++] = static_cast<int64_t>(v);
        for (const auto v : t->ni[k]) ni[b++] = static_cast<int64_t>(v);
        for (const auto v : t->nb[k]) nb[c++] = static_cast<int64_t>(v);
    }
    ap_off[t->ap.size()] = a;
    ni_off[t->ap.size()] = b;
    nb_off[t->ap.size()] = c;
}

// SolverRef::acceleration (solver_ref.hpp:164-177) on a built tree; M and L
// exported per cell as (p+1)(2p+1) complex numbers [cell][n][p+m].
int ref_tree_fmm(ref_tree *t, double m, double G, double eps, uint32_t p, double *a, double *M, double *L) {
    return guarded([&] {
        if (t->ap.empty()) FMMTest::link(*t->tree, t->ap, t->ni, t->nb);
        std::vector<Vector<D>> acc(t->x.size(), Vector<D>{});
        FMMTest::fmm(t->x, G, m, eps, acc, p, *t->tree, t->ap, t->ni);
        from_vec(acc, a);
        const auto &cells = t->tree->cells();
        const size_t stride = (p + 1) * (2 * p + 1);
        for (size_t c = 0; c < cells.size(); ++c)
            for (unsigned k = 0; k <= p; ++k)
                for (int mm = -static_cast<int>(p); mm <= static_cast<int>(p); ++mm) {
                    const size_t o = c * stride + k * (2 * p + 1) + static_cast<size_t>(static_cast<int>(p) + mm);
                    const auto mv = std::abs(mm) <= static_cast<int>(k) ? FMMTest::M(cells[c], p, k, mm) : std::complex<D>{};
                    const auto lv = std::abs(mm) <= static_cast<int>(k) ? FMMTest::L(cells[c], p, k, mm) : std::complex<D>{};
                    if (M) {
                        M[2 * o] = mv.real();
                        M[2 * o + 1] = mv.imag();
                    }
                    if (L) {
                        L[2 * o] = lv.real();
                        L[2 * o + 1] = lv.imag();
                    }
                }
    });
}

int ref_rk4_fmm(double *x, double *v, int64_t n, double m, double G, double eps, double dt, int64_t s, uint32_t p) {
    return guarded([&] {
        auto xv = to_vec(x, n), vv = to_vec(v, n);
        SolverRefTest::rk4(xv, vv, G, m, eps, dt, static_cast<std::size_t>(s), p);
        from_vec(xv, x);
        from_vec(vv, v);
    });
}

int ref_rk4_direct(double *x, double *v, int64_t n, double m, double G, double eps, double dt) {
    return guarded([&] {
        auto xv = to_vec(x, n), vv = to_vec(v, n);
        SolverRefNaiveTest::rk4(xv, vv, G, m, eps, dt);
        from_vec(xv, x);
        from_vec(vv, v);
    });
}

void ref_direct_accel(const double *x, int64_t n, double m, double G, double eps, double *a) {
    const auto xv = to_vec(x, n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const auto ai = SolverRefNaiveTest::acc(xv, G, m, eps, static_cast<std::size_t>(i));
        a[3 * i] = ai.x();
        a[3 * i + 1] = ai.y();
        a[3 * i + 2] = ai.z();
    }
}

void ref_P(uint32_t nmax, double x, double *out) {
    const P<D> t{nmax, nmax, x};
    for (unsigned n = 0; n <= nmax; ++n)
        for (unsigned m = 0; m <= nmax; ++m) out[n * (nmax + 1) + m] = m <= n ? t(n, m) : 0.0;
}

void ref_DP(uint32_t nmax, double theta, double *out) {
    const DP<D> t{nmax, nmax, theta};
    for (unsigned n = 0; n <= nmax; ++n)
        for (unsigned m = 0; m <= nmax; ++m) out[n * (nmax + 1) + m] = m <= n ? t(n, m) : 0.0;
}

void ref_Y(uint32_t kmax, double theta, double phi, double *out) {
    const Y<D> y{kmax, kmax, theta, phi};
    for (unsigned k = 0; k <= kmax; ++k)
        for (int m = -static_cast<int>(kmax); m <= static_cast<int>(kmax); ++m) {
            const size_t o = k * (2 * kmax + 1) + static_cast<size_t>(static_cast<int>(kmax) + m);
            const auto v = std::abs(m) <= static_cast<int>(k) ? y(k, m) : std::complex<D>{};
            out[2 * o] = v.real();
            out[2 * o + 1] = v.imag();
        }
}

double ref_A(uint32_t n, int32_t m) { return A_<D, D, D, D, D, D>(n, m); }

double ref_problem_initialize(int64_t n, uint64_t seed, double eps, double *x, double *v) {
    const auto opt = eps < 0 ? std::optional<D>{} : std::optional<D>{eps};
    const ProblemD prob{static_cast<std::size_t>(n), 10, opt, 1, 1, 1, 4, seed};
    std::vector<Vector<D>> xv, vv;
    prob.initialize(xv, vv);
    from_vec(xv, x);
    from_vec(vv, v);
    return prob.epsilon();
}

double ref_problem_delta_t(double eps) { return ProblemD{1, 10, std::optional<D>{eps}}.delta_t(); }

void ref_curve_fit(const double *e_t, const double *t, int64_t k, double *alpha, double *c) {
    const std::vector<D> e(e_t, e_t + k), tt(t, t + k);
    const auto [a, cc] = BenchmarkTest::fit(e, tt);
    *alpha = a;
    *c = cc;
}

// CPU-baseline timing of the reference FMM (SolverRef::acceleration,
// solver_ref.hpp:164-232) on a bounded sample of one force evaluation.
// The tree (Octree ctor) and calc_link run in full.  Every other phase runs
// on an evenly strided sample of its work items and is scaled up by the exact
// ratio of work units (particles for P2M/L2P, cells for M2M/L2L, AP entries for
// M2L, particle pairs for P2P).  Expansion storage is sized up front (zeros;
// the operators' cost does not depend on the values).
// Threads follow the reference's own OpenMP structure: the Octree ctor is
// serial (octree.hpp:79-191), calc_link forks at its root (solver_ref.hpp:155-
// 161), and the sampled items of P2M / M2M (solver_ref.hpp:192), P2P
// (solver_ref.hpp:203-212) and L2L / L2P (solver_ref.hpp:225) run on `threads`
// threads, each item writing only its own cell or points.  M2L runs on one
// thread: the reference's flat loop copies each source Cell (fmm.hpp:281)
// while other threads resize and write that cell's local expansion
// (fmm.hpp:279,296), a data race; its time is reported for one core.
//   out[0..6]  extrapolated seconds: tree, link, P2M, M2M, M2L (one thread),
//              P2P, L2L+L2P
//   info[0..6] sampled fraction of each phase (1 for the full phases)
int ref_time_sample(const double *x, int64_t n, int64_t s, uint32_t p, double m, double G, double eps, int64_t per_phase, int32_t threads,
                    double *out, double *info) {
    return guarded([&] {
        using clk = std::chrono::steady_clock;
        auto el = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
        const auto xv = to_vec(x, n);
        auto t0 = clk::now();
        TreeD tree{xv, static_cast<std::size_t>(s)};
        auto t1 = clk::now();
        Lists ap, ni, nb;
        FMMTest::link(tree, ap, ni, nb);
        auto t2 = clk::now();
        out[0] = el(t0, t1);
        out[1] = el(t1, t2);
        info[0] = info[1] = 1;
        auto &cells = FMMTest::cells(tree);
        const size_t nc = cells.size();
        std::vector<int> level(nc, 0);
        for (size_t c = 1; c < nc; ++c) level[c] = level[cells[c].parent().value()] + 1;
        for (auto &c : cells) FMMTest::presize(c, p);
        std::vector<Vector<D>> acc(xv.size(), Vector<D>{});
        // select every k-th of `items` (weights w) up to `per` items, run f on
        // `nth` threads, scale by weight ratio
        auto sampled = [&](const std::vector<size_t> &items, const std::vector<double> &w, auto &&f, double &secs, double &frac, int64_t per,
                           int nth) {
            double tot = 0, done = 0;
            for (double q : w) tot += q;
            if (items.empty() || tot == 0) {
                secs = 0;
                frac = 1;
                return;
            }
            const size_t stride = std::max<size_t>(1, items.size() / static_cast<size_t>(std::max<int64_t>(1, per)));
            std::vector<size_t> pick;
            for (size_t k = 0; k < items.size(); k += stride) {
                pick.push_back(items[k]);
                done += w[k];
            }
            auto a = clk::now();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth)
            for (size_t k = 0; k < pick.size(); ++k) f(pick[k]);
            auto b = clk::now();
            frac = done / tot;
            secs = el(a, b) / frac;
        };
        const int nth = std::max(1, threads);
        const int64_t per_par = per_phase * nth;  // the same sampled work per thread
        std::vector<size_t> leaves2, inner2, targets, leavesall, l2cells;
        std::vector<double> wl2, wi2, wt, wla, wl2c;
        auto subtree_count = [&](auto &&self, size_t c) -> double {
            if (cells[c].is_leaf()) return static_cast<double>(cells[c].points().size());
            double r = 0;
            for (const auto ch : cells[c].cells()) r += self(self, ch);
            return r;
        };
        for (size_t c = 0; c < nc; ++c) {
            const bool leaf = cells[c].is_leaf();
            if (level[c] >= 2 && leaf && !cells[c].points().empty()) {
                leaves2.push_back(c);
                wl2.push_back(static_cast<double>(cells[c].points().size()));
            }
            if (level[c] >= 2 && !leaf) {
                inner2.push_back(c);
                wi2.push_back(1);
            }
            if (c > 8 && !ap[c].empty()) {
                targets.push_back(c);
                wt.push_back(static_cast<double>(ap[c].size()));
            }
            if (leaf && !cells[c].points().empty()) {
                double src = 0;
                for (const auto q : ni[c]) src += subtree_count(subtree_count, q);
                leavesall.push_back(c);
                wla.push_back(src * static_cast<double>(cells[c].points().size()));
            }
            if (level[c] >= 3) {
                l2cells.push_back(c);
                wl2c.push_back(1);
            }
        }
        sampled(leaves2, wl2, [&](size_t c) { FMMTest::p2m(cells[c], p, xv, m); }, out[2], info[2], per_par, nth);
        sampled(inner2, wi2, [&](size_t c) { FMMTest::m2m(cells[c], p, cells); }, out[3], info[3], per_par, nth);
        sampled(targets, wt, [&](size_t c) { FMMTest::m2l(cells[c], p, cells, ap[c]); }, out[4], info[4], per_phase, 1);
        sampled(leavesall, wla, [&](size_t c) { FMMTest::p2p(cells[c], xv, acc, cells, ni[c], G, m, eps); }, out[5], info[5], per_par, nth);
        double s1 = 0, f1 = 1, s2 = 0, f2 = 1;
        sampled(l2cells, wl2c, [&](size_t c) { FMMTest::l2l(cells[c], p, cells); }, s1, f1, per_par, nth);
        sampled(leaves2, wl2, [&](size_t c) { FMMTest::l2p(cells[c], p, xv, acc); }, s2, f2, per_par, nth);
        out[6] = s1 + s2;
        info[6] = std::min(f1, f2);
    });
}

}  // extern "C"
