// oracle.cpp - CPU restatement of the VIPP gravity FMM (double precision).
//
// TEST INFRASTRUCTURE ONLY: this is the checker for the HIP path.  It is
// never measured as the product and never shipped; the product library does
// not link it.  Each routine cites the reference file:line it restates
// (paths relative to the reference repository).
//
// The restatement keeps the reference's algorithm: sequential first-fit octree
// insertion, the calc_link list rule, the explicit-factorial special functions
// and the (p+1)^4 operator loops.  It differs only in bookkeeping (tables are
// built once per call instead of once per term, i^e is taken as the exact
// +-1 it always is, and per-particle masses / a G on the far field are
// optional extensions).
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

using cplx = std::complex<double>;

struct V3 {
    double x, y, z;
};

inline V3 sub(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

// problem.hpp:108-110 (norm2 = x*x + y*y + z*z, evaluated left to right)
inline double norm2(const V3 &a) { return a.x * a.x + a.y * a.y + a.z * a.z; }

// problem.hpp:113-119
inline void to_polar(const V3 &a, double &r, double &th, double &ph) {
    r = std::sqrt(norm2(a));
    th = std::abs(r) < DBL_MIN ? 0.0 : std::acos(a.z / r);
    const double rxy = std::sqrt(a.x * a.x + a.y * a.y);
    ph = (a.y < 0 ? -1 : 1) * (rxy < DBL_MIN ? 0.0 : std::acos(a.x / rxy));
}

// fmm.hpp:9-16
inline double factorial(unsigned n) {
    double r = 1;
    for (unsigned i = 2; i <= n; ++i) r *= i;
    return r;
}

// ---------------------------------------------------------------------------
// Special functions
// ---------------------------------------------------------------------------

// Associated Legendre P_n^m(x) with the Condon-Shortley phase, by the explicit
// alternating factorial sum of fmm.hpp:23-60.  Table [n][m], (nmax+1)^2.
std::vector<double> legendre_table(unsigned nmax, unsigned mmax, double x) {
    std::vector<double> p((nmax + 1) * (nmax + 1), 0.0);
    double x1m2 = 1;
    for (unsigned m = 0; m <= mmax; ++m) {
        double n2 = std::pow(2, m);
        for (unsigned n = m; n <= nmax; ++n) {
            const unsigned kmax = (n - m) / 2;
            double sign = 1, f2nk = factorial(2 * n), fk = 1, fnk = factorial(n), fn2km = factorial(n - m);
            double acc = 0;
            for (unsigned k = 0; k <= kmax; ++k) {
                const double xp = std::pow(x, static_cast<int>(n - 2 * k - m));
                acc += sign * f2nk * xp / (fk * fnk * fn2km);
                if (k == kmax) break;
                sign *= -1;
                f2nk /= 2.0 * (n - k) * (2.0 * (n - k) - 1);
                fk *= k + 1;
                fnk /= n - k;
                fn2km /= static_cast<double>(n - m - 2 * k) * (n - m - 2 * k - 1);
            }
            p[n * (nmax + 1) + m] = acc * (x1m2 / n2 * (m % 2 ? -1 : 1));
            n2 *= 2;
        }
        x1m2 *= std::sqrt(1 - x * x);
    }
    return p;
}

// -sin(theta) * dP_n^m/dx at x = cos(theta), explicit form of fmm.hpp:80-152.
std::vector<double> dlegendre_table(unsigned nmax, unsigned mmax, double theta) {
    std::vector<double> dp((nmax + 1) * (nmax + 1), 0.0);
    const double x = std::cos(theta), y = std::sin(theta);
    {
        double n2 = 1;
        for (unsigned n = 0; n <= nmax; ++n) {
            const unsigned kmax = n / 2;
            double f2nk = factorial(2 * n), fk = 1, fnk = factorial(n), fn2km = factorial(n);
            double acc = 0;
            for (unsigned k = 0; k <= kmax; ++k) {
                if (n != k * 2) {
                    const double xp = std::pow(x, static_cast<int>(n - 2 * k - 1));
                    acc -= y / n2 * (k % 2 ? -1 : 1) * f2nk * (n - 2 * k) / (fk * fnk * fn2km) * xp;
                }
                if (k == kmax) break;
                f2nk /= 2.0 * (n - k) * (2.0 * (n - k) - 1);
                fk *= k + 1;
                fnk /= n - k;
                fn2km /= static_cast<double>(n - 2 * k) * (n - 2 * k - 1);
            }
            dp[n * (nmax + 1)] = acc;
            n2 *= 2;
        }
    }
    double ym1 = 1;
    for (unsigned m = 1; m <= mmax; ++m) {
        double n2 = std::pow(2, m);
        for (unsigned n = m; n <= nmax; ++n) {
            const unsigned kmax = (n - m) / 2;
            double f2nk = factorial(2 * n), fk = 1, fnk = factorial(n), fn2km = factorial(n - m);
            double acc = 0;
            for (unsigned k = 0; k <= kmax; ++k) {
                if (n - m != k * 2) {
                    const double xp = std::pow(x, static_cast<int>(n - 2 * k - m - 1));
                    acc -= ym1 / n2 * (k % 2 ? -1 : 1) * f2nk / (fk * fnk * fn2km) * xp *
                           (-static_cast<double>(m) * x * x + y * y * (n - 2 * k - m));
                } else {
                    const double xp = std::pow(x, static_cast<int>(n - 2 * k - m));
                    acc -= (-static_cast<double>(m) * x * ym1) / n2 * (k % 2 ? -1 : 1) * f2nk / (fk * fnk * fn2km) * xp;
                }
                if (k == kmax) break;
                f2nk /= 2.0 * (n - k) * (2.0 * (n - k) - 1);
                fk *= k + 1;
                fnk /= n - k;
                fn2km /= static_cast<double>(n - m - 2 * k) * (n - m - 2 * k - 1);
            }
            dp[n * (nmax + 1) + m] = acc * (m % 2 ? -1 : 1);
            n2 *= 2;
        }
        ym1 *= y;
    }
    return dp;
}

// sqrt((k-|m|)!/(k+|m|)!) as fmm.hpp:183-187 forms it.
inline double ynorm(unsigned k, unsigned am) {
    double c = 1;
    for (unsigned i = k - am + 1; i <= k + am; ++i) c /= i;
    return std::sqrt(c);
}

// Spherical harmonics of fmm.hpp:169-194: Y_k^m = ynorm * P_k^{|m|}(cos t) e^{i m phi}.
struct YTable {
    unsigned kmax;
    std::vector<cplx> v;  // [k][kmax + m]
    YTable(unsigned kmax_, double theta, double phi) : kmax(kmax_), v((kmax_ + 1) * (2 * kmax_ + 1)) {
        const auto P = legendre_table(kmax, kmax, std::cos(theta));
        for (unsigned k = 0; k <= kmax; ++k)
            for (int m = -static_cast<int>(k); m <= static_cast<int>(k); ++m) {
                const unsigned am = static_cast<unsigned>(std::abs(m));
                const cplx e{std::cos(m * phi), std::sin(m * phi)};
                v[k * (2 * kmax + 1) + kmax + m] = ynorm(k, am) * P[k * (kmax + 1) + am] * e;
            }
    }
    cplx operator()(unsigned k, int m) const { return v[k * (2 * kmax + 1) + kmax + m]; }
};

// A_n^m = (-1)^n / sqrt((n-|m|)!(n+|m|)!) (fmm.hpp:201-221)
double A_coef(unsigned n, int m) {
    const unsigned am = static_cast<unsigned>(std::abs(m));
    if (n < am) throw std::runtime_error("n must greater than or equal to |m|");
    double d = 1;
    for (unsigned i = n - am + 1; i <= n + am; ++i) d *= i;
    d = std::sqrt(d);
    for (unsigned i = 1; i <= n - am; ++i) d *= i;
    return n % 2 == 0 ? 1.0 / d : -1.0 / d;
}

struct ATable {
    unsigned nmax;
    std::vector<double> v;  // [n][nmax + m]
    explicit ATable(unsigned nmax_) : nmax(nmax_), v((nmax_ + 1) * (2 * nmax_ + 1), 0.0) {
        for (unsigned n = 0; n <= nmax; ++n)
            for (int m = -static_cast<int>(n); m <= static_cast<int>(n); ++m) v[n * (2 * nmax + 1) + nmax + m] = A_coef(n, m);
    }
    double operator()(unsigned n, int m) const { return v[n * (2 * nmax + 1) + nmax + m]; }
};

// i^e for even e (the only case the operators produce): exact +-1.
inline double ipow_even(int e) {
    const int h = e / 2;
    return (h % 2 == 0) ? 1.0 : -1.0;
}

// ---------------------------------------------------------------------------
// Octree (octree.hpp:70-195) with the Cell geometry of fmm.hpp:388-493
// ---------------------------------------------------------------------------
struct Cell {
    int64_t parent = -1;
    int32_t level = 0;
    V3 mn{0, 0, 0}, mx{0, 0, 0};
    bool leaf = true;
    int64_t child_base = -1;
    std::vector<int64_t> pts;
    V3 center() const { return {(mn.x + mx.x) / 2, (mn.y + mx.y) / 2, (mn.z + mx.z) / 2}; }
};

constexpr double PADDING = static_cast<double>(1E-5L);  // problem.hpp:167
constexpr int32_t RUNAWAY_DEPTH = 200;                  // guard against coincident points

}  // namespace

struct orc_tree {
    std::vector<V3> x;
    std::vector<Cell> cells;
    size_t thr = 0;
    bool has_lists = false;
    std::vector<std::vector<int64_t>> ap, ni, nb;
};

namespace {

bool tree_insert(orc_tree &t, size_t idx, size_t i);

// octree.hpp:166-191
void tree_subdivide(orc_tree &t, size_t idx) {
    const V3 mn = t.cells[idx].mn, mx = t.cells[idx].mx;
    const V3 mid{(mx.x + mn.x) / 2, (mx.y + mn.y) / 2, (mx.z + mn.z) / 2};
    if (t.cells[idx].level >= RUNAWAY_DEPTH) throw std::runtime_error("octree depth runaway (coincident points)");
    const size_t base = t.cells.size();
    for (size_t off = 0; off < 8; ++off) {
        Cell c;
        c.parent = static_cast<int64_t>(idx);
        c.level = t.cells[idx].level + 1;
        c.mn = {off % 2 == 0 ? mid.x : mn.x, off % 4 < 2 ? mid.y : mn.y, off < 4 ? mid.z : mn.z};
        c.mx = {off % 2 == 0 ? mx.x : mid.x, off % 4 < 2 ? mx.y : mid.y, off < 4 ? mx.z : mid.z};
        t.cells.push_back(std::move(c));
    }
    const std::vector<int64_t> moved = std::move(t.cells[idx].pts);
    t.cells[idx].pts.clear();
    t.cells[idx].leaf = false;
    t.cells[idx].child_base = static_cast<int64_t>(base);
    for (const auto i : moved) {
        bool ok = false;
        for (size_t ci = base; ci < base + 8; ++ci)
            if (tree_insert(t, ci, static_cast<size_t>(i))) {
                ok = true;
                break;
            }
        if (!ok) throw std::runtime_error("failed to insert point");
    }
}

// octree.hpp:131-161
bool tree_insert(orc_tree &t, size_t idx, size_t i) {
    const V3 &p = t.x[i];
    {
        const Cell &c = t.cells[idx];
        if (p.x < c.mn.x || c.mx.x < p.x || p.y < c.mn.y || c.mx.y < p.y || p.z < c.mn.z || c.mx.z < p.z) return false;
        if (c.leaf && c.pts.size() < t.thr) {
            t.cells[idx].pts.push_back(static_cast<int64_t>(i));
            return true;
        }
    }
    if (t.cells[idx].leaf) tree_subdivide(t, idx);
    const size_t base = static_cast<size_t>(t.cells[idx].child_base);
    for (size_t ci = base; ci < base + 8; ++ci)
        if (tree_insert(t, ci, i)) return true;
    return false;
}

struct U3 {
    uint64_t x, y, z;
};
// fmm.hpp:464-471
U3 child_position(const orc_tree &t, const Cell &c, size_t index) {
    const Cell &par = t.cells[static_cast<size_t>(c.parent)];
    const uint64_t ci = index - static_cast<uint64_t>(par.child_base);
    return {ci & 1, (ci & 2) >> 1, (ci & 4) >> 2};
}

// fmm.hpp:475-493 (unsigned, wrap-safe norm2 <= 3)
bool adjacent_or_equal(const orc_tree &t, size_t a, size_t b) {
    const Cell &ca = t.cells[a], &cb = t.cells[b];
    if (ca.parent == cb.parent) return true;
    U3 ia = child_position(t, ca, a), ib = child_position(t, cb, b);
    size_t anc_a = static_cast<size_t>(ca.parent), anc_b = static_cast<size_t>(cb.parent);
    for (uint64_t level = 1;; ++level) {
        const U3 pa = child_position(t, t.cells[anc_a], anc_a), pb = child_position(t, t.cells[anc_b], anc_b);
        ia = {(uint64_t{1} << level) * pa.x + ia.x, (uint64_t{1} << level) * pa.y + ia.y, (uint64_t{1} << level) * pa.z + ia.z};
        ib = {(uint64_t{1} << level) * pb.x + ib.x, (uint64_t{1} << level) * pb.y + ib.y, (uint64_t{1} << level) * pb.z + ib.z};
        if (t.cells[anc_a].parent == t.cells[anc_b].parent) {
            const uint64_t dx = ia.x - ib.x, dy = ia.y - ib.y, dz = ia.z - ib.z;
            return dx * dx + dy * dy + dz * dz <= 3;
        }
        anc_a = static_cast<size_t>(t.cells[anc_a].parent);
        anc_b = static_cast<size_t>(t.cells[anc_b].parent);
    }
}

// solver_ref.hpp:115-162 (the recursion is done breadth-first; lists of a cell
// depend only on its parent's lists, so the visiting order does not matter)
void calc_link(orc_tree &t) {
    const size_t nc = t.cells.size();
    t.ap.assign(nc, {});
    t.ni.assign(nc, {});
    t.nb.assign(nc, {});
    if (t.cells[0].leaf)
        t.ni[0].push_back(0);
    else
        t.nb[0].push_back(0);
    // cells are created parent-before-child, so index order is a valid top-down order
    for (size_t c = 1; c < nc; ++c) {
        const Cell &cell = t.cells[c];
        const size_t par = static_cast<size_t>(cell.parent);
        t.ni[c] = t.ni[par];
        for (const auto pn : t.nb[par]) {
            const int64_t base = t.cells[static_cast<size_t>(pn)].child_base;
            for (int64_t cous = base; cous < base + 8; ++cous) {
                const Cell &cc = t.cells[static_cast<size_t>(cous)];
                if (adjacent_or_equal(t, c, static_cast<size_t>(cous))) {
                    if (!cell.leaf && !cc.leaf)
                        t.nb[c].push_back(cous);
                    else
                        t.ni[c].push_back(cous);
                } else {
                    t.ap[c].push_back(cous);
                }
            }
        }
    }
    t.has_lists = true;
}

// ---------------------------------------------------------------------------
// FMM operators (fmm.hpp:224-385)
// ---------------------------------------------------------------------------
struct Expansions {
    unsigned p;
    size_t stride;  // (p+1)(2p+1)
    std::vector<cplx> M, L;
    std::vector<char> hasM, hasL;
    Expansions(size_t nc, unsigned p_) : p(p_), stride((p_ + 1) * (2 * p_ + 1)), M(nc * stride), L(nc * stride), hasM(nc, 0), hasL(nc, 0) {}
    cplx &m(size_t c, unsigned n, int mm) { return M[c * stride + n * (2 * p + 1) + p + mm]; }
    cplx &l(size_t c, unsigned n, int mm) { return L[c * stride + n * (2 * p + 1) + p + mm]; }
};

inline double mass_of(const double *masses, double m, int64_t i) { return masses ? masses[i] : m; }

// fmm.hpp:224-244
void p2m(const orc_tree &t, Expansions &E, size_t c, const double *masses, double m) {
    const unsigned p = E.p;
    const Cell &cell = t.cells[c];
    const V3 cen = cell.center();
    E.hasM[c] = 1;
    for (const auto i : cell.pts) {
        double rho, al, be;
        to_polar(sub(t.x[static_cast<size_t>(i)], cen), rho, al, be);
        const YTable y(p, al, be);
        double rk = 1;
        const double mi = mass_of(masses, m, i);
        for (unsigned k = 0; k <= p; ++k) {
            for (unsigned mm = 0; mm <= k; ++mm) E.m(c, k, static_cast<int>(mm)) += (mi * rk) * y(k, -static_cast<int>(mm));
            rk *= rho;
        }
        for (unsigned k = 1; k <= p; ++k)
            for (unsigned mm = 1; mm <= k; ++mm) E.m(c, k, -static_cast<int>(mm)) = std::conj(E.m(c, k, static_cast<int>(mm)));
    }
}

// fmm.hpp:247-274
void m2m(const orc_tree &t, Expansions &E, const ATable &A, size_t c) {
    const unsigned p = E.p;
    const Cell &cell = t.cells[c];
    const V3 cen = cell.center();
    E.hasM[c] = 1;
    for (int64_t ch = cell.child_base; ch < cell.child_base + 8; ++ch) {
        double rho, al, be;
        to_polar(sub(t.cells[static_cast<size_t>(ch)].center(), cen), rho, al, be);
        const YTable y(p, al, be);
        for (unsigned j = 0; j <= p; ++j)
            for (int k = -static_cast<int>(j); k <= static_cast<int>(j); ++k) {
                const double ajk = A(j, k);
                double rn = 1;
                for (unsigned n = 0; n <= j; ++n) {
                    const int lo = std::max(-static_cast<int>(n), static_cast<int>(n) + k - static_cast<int>(j));
                    const int hi = std::min(static_cast<int>(j) + k - static_cast<int>(n), static_cast<int>(n));
                    for (int mm = lo; mm <= hi; ++mm) {
                        const cplx O = E.m(static_cast<size_t>(ch), j - n, k - mm);
                        const double im = ipow_even(std::abs(k) - std::abs(mm) - std::abs(k - mm));
                        E.m(c, j, k) += (O * im * A(n, mm) * A(j - n, k - mm) * rn * y(n, -mm)) / ajk;
                    }
                    rn *= rho;
                }
            }
    }
}

// fmm.hpp:277-304
void m2l(const orc_tree &t, Expansions &E, const ATable &A, size_t c, const std::vector<int64_t> &ap) {
    const unsigned p = E.p;
    const V3 cen = t.cells[c].center();
    E.hasL[c] = 1;
    for (const auto s : ap) {
        double rho, al, be;
        to_polar(sub(t.cells[static_cast<size_t>(s)].center(), cen), rho, al, be);
        const YTable y(2 * p, al, be);
        double rj1 = rho;
        for (unsigned j = 0; j <= p; ++j) {
            for (int k = -static_cast<int>(j); k <= static_cast<int>(j); ++k) {
                const double ajk = A(j, k);
                double rn = 1;
                cplx acc = 0;
                for (unsigned n = 0; n <= p; ++n) {
                    for (int mm = -static_cast<int>(n); mm <= static_cast<int>(n); ++mm) {
                        const cplx Mv = E.m(static_cast<size_t>(s), n, mm);
                        const double im = ipow_even(std::abs(k - mm) - std::abs(k) - std::abs(mm));
                        acc += (Mv * im * A(n, mm) * ajk * y(j + n, mm - k)) / ((n % 2 ? -1 : 1) * A(j + n, mm - k) * rj1 * rn);
                    }
                    rn *= rho;
                }
                E.l(c, j, k) += acc;
            }
            rj1 *= rho;
        }
    }
}

// fmm.hpp:307-329
void l2l(const orc_tree &t, Expansions &E, const ATable &A, size_t c) {
    const unsigned p = E.p;
    const Cell &cell = t.cells[c];
    const size_t par = static_cast<size_t>(cell.parent);
    double rho, al, be;
    to_polar(sub(t.cells[par].center(), cell.center()), rho, al, be);
    const YTable y(p, al, be);
    E.hasL[c] = 1;
    for (unsigned j = 0; j <= p; ++j)
        for (int k = -static_cast<int>(j); k <= static_cast<int>(j); ++k) {
            const double ajk = A(j, k);
            double rnj = 1;
            for (unsigned n = j; n <= p; ++n) {
                const int lo = std::max(-static_cast<int>(n), k - static_cast<int>(n - j));
                const int hi = std::min(static_cast<int>(n), static_cast<int>(n - j) + k);
                for (int mm = lo; mm <= hi; ++mm) {
                    const cplx Pm = E.l(par, n, mm);
                    const double im = ipow_even(std::abs(mm) - std::abs(mm - k) - std::abs(k));
                    E.l(c, j, k) += (Pm * im * A(n - j, mm - k) * ajk * y(n - j, mm - k) * rnj) / (((n + j) % 2 ? -1 : 1) * A(n, mm));
                }
                rnj *= rho;
            }
        }
}

// fmm.hpp:332-374; also returns the local potential sum L r^l Y (extension)
void l2p(const orc_tree &t, Expansions &E, size_t c, double gscale, double *a, double *pot) {
    const unsigned p = E.p;
    const Cell &cell = t.cells[c];
    const V3 cen = cell.center();
    for (const auto i : cell.pts) {
        double r, th, ph;
        to_polar(sub(t.x[static_cast<size_t>(i)], cen), r, th, ph);
        const YTable y(p, th, ph);
        const auto dp = dlegendre_table(p, p, th);
        double rl = 1, dr = 0, dth = 0, dph = 0, phi = 0;
        for (unsigned l = 0; l <= p; ++l) {
            for (int mm = -static_cast<int>(l); mm <= static_cast<int>(l); ++mm) {
                const cplx Llm = E.l(c, l, mm);
                const cplx ylm = y(l, mm);
                const unsigned am = static_cast<unsigned>(std::abs(mm));
                const double cn = ynorm(l, am);
                const cplx e{std::cos(mm * ph), std::sin(mm * ph)};
                dr += (Llm * rl * ylm * (l / r)).real();
                dth += (Llm * rl * cn * dp[l * (p + 1) + am] * e).real();
                dph += (Llm * rl * ylm * cplx{0, static_cast<double>(mm)}).real();
                phi += (Llm * rl * ylm).real();
            }
            rl *= r;
        }
        const double ct = std::cos(th), st = std::sin(th), cp = std::cos(ph), sp = std::sin(ph);
        double *ai = a + 3 * i;
        ai[0] += gscale * (st * cp * dr + ct * cp / r * dth - sp / (r * st) * dph);
        ai[1] += gscale * (st * sp * dr + ct * sp / r * dth + cp / (r * st) * dph);
        ai[2] += gscale * (ct * dr - st / r * dth);
        if (pot) pot[i] += phi;
    }
}

// fmm.hpp:513-540 (acceleration_i / acceleration_ij); pot: softened sum m/r
void acc_cell(const orc_tree &t, size_t c, int64_t i, const double *masses, double m, double G, double eps, double acc[3], double &pot) {
    const Cell &cell = t.cells[c];
    if (cell.leaf) {
        const V3 xi = t.x[static_cast<size_t>(i)];
        for (const auto j : cell.pts) {
            if (i == j) continue;
            const V3 d = sub(t.x[static_cast<size_t>(j)], xi);
            const double gm = G * mass_of(masses, m, j);
            const double den = std::pow(norm2(d) + eps * eps, 3.0 / 2);
            acc[0] += gm * d.x / den;
            acc[1] += gm * d.y / den;
            acc[2] += gm * d.z / den;
            pot += mass_of(masses, m, j) / std::sqrt(norm2(d) + eps * eps);
        }
    } else {
        for (int64_t ch = cell.child_base; ch < cell.child_base + 8; ++ch) {
            double sub_acc[3] = {0, 0, 0};
            acc_cell(t, static_cast<size_t>(ch), i, masses, m, G, eps, sub_acc, pot);
            acc[0] += sub_acc[0];
            acc[1] += sub_acc[1];
            acc[2] += sub_acc[2];
        }
    }
}

void run_fmm(orc_tree &t, const double *masses, double m, double G, double eps, unsigned p, int far_g, double *a,
             double *a_near, double *a_far, double *pot, double *Mout, double *Lout) {
    if (!t.has_lists) calc_link(t);
    const size_t nc = t.cells.size(), n = t.x.size();
    Expansions E(nc, p);
    const ATable A(2 * p);
    int32_t maxlev = 0;
    for (const auto &c : t.cells) maxlev = std::max(maxlev, c.level);
    std::vector<std::vector<size_t>> bylev(static_cast<size_t>(maxlev) + 1);
    for (size_t c = 0; c < nc; ++c) bylev[static_cast<size_t>(t.cells[c].level)].push_back(c);

    // bottom-up sweep: P2M on leaves, M2M on internal cells, level >= 2 (solver_ref.hpp:184-201)
    for (int lev = maxlev; lev >= 2; --lev) {
        const auto &cs = bylev[static_cast<size_t>(lev)];
#pragma omp parallel for schedule(dynamic, 4)
        for (size_t k = 0; k < cs.size(); ++k) {
            const size_t c = cs[k];
            if (t.cells[c].leaf)
                p2m(t, E, c, masses, m);
            else
                m2m(t, E, A, c);
        }
    }
    std::vector<double> near(3 * n, 0.0), far(3 * n, 0.0), pnear(n, 0.0), pfar(n, 0.0);
    // flat sweep: M2L for index > 8, P2P on leaves (solver_ref.hpp:203-212)
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t c = 0; c < nc; ++c) {
        if (c > 8) m2l(t, E, A, c, t.ap[c]);
        if (t.cells[c].leaf) {
            for (const auto i : t.cells[c].pts) {
                double acc[3] = {0, 0, 0};
                double pt = 0;
                for (const auto nbc : t.ni[c]) {
                    double sub_acc[3] = {0, 0, 0};
                    acc_cell(t, static_cast<size_t>(nbc), i, masses, m, G, eps, sub_acc, pt);
                    acc[0] += sub_acc[0];
                    acc[1] += sub_acc[1];
                    acc[2] += sub_acc[2];
                }
                near[3 * i] = acc[0];
                near[3 * i + 1] = acc[1];
                near[3 * i + 2] = acc[2];
                pnear[i] = pt;
            }
        }
    }
    // top-down sweep: L2L for level >= 3, L2P on leaves of level >= 2 (solver_ref.hpp:214-231)
    const double gscale = far_g ? G : 1.0;
    for (int lev = 2; lev <= maxlev; ++lev) {
        const auto &cs = bylev[static_cast<size_t>(lev)];
#pragma omp parallel for schedule(dynamic, 4)
        for (size_t k = 0; k < cs.size(); ++k) {
            const size_t c = cs[k];
            if (lev >= 3) l2l(t, E, A, c);
            if (t.cells[c].leaf) l2p(t, E, c, gscale, far.data(), pfar.data());
        }
    }
    for (size_t i = 0; i < 3 * n; ++i) a[i] = near[i] + far[i];
    if (a_near) std::memcpy(a_near, near.data(), sizeof(double) * 3 * n);
    if (a_far) std::memcpy(a_far, far.data(), sizeof(double) * 3 * n);
    if (pot)
        for (size_t i = 0; i < n; ++i) pot[i] = pnear[i] + pfar[i];
    if (Mout)
        for (size_t k = 0; k < E.M.size(); ++k) {
            Mout[2 * k] = E.M[k].real();
            Mout[2 * k + 1] = E.M[k].imag();
        }
    if (Lout)
        for (size_t k = 0; k < E.L.size(); ++k) {
            Lout[2 * k] = E.L[k].real();
            Lout[2 * k + 1] = E.L[k].imag();
        }
}

orc_tree *build_tree(const double *x, int64_t n, int64_t s) {
    if (s <= 0) throw std::runtime_error("s must not be 0");
    auto *t = new orc_tree;
    try {
        t->x.resize(static_cast<size_t>(n));
        for (int64_t i = 0; i < n; ++i) t->x[static_cast<size_t>(i)] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
        t->thr = static_cast<size_t>(std::floor(std::sqrt(8) * static_cast<double>(s)));  // octree.hpp:83
        Cell root;
        if (n != 0) {  // fmm.hpp:403-405 with problem.hpp:122-141
            V3 mn{DBL_MAX, DBL_MAX, DBL_MAX}, mx{-DBL_MAX, -DBL_MAX, -DBL_MAX};
            for (const auto &q : t->x) {
                mn = {std::min(mn.x, q.x), std::min(mn.y, q.y), std::min(mn.z, q.z)};
                mx = {std::max(mx.x, q.x), std::max(mx.y, q.y), std::max(mx.z, q.z)};
            }
            root.mn = {mn.x - PADDING, mn.y - PADDING, mn.z - PADDING};
            root.mx = {mx.x + PADDING, mx.y + PADDING, mx.z + PADDING};
        }
        t->cells.push_back(root);
        for (int64_t i = 0; i < n; ++i)
            if (!tree_insert(*t, 0, static_cast<size_t>(i))) throw std::runtime_error("failed to insert point");
    } catch (...) {
        delete t;
        throw;
    }
    return t;
}

// solver_ref.hpp:36-104 / solver_ref_naive.hpp:33-100
template <class Accel>
void rk4(double *x, double *v, int64_t n, double dt, Accel &&accel) {
    const size_t N3 = static_cast<size_t>(3 * n);
    const std::vector<double> x0(x, x + N3), v0(v, v + N3);
    std::vector<double> xo(x0), vo(v0), a(N3);
    auto accum = [&](double c) {
        for (size_t k = 0; k < N3; ++k) {
            xo[k] += c * v[k];
            vo[k] += c * a[k];
        }
    };
    auto euler = [&](double h) {
        for (size_t k = 0; k < N3; ++k) {
            x[k] = x0[k] + v[k] * h;
            v[k] = v0[k] + a[k] * h;
        }
    };
    accel(x, a.data());
    accum(dt / 6);
    euler(dt / 2);
    accel(x, a.data());
    accum(dt / 3);
    euler(dt / 2);
    accel(x, a.data());
    accum(dt / 3);
    euler(dt);
    accel(x, a.data());
    accum(dt / 6);
    std::memcpy(x, xo.data(), sizeof(double) * N3);
    std::memcpy(v, vo.data(), sizeof(double) * N3);
}

void direct_one(const double *x, int64_t n, const double *masses, double m, double G, double eps, int64_t i, double *ai, double *pi) {
    const V3 xi{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    double acc[3] = {0, 0, 0}, pt = 0;
    for (int64_t j = 0; j < n; ++j) {
        if (i == j) continue;
        const V3 d = sub(V3{x[3 * j], x[3 * j + 1], x[3 * j + 2]}, xi);
        const double gm = G * mass_of(masses, m, j);
        const double den = std::pow(norm2(d) + eps * eps, 3.0 / 2);
        acc[0] += gm * d.x / den;
        acc[1] += gm * d.y / den;
        acc[2] += gm * d.z / den;
        pt += mass_of(masses, m, j) / std::sqrt(norm2(d) + eps * eps);
    }
    ai[0] = acc[0];
    ai[1] = acc[1];
    ai[2] = acc[2];
    if (pi) *pi = pt;
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

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

int orc_tree_build(const double *x, int64_t n, int64_t s, orc_tree **out) {
    return guarded([&] { *out = build_tree(x, n, s); });
}

void orc_tree_free(orc_tree *t) { delete t; }

int64_t orc_tree_ncells(const orc_tree *t) { return static_cast<int64_t>(t->cells.size()); }

void orc_tree_export(const orc_tree *t, int64_t *parent, int32_t *level, double *bmin, double *bmax, int64_t *child_base, int64_t *pt_off,
                     int64_t *pts) {
    int64_t off = 0;
    for (size_t c = 0; c < t->cells.size(); ++c) {
        const Cell &cell = t->cells[c];
        parent[c] = cell.parent;
        level[c] = cell.level;
        bmin[3 * c] = cell.mn.x;
        bmin[3 * c + 1] = cell.mn.y;
        bmin[3 * c + 2] = cell.mn.z;
        bmax[3 * c] = cell.mx.x;
        bmax[3 * c + 1] = cell.mx.y;
        bmax[3 * c + 2] = cell.mx.z;
        child_base[c] = cell.leaf ? -1 : cell.child_base;
        pt_off[c] = off;
        for (const auto i : cell.pts) pts[off++] = i;
    }
    pt_off[t->cells.size()] = off;
}

// Manual subdivision of one leaf, as the reference's tests do through the
// friend hook (test_fmm.cpp:97-99 calling octree.hpp:166-191; used at test_fmm.cpp:346,364-365).
int orc_tree_subdivide(orc_tree *t, int64_t idx, int64_t threshold) {
    return guarded([&] {
        if (idx < 0 || static_cast<size_t>(idx) >= t->cells.size() || !t->cells[static_cast<size_t>(idx)].leaf)
            throw std::runtime_error("subdivide: not a leaf");
        const size_t saved = t->thr;
        t->thr = static_cast<size_t>(threshold);
        try {
            tree_subdivide(*t, static_cast<size_t>(idx));
        } catch (...) {
            t->thr = saved;
            throw;
        }
        t->thr = saved;
        t->has_lists = false;
    });
}

int orc_tree_lists(orc_tree *t, int64_t *ap_total, int64_t *ni_total, int64_t *nb_total) {
    return guarded([&] {
        calc_link(*t);
        int64_t a = 0, b = 0, c = 0;
        for (size_t k = 0; k < t->cells.size(); ++k) {
            a += static_cast<int64_t>(t->ap[k].size());
            b += static_cast<int64_t>(t->ni[k].size());
            c += static_cast<int64_t>(t->nb[k].size());
        }
        *ap_total = a;
        *ni_total = b;
        *nb_total = c;
    });
}

void orc_tree_lists_export(const orc_tree *t, int64_t *ap_off, int64_t *ap, int64_t *ni_off, int64_t *ni, int64_t *nb_off, int64_t *nb) {
    int64_t a = 0, b = 0, c = 0;
    for (size_t k = 0; k < t->cells.size(); ++k) {
        ap_off[k] = a;
        ni_off[k] = b;
        nb_off[k] = c;
        for (const auto v : t->ap[k]) ap[a++] = v;
        for (const auto v : t->ni[k]) ni[b++] = v;
        for (const auto v : t->nb[k]) nb[c++] = v;
    }
    ap_off[t->cells.size()] = a;
    ni_off[t->cells.size()] = b;
    nb_off[t->cells.size()] = c;
}

int orc_tree_fmm(orc_tree *t, const double *x, const double *masses, double m, double G, double eps, uint32_t p, int far_g, double *a,
                 double *a_near, double *a_far, double *pot, double *M, double *L) {
    return guarded([&] {
        if (x)
            for (size_t i = 0; i < t->x.size(); ++i) t->x[i] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
        run_fmm(*t, masses, m, G, eps, p, far_g, a, a_near, a_far, pot, M, L);
    });
}

// Upward pass (P2M on leaves, M2M on internal cells) for every cell at level
// >= min_level; the reference gates it at level 2 (solver_ref.hpp:187,196),
// its tests call p2m/m2m on the root directly (test_fmm.cpp:326-355).
int orc_tree_upward(orc_tree *t, const double *masses, double m, uint32_t p, int32_t min_level, double *Mout) {
    return guarded([&] {
        Expansions E(t->cells.size(), p);
        const ATable A(2 * p);
        int32_t maxlev = 0;
        for (const auto &c : t->cells) maxlev = std::max(maxlev, c.level);
        for (int32_t lev = maxlev; lev >= min_level; --lev)
            for (size_t c = 0; c < t->cells.size(); ++c)
                if (t->cells[c].level == lev) {
                    if (t->cells[c].leaf)
                        p2m(*t, E, c, masses, m);
                    else
                        m2m(*t, E, A, c);
                }
        for (size_t k = 0; k < E.M.size(); ++k) {
            Mout[2 * k] = E.M[k].real();
            Mout[2 * k + 1] = E.M[k].imag();
        }
    });
}

void orc_direct_accel(const double *x, int64_t n, const double *masses, double m, double G, double eps, double *a, double *pot) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) direct_one(x, n, masses, m, G, eps, i, a + 3 * i, pot ? pot + i : nullptr);
}

void orc_direct_accel_subset(const double *x, int64_t n, const double *masses, double m, double G, double eps, const int64_t *idx, int64_t nidx,
                             double *a, double *pot) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nidx; ++k) direct_one(x, n, masses, m, G, eps, idx[k], a + 3 * k, pot ? pot + k : nullptr);
}

int orc_rk4_fmm(double *x, double *v, int64_t n, const double *masses, double m, double G, double eps, double dt, int64_t s, uint32_t p,
                int far_g) {
    return guarded([&] {
        rk4(x, v, n, dt, [&](const double *xs, double *a) {
            orc_tree *t = build_tree(xs, n, s);
            try {
                run_fmm(*t, masses, m, G, eps, p, far_g, a, nullptr, nullptr, nullptr, nullptr, nullptr);
            } catch (...) {
                delete t;
                throw;
            }
            delete t;
        });
    });
}

int orc_rk4_direct(double *x, double *v, int64_t n, const double *masses, double m, double G, double eps, double dt) {
    return guarded([&] { rk4(x, v, n, dt, [&](const double *xs, double *a) { orc_direct_accel(xs, n, masses, m, G, eps, a, nullptr); }); });
}

void orc_P(uint32_t nmax, double x, double *out) {
    const auto t = legendre_table(nmax, nmax, x);
    std::memcpy(out, t.data(), sizeof(double) * t.size());
}

void orc_DP(uint32_t nmax, double theta, double *out) {
    const auto t = dlegendre_table(nmax, nmax, theta);
    std::memcpy(out, t.data(), sizeof(double) * t.size());
}

void orc_Y(uint32_t kmax, double theta, double phi, double *out) {
    const YTable y(kmax, theta, phi);
    for (size_t k = 0; k < y.v.size(); ++k) {
        out[2 * k] = y.v[k].real();
        out[2 * k + 1] = y.v[k].imag();
    }
}

double orc_A(uint32_t n, int32_t m) { return A_coef(n, m); }

// problem.hpp:188-211 and the helpers of problem.hpp:250-292
double orc_problem_initialize(int64_t n, uint64_t seed, double eps, double *x, double *v) {
    const double epsilon = eps < 0 ? 63.0 / 100 * std::pow(static_cast<double>(n), -22.0 / 100) : eps;
    std::mt19937_64 engine(seed);
    std::uniform_real_distribution<double> dpos(0, 1.0 - epsilon), upos(0, 1), uvel(0, 1);
    const double pmax = std::sqrt(7.0) * 686 / 19683;
    const double pi = 3.141592653589793;
    auto calc_xyz = [&](double norm, double r, double X_, double *out) {
        const double xx = (1.0 - r * 2) * norm;
        const double yz2 = std::sqrt(norm * norm - xx * xx);
        out[0] = xx;
        out[1] = yz2 * std::cos(pi * 2 * X_);
        out[2] = yz2 * std::sin(pi * 2 * X_);
    };
    for (int64_t i = 0; i < n; ++i) {
        const double x0 = dpos(engine);
        const double ri = std::pow(x0, 1.0 / 3) / std::pow(1.0 - std::pow(x0, 2.0 / 3), 1.0 / 2);
        const double ve = std::sqrt(2.0) * std::pow(ri * ri + 1, -1.0 / 4);
        double q;
        while (true) {
            const double x1 = uvel(engine), x2 = uvel(engine);
            if (pmax * x1 < x2 * x2 * std::pow(1.0 - x2 * x2, 7.0 / 2)) {
                q = x2;
                break;
            }
        }
        const double vi = q * ve;
        const double x3 = upos(engine), x4 = upos(engine);
        calc_xyz(ri, x3, x4, x + 3 * i);
        const double x5 = uvel(engine), x6 = uvel(engine);
        calc_xyz(vi, x5, x6, v + 3 * i);
    }
    return epsilon;
}

// problem.hpp:235-243
double orc_problem_delta_t(double eps) { return eps / std::sqrt(2.0); }

// benchmark.hpp:198-241
void orc_curve_fit(const double *e_t, const double *t, int64_t k, double *alpha, double *c_out) {
    const double cmin0 = -10, cmax0 = 10, ceps = 1e-10;
    double cmin = cmin0, cmax = cmax0, c = (cmin + cmax) / 2;
    const auto iters = static_cast<size_t>(std::ceil(std::log2((cmax0 - cmin0) / ceps)));
    for (size_t it = 0; it < iters; ++it) {
        double etx = 0, x2 = 0, etxl = 0, x2l = 0;
        for (int64_t q = 0; q < k; ++q) {
            const double xi = std::pow(t[q], c), lt = std::log(t[q]);
            etx += e_t[q] * xi;
            x2 += xi * xi;
            etxl += e_t[q] * xi * lt;
            x2l += xi * xi * lt;
        }
        if (etx * x2l - x2 * etxl >= 0)
            cmax = c;
        else
            cmin = c;
        c = (cmin + cmax) / 2;
    }
    double num = 0, den = 0;
    for (int64_t q = 0; q < k; ++q) {
        const double xi = std::pow(t[q], c);
        num += e_t[q] * xi;
        den += xi * xi;
    }
    *alpha = num / den;
    *c_out = c;
}

}  // extern "C"
