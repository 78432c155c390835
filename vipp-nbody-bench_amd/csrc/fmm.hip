// fmm.hip - expansions and translations of the gravity FMM.
//
// Conventions follow the reference (fmm.hpp:167-374): Y_n^m is
// sqrt((n-|m|)!/(n+|m|)!) P_n^{|m|}(cos a) e^{i m b} with the Condon-Shortley
// phase, A_n^m = (-1)^n / sqrt((n-|m|)!(n+|m|)!), and
//   P2M  M_n^m = sum_i m_i rho_i^n Y_n^{-m}                       (fmm.hpp:224-244)
//   M2M  Greengard's shift with i^{|k|-|m|-|k-m|}                  (fmm.hpp:247-274)
//   M2L  L_j^k = sum M_n^m i^{|k-m|-|k|-|m|} A_n^m A_j^k Y_{j+n}^{m-k}
//              / ((-1)^n A_{j+n}^{m-k} rho^{j+n+1})                (fmm.hpp:277-304)
//   L2L  with i^{|m|-|m-k|-|k|}                                   (fmm.hpp:307-329)
//   L2P  gradient of sum L_l^m r^l Y_l^m                           (fmm.hpp:332-374)
// Storage differs: only m >= 0 is kept (M_n^{-m} = conj M_n^m), as (p+1)^2
// reals per cell, and every expansion is scaled by the cell width w of its
// level (M~_n = M_n / w^n, L~_j = L_j w^{j+1}).  Then every operator depends
// only on the cell offset in units of w, which is the same at every level
// (all levels share the root box's aspect ratio), so one set of operators
// (8 M2M, 8 L2L, 316 M2L) serves the whole tree.  Operators are built in FP64
// on the device for the current root box before each evaluation.
// Solid harmonics R_n^m = r^n Y_n^m are evaluated with Cartesian recurrences
// (no trigonometry), gradients with the ladder relations
//   dR_n^m/dz = sqrt((n+m)(n-m)) R_{n-1}^m,
//   (d/dx + i d/dy) R_n^m = sqrt((n-m)(n-m-1)) R_{n-1}^{m+1},
//   (d/dx - i d/dy) R_n^m = -sqrt((n+m)(n+m-1)) R_{n-1}^{m-1}   (m >= 1).
#include "gravity.hpp"

#include <cmath>
#include <mutex>

namespace vg {
namespace {

constexpr int AL = 2 * MAXP;        // A_n^m needed up to n = 2p
constexpr int AW = 2 * AL + 1;
__constant__ double cA[(AL + 1) * AW];
// R recurrences: diag[m], sub[m], r1[n][m], r2[n][m]
__constant__ double cDiag[MAXP + 1], cSub[MAXP + 1], cR1[(MAXP + 1) * (MAXP + 1)], cR2[(MAXP + 1) * (MAXP + 1)];
// ladder coefficients of the gradient: dz[n][m] = sqrt((n+m)(n-m)), am[n][m] = sqrt((n+m)(n+m-1)), bm[n][m] = sqrt((n-m)(n-m-1))
__constant__ double cLz[(MAXP + 2) * (MAXP + 2)], cLa[(MAXP + 2) * (MAXP + 2)], cLb[(MAXP + 2) * (MAXP + 2)];

struct dc {
    double re, im;
};
__device__ inline dc cmul(dc a, dc b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline dc cconj(dc a) { return {a.re, -a.im}; }
__device__ inline dc cscale(dc a, double s) { return {a.re * s, a.im * s}; }
__device__ inline dc cadd(dc a, dc b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline dc csub(dc a, dc b) { return {a.re - b.re, a.im - b.im}; }

__device__ inline double Acoef(int n, int m) { return cA[n * AW + AL + m]; }
__device__ inline double ipow_even(int e) { return ((e / 2) % 2 == 0) ? 1.0 : -1.0; }

// Y_l^mu of the reference (fmm.hpp:169-194) for one direction; tab holds m >= 0
// at l(l+1)/2 + m as (re, im).
__device__ inline dc Yget(const double *tab, int l, int mu) {
    const int am = mu < 0 ? -mu : mu;
    const dc v{tab[2 * (l * (l + 1) / 2 + am)], tab[2 * (l * (l + 1) / 2 + am) + 1]};
    return mu < 0 ? cconj(v) : v;
}

constexpr int HSTRIDE = 2 * ((AL + 1) * (AL + 2) / 2) + 2;  // per offset: Y table + rho

// Y_l^m(direction) for l <= lmax and |v| of each offset vector.
// One thread per (offset, m): P_l^m by the upward recurrence in l, with
// P_m^m = (-1)^m (2m-1)!! sin^m, and sqrt((l-m)!/(l+m)!) carried along l.
// Offset vectors in units of a level's x width, the cell aspect ratio being
// (1, ry, rz): 0..7 M2M (child centre - parent centre of child octant o),
// 8..15 L2L (their negatives), 16 + o M2L (source - target cell, o in base 7
// holding (dx + 3, dy + 3, dz + 3)).
__global__ __launch_bounds__(64) void k_harmonics(const int16_t *__restrict__ sel, int nvec, double ry, double rz, int lmax, double *__restrict__ out) {
    const int o = sel[blockIdx.x];  // the vectors any kernel reads (see operators_build)
    const int m = threadIdx.x;
    if (o >= nvec || m > lmax) return;
    double vx, vy, vz;
    if (o < 16) {
        const int q = o & 7;
        const double cx = (q & 1) ? -0.25 : 0.25, cy = ((q >> 1) & 1) ? -0.25 * ry : 0.25 * ry, cz = ((q >> 2) & 1) ? -0.25 * rz : 0.25 * rz;
        vx = o < 8 ? cx : -cx;
        vy = o < 8 ? cy : -cy;
        vz = o < 8 ? cz : -cz;
    } else {
        const int e = o - 16;
        vx = e % 7 - 3;
        vy = ((e / 7) % 7 - 3) * ry;
        vz = (e / 49 - 3) * rz;
    }
    double *tab = out + static_cast<size_t>(o) * HSTRIDE;
    const double rho = sqrt(vx * vx + vy * vy + vz * vz);
    if (m == 0) {
        tab[HSTRIDE - 2] = rho;
        tab[HSTRIDE - 1] = 0;
    }
    const double rxy = sqrt(vx * vx + vy * vy);
    const double ct = rho > 0 ? vz / rho : 1.0, st = rho > 0 ? rxy / rho : 0.0;
    const dc eph = rxy > 0 ? dc{vx / rxy, vy / rxy} : dc{1.0, 0.0};
    dc em{1.0, 0.0};  // e^{i m phi}
    double pmm = 1.0, f = 1.0;  // P_m^m, (l-m)!/(l+m)! at l = m: 1/(2m)!
    for (int i = 1; i <= m; ++i) {
        pmm *= -(2.0 * i - 1) * st;
        em = cmul(em, eph);
        f /= (2.0 * i - 1) * (2.0 * i);
    }
    double p0 = 0, p1 = pmm;
    for (int l = m; l <= lmax; ++l) {
        double pl;
        if (l == m)
            pl = pmm;
        else if (l == m + 1)
            pl = ct * (2.0 * m + 1) * pmm;
        else
            pl = ((2.0 * l - 1) * ct * p1 - (l + m - 1.0) * p0) / (l - m);
        if (l > m) {
            p0 = p1;
            p1 = pl;
            f *= static_cast<double>(l - m) / static_cast<double>(l + m);
        }
        const double nrm = sqrt(f) * pl;
        tab[2 * (l * (l + 1) / 2 + m)] = nrm * em.re;
        tab[2 * (l * (l + 1) / 2 + m) + 1] = nrm * em.im;
    }
}

// Row r of an operator in the real-packed basis: degree j, order k, part
// 0 = Re (cosine) or 1 = Im (sine).  Column pair (n, m): Re at kidx_re, Im at
// kidx_im, from the complex coefficients T(+m), T(-m).
__device__ inline void row_of(int r, int &j, int &k, int &part) {
    j = static_cast<int>(sqrt(static_cast<double>(r)));
    while (j * j > r) --j;
    while ((j + 1) * (j + 1) <= r) ++j;
    const int q = r - j * j;
    if (q == 0) {
        k = 0;
        part = 0;
    } else {
        k = (q + 1) / 2;
        part = (q % 2 == 1) ? 0 : 1;
    }
}

// op block layout: [M2M 0..7][L2L 8..15][M2L 16..16+342], each K*K; M2L
// blocks row-major [row][col], M2M and L2L blocks column-major [col][row] (their
// kernels give each lane one output row, so lanes then read consecutive words).
// One block per (operator, OPS_ROWS rows), 128 threads per row r; thread t of
// a row takes the column pair (n, m >= 0) number t, i.e. the real-packed
// columns of degree n, order m.
constexpr int OPS_ROWS = 4;
template <class R>
__global__ __launch_bounds__(128 * OPS_ROWS) void k_ops(const int16_t *__restrict__ sel, const double *__restrict__ harm, int p, int K, R *__restrict__ op) {
    const int blk = sel[blockIdx.x];  // operator index
    const int r = blockIdx.y * OPS_ROWS + static_cast<int>(threadIdx.x >> 7);
    const int t = threadIdx.x & 127;
    const bool cmaj = blk < 16;
    R *dst = cmaj ? op + static_cast<size_t>(blk) * K * K + r : op + (static_cast<size_t>(blk) * K + r) * K;
    const size_t cs = cmaj ? static_cast<size_t>(K) : 1;  // column stride
    if (r < K)
        for (int c = t; c < K; c += 128) dst[c * cs] = R(0);
    __syncthreads();
    if (r >= K) return;
    int j, k, part;
    row_of(r, j, k, part);
    if (t >= (p + 1) * (p + 2) / 2) return;
    int n = 0;
    while ((n + 1) * (n + 2) / 2 <= t) ++n;
    const int mp = t - n * (n + 1) / 2;  // column pair (n, mp)
    const double *tab = harm + static_cast<size_t>(blk) * HSTRIDE;
    const double rho = tab[HSTRIDE - 2];
    dc tpm[2] = {{0, 0}, {0, 0}};
    if (blk < 8) {
        // M2M, child offset q: column (j', k') = (n, mp) of the child expansion, scaled by 2^{-j'} (fmm.hpp:256-270)
        const int jp = n, kp = mp;
        if (jp > j) return;
        const int nn = j - jp;
        const double rn = pow(rho, nn);
        const double scale = ldexp(1.0, -jp);
        for (int sgn = 0; sgn < 2; ++sgn) {
            const int kk = sgn == 0 ? kp : -kp;
            const int m = k - kk;
            if (m < -nn || m > nn) continue;
            const double coef = ipow_even(abs(k) - abs(m) - abs(k - m)) * Acoef(nn, m) * Acoef(jp, kk) * rn / Acoef(j, k) * scale;
            tpm[sgn] = cscale(Yget(tab, nn, -m), coef);
        }
    } else if (blk < 16) {
        // L2L, child offset q: column (n, m) of the parent expansion, scaled by 2^{-(j+1)} (fmm.hpp:313-327)
        if (n < j) return;
        const double scale = ldexp(1.0, -(j + 1));
        const double rn = pow(rho, n - j);
        for (int sgn = 0; sgn < 2; ++sgn) {
            const int m = sgn == 0 ? mp : -mp;
            if (m - k < -(n - j) || m - k > n - j) continue;
            const double coef = ipow_even(abs(m) - abs(m - k) - abs(k)) * Acoef(n - j, m - k) * Acoef(j, k) * rn /
                                (((n + j) % 2 ? -1.0 : 1.0) * Acoef(n, m)) * scale;
            tpm[sgn] = cscale(Yget(tab, n - j, m - k), coef);
        }
    } else {
        // M2L, offset o (fmm.hpp:287-302); rows of the 27 near offsets stay zero
        const int o = blk - 16;
        const int ox = o % 7 - 3, oy = (o / 7) % 7 - 3, oz = o / 49 - 3;
        if (ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1 && oz >= -1 && oz <= 1) return;
        const double rj = pow(rho, j + n + 1);
        for (int sgn = 0; sgn < 2; ++sgn) {
            if (mp == 0 && sgn == 1) continue;
            const int m = sgn == 0 ? mp : -mp;
            const double coef = ipow_even(abs(k - m) - abs(k) - abs(m)) * Acoef(n, m) * Acoef(j, k) / ((n % 2 ? -1.0 : 1.0) * Acoef(j + n, m - k) * rj);
            tpm[sgn] = cscale(Yget(tab, j + n, m - k), coef);
        }
    }
    // the pair's real-packed columns (each written by this thread only)
    if (mp == 0) {
        dst[kidx_re(n, 0) * cs] = static_cast<R>(part == 0 ? tpm[0].re : tpm[0].im);
    } else {
        const dc sum = cadd(tpm[0], tpm[1]), dif = csub(tpm[0], tpm[1]);
        dst[kidx_re(n, mp) * cs] = static_cast<R>(part == 0 ? sum.re : sum.im);
        dst[kidx_im(n, mp) * cs] = static_cast<R>(part == 0 ? -dif.im : dif.re);
    }
}

// ---------------------------------------------------------------------------
// P2M: one wavefront per leaf.  Lane = g NM + i, i < NM = p / 2 + 1: particle
// group g (G = 64 / NM groups, particles g, g + G, ...) and the order pair
// (m1, m2) = (i, p - i), so that every lane walks the same number of
// recurrence steps: R_{m1}^{m1} .. R_p^{m1} (p - m1 + 1 values), then
// R_{m2}^{m2} .. R_p^{m2} (p - m2 + 1; none when m2 = m1), p + 2 in all.
// Every lane forms the diagonal R_q^q, q <= p, and keeps the two it starts
// from; step s accumulates m_i conj(R) into accumulator s.  The groups are
// then summed in fixed order through LDS by all 64 lanes.
// ---------------------------------------------------------------------------
template <class R, int P>
__global__ __launch_bounds__(256) void k_p2m(const int32_t *__restrict__ leaves, int nleaves, const int8_t *__restrict__ level,
                                             const uint32_t *__restrict__ start, const uint32_t *__restrict__ count, const double *__restrict__ bmin,
                                             const double *__restrict__ bmax, const double *__restrict__ xs64, double wx0, int K, int ldk,
                                             R *__restrict__ M) {
    constexpr int p = P;       // the order is a template argument, so the recurrence unrolls fully
    constexpr int NS = P + 2;  // recurrence steps per lane
    __shared__ R red[4][64][2 * NS];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = blockIdx.x * 4 + w;
    if (li >= nleaves) return;  // whole wavefront: no barrier below spans wavefronts
    const int32_t c = leaves[li];
    if (level[c] < 2) return;  // solver_ref.hpp:187-188
    const int NM = p / 2 + 1, G = 64 / NM;
    const int i = lane % NM, g = lane / NM;
    const bool act = g < G;
    const int m1 = i, m2 = p - i;
    const int len1 = p - m1 + 1, len2 = m2 != m1 ? p - m2 + 1 : 0, tot = len1 + len2;
    const uint32_t s = start[c], cnt = count[c];
    const double iw = 1.0 / ldexp(wx0, -level[c]);
    const double cx = (bmin[3 * c] + bmax[3 * c]) / 2, cy = (bmin[3 * c + 1] + bmax[3 * c + 1]) / 2, cz = (bmin[3 * c + 2] + bmax[3 * c + 2]) / 2;
    // R_m^m = (prod_{q<=m} cDiag[q]) (x + i y)^m: the products for m1 and m2
    double sc1 = 1.0, sc2 = 1.0;
    for (int q = 1; q <= m2; ++q) {
        if (q <= m1) sc1 *= cDiag[q];
        sc2 *= cDiag[q];
    }
    const R s1 = static_cast<R>(sc1), s2 = static_cast<R>(sc2);
    // recurrence coefficients of each step (unused at the two chain starts)
    R c1[NS], c2[NS], are[NS], aim[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int n = k < len1 ? m1 + k : m2 + (k - len1), m = k < len1 ? m1 : m2;
        const int idx = (k < tot && n <= MAXP) ? n * (MAXP + 1) + m : 0;
        c1[k] = static_cast<R>(cR1[idx]);
        c2[k] = static_cast<R>(cR2[idx]);
        are[k] = 0;
        aim[k] = 0;
    }
    if (act) {
        // the next particle of this lane group is loaded before the current
        // one's harmonics are formed, so its latency hides behind them
        const double2 *xs2 = reinterpret_cast<const double2 *>(xs64);
        double2 nxy = make_double2(0, 0), nzm = make_double2(0, 0);
        if (g < cnt) {
            nxy = xs2[2 * static_cast<size_t>(s + g)];
            nzm = xs2[2 * static_cast<size_t>(s + g) + 1];
        }
        for (uint32_t j = g; j < cnt; j += G) {
            const double2 cxy = nxy, czm = nzm;
            if (j + G < cnt) {
                nxy = xs2[2 * static_cast<size_t>(s + j + G)];
                nzm = xs2[2 * static_cast<size_t>(s + j + G) + 1];
            }
            const R x = static_cast<R>((cxy.x - cx) * iw), y = static_cast<R>((cxy.y - cy) * iw), z = static_cast<R>((czm.x - cz) * iw);
            const R mi = static_cast<R>(czm.y);
            const R r2 = x * x + y * y + z * z;
            // (x + i y)^m1 and ^m2 by repeated squaring: w^(2^b) once, a
            // multiply per set bit of each exponent (selects, no branches)
            R d1re = 1, d1im = 0, d2re = 1, d2im = 0;
            {
                R wre = x, wim = y;  // w^(2^b)
                for (int b = 0; (1 << b) <= p; ++b) {
                    const R t1re = d1re * wre - d1im * wim, t1im = d1re * wim + d1im * wre;
                    const R t2re = d2re * wre - d2im * wim, t2im = d2re * wim + d2im * wre;
                    const bool u1 = (m1 >> b) & 1, u2 = (m2 >> b) & 1;
                    d1re = u1 ? t1re : d1re;
                    d1im = u1 ? t1im : d1im;
                    d2re = u2 ? t2re : d2re;
                    d2im = u2 ? t2im : d2im;
                    const R sre = wre * wre - wim * wim, sim = 2 * wre * wim;
                    wre = sre;
                    wim = sim;
                }
            }
            d1re *= s1;
            d1im *= s1;
            d2re *= s2;
            d2im *= s2;
            R bre = 0, bim = 0, cre = 0, cim = 0;  // R_{n-2}^m, R_{n-1}^m
#pragma unroll
            for (int st = 0; st < NS; ++st) {
                if (st > p + 1) break;
                if (st == 0) {
                    cre = d1re;
                    cim = d1im;
                } else {
                    // the second chain starts at step len1: R_{m2}^{m2}, with R_{m2-1}^{m2} = 0
                    const bool rs = st == len1;
                    const R nre = c1[st] * z * cre - c2[st] * r2 * bre, nim = c1[st] * z * cim - c2[st] * r2 * bim;
                    bre = rs ? R(0) : cre;
                    bim = rs ? R(0) : cim;
                    cre = rs ? d2re : nre;
                    cim = rs ? d2im : nim;
                }
                if (st < tot) {
                    are[st] += mi * cre;
                    aim[st] -= mi * cim;  // conj
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        red[w][lane][2 * k] = are[k];
        red[w][lane][2 * k + 1] = aim[k];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // output (i, step, part) summed over the groups in fixed order
    R *out = M + static_cast<size_t>(c) * ldk;
    for (int o = lane; o < NM * NS * 2; o += 64) {
        const int oi = o / (2 * NS), ost = (o / 2) % NS, part = o & 1;
        const int om1 = oi, om2 = p - oi, ol1 = p - om1 + 1, ol2 = om2 != om1 ? p - om2 + 1 : 0;
        if (ost >= ol1 + ol2) continue;
        const int n = ost < ol1 ? om1 + ost : om2 + (ost - ol1), m = ost < ol1 ? om1 : om2;
        if (part == 1 && m == 0) continue;
        R sum = 0;
        for (int gg = 0; gg < G; ++gg) sum += red[w][gg * NM + oi][2 * ost + part];
        out[part == 0 ? kidx_re(n, m) : kidx_im(n, m)] = sum;
    }
    for (int q = K + lane; q < ldk; q += 64) out[q] = R(0);
}

// M2M for the internal cells of one level (solver_ref.hpp:196-197).
// M2M / L2L of one tree level, PB parents per block: every operator word read
// from L2 serves PB parents, and PB is chosen per level (pick_pb) so that the
// launch still has enough blocks to fill the chip.  Operator blocks are
// column-major, so the lanes (output rows) read consecutive words; the
// staged expansions are [col][parent] so a lane reads its PB inputs as vectors.
template <class R, int PB>
__global__ __launch_bounds__(256) void k_m2m(const int32_t *__restrict__ ids, int nids, const int32_t *__restrict__ child,
                                             const uint32_t *__restrict__ count, const R *__restrict__ op, int K, int ldk, R *__restrict__ M) {
    __shared__ R cm[8 * MAXK][PB];  // [child q][col][parent]
    const int p0 = blockIdx.x * PB;
    const int np = min(PB, nids - p0);
    for (int i = threadIdx.x; i < PB * 8 * K; i += 256) {
        const int u = i % PB, qc = i / PB;  // qc = q K + col
        const int q = qc / K, col = qc % K;
        R v = 0;
        if (u < np) {
            const int32_t cb = child[ids[p0 + u]] + q;
            if (count[cb]) v = M[static_cast<size_t>(cb) * ldk + col];
        }
        cm[qc][u] = v;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < K; r += 256) {
        int j = 0;
        while ((j + 1) * (j + 1) <= r) ++j;
        const int cend = (j + 1) * (j + 1);  // output degree j uses child degrees <= j
        R acc[PB];
#pragma unroll
        for (int u = 0; u < PB; ++u) acc[u] = 0;
        for (int q = 0; q < 8; ++q) {
            const R *ocol = op + static_cast<size_t>(q) * K * K + r;
            for (int col = 0; col < cend; ++col) {
                const R a = ocol[static_cast<size_t>(col) * K];
#pragma unroll
                for (int u = 0; u < PB; ++u) acc[u] += a * cm[q * K + col][u];  // empty children are zero
            }
        }
#pragma unroll
        for (int u = 0; u < PB; ++u)
            if (u < np) M[static_cast<size_t>(ids[p0 + u]) * ldk + r] = acc[u];
    }
    for (int i = K + threadIdx.x; i < ldk; i += 256)
        for (int u = 0; u < np; ++u) M[static_cast<size_t>(ids[p0 + u]) * ldk + i] = R(0);
}

template <class R, int PB>
__global__ __launch_bounds__(256) void k_l2l(const int32_t *__restrict__ ids, int nids, const int32_t *__restrict__ child,
                                             const uint32_t *__restrict__ count, const R *__restrict__ op, int K, int ldk, R *__restrict__ L) {
    __shared__ R pl[MAXK][PB];  // [col][parent]
    const int p0 = blockIdx.x * PB;
    const int np = min(PB, nids - p0);
    for (int i = threadIdx.x; i < PB * K; i += 256) {
        const int u = i % PB, col = i / PB;
        pl[col][u] = u < np ? L[static_cast<size_t>(ids[p0 + u]) * ldk + col] : R(0);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 8 * K; i += 256) {
        const int q = i / K, r = i % K;
        const R *ocol = op + static_cast<size_t>(8 + q) * K * K + r;
        int j = 0;
        while ((j + 1) * (j + 1) <= r) ++j;
        R acc[PB];
#pragma unroll
        for (int u = 0; u < PB; ++u) acc[u] = 0;
        for (int col = j * j; col < K; ++col) {  // parent degrees >= j
            const R a = ocol[static_cast<size_t>(col) * K];
#pragma unroll
            for (int u = 0; u < PB; ++u) acc[u] += a * pl[col][u];
        }
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            if (u >= np) continue;
            const int32_t cc = child[ids[p0 + u]] + q;
            if (count[cc]) L[static_cast<size_t>(cc) * ldk + r] += acc[u];
        }
    }
}



// ---------------------------------------------------------------------------
// Plain M2L: targets are the children of octant qt of a chunk of TB internal
// parents; for each well-separated (parent neighbour, source octant) pair the
// (static) cell offset selects one operator, applied as a K x K by K x TB
// product with the gathered source multipoles (offset-grouped GEMM on VALU).
// As on the matrix cores, every block walks octant 0's 189 offsets (m2l_walk)
// and the mirror signs D_qt go on the gathered multipoles and on the result,
// so all blocks share 189 operators.  Blocks are placed XCD-aware: the eight
// octant blocks of a parent chunk run side by side on one XCD and share its
// L2.  The prologue tabulates, per parent neighbour and target, the first
// child and the mask of non-empty children, and from them the offsets that
// hold any source; only those are walked.  With nsplit > 1 the walked offsets
// are cut into nsplit consecutive parts, one block each, which write partial
// sums per target slot (parent index * 8 + octant) to part; k_m2l_sum adds
// them in part order.  A block has RG groups of 256
// threads; a thread owns the rows tr + 32 (rg + RG i), i < RPT / RG, and CPT
// consecutive targets; the operator streams through LDS in chunks of KC
// columns, double-buffered through registers (one barrier per chunk), and the
// next offset's first chunk is fetched during the last.
// ---------------------------------------------------------------------------
template <class R, int TB, int RPT, int RG>
__global__ __launch_bounds__(256 * RG) void k_m2l_plain(const int32_t *__restrict__ parents, int nparents, const int32_t *__restrict__ child,
                                                   const uint32_t *__restrict__ count, const int32_t *__restrict__ nbr27, const R *__restrict__ Mx,
                                                   R *__restrict__ Lx, const R *__restrict__ op, int K, int ldk, int nsplit,
                                                   R *__restrict__ part) {
    constexpr int KC = sizeof(R) == 4 ? 16 : 8;     // operator columns per chunk
    constexpr int CPT = TB / 8;                      // targets per thread
    constexpr int VW = 16 / sizeof(R);               // elements per 16-byte LDS read
    constexpr int SMS = TB + VW;                     // multipole column stride (16-byte rows)
    constexpr int SOS = 32 * RPT + 1;                // operator column stride (odd: no bank conflicts)
    constexpr int NTH = 256 * RG;                    // threads
    constexpr int RPL = RPT / RG;                    // rows per thread
    constexpr int NPF = KC * 32 * RPT / NTH;         // prefetched operator words per thread
    static_assert(CPT % VW == 0 && (KC * 32 * RPT) % NTH == 0 && RPT % RG == 0, "tile shape");
    extern __shared__ __align__(16) unsigned char m2l_smem[];
    R *sM = reinterpret_cast<R *>(m2l_smem);         // K x SMS: column k of the gathered, signed multipoles
    R *sOp = sM + K * SMS;                           // 2 x KC x SOS: operator columns, transposed
    __shared__ R sg[32 * RPT];                       // D_qt
    __shared__ int32_t s_nc[27 * TB];                // first child of parent neighbour d of target t, or -1
    __shared__ uint8_t s_nm[27 * TB];                // its non-empty children
    __shared__ int32_t tgt[TB];
    __shared__ int16_t s_o[189], s_dn[189], s_qs[189], s_live[189];
    __shared__ int s_nlive;
    const int nchunks = (nparents + TB - 1) / TB;
    const int per = nchunks * nsplit;                          // work items per XCD
    const int wi = (blockIdx.x & 7) * per + (blockIdx.x >> 3);  // XCD x takes work items [x W / 8, (x + 1) W / 8)
    if (wi >= 8 * per) return;
    const int sp = wi % nsplit, w = wi / nsplit;  // the parts of one (chunk, octant) are neighbours
    const int chunk = w >> 3, qt = w & 7;
    const int tid = threadIdx.x, tr = tid & 31, tc = (tid >> 5) & 7, rg = tid >> 8;
    if (tid == 0) m2l_walk(qt, s_o, s_dn, s_qs);
    for (int k = tid; k < 32 * RPT; k += NTH) sg[k] = static_cast<R>(k < K ? m2l_mirror_sign(qt, k) : 0);
    if (tid < TB) {
        const int pi = chunk * TB + tid;
        const int32_t P = pi < nparents ? parents[pi] : -1;
        int32_t t = P >= 0 ? child[P] + qt : -1;
        if (t >= 0 && count[t] == 0) t = -1;
        tgt[tid] = t;
    }
    __syncthreads();
    for (int i = tid; i < 27 * TB; i += NTH) {
        const int d = i / TB, t = i % TB;
        const int pi = chunk * TB + t;
        int32_t cb = -1;
        uint32_t msk = 0;
        if (tgt[t] >= 0) {
            const int32_t sp = nbr27[27 * static_cast<int64_t>(parents[pi]) + d];
            cb = sp >= 0 ? child[sp] : -1;
            if (cb >= 0)
                for (int q = 0; q < 8; ++q) msk |= (count[cb + q] != 0 ? 1u : 0u) << q;
        }
        s_nc[i] = cb;
        s_nm[i] = static_cast<uint8_t>(msk);
    }
    __syncthreads();
    if (tid < 189) {
        const int dn = s_dn[tid], qs = s_qs[tid];
        int live = 0;
        for (int t = 0; t < TB; ++t) live |= (s_nm[dn * TB + t] >> qs) & 1;
        s_live[tid] = static_cast<int16_t>(live);
    }
    __syncthreads();
    if (tid == 0) {
        int nl = 0;
        for (int st = 0; st < 189; ++st)
            if (s_live[st]) s_live[nl++] = static_cast<int16_t>(st);
        s_nlive = nl;
    }
    __syncthreads();
    const int nlive = s_nlive;
    const int lo = sp * nlive / nsplit, hi = (sp + 1) * nlive / nsplit;  // this block's walked offsets
    R acc[RPL][CPT];
#pragma unroll
    for (int i = 0; i < RPL; ++i)
#pragma unroll
        for (int jj = 0; jj < CPT; ++jj) acc[i][jj] = 0;
    const int nkc = (K + KC - 1) / KC;
    const size_t KK = static_cast<size_t>(K) * K;
    R pf[NPF];
    // operator chunk (offset step li, columns [c0, c0 + KC)) into registers:
    // word u of thread tid is row r = i / KC, column c0 + i % KC, i = tid + NTH u
    auto load_op = [&](int li, int c0) {
        const R *O = op + (16 + s_o[s_live[li]]) * KK;
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int i = tid + NTH * u, r = i / KC, c = c0 + i % KC;
            pf[u] = (r < K && c < K) ? O[static_cast<size_t>(r) * K + c] : R(0);
        }
    };
    auto store_op = [&](int b) {
        R *S = sOp + b * KC * SOS;
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int i = tid + NTH * u;
            S[(i % KC) * SOS + i / KC] = pf[u];
        }
    };
    auto fma_col = [&](const R *S, int col) {
        R mv[CPT];
#pragma unroll
        for (int v = 0; v < CPT / VW; ++v) {
            typedef R vec __attribute__((ext_vector_type(VW)));
            const vec x = *reinterpret_cast<const vec *>(sM + col * SMS + tc * CPT + v * VW);
#pragma unroll
            for (int e = 0; e < VW; ++e) mv[v * VW + e] = x[e];
        }
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            const R ov = S[tr + 32 * (rg + RG * i)];
#pragma unroll
            for (int jj = 0; jj < CPT; ++jj) acc[i][jj] += ov * mv[jj];
        }
    };
    if (hi > lo) load_op(lo, 0);
    const int wv = tid >> 6, lane = tid & 63;
    for (int li = lo; li < hi; ++li) {
        const int st = s_live[li], dn = s_dn[st], qs = s_qs[st];
        // gather: wave wv takes targets wv, wv + 4 RG, ..., lanes the columns
        for (int t = wv; t < TB; t += 4 * RG) {
            const int32_t cb = s_nc[dn * TB + t];
            const bool has = (s_nm[dn * TB + t] >> qs) & 1;
            const R *src = Mx + static_cast<size_t>(has ? cb + qs : 0) * ldk;
            for (int col = lane; col < K; col += 64) sM[col * SMS + t] = has ? sg[col] * src[col] : R(0);
        }
        store_op(0);
        if (nkc > 1) load_op(li, KC);
        else if (li + 1 < hi) load_op(li + 1, 0);
        __syncthreads();  // multipoles and chunk 0 are in LDS
        for (int kc = 0; kc < nkc; ++kc) {
            const int c0 = kc * KC, b = kc & 1;
            const R *S = sOp + b * KC * SOS;
            if (c0 + KC <= K) {
#pragma unroll
                for (int cc = 0; cc < KC; ++cc) fma_col(S + cc * SOS, c0 + cc);
            } else {
                for (int cc = 0; cc < K - c0; ++cc) fma_col(S + cc * SOS, c0 + cc);
            }
            if (kc + 1 < nkc) {
                store_op(b ^ 1);
                if (kc + 2 < nkc) load_op(li, c0 + 2 * KC);
                else if (li + 1 < hi) load_op(li + 1, 0);
            }
            __syncthreads();  // chunk kc is read; chunk kc + 1 is in LDS
        }
    }
#pragma unroll
    for (int jj = 0; jj < CPT; ++jj) {
        const int32_t t = tgt[tc * CPT + jj];
        if (t < 0) continue;
        R *dst = nsplit == 1 ? Lx + static_cast<size_t>(t) * ldk
                             : part + (static_cast<size_t>(sp) * 8 * nparents + static_cast<size_t>(chunk * TB + tc * CPT + jj) * 8 + qt) * ldk;
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            const int r = tr + 32 * (rg + RG * i);
            if (r < K) dst[r] = sg[r] * acc[i][jj];
        }
    }
}

// The nsplit partial sums of each non-empty M2L target into its local
// expansion, in part order (one thread per target slot and coefficient).
template <class R>
__global__ __launch_bounds__(256) void k_m2l_sum(const int32_t *__restrict__ parents, int nparents, const int32_t *__restrict__ child,
                                                 const uint32_t *__restrict__ count, int nsplit, const R *__restrict__ part, R *__restrict__ Lx,
                                                 int K, int ldk) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= static_cast<int64_t>(8) * nparents * K) return;
    const int64_t slot = i / K;
    const int r = static_cast<int>(i % K);
    const int32_t t = child[parents[slot >> 3]] + static_cast<int32_t>(slot & 7);
    if (count[t] == 0) return;
    const size_t stride = static_cast<size_t>(8) * nparents * ldk;
    R v = part[static_cast<size_t>(slot) * ldk + r];
    for (int q = 1; q < nsplit; ++q) v += part[q * stride + static_cast<size_t>(slot) * ldk + r];
    Lx[static_cast<size_t>(t) * ldk + r] = v;
}

// ---------------------------------------------------------------------------
// L2P: one thread per particle of a leaf; the R_n^m recurrence feeds the
// potential, d/dz and (d/dx - i d/dy) directly, so no table is stored.
// ---------------------------------------------------------------------------
// Recurrence and ladder coefficients as functions of (n, m): with p a template
// parameter and the loops unrolled they fold to immediates (the same values
// as the __constant__ tables, which serve the kernels with a runtime order).
__device__ __forceinline__ double k_diag(int m) { return -sqrt((2.0 * m - 1) / (2.0 * m)); }
__device__ __forceinline__ double k_r1(int n, int m) { return (2.0 * n - 1) / sqrt(static_cast<double>(n - m) * (n + m)); }
__device__ __forceinline__ double k_r2(int n, int m) { return sqrt(static_cast<double>(n + m - 1) * (n - m - 1) / (static_cast<double>(n - m) * (n + m))); }
__device__ __forceinline__ double k_lz(int l, int m) { return m <= l ? sqrt(static_cast<double>(l + m) * (l - m)) : 0.0; }
__device__ __forceinline__ double k_la(int l, int m) { return m <= l ? sqrt(static_cast<double>(l + m) * (l + m - 1)) : 0.0; }
__device__ __forceinline__ double k_lb(int l, int m) { return m + 1 <= l ? sqrt(static_cast<double>(l - m) * (l - m - 1)) : 0.0; }

template <class R, int P>
__global__ __launch_bounds__(64) void k_l2p(const int32_t *__restrict__ leaves, const int8_t *__restrict__ level, const uint32_t *__restrict__ start,
                                            const uint32_t *__restrict__ count, const double *__restrict__ bmin, const double *__restrict__ bmax,
                                            const double *__restrict__ xs64, double wx0, int ldk, const R *__restrict__ Lx,
                                            double gscale, R *__restrict__ acc4, const uint8_t *__restrict__ skip) {
    __shared__ R sl[MAXK];
    if (skip && skip[blockIdx.x]) return;  // the leaf takes the closure's fallback instead
    const int32_t c = leaves[blockIdx.x];
    if (level[c] < 2) return;  // solver_ref.hpp:220-221
    constexpr int K = (P + 1) * (P + 1);
    for (int i = threadIdx.x; i < K; i += 64) sl[i] = Lx[static_cast<size_t>(c) * ldk + i];
    __syncthreads();
    const uint32_t s = start[c], cnt = count[c];
    const double w = ldexp(wx0, -level[c]);
    const double iw = 1.0 / w;
    const double cx = (bmin[3 * c] + bmax[3 * c]) / 2, cy = (bmin[3 * c + 1] + bmax[3 * c + 1]) / 2, cz = (bmin[3 * c + 2] + bmax[3 * c + 2]) / 2;
    for (uint32_t i = threadIdx.x; i < cnt; i += 64) {
        // read the local coefficients from LDS inside the loop rather than hold all of them in registers
        int z0 = 0;
        asm volatile("" : "+v"(z0));
        const R *slv = sl + z0;
        const uint32_t k = s + i;
        const R x = static_cast<R>((xs64[4 * k] - cx) * iw), y = static_cast<R>((xs64[4 * k + 1] - cy) * iw), z = static_cast<R>((xs64[4 * k + 2] - cz) * iw);
        const R r2 = x * x + y * y + z * z;
        R phi = 0, gz = 0, gre = 0, gim = 0;  // potential, d/dz, (d/dx - i d/dy)
        R dre = 1, dim = 0;
#pragma unroll
        for (int m = 0; m <= P; ++m) {
            if (m > 0) {
                const R d = static_cast<R>(k_diag(m));
                const R nre = d * (x * dre - y * dim), nim = d * (x * dim + y * dre);
                dre = nre;
                dim = nim;
            }
            R are = 0, aim = 0, cre = dre, cim = dim;  // R_{n-2}^m, R_n^m
#pragma unroll
            for (int n = m; n <= P; ++n) {
                if (n > m) {
                    const R c1 = static_cast<R>(k_r1(n, m)), c2 = static_cast<R>(k_r2(n, m));
                    const R nre = c1 * z * cre - c2 * r2 * are, nim = c1 * z * cim - c2 * r2 * aim;
                    are = cre;
                    aim = cim;
                    cre = nre;
                    cim = nim;
                }
                // potential: sum L_n^0 R_n^0 + 2 Re(L_n^m R_n^m)
                if (m == 0)
                    phi += slv[kidx_re(n, 0)] * cre;
                else
                    phi += 2 * (slv[kidx_re(n, m)] * cre - slv[kidx_im(n, m)] * cim);
                if (n + 1 <= P) {
                    const int l = n + 1;  // R_n^m enters the derivatives of degree-l terms
                    const R lz = static_cast<R>(k_lz(l, m));
                    if (m == 0)
                        gz += slv[kidx_re(l, 0)] * lz * cre;
                    else
                        gz += 2 * lz * (slv[kidx_re(l, m)] * cre - slv[kidx_im(l, m)] * cim);
                    {  // -sqrt((l+m+1)(l+m)) L_l^{m+1} R_n^m
                        const R la = static_cast<R>(k_la(l, m + 1));
                        const R Lre = slv[kidx_re(l, m + 1)], Lim = slv[kidx_im(l, m + 1)];
                        gre -= la * (Lre * cre - Lim * cim);
                        gim -= la * (Lre * cim + Lim * cre);
                    }
                    if (m >= 1) {  // sqrt((l-m+1)(l-m)) conj(L_l^{m-1} R_n^m), with Re(L_l^0) for m = 1
                        const R lb = static_cast<R>(k_lb(l, m - 1));
                        const R Lre = slv[kidx_re(l, m - 1)], Lim = (m - 1) == 0 ? R(0) : slv[kidx_im(l, m - 1)];
                        gre += lb * (Lre * cre - Lim * cim);
                        gim -= lb * (Lre * cim + Lim * cre);
                    }
                }
            }
        }
        const R gs = static_cast<R>(gscale * iw * iw);
        acc4[4 * k] += gs * gre;
        acc4[4 * k + 1] -= gs * gim;
        acc4[4 * k + 2] += gs * gz;
        acc4[4 * k + 3] += static_cast<R>(iw) * phi;
    }
}

template <class R>
struct Bufs;
template <>
struct Bufs<double> {
    static double *M(Ctx &c) { return c.M64; }
    static double *L(Ctx &c) { return c.L64; }
    static double *op(Ctx &c) { return c.op64; }
    static double *acc(Ctx &c) { return c.as64; }
    static void ensure(Ctx &c, size_t nc) {
        c.M64.ensure(nc * c.ldk + 16);
        c.L64.ensure(nc * c.ldk + 16);
    }
};
template <>
struct Bufs<float> {
    static float *M(Ctx &c) { return c.M32; }
    static float *L(Ctx &c) { return c.L32; }
    static float *op(Ctx &c) { return c.op32; }
    static float *acc(Ctx &c) { return c.as32; }
    static void ensure(Ctx &c, size_t nc) {
        c.M32.ensure(nc * c.ldk + 16);
        c.L32.ensure(nc * c.ldk + 16);
    }
};

// One parent per 1024-thread block, for the small levels near the root: the
// (child q, row r) items are spread over the threads; M2M then adds the 8
// children's partial rows through LDS.
template <class R>
__global__ __launch_bounds__(1024) void k_m2m_wide(const int32_t *__restrict__ ids, const int32_t *__restrict__ child, const uint32_t *__restrict__ count,
                                                   const R *__restrict__ op, int K, int ldk, R *__restrict__ M) {
    __shared__ R cm[8 * MAXK];
    __shared__ R part[8 * MAXK];
    const int32_t P = ids[blockIdx.x];
    const int32_t cb = child[P];
    for (int i = threadIdx.x; i < 8 * K; i += 1024) {
        const int q = i / K, col = i % K;
        cm[i] = count[cb + q] ? M[static_cast<size_t>(cb + q) * ldk + col] : R(0);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 8 * K; i += 1024) {
        const int q = i / K, r = i % K;
        int j = 0;
        while ((j + 1) * (j + 1) <= r) ++j;
        const R *ocol = op + static_cast<size_t>(q) * K * K + r;
        R sum = 0;
        for (int col = 0; col < (j + 1) * (j + 1); ++col) sum += ocol[static_cast<size_t>(col) * K] * cm[q * K + col];
        part[i] = sum;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < K; r += 1024) {
        R sum = 0;
        for (int q = 0; q < 8; ++q) sum += part[q * K + r];  // children in order
        M[static_cast<size_t>(P) * ldk + r] = sum;
    }
    for (int i = K + threadIdx.x; i < ldk; i += 1024) M[static_cast<size_t>(P) * ldk + i] = R(0);
}

template <class R>
__global__ __launch_bounds__(1024) void k_l2l_wide(const int32_t *__restrict__ ids, const int32_t *__restrict__ child, const uint32_t *__restrict__ count,
                                                   const R *__restrict__ op, int K, int ldk, R *__restrict__ L) {
    __shared__ R pl[MAXK];
    const int32_t P = ids[blockIdx.x];
    const int32_t cb = child[P];
    for (int i = threadIdx.x; i < K; i += 1024) pl[i] = L[static_cast<size_t>(P) * ldk + i];
    __syncthreads();
    for (int i = threadIdx.x; i < 8 * K; i += 1024) {
        const int q = i / K, r = i % K;
        if (!count[cb + q]) continue;
        int j = 0;
        while ((j + 1) * (j + 1) <= r) ++j;
        const R *ocol = op + static_cast<size_t>(8 + q) * K * K + r;
        R sum = 0;
        for (int col = j * j; col < K; ++col) sum += ocol[static_cast<size_t>(col) * K] * pl[col];
        L[static_cast<size_t>(cb + q) * ldk + r] += sum;
    }
}

inline int pick_pb(int64_t nparents, int maxpb) {
    int pb = 1;
    while (pb < maxpb && nparents >= static_cast<int64_t>(pb) * 2 * 512) pb *= 2;
    return pb;
}

template <class R>
void m2m_level(Ctx &c, const int32_t *ids, int64_t n) {
    const TreeDev &t = c.t;
    if (sizeof(R) == 4 && c.prec == 2) {  // split mode: on the matrix cores
        ml_split_level(c, ids, n, true);
        return;
    }
    const int pb = pick_pb(n, 8);  // 8 children x PB staged expansions in LDS
    if (n < 1024) {  // few parents: one wide block each
        k_m2m_wide<R><<<static_cast<unsigned>(n), 1024, 0, c.stream>>>(ids, t.child, t.count, Bufs<R>::op(c), c.K, c.ldk, Bufs<R>::M(c));
        return;
    }
#define VG_M2M(PB_) \
    k_m2m<R, PB_><<<ceil_div(n, PB_), 256, 0, c.stream>>>(ids, static_cast<int>(n), t.child, t.count, Bufs<R>::op(c), c.K, c.ldk, Bufs<R>::M(c))
    if (pb == 1) VG_M2M(1);
    else if (pb == 2) VG_M2M(2);
    else if (pb == 4) VG_M2M(4);
    else VG_M2M(8);
#undef VG_M2M
}

template <class R>
void l2l_level(Ctx &c, const int32_t *ids, int64_t n) {
    const TreeDev &t = c.t;
    if (sizeof(R) == 4 && c.prec == 2) {
        ml_split_level(c, ids, n, false);
        return;
    }
    const int pb = pick_pb(n, 16);
    if (n < 1024) {
        k_l2l_wide<R><<<static_cast<unsigned>(n), 1024, 0, c.stream>>>(ids, t.child, t.count, Bufs<R>::op(c), c.K, c.ldk, Bufs<R>::L(c));
        return;
    }
#define VG_L2L(PB_) \
    k_l2l<R, PB_><<<ceil_div(n, PB_), 256, 0, c.stream>>>(ids, static_cast<int>(n), t.child, t.count, Bufs<R>::op(c), c.K, c.ldk, Bufs<R>::L(c))
    if (pb == 1) VG_L2L(1);
    else if (pb == 2) VG_L2L(2);
    else if (pb == 4) VG_L2L(4);
    else if (pb == 8) VG_L2L(8);
    else VG_L2L(16);
#undef VG_L2L
}

template <class R>
void upward_t(Ctx &c) {
    TreeDev &t = c.t;
    Bufs<R>::ensure(c, t.ncells);
    const double wx0 = t.box[3] - t.box[0];
    if (t.nleaves) {
#define VG_P2M(P_)                                                                                                                          \
    k_p2m<R, P_><<<ceil_div(t.nleaves, 4), 256, 0, c.stream>>>(t.leaf_ids, static_cast<int>(t.nleaves), t.level, t.start, t.count, t.bmin, \
                                                               t.bmax, c.xs64, wx0, c.K, c.ldk, Bufs<R>::M(c))
        switch (c.p) {
            case 1: VG_P2M(1); break;
            case 2: VG_P2M(2); break;
            case 3: VG_P2M(3); break;
            case 4: VG_P2M(4); break;
            case 5: VG_P2M(5); break;
            case 6: VG_P2M(6); break;
            case 7: VG_P2M(7); break;
            case 8: VG_P2M(8); break;
            case 9: VG_P2M(9); break;
            case 10: VG_P2M(10); break;
            case 11: VG_P2M(11); break;
            case 12: VG_P2M(12); break;
            default: throw Error(1, "P2M supports 1 <= p <= 12");
        }
#undef VG_P2M
    }
    for (int d = t.depth; d >= 2; --d) {
        const int64_t a = t.int_off[d], b = t.int_off[d + 1];
        if (b > a) m2m_level<R>(c, t.int_ids.p + a, b - a);
    }
    VG_LAUNCH_CHECK();
}

template <class R, int TB, int RPT, int RG = 1>
void m2l_plain_launch(Ctx &c, const int32_t *parents, int np, unsigned grid, int nsplit, R *part) {
    constexpr int KC = sizeof(R) == 4 ? 16 : 8;
    const size_t lds = (static_cast<size_t>(c.K) * (TB + 16 / sizeof(R)) + 2 * KC * (32 * RPT + 1)) * sizeof(R);
    // dynamic LDS above 64 KiB is allowed per kernel and device, here for the
    // largest K of this RPT (set once per device, under a lock: contexts of
    // different host threads may launch concurrently)
    static std::mutex mu;
    static uint64_t done = 0;  // bit d: set on device d
    {
        std::lock_guard<std::mutex> lock(mu);
        const uint64_t bit = c.device < 64 ? (1ull << c.device) : 0;
        if (!bit || !(done & bit)) {
            const size_t cap = (static_cast<size_t>(32 * RPT) * (TB + 16 / sizeof(R)) + 2 * KC * (32 * RPT + 1)) * sizeof(R);
            VG_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_m2l_plain<R, TB, RPT, RG>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(cap)));
            done |= bit;
        }
    }
    const TreeDev &t = c.t;
    k_m2l_plain<R, TB, RPT, RG><<<grid * nsplit, 256 * RG, lds, c.stream>>>(parents, np, t.child, t.count, t.nbr27, Bufs<R>::M(c), Bufs<R>::L(c),
                                                                           Bufs<R>::op(c), c.K, c.ldk, nsplit, part);
    if (nsplit > 1)
        k_m2l_sum<R><<<ceil_div(static_cast<size_t>(8) * np * c.K, 256), 256, 0, c.stream>>>(parents, np, t.child, t.count, nsplit, part,
                                                                                          Bufs<R>::L(c), c.K, c.ldk);
    VG_LAUNCH_CHECK();
}

template <class R, int TB>
void m2l_plain_rows(Ctx &c, const int32_t *parents, int np) {
    // FP32 with an even number of 32-row groups: two thread groups split the
    // rows (512 threads, twice the wavefronts per CU within the same LDS)
    constexpr int RG2 = sizeof(R) == 4 ? 2 : 1;
    const int rpt = ceil_div(c.K, 32);
    const unsigned grid = 8 * ceil_div(np, TB);  // one block per (parent chunk, child octant)
    // still under four blocks per CU: the walked offsets of a (chunk, octant)
    // are cut into parts, one block each (at most 32 parts, so a part keeps
    // about six of the at most 189 offsets), and a second kernel adds the
    // partial sums (the 16k Plummer sphere at p = 4: 40 blocks, 26 parts,
    // FP64 M2L 0.46 -> 0.05 ms; cutting grids of up to eight blocks per CU
    // as well gained nothing on the 2^20 cube or the 4M sphere)
    int nsplit = 1;
    if (c.ncu > 0 && grid < 4u * static_cast<unsigned>(c.ncu)) nsplit = std::min<int>(32, static_cast<int>(ceil_div(4 * c.ncu, grid)));
    R *part = nullptr;
    if (nsplit > 1) {
        c.m2l_part.ensure(static_cast<size_t>(nsplit) * 8 * np * c.ldk * sizeof(R));
        part = reinterpret_cast<R *>(c.m2l_part.p);
    }
    switch (rpt) {
    case 1: m2l_plain_launch<R, TB, 1>(c, parents, np, grid, nsplit, part); break;
    case 2: m2l_plain_launch<R, TB, 2, RG2>(c, parents, np, grid, nsplit, part); break;
    case 3: m2l_plain_launch<R, TB, 3>(c, parents, np, grid, nsplit, part); break;
    case 4: m2l_plain_launch<R, TB, 4, RG2>(c, parents, np, grid, nsplit, part); break;
    case 5:  // FP32: padded to 6 row groups on two thread groups (p = 11 on the 4M Plummer input: 19.9 -> 16.7 ms;
             // padding 3 to 4 at p = 8 was slower, 3.0 -> 4.1 ms)
        if (sizeof(R) == 4)
            m2l_plain_launch<R, TB, 6, RG2>(c, parents, np, grid, nsplit, part);
        else
            m2l_plain_launch<R, TB, 5>(c, parents, np, grid, nsplit, part);
        break;
    case 6: m2l_plain_launch<R, TB, 6, RG2>(c, parents, np, grid, nsplit, part); break;
    default: throw std::runtime_error("plain M2L: expansion order out of range");
    }
}

template <class R>
void m2l_plain_list_t(Ctx &c, const int32_t *parents, int np) {
    if (np <= 0) return;
    constexpr int TB = sizeof(R) == 4 ? 64 : 32;
    // Fewer than four blocks per CU at TB targets (2^20 uniform points have
    // 4,680 parents: 592 blocks): half the targets per block, twice the
    // blocks, which keeps more wavefronts in flight per CU (FP32 p = 8 on that
    // input: 2.97 against 3.35 ms).
    if (8 * static_cast<int64_t>(ceil_div(np, TB)) < 4 * static_cast<int64_t>(c.ncu))
        m2l_plain_rows<R, TB / 2>(c, parents, np);
    else
        m2l_plain_rows<R, TB>(c, parents, np);
}

template <class R>
void m2l_plain_t(Ctx &c) {
    TreeDev &t = c.t;
    // parents: internal cells of levels >= 1 (their children are the targets at level >= 2)
    if (t.depth < 1) return;
    const int64_t a = t.int_off[1], b = t.int_off.back();
    if (b <= a) return;
    m2l_plain_list_t<R>(c, t.int_ids.p + a, static_cast<int>(b - a));
}

template <class R>
void l2p_t(Ctx &c, const int32_t *leaves, int64_t n, const uint8_t *skip) {
    const TreeDev &t = c.t;
    const double wx0 = t.box[3] - t.box[0];
    const double gscale = (c.flags & 1u) ? 1.0 : c.G;
    if (!n) return;
    const unsigned grid = static_cast<unsigned>(n);
#define VG_L2P(P_)                                                                                                                            \
    k_l2p<R, P_><<<grid, 64, 0, c.stream>>>(leaves, t.level, t.start, t.count, t.bmin, t.bmax, c.xs64, wx0, c.ldk, Bufs<R>::L(c), gscale, \
                                             Bufs<R>::acc(c), skip)
    switch (c.p) {
        case 1: VG_L2P(1); break;
        case 2: VG_L2P(2); break;
        case 3: VG_L2P(3); break;
        case 4: VG_L2P(4); break;
        case 5: VG_L2P(5); break;
        case 6: VG_L2P(6); break;
        case 7: VG_L2P(7); break;
        case 8: VG_L2P(8); break;
        case 9: VG_L2P(9); break;
        case 10: VG_L2P(10); break;
        case 11: VG_L2P(11); break;
        case 12: VG_L2P(12); break;
        default: throw Error(1, "L2P supports 1 <= p <= 12");
    }
#undef VG_L2P
    VG_LAUNCH_CHECK();
}

template <class R>
void downward_t(Ctx &c) {
    TreeDev &t = c.t;
    for (int d = 2; d < t.depth; ++d) {  // parents at level d >= 2 feed children at level >= 3
        const int64_t a = t.int_off[d], b = t.int_off[d + 1];
        if (b > a) l2l_level<R>(c, t.int_ids.p + a, b - a);
    }
    join_p2p(c);  // L2P adds to the P2P accelerations
    bool fallback = false;
    if constexpr (sizeof(R) == 4) {
        if (c.cl.on) {
            closure_apply(c);                         // learned correction of the leaf locals before L2P
            fallback = closure_fallback_prepare(c);   // order-p_hi locals where the estimate exceeds the bound
        }
    }
    l2p_t<R>(c, t.leaf_ids, t.nleaves, fallback ? c.cl.flag.p : nullptr);
    if constexpr (sizeof(R) == 4) {
        if (fallback) closure_fallback_l2p(c);
    }
}

}  // namespace

void m2l_split_operators(Ctx &c, int nvec);

// The tables live in each device's constant memory: upload them once per
// device (the caller has made the context's device current).
void fmm_constants_init() {
    static std::mutex mu;
    static std::vector<bool> done;
    int dev = 0;
    VG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if (static_cast<size_t>(dev) < done.size() && done[static_cast<size_t>(dev)]) return;
    std::vector<double> A((AL + 1) * AW, 0.0);
    for (int n = 0; n <= AL; ++n)
        for (int m = -n; m <= n; ++m) {
            const int am = m < 0 ? -m : m;
            double d = 1;
            for (int i = n - am + 1; i <= n + am; ++i) d *= i;
            d = std::sqrt(d);
            for (int i = 1; i <= n - am; ++i) d *= i;
            A[n * AW + AL + m] = (n % 2 == 0 ? 1.0 : -1.0) / d;  // fmm.hpp:201-221
        }
    std::vector<double> diag(MAXP + 1, 0), sub(MAXP + 1, 0), r1((MAXP + 1) * (MAXP + 1), 0), r2((MAXP + 1) * (MAXP + 1), 0);
    for (int m = 1; m <= MAXP; ++m) diag[m] = -std::sqrt((2.0 * m - 1) / (2.0 * m));
    for (int m = 0; m <= MAXP; ++m) sub[m] = std::sqrt(2.0 * m + 1);
    for (int n = 1; n <= MAXP; ++n)
        for (int m = 0; m < n; ++m) {
            r1[n * (MAXP + 1) + m] = (2.0 * n - 1) / std::sqrt(static_cast<double>(n - m) * (n + m));
            r2[n * (MAXP + 1) + m] = std::sqrt(static_cast<double>(n + m - 1) * (n - m - 1) / (static_cast<double>(n - m) * (n + m)));
        }
    const int LW = MAXP + 2;
    std::vector<double> lz(LW * LW, 0), la(LW * LW, 0), lb(LW * LW, 0);
    for (int l = 0; l < LW; ++l)
        for (int m = 0; m < LW; ++m) {
            lz[l * LW + m] = (m <= l) ? std::sqrt(static_cast<double>(l + m) * (l - m)) : 0.0;
            la[l * LW + m] = (m <= l) ? std::sqrt(static_cast<double>(l + m) * (l + m - 1)) : 0.0;
            lb[l * LW + m] = (m + 1 <= l) ? std::sqrt(static_cast<double>(l - m) * (l - m - 1)) : 0.0;
        }
    VG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cA), A.data(), A.size() * sizeof(double)));
    VG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cDiag), diag.data(), diag.size() * sizeof(double)));
    VG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cSub), sub.data(), sub.size() * sizeof(double)));
    VG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cR1), r1.data(), r1.size() * sizeof(double)));
    VG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cR2), r2.data(), r2.size() * sizeof(double)));
    VG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cLz), lz.data(), lz.size() * sizeof(double)));
    VG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cLa), la.data(), la.size() * sizeof(double)));
    VG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cLb), lb.data(), lb.size() * sizeof(double)));
    if (done.size() <= static_cast<size_t>(dev)) done.resize(static_cast<size_t>(dev) + 1, false);
    done[static_cast<size_t>(dev)] = true;
}

// The operator (and offset vector) indices that operators_build fills: the
// 16 M2M and L2L ones, then 16 + o for the 189 M2L offsets o of octant 0's
// walk.  A device table, written once per context.
void ops_selection(Ctx &c) {
    if (c.ops_sel.p) return;
    int16_t o[189], dn[189], qs[189];
    m2l_walk(0, o, dn, qs);
    std::vector<int16_t> h(NOPS_USED);
    for (int i = 0; i < 16; ++i) h[static_cast<size_t>(i)] = static_cast<int16_t>(i);
    for (int i = 0; i < 189; ++i) h[static_cast<size_t>(16 + i)] = static_cast<int16_t>(16 + o[i]);
    c.ops_sel.ensure(NOPS_USED);
    VG_HIP(hipMemcpy(c.ops_sel.p, h.data(), NOPS_USED * sizeof(int16_t), hipMemcpyHostToDevice));
}

void operators_build(Ctx &c) {
    fmm_constants_init();
    const TreeDev &t = c.t;
    const double wx = t.box[3] - t.box[0], wy = t.box[4] - t.box[1], wz = t.box[5] - t.box[2];
    const double ry = wx > 0 ? wy / wx : 1.0, rz = wx > 0 ? wz / wx : 1.0;
    // 8 M2M (child - parent), 8 L2L (parent - child) and 343 M2L (source -
    // target) offsets, formed on the device from the aspect ratio.  Only the
    // 189 M2L offsets of octant 0's walk are ever read (every octant uses them
    // through the mirror signs), so only those and the 16 M2M / L2L operators
    // are built; the other slots stay unwritten.
    const int nvec = 16 + 343;
    ops_selection(c);
    c.harm.ensure(static_cast<size_t>(nvec) * HSTRIDE);
    k_harmonics<<<NOPS_USED, 64, 0, c.stream>>>(c.ops_sel, nvec, ry, rz, 2 * c.p, c.harm);
    dim3 grid(NOPS_USED, ceil_div(c.K, OPS_ROWS));
    if (c.prec == 0) {
        c.op64.ensure(static_cast<size_t>(nvec) * c.K * c.K);
        k_ops<double><<<grid, 128 * OPS_ROWS, 0, c.stream>>>(c.ops_sel, c.harm, c.p, c.K, c.op64);
    } else if (c.prec == 2) {
        c.op64.ensure(static_cast<size_t>(nvec) * c.K * c.K);
        k_ops<double><<<grid, 128 * OPS_ROWS, 0, c.stream>>>(c.ops_sel, c.harm, c.p, c.K, c.op64);
        c.op32.ensure(static_cast<size_t>(nvec) * c.K * c.K);
        m2l_split_operators(c, nvec);
    } else {
        c.op32.ensure(static_cast<size_t>(nvec) * c.K * c.K);
        k_ops<float><<<grid, 128 * OPS_ROWS, 0, c.stream>>>(c.ops_sel, c.harm, c.p, c.K, c.op32);
    }
    VG_LAUNCH_CHECK();
}

void fmm_upward(Ctx &c) {
    if (c.prec == 0)
        upward_t<double>(c);
    else
        upward_t<float>(c);
}

void fmm_m2l_plain(Ctx &c) {
    if (c.prec == 0)
        m2l_plain_t<double>(c);
    else
        m2l_plain_t<float>(c);
}

void fmm_m2l_plain_list(Ctx &c, const int32_t *parents, int np) {
    if (c.prec == 0)
        m2l_plain_list_t<double>(c, parents, np);
    else
        m2l_plain_list_t<float>(c, parents, np);
}

void l2l_list(Ctx &c, const int32_t *parents, int64_t np) {
    if (np > 0) l2l_level<float>(c, parents, np);
    VG_LAUNCH_CHECK();
}

void l2p_list(Ctx &c, const int32_t *leaves, int64_t n, const uint8_t *skip) { l2p_t<float>(c, leaves, n, skip); }

void fmm_downward(Ctx &c, bool) {
    if (c.prec == 0)
        downward_t<double>(c);
    else
        downward_t<float>(c);
}

}  // namespace vg
