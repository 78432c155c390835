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

// One thread per offset vector: Y_l^m(direction) for l <= lmax and |v|.
__global__ void k_harmonics(const double *__restrict__ vec, int nvec, int lmax, double *__restrict__ out) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nvec) return;
    const double vx = vec[3 * o], vy = vec[3 * o + 1], vz = vec[3 * o + 2];
    double *tab = out + static_cast<size_t>(o) * HSTRIDE;
    const double rho = sqrt(vx * vx + vy * vy + vz * vz);
    tab[HSTRIDE - 2] = rho;
    tab[HSTRIDE - 1] = 0;
    const double rxy = sqrt(vx * vx + vy * vy);
    const double ct = rho > 0 ? vz / rho : 1.0, st = rho > 0 ? rxy / rho : 0.0;
    const dc eph = rxy > 0 ? dc{vx / rxy, vy / rxy} : dc{1.0, 0.0};
    dc em{1.0, 0.0};  // e^{i m phi}
    double pmm = 1.0;
    for (int m = 0; m <= lmax; ++m) {
        if (m > 0) {
            pmm *= -(2.0 * m - 1) * st;
            em = cmul(em, eph);
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
            }
            double f = 1.0;
            for (int i = l - m + 1; i <= l + m; ++i) f /= i;
            const double nrm = sqrt(f) * pl;
            tab[2 * (l * (l + 1) / 2 + m)] = nrm * em.re;
            tab[2 * (l * (l + 1) / 2 + m) + 1] = nrm * em.im;
        }
    }
}

// Fill one real operator row from complex coefficients T(+m) and T(-m).
// row part 0 = Re, 1 = Im; column (n, m) -> Re at kidx_re, Im at kidx_im.
struct RowSink {
    double *row;
    int part;
    __device__ void put(int n, int m, dc tp, dc tm) {
        if (m == 0) {
            row[kidx_re(n, 0)] += part == 0 ? tp.re : tp.im;
        } else {
            const dc s = cadd(tp, tm), d = csub(tp, tm);
            row[kidx_re(n, m)] += part == 0 ? s.re : s.im;
            row[kidx_im(n, m)] += part == 0 ? -d.im : d.re;
        }
    }
};

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

// op block layout: [M2M 0..7][L2L 8..15][M2L 16..16+342], each K*K row-major [row][col].
template <class R>
__global__ __launch_bounds__(64) void k_ops(const double *__restrict__ harm, int p, int K, R *__restrict__ op) {
    const int blk = blockIdx.x;  // operator index
    const int r = blockIdx.y * 64 + threadIdx.x;
    if (r >= K) return;
    int j, k, part;
    row_of(r, j, k, part);
    double row[MAXK];
    for (int c = 0; c < K; ++c) row[c] = 0;
    RowSink sink{row, part};
    const double *tab = harm + static_cast<size_t>(blk) * HSTRIDE;
    const double rho = tab[HSTRIDE - 2];
    if (blk < 8) {
        // M2M, child offset q: column (j', k') of the child expansion, scaled by 2^{-j'} (fmm.hpp:256-270)
        for (int jp = 0; jp <= j; ++jp) {
            const int n = j - jp;
            const double rn = pow(rho, n);
            const double scale = ldexp(1.0, -jp);
            for (int kp = 0; kp <= jp; ++kp) {
                dc tpm[2] = {{0, 0}, {0, 0}};
                for (int sgn = 0; sgn < 2; ++sgn) {
                    const int kk = sgn == 0 ? kp : -kp;
                    const int m = k - kk;
                    if (m < -n || m > n) continue;
                    const double coef = ipow_even(abs(k) - abs(m) - abs(k - m)) * Acoef(n, m) * Acoef(jp, kk) * rn / Acoef(j, k) * scale;
                    tpm[sgn] = cscale(Yget(tab, n, -m), coef);
                }
                sink.put(jp, kp, tpm[0], tpm[1]);
            }
        }
    } else if (blk < 16) {
        // L2L, child offset q: column (n, m) of the parent expansion, scaled by 2^{-(j+1)} (fmm.hpp:313-327)
        const double scale = ldexp(1.0, -(j + 1));
        for (int n = j; n <= p; ++n) {
            const double rn = pow(rho, n - j);
            for (int mp = 0; mp <= n; ++mp) {
                dc tpm[2] = {{0, 0}, {0, 0}};
                for (int sgn = 0; sgn < 2; ++sgn) {
                    const int m = sgn == 0 ? mp : -mp;
                    if (m - k < -(n - j) || m - k > n - j) continue;
                    const double coef = ipow_even(abs(m) - abs(m - k) - abs(k)) * Acoef(n - j, m - k) * Acoef(j, k) * rn /
                                        (((n + j) % 2 ? -1.0 : 1.0) * Acoef(n, m)) * scale;
                    tpm[sgn] = cscale(Yget(tab, n - j, m - k), coef);
                }
                sink.put(n, mp, tpm[0], tpm[1]);
            }
        }
    } else {
        // M2L, offset o (fmm.hpp:287-302); rows of the 27 near offsets stay zero
        const int o = blk - 16;
        const int ox = o % 7 - 3, oy = (o / 7) % 7 - 3, oz = o / 49 - 3;
        const bool near = ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1 && oz >= -1 && oz <= 1;
        if (!near) {
            for (int n = 0; n <= p; ++n) {
                for (int mp = 0; mp <= n; ++mp) {
                    dc tpm[2] = {{0, 0}, {0, 0}};
                    for (int sgn = 0; sgn < 2; ++sgn) {
                        if (mp == 0 && sgn == 1) continue;
                        const int m = sgn == 0 ? mp : -mp;
                        const double coef = ipow_even(abs(k - m) - abs(k) - abs(m)) * Acoef(n, m) * Acoef(j, k) /
                                            ((n % 2 ? -1.0 : 1.0) * Acoef(j + n, m - k) * pow(rho, j + n + 1));
                        tpm[sgn] = cscale(Yget(tab, j + n, m - k), coef);
                    }
                    sink.put(n, mp, tpm[0], tpm[1]);
                }
            }
        }
    }
    R *dst = op + (static_cast<size_t>(blk) * K + r) * K;
    for (int c = 0; c < K; ++c) dst[c] = static_cast<R>(row[c]);
}

// ---------------------------------------------------------------------------
// P2M: one wavefront per leaf; lane i evaluates R_n^m at its particle, the
// wavefront then sums the 64 rows column by column.
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(64) void k_p2m(const int32_t *__restrict__ leaves, const int8_t *__restrict__ level, const uint32_t *__restrict__ start,
                                            const uint32_t *__restrict__ count, const double *__restrict__ bmin, const double *__restrict__ bmax,
                                            const double *__restrict__ xs64, double wx0, int p, int K, int ldk, R *__restrict__ M) {
    constexpr int ROWS = 32;
    __shared__ R rows[ROWS * MAXK];
    const int32_t c = leaves[blockIdx.x];
    if (level[c] < 2) return;  // solver_ref.hpp:187-188
    const uint32_t s = start[c], cnt = count[c];
    const double iw = 1.0 / ldexp(wx0, -level[c]);
    const double cx = (bmin[3 * c] + bmax[3 * c]) / 2, cy = (bmin[3 * c + 1] + bmax[3 * c + 1]) / 2, cz = (bmin[3 * c + 2] + bmax[3 * c + 2]) / 2;
    R acc[3] = {0, 0, 0};
    const int lane = threadIdx.x;
    for (uint32_t base = 0; base < cnt; base += ROWS) {
        if (lane < ROWS) {
            R *row = rows + lane * K;
            if (base + lane < cnt) {
                const uint32_t k = s + base + lane;
                const R x = static_cast<R>((xs64[4 * k] - cx) * iw), y = static_cast<R>((xs64[4 * k + 1] - cy) * iw),
                        z = static_cast<R>((xs64[4 * k + 2] - cz) * iw);
                const R mi = static_cast<R>(xs64[4 * k + 3]);
                const R r2 = x * x + y * y + z * z;
                R dre = 1, dim = 0;  // R_m^m
                for (int m = 0; m <= p; ++m) {
                    if (m > 0) {
                        const R d = static_cast<R>(cDiag[m]);
                        const R nre = d * (x * dre - y * dim), nim = d * (x * dim + y * dre);
                        dre = nre;
                        dim = nim;
                    }
                    R are = 0, aim = 0, bre = dre, bim = dim;  // R_{n-2}^m, R_{n-1}^m
                    for (int n = m; n <= p; ++n) {
                        if (n > m) {
                            const R c1 = static_cast<R>(cR1[n * (MAXP + 1) + m]), c2 = static_cast<R>(cR2[n * (MAXP + 1) + m]);
                            const R nre = c1 * z * bre - c2 * r2 * are, nim = c1 * z * bim - c2 * r2 * aim;
                            are = bre;
                            aim = bim;
                            bre = nre;
                            bim = nim;
                        }
                        // M_n^m += m_i conj(R_n^m)
                        row[kidx_re(n, m)] = mi * bre;
                        if (m > 0) row[kidx_im(n, m)] = -mi * bim;
                    }
                }
            } else {
                for (int q = 0; q < K; ++q) row[q] = 0;
            }
        }
        __syncthreads();
        for (int q = lane, slot = 0; q < K; q += 64, ++slot) {
            R sum = 0;
            for (int i = 0; i < ROWS; ++i) sum += rows[i * K + q];
            acc[slot] += sum;
        }
        __syncthreads();
    }
    for (int q = lane, slot = 0; q < K; q += 64, ++slot) M[static_cast<size_t>(c) * ldk + q] = acc[slot];
    for (int q = K + lane; q < ldk; q += 64) M[static_cast<size_t>(c) * ldk + q] = R(0);
}

// M2M for the internal cells of one level (solver_ref.hpp:196-197).
template <class R>
__global__ __launch_bounds__(256) void k_m2m(const int32_t *__restrict__ ids, const int32_t *__restrict__ child, const uint32_t *__restrict__ count,
                                             const R *__restrict__ op, int K, int ldk, R *__restrict__ M) {
    __shared__ R cm[8 * MAXK];
    const int32_t P = ids[blockIdx.x];
    const int32_t base = child[P];
    for (int i = threadIdx.x; i < 8 * K; i += 256) {
        const int q = i / K, col = i % K;
        cm[i] = count[base + q] ? M[static_cast<size_t>(base + q) * ldk + col] : R(0);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < K; r += 256) {
        int j = 0;
        while ((j + 1) * (j + 1) <= r) ++j;
        const int cend = (j + 1) * (j + 1);  // output degree j uses child degrees <= j
        R sum = 0;
        for (int q = 0; q < 8; ++q) {
            if (!count[base + q]) continue;
            const R *orow = op + (static_cast<size_t>(q) * K + r) * K;
            const R *mv = cm + q * K;
            for (int col = 0; col < cend; ++col) sum += orow[col] * mv[col];
        }
        M[static_cast<size_t>(P) * ldk + r] = sum;
    }
    for (int q = K + threadIdx.x; q < ldk; q += 256) M[static_cast<size_t>(P) * ldk + q] = R(0);
}

// L2L from the internal cells of one level to their children (solver_ref.hpp:216-217).
template <class R>
__global__ __launch_bounds__(256) void k_l2l(const int32_t *__restrict__ ids, const int32_t *__restrict__ child, const uint32_t *__restrict__ count,
                                             const R *__restrict__ op, int K, int ldk, R *__restrict__ L) {
    __shared__ R pl[MAXK];
    const int32_t P = ids[blockIdx.x];
    const int32_t base = child[P];
    for (int i = threadIdx.x; i < K; i += 256) pl[i] = L[static_cast<size_t>(P) * ldk + i];
    __syncthreads();
    for (int i = threadIdx.x; i < 8 * K; i += 256) {
        const int q = i / K, r = i % K;
        if (!count[base + q]) continue;
        const R *orow = op + (static_cast<size_t>(8 + q) * K + r) * K;
        int j = 0;
        while ((j + 1) * (j + 1) <= r) ++j;
        R sum = 0;
        for (int col = j * j; col < K; ++col) sum += orow[col] * pl[col];  // parent degrees >= j
        L[static_cast<size_t>(base + q) * ldk + r] += sum;
    }
}

// ---------------------------------------------------------------------------
// Plain M2L: targets are the children of octant qt of a block of internal
// parents; for each of the 27 parent-neighbour offsets and 8 source octants
// the (static) cell offset selects one operator, applied as a K x K by K x TB
// product with the gathered source multipoles (offset-grouped GEMM on VALU).
// ---------------------------------------------------------------------------
template <class R, int TB>
__global__ __launch_bounds__(256) void k_m2l_plain(const int32_t *__restrict__ parents, int nparents, const int32_t *__restrict__ child,
                                                   const uint32_t *__restrict__ count, const int32_t *__restrict__ nbr27, const R *__restrict__ Mx,
                                                   R *__restrict__ Lx, const R *__restrict__ op, int K, int ldk) {
    constexpr int KC = 16;
    constexpr int CPT = TB / 8;  // targets per thread
    constexpr int SMS = TB + 1, SOS = 193;  // padded strides: no bank conflicts
    __shared__ R sM[MAXK * SMS];
    __shared__ R sOp[KC * SOS];
    __shared__ int32_t par[TB], tgt[TB], src[TB];
    __shared__ int any;
    const int qt = blockIdx.y;
    const int tr = threadIdx.x & 31, tc = threadIdx.x >> 5;
    if (threadIdx.x < TB) {
        const int pi = blockIdx.x * TB + threadIdx.x;
        int32_t P = pi < nparents ? parents[pi] : -1;
        int32_t t = P >= 0 ? child[P] + qt : -1;
        if (t >= 0 && count[t] == 0) t = -1;
        par[threadIdx.x] = t >= 0 ? P : -1;
        tgt[threadIdx.x] = t;
    }
    R acc[6][CPT];
    for (int i = 0; i < 6; ++i)
        for (int jj = 0; jj < CPT; ++jj) acc[i][jj] = 0;
    const int tx = 1 - (qt & 1), ty = 1 - ((qt >> 1) & 1), tz = 1 - ((qt >> 2) & 1);
    for (int dn = 0; dn < 27; ++dn) {
        const int dx = dn % 3 - 1, dy = (dn / 3) % 3 - 1, dz = dn / 9 - 1;
        for (int qs = 0; qs < 8; ++qs) {
            const int ex = 2 * dx + (1 - (qs & 1)) - tx, ey = 2 * dy + (1 - ((qs >> 1) & 1)) - ty, ez = 2 * dz + (1 - ((qs >> 2) & 1)) - tz;
            if (ex >= -1 && ex <= 1 && ey >= -1 && ey <= 1 && ez >= -1 && ez <= 1) continue;  // adjacent: P2P
            const int o = (ex + 3) + 7 * (ey + 3) + 49 * (ez + 3);
            __syncthreads();
            if (threadIdx.x == 0) any = 0;
            __syncthreads();
            if (threadIdx.x < TB) {
                const int32_t P = par[threadIdx.x];
                int32_t s = -1;
                if (P >= 0) {
                    const int32_t sp = nbr27[27 * static_cast<int64_t>(P) + dn];
                    if (sp >= 0) {
                        s = child[sp] + qs;
                        if (count[s] == 0) s = -1;
                    }
                }
                src[threadIdx.x] = s;
                if (s >= 0) any = 1;
            }
            __syncthreads();
            if (!any) continue;
            for (int i = threadIdx.x; i < K * TB; i += 256) {
                const int t = i / K, col = i % K;
                const int32_t s = src[t];
                sM[col * SMS + t] = s >= 0 ? Mx[static_cast<size_t>(s) * ldk + col] : R(0);
            }
            const R *O = op + static_cast<size_t>(16 + o) * K * K;
            for (int c0 = 0; c0 < K; c0 += KC) {
                __syncthreads();
                for (int i = threadIdx.x; i < KC * K; i += 256) {
                    const int r = i / KC, cc = i % KC;
                    sOp[cc * SOS + r] = (c0 + cc < K) ? O[static_cast<size_t>(r) * K + c0 + cc] : R(0);
                }
                __syncthreads();
                const int ce = min(KC, K - c0);
                for (int cc = 0; cc < ce; ++cc) {
                    R mv[CPT];
                    for (int jj = 0; jj < CPT; ++jj) mv[jj] = sM[(c0 + cc) * SMS + tc + 8 * jj];
#pragma unroll
                    for (int i = 0; i < 6; ++i) {
                        const R ov = sOp[cc * SOS + tr + 32 * i];
#pragma unroll
                        for (int jj = 0; jj < CPT; ++jj) acc[i][jj] += ov * mv[jj];
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int jj = 0; jj < CPT; ++jj) {
        const int32_t t = tgt[tc + 8 * jj];
        if (t < 0) continue;
        for (int i = 0; i < 6; ++i) {
            const int r = tr + 32 * i;
            if (r < K) Lx[static_cast<size_t>(t) * ldk + r] = acc[i][jj];
        }
    }
}

// ---------------------------------------------------------------------------
// L2P: one thread per particle of a leaf; the R_n^m recurrence feeds the
// potential, d/dz and (d/dx - i d/dy) directly, so no table is stored.
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(64) void k_l2p(const int32_t *__restrict__ leaves, const int8_t *__restrict__ level, const uint32_t *__restrict__ start,
                                            const uint32_t *__restrict__ count, const double *__restrict__ bmin, const double *__restrict__ bmax,
                                            const double *__restrict__ xs64, double wx0, int p, int K, int ldk, const R *__restrict__ Lx,
                                            double gscale, R *__restrict__ acc4, const uint8_t *__restrict__ skip) {
    __shared__ R sl[MAXK];
    if (skip && skip[blockIdx.x]) return;  // the leaf takes the closure's fallback instead
    const int32_t c = leaves[blockIdx.x];
    if (level[c] < 2) return;  // solver_ref.hpp:220-221
    for (int i = threadIdx.x; i < K; i += 64) sl[i] = Lx[static_cast<size_t>(c) * ldk + i];
    __syncthreads();
    const uint32_t s = start[c], cnt = count[c];
    const double w = ldexp(wx0, -level[c]);
    const double iw = 1.0 / w;
    const double cx = (bmin[3 * c] + bmax[3 * c]) / 2, cy = (bmin[3 * c + 1] + bmax[3 * c + 1]) / 2, cz = (bmin[3 * c + 2] + bmax[3 * c + 2]) / 2;
    for (uint32_t i = threadIdx.x; i < cnt; i += 64) {
        const uint32_t k = s + i;
        const R x = static_cast<R>((xs64[4 * k] - cx) * iw), y = static_cast<R>((xs64[4 * k + 1] - cy) * iw), z = static_cast<R>((xs64[4 * k + 2] - cz) * iw);
        const R r2 = x * x + y * y + z * z;
        R phi = 0, gz = 0, gre = 0, gim = 0;  // potential, d/dz, (d/dx - i d/dy)
        R dre = 1, dim = 0;
        for (int m = 0; m <= p; ++m) {
            if (m > 0) {
                const R d = static_cast<R>(cDiag[m]);
                const R nre = d * (x * dre - y * dim), nim = d * (x * dim + y * dre);
                dre = nre;
                dim = nim;
            }
            R are = 0, aim = 0, cre = dre, cim = dim;  // R_{n-2}^m, R_n^m
            for (int n = m; n <= p; ++n) {
                if (n > m) {
                    const R c1 = static_cast<R>(cR1[n * (MAXP + 1) + m]), c2 = static_cast<R>(cR2[n * (MAXP + 1) + m]);
                    const R nre = c1 * z * cre - c2 * r2 * are, nim = c1 * z * cim - c2 * r2 * aim;
                    are = cre;
                    aim = cim;
                    cre = nre;
                    cim = nim;
                }
                // potential: sum L_n^0 R_n^0 + 2 Re(L_n^m R_n^m)
                if (m == 0)
                    phi += sl[kidx_re(n, 0)] * cre;
                else
                    phi += 2 * (sl[kidx_re(n, m)] * cre - sl[kidx_im(n, m)] * cim);
                if (n + 1 <= p) {
                    const int l = n + 1;  // R_n^m enters the derivatives of degree-l terms
                    const R lz = static_cast<R>(cLz[l * (MAXP + 2) + m]);
                    if (m == 0)
                        gz += sl[kidx_re(l, 0)] * lz * cre;
                    else
                        gz += 2 * lz * (sl[kidx_re(l, m)] * cre - sl[kidx_im(l, m)] * cim);
                    {  // -sqrt((l+m+1)(l+m)) L_l^{m+1} R_n^m
                        const R la = static_cast<R>(cLa[l * (MAXP + 2) + m + 1]);
                        const R Lre = sl[kidx_re(l, m + 1)], Lim = sl[kidx_im(l, m + 1)];
                        gre -= la * (Lre * cre - Lim * cim);
                        gim -= la * (Lre * cim + Lim * cre);
                    }
                    if (m >= 1) {  // sqrt((l-m+1)(l-m)) conj(L_l^{m-1} R_n^m), with Re(L_l^0) for m = 1
                        const R lb = static_cast<R>(cLb[l * (MAXP + 2) + m - 1]);
                        const R Lre = sl[kidx_re(l, m - 1)], Lim = (m - 1) == 0 ? R(0) : sl[kidx_im(l, m - 1)];
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

template <class R>
void upward_t(Ctx &c) {
    TreeDev &t = c.t;
    Bufs<R>::ensure(c, t.ncells);
    const double wx0 = t.box[3] - t.box[0];
    if (t.nleaves)
        k_p2m<R><<<static_cast<unsigned>(t.nleaves), 64, 0, c.stream>>>(t.leaf_ids, t.level, t.start, t.count, t.bmin, t.bmax, c.xs64, wx0, c.p,
                                                                       c.K, c.ldk, Bufs<R>::M(c));
    for (int d = t.depth; d >= 2; --d) {
        const int64_t a = t.int_off[d], b = t.int_off[d + 1];
        if (b > a) k_m2m<R><<<static_cast<unsigned>(b - a), 256, 0, c.stream>>>(t.int_ids.p + a, t.child, t.count, Bufs<R>::op(c), c.K, c.ldk, Bufs<R>::M(c));
    }
    VG_LAUNCH_CHECK();
}

template <class R>
void m2l_plain_list_t(Ctx &c, const int32_t *parents, int np) {
    const TreeDev &t = c.t;
    if (np <= 0) return;
    constexpr int TB = sizeof(R) == 4 ? 32 : 16;
    dim3 grid(ceil_div(np, TB), 8);
    k_m2l_plain<R, TB><<<grid, 256, 0, c.stream>>>(parents, np, t.child, t.count, t.nbr27, Bufs<R>::M(c), Bufs<R>::L(c), Bufs<R>::op(c), c.K, c.ldk);
    VG_LAUNCH_CHECK();
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
    if (n)
        k_l2p<R><<<static_cast<unsigned>(n), 64, 0, c.stream>>>(leaves, t.level, t.start, t.count, t.bmin, t.bmax, c.xs64, wx0, c.p, c.K, c.ldk,
                                                                Bufs<R>::L(c), gscale, Bufs<R>::acc(c), skip);
    VG_LAUNCH_CHECK();
}

template <class R>
void downward_t(Ctx &c) {
    TreeDev &t = c.t;
    for (int d = 2; d < t.depth; ++d) {  // parents at level d >= 2 feed children at level >= 3
        const int64_t a = t.int_off[d], b = t.int_off[d + 1];
        if (b > a) k_l2l<R><<<static_cast<unsigned>(b - a), 256, 0, c.stream>>>(t.int_ids.p + a, t.child, t.count, Bufs<R>::op(c), c.K, c.ldk, Bufs<R>::L(c));
    }
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

void fmm_constants_init() {
    static bool done = false;
    if (done) return;
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
    done = true;
}

void operators_build(Ctx &c) {
    fmm_constants_init();
    const TreeDev &t = c.t;
    const double wx = t.box[3] - t.box[0], wy = t.box[4] - t.box[1], wz = t.box[5] - t.box[2];
    const double ry = wx > 0 ? wy / wx : 1.0, rz = wx > 0 ? wz / wx : 1.0;
    // offset vectors in units of the x width of the level: 8 M2M (child - parent), 8 L2L (parent - child), 343 M2L (source - target)
    std::vector<double> vec(3 * (16 + 343));
    for (int q = 0; q < 8; ++q) {
        const double vx = (q & 1) ? -0.25 : 0.25, vy = ((q >> 1) & 1) ? -0.25 * ry : 0.25 * ry, vz = ((q >> 2) & 1) ? -0.25 * rz : 0.25 * rz;
        vec[3 * q] = vx;
        vec[3 * q + 1] = vy;
        vec[3 * q + 2] = vz;
        vec[3 * (8 + q)] = -vx;
        vec[3 * (8 + q) + 1] = -vy;
        vec[3 * (8 + q) + 2] = -vz;
    }
    for (int o = 0; o < 343; ++o) {
        vec[3 * (16 + o)] = o % 7 - 3;
        vec[3 * (16 + o) + 1] = ((o / 7) % 7 - 3) * ry;
        vec[3 * (16 + o) + 2] = (o / 49 - 3) * rz;
    }
    const int nvec = 16 + 343;
    DevBuf<double> dvec;
    dvec.ensure(vec.size());
    VG_HIP(hipMemcpyAsync(dvec.p, vec.data(), vec.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    c.harm.ensure(static_cast<size_t>(nvec) * HSTRIDE);
    k_harmonics<<<ceil_div(nvec, 64), 64, 0, c.stream>>>(dvec, nvec, 2 * c.p, c.harm);
    dim3 grid(nvec, ceil_div(c.K, 64));
    if (c.prec == 0) {
        c.op64.ensure(static_cast<size_t>(nvec) * c.K * c.K);
        k_ops<double><<<grid, 64, 0, c.stream>>>(c.harm, c.p, c.K, c.op64);
    } else if (c.prec == 2) {
        c.op64.ensure(static_cast<size_t>(nvec) * c.K * c.K);
        k_ops<double><<<grid, 64, 0, c.stream>>>(c.harm, c.p, c.K, c.op64);
        c.op32.ensure(static_cast<size_t>(nvec) * c.K * c.K);
        m2l_split_operators(c, nvec);
    } else {
        c.op32.ensure(static_cast<size_t>(nvec) * c.K * c.K);
        k_ops<float><<<grid, 64, 0, c.stream>>>(c.harm, c.p, c.K, c.op32);
    }
    VG_LAUNCH_CHECK();
    VG_HIP(hipStreamSynchronize(c.stream));  // dvec is released on return
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
    if (np > 0) k_l2l<float><<<static_cast<unsigned>(np), 256, 0, c.stream>>>(parents, c.t.child, c.t.count, c.op32, c.K, c.ldk, c.L32);
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
