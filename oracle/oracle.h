/*
 * oracle.h - C-ABI of the CPU oracle for the VIPP gravity FMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU and in
 * double precision, the algorithm of the reference's gravity module
 * (gravity/include/{octree,fmm,solver_ref,solver_ref_naive,problem,benchmark}.hpp).
 * It is the checker that the HIP path is compared against; it is never the
 * thing measured or shipped.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * Every function returns 0 on success and a nonzero code on error; the text of
 * the last error is available from orc_last_error().
 */
#ifndef VIPP_ORACLE_H
#define VIPP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_tree orc_tree;

const char *orc_last_error(void);

/* Octree by sequential first-fit insertion (octree.hpp:79-191, root box
 * fmm.hpp:403-405).  Split threshold floor(sqrt(8) * s) (octree.hpp:83). */
int orc_tree_build(const double *x, int64_t n, int64_t s, orc_tree **out);
void orc_tree_free(orc_tree *t);
int64_t orc_tree_ncells(const orc_tree *t);
/* Cells in reference numbering.  parent[0] = -1; child_base = -1 for leaves;
 * leaf points as CSR (pt_off has ncells+1 entries; internal cells are empty). */
void orc_tree_export(const orc_tree *t, int64_t *parent, int32_t *level, double *bmin, double *bmax,
                     int64_t *child_base, int64_t *pt_off, int64_t *pts);

/* Subdivide leaf idx with the given reinsertion threshold (test hook). */
int orc_tree_subdivide(orc_tree *t, int64_t idx, int64_t threshold);

/* Interaction lists of calc_link (solver_ref.hpp:115-162): returns total sizes. */
int orc_tree_lists(orc_tree *t, int64_t *ap_total, int64_t *ni_total, int64_t *nb_total);
void orc_tree_lists_export(const orc_tree *t, int64_t *ap_off, int64_t *ap, int64_t *ni_off, int64_t *ni,
                           int64_t *nb_off, int64_t *nb);

/* One FMM force evaluation on a built tree (solver_ref.hpp:164-232).
 *   masses: per-particle masses or NULL for the scalar m (problem.hpp:223).
 *   far_g:  0 reproduces the reference (G only in P2P, fmm.hpp:539); 1 applies G
 *           to the far field too.
 *   a:      3n accelerations (P2P first, then L2P, as the reference).
 *   a_near, a_far, pot (n), M, L: optional (NULL) diagnostics.  M and L are
 *   per cell (p+1)*(2p+1) complex numbers (re, im) laid out [cell][n][p+m].
 *   pot is the far + near potential sum_j m_j / r (softened near field),
 *   without G and with the sign of the reference's expansions. */
int orc_tree_fmm(orc_tree *t, const double *x, const double *masses, double m, double G, double eps,
                 uint32_t p, int far_g, double *a, double *a_near, double *a_far, double *pot, double *M,
                 double *L);

/* Upward pass only, for every cell at level >= min_level (test hook). */
int orc_tree_upward(orc_tree *t, const double *masses, double m, uint32_t p, int32_t min_level, double *M);

/* FP64 direct sum of solver_ref_naive.hpp:103-119 (pot optional). */
void orc_direct_accel(const double *x, int64_t n, const double *masses, double m, double G, double eps,
                      double *a, double *pot);
/* Direct sum for a subset of targets (idx[k]) against all sources. */
void orc_direct_accel_subset(const double *x, int64_t n, const double *masses, double m, double G,
                             double eps, const int64_t *idx, int64_t nidx, double *a, double *pot);

/* One RK4 step (solver_ref.hpp:36-104 / solver_ref_naive.hpp:33-100), in place. */
int orc_rk4_fmm(double *x, double *v, int64_t n, const double *masses, double m, double G, double eps,
                double dt, int64_t s, uint32_t p, int far_g);
int orc_rk4_direct(double *x, double *v, int64_t n, const double *masses, double m, double G, double eps,
                   double dt);

/* Special functions of fmm.hpp:20-221.  Tables are [n][m] with (nmax+1)^2 entries
 * (entries with m > n are zero); Y is [k][k_max + m] complex with (kmax+1)(2kmax+1). */
void orc_P(uint32_t nmax, double x, double *out);
void orc_DP(uint32_t nmax, double theta, double *out);
void orc_Y(uint32_t kmax, double theta, double phi, double *out);
double orc_A(uint32_t n, int32_t m);

/* Problem::initialize of problem.hpp:188-211 (mt19937_64 stream). eps<0 selects
 * the default calc_epsilon(N) (problem.hpp:290-292). Returns the epsilon used. */
double orc_problem_initialize(int64_t n, uint64_t seed, double eps, double *x, double *v);
double orc_problem_delta_t(double eps);

/* Curve fit E(T) = alpha T^c of benchmark.hpp:198-241. */
void orc_curve_fit(const double *e_t, const double *t, int64_t k, double *alpha, double *c);

#ifdef __cplusplus
}
#endif
#endif
