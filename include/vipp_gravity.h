/*
 * vipp_gravity.h - C ABI of the MI355X gravity FMM library (libvipp_gravity.so).
 *
 * This is the drop-in boundary for the reference's gravity hot path.  The
 * reference selects its solver through the alias gravity::Solver
 * (gravity/include/solver.hpp:7-8); a solver exposes preprocess / solve /
 * postprocess (gravity/include/solver_ref.hpp:11-27) over
 * std::vector<Vector<X>> positions and velocities, whose element type is three
 * contiguous coordinates (gravity/include/problem.hpp:143-146), so a
 * std::vector<Vector<double>> is exactly a double[3N].  The header-only C++
 * wrapper include/gravity/solver_gpu.hpp (gravity::GpuSolver) maps those three
 * calls onto the functions below.
 *
 * Conventions: plain pointers and sizes; every function returns VIPP_GRAV_OK
 * (0) or an error code, and vipp_grav_last_error() gives the message (errors
 * never throw across the ABI).  Host pointers are read or written before the
 * call returns; none is retained.  Calls on one context must not overlap.
 */
#ifndef VIPP_GRAVITY_H
#define VIPP_GRAVITY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VIPP_GRAV_OK 0
#define VIPP_GRAV_E_INVALID 1   /* bad argument or configuration */
#define VIPP_GRAV_E_INPUT 2     /* non-finite coordinates (the reference would recurse without bound) */
#define VIPP_GRAV_E_DEPTH 3     /* more than floor(sqrt(8) s) points closer than the 21-level resolution */
#define VIPP_GRAV_E_DEVICE 4    /* HIP runtime error */
#define VIPP_GRAV_E_STATE 5     /* call out of order (e.g. step before upload) */

/* Arithmetic of the force evaluation.  The RK4 state (x, v) is FP64 in every mode. */
#define VIPP_GRAV_F64 0         /* everything FP64: parity mode against the reference's double FMM */
#define VIPP_GRAV_F32 1         /* FP32 P2P, FP32 expansions and plain FP32 M2L */
#define VIPP_GRAV_F32_SPLIT 2   /* FP32 P2P; M2L on MFMA with BF16 hi/lo split operands, M2M and L2L with a
                                   three-way BF16 split, FP32 accumulation */

/* Flags */
#define VIPP_GRAV_FLAG_REF_FAR_G 1u  /* omit G from the far field exactly as the reference does (fmm.hpp:234,370-372) */

typedef struct vipp_grav_config {
    uint64_t n;          /* number of particles: Problem::N() (problem.hpp:214) */
    uint64_t s;          /* leaf parameter: Problem::s() (problem.hpp:232); split above floor(sqrt(8) s) points (octree.hpp:83) */
    uint32_t p;          /* expansion order: Problem::p() (problem.hpp:229); 1 <= p <= 12 */
    uint32_t precision;  /* VIPP_GRAV_F64 / VIPP_GRAV_F32 / VIPP_GRAV_F32_SPLIT */
    double epsilon;      /* Plummer softening: Problem::epsilon() (problem.hpp:217) */
    double G;            /* gravitational constant: Problem::get_G() (problem.hpp:220) */
    double mass;         /* per-particle mass when no mass array is uploaded: Problem::m() (problem.hpp:223) */
    int32_t device;      /* HIP device ordinal */
    uint32_t flags;      /* VIPP_GRAV_FLAG_* */
} vipp_grav_config;

typedef struct vipp_grav_ctx vipp_grav_ctx;

/* Context lifetime (SolverRef is stateless, solver_ref.hpp:8-9; the context
 * owns device memory between preprocess and postprocess). */
int vipp_grav_create(const vipp_grav_config *cfg, vipp_grav_ctx **out);
int vipp_grav_destroy(vipp_grav_ctx *ctx);
/* Message of the last error on ctx (or of the last failed create when ctx is NULL). */
const char *vipp_grav_last_error(const vipp_grav_ctx *ctx);

/* State transfer: SolverRef::preprocess / postprocess (solver_ref.hpp:11-14).
 * x, v: 3n doubles each (AoS x,y,z); masses: n doubles or NULL (scalar cfg.mass). */
int vipp_grav_upload(vipp_grav_ctx *ctx, const double *x, const double *v, const double *masses);
int vipp_grav_download(vipp_grav_ctx *ctx, double *x, double *v);

/* nsteps classic RK4 steps of size dt (SolverRef::solve -> runge_kutta,
 * solver_ref.hpp:17-27,36-104); each step rebuilds tree and lists at every
 * stage as SolverRef::acceleration does (solver_ref.hpp:164-177).  Returns after
 * the device work has finished.  If a stage fails (E_INPUT for a non-finite
 * stage position, E_DEPTH), x and v are put back to the state at the start of
 * the failed step (the steps before it are kept) and tree, lists and
 * expansions are rebuilt from that state by the next call that needs them. */
int vipp_grav_step(vipp_grav_ctx *ctx, double dt, uint32_t nsteps);

/* One force evaluation at the current positions (SolverRef::acceleration,
 * solver_ref.hpp:164-177): a = 3n doubles (may be NULL to keep it on device). */
int vipp_grav_accelerations(vipp_grav_ctx *ctx, double *a);

/* Total kinetic and potential energy in FP64 at the current state (no
 * counterpart in the reference; potential = FMM far field + softened near field). */
int vipp_grav_energy(vipp_grav_ctx *ctx, double *kinetic, double *potential);

/* Tree and lists of the last force evaluation (or built now if none), in the
 * reference's cell numbering (octree.hpp:166-191): cells as in Octree::cells(),
 * lists as calc_link builds them (solver_ref.hpp:115-162). */
int vipp_grav_tree_info(vipp_grav_ctx *ctx, int64_t *ncells, int64_t *nleaves, int32_t *depth);
int vipp_grav_export_tree(vipp_grav_ctx *ctx, int64_t *parent, int32_t *level, double *bmin, double *bmax,
                          int64_t *child_base, int64_t *pt_off, int64_t *pts);
int vipp_grav_export_lists_sizes(vipp_grav_ctx *ctx, int64_t *ap_total, int64_t *ni_total, int64_t *nb_total);
int vipp_grav_export_lists(vipp_grav_ctx *ctx, int64_t *ap_off, int64_t *ap, int64_t *ni_off, int64_t *ni,
                           int64_t *nb_off, int64_t *nb);
/* Multipole and local expansions of the last evaluation in the reference's
 * layout (Cell::_multipole/_local, fmm.hpp:563-566): per cell (p+1)(2p+1)
 * complex values (re, im) indexed [n][p+m], reference numbering, unscaled. */
int vipp_grav_export_expansions(vipp_grav_ctx *ctx, double *M, double *L);

/* Work counts of the last evaluation: [0] P2P pairs (i != j counted as the
 * reference sums them, self pairs excluded), [1] M2L cell pairs (every AP
 * entry of every cell at level >= 2, as the reference evaluates them),
 * [2] cells, [3] non-empty leaves, [4] depth, [5] M2L pairs whose source and
 * target both hold particles (the pairs that carry work), [6] leaves that took
 * the learned closure's fallback, [7] interaction-list entries stored (every
 * cell's AP and NB entries and its own NI entries, the inherited ones not
 * repeated). */
int vipp_grav_counts(vipp_grav_ctx *ctx, double *out, int32_t n);

/* Per-phase device milliseconds accumulated since the last reset, measured with
 * HIP events on the context's stream: [0] tree, [1] lists, [2] upward (P2M,
 * M2M), [3] M2L, [4] P2P, [5] downward (L2L, L2P), [6] RK4 updates,
 * [7] operators.  Pass NULL to reset.  Enable timing with vipp_grav_set_timing. */
int vipp_grav_phase_times(vipp_grav_ctx *ctx, double *ms, int32_t n);
int vipp_grav_set_timing(vipp_grav_ctx *ctx, int32_t enabled);

/* FP64 direct-sum accelerations of the particles idx[0..nidx) (caller
 * indices) against all N at the current positions: the O(N^2) force of
 * SolverRefNaive::acceleration (solver_ref_naive.hpp:103-119), with the
 * context's G, softening and masses.  a receives 3 nidx values; pot (may be
 * NULL) receives sum_{j != i} m_j / sqrt(r^2 + eps^2) without G.  Runs on the
 * device; meant for force-error samples of large N. */
int vipp_grav_direct(vipp_grav_ctx *ctx, const int64_t *idx, int64_t nidx, double *a, double *pot);

/* vipp_grav_direct with idx (nidx int64), a (3 nidx doubles) and pot (nidx
 * doubles, may be NULL) in device memory of the context's device, for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation); no counterpart in the reference.  It
 * allocates nothing: its partial sums live in a static device array of the
 * library (16 MiB per device, held until the process ends), so nidx is at
 * most 524288 (E_INVALID above).  Results agree with vipp_grav_direct to FP64
 * rounding (the source range may be cut into fewer slices).  Calls on one
 * device share that array: calls that overlap on one device race.  An index
 * outside [0, N) gives E_INVALID after the call (outputs unspecified).
 * Returns after the device work has finished. */
int vipp_grav_direct_noalloc(vipp_grav_ctx *ctx, const int64_t *idx, int64_t nidx, double *a, double *pot);

/* ---- Learned closure -------------------------------------------------------
 * A low-order FMM (p >= 4, FP32 expansions) whose leaf local expansions are
 * corrected before L2P by a small network: 28 inputs (the leaf's first 25
 * scaled local coefficients, orders <= 4, divided by their norm, then
 * level / 16, 2^-level and the distance of the leaf centre from the root
 * centre in root widths) -> 128 -> 128 (GELU) -> 25 outputs (the residual of
 * those 25 coefficients against an order-12 evaluation, in the same
 * normalisation).  FP16 weights and activations with FP32 accumulation in the
 * forward pass (MFMA), FP32 backward pass and Adam.  Parameters are one FP32
 * vector of VIPP_MLP_NPARAMS values, row-major [out][in]: W1 (128 x 28), b1,
 * W2 (128 x 128), b2, W3 (25 x 128), b3.  No counterpart in the reference. */
#define VIPP_MLP_IN 28
#define VIPP_MLP_HIDDEN 128
#define VIPP_MLP_OUT 25
#define VIPP_MLP_NPARAMS 23449

/* y (n x 25) = network(x (n x 28)) on the given device. */
int vipp_mlp_forward(int32_t device, const float *params, const float *x, int64_t n, float *y);
/* Mean squared error over the n x 25 outputs against t, each sample weighted
 * by w (NULL: all 1; the loss is sum_j w_j |y_j - t_j|^2 / (25 sum_j w_j)),
 * and its gradient with respect to the parameters (VIPP_MLP_NPARAMS values). */
int vipp_mlp_loss_grad(int32_t device, const float *params, const float *x, const float *t, const float *w, int64_t n, float *loss,
                       float *grad);
/* Adam (beta 0.9 / 0.999, eps 1e-8) on that loss, minibatches of `batch` in an
 * order shuffled per epoch from `seed`, learning rate decayed from lr to
 * 0.05 lr by a half cosine over the epochs.  params is updated in place;
 * epoch_loss (may be NULL) receives the weighted mean loss of each epoch. */
int vipp_mlp_train(int32_t device, float *params, const float *x, const float *t, const float *w, int64_t n, uint32_t epochs, uint32_t batch,
                   float lr, uint64_t seed, float *epoch_loss);

/* The closure's view of the last evaluation, one row per non-empty leaf in the
 * library's leaf order (the same for two contexts holding the same positions):
 * x (n x 28) network inputs, norm (n) the norm of the first 25 scaled local
 * coefficients, l25 (n x 25) those coefficients, level (n).  Leaves above
 * level 2 (no far field) have norm 0.  Any output pointer may be NULL; with all
 * NULL only *n is set.  Needs precision f32 or split and p >= 4. */
int vipp_grav_leaf_inputs(vipp_grav_ctx *ctx, int64_t *n, float *x, float *norm, float *l25, int32_t *level);
/* Enable the closure with the given parameters (NULL disables it).  io (81
 * values, or NULL for the identity) standardises the network's input and
 * scales its output: the network sees (x - io[0..27]) * io[28..55] and the
 * correction is |L~_25| * io[56..80] * network(x).  Leaves whose predicted
 * relative correction |io[56..80] * network(x)| exceeds `bound` take the
 * order-p_fallback far field instead (INFINITY: never); they are counted in
 * vipp_grav_counts()[6]. */
int vipp_grav_set_closure(vipp_grav_ctx *ctx, const float *params, const float *io, double bound, uint32_t p_fallback);

/* The reference's initial conditions, Problem::initialize
 * (gravity/include/problem.hpp:188-211) for the double type set: n particles
 * of a Plummer sphere from one mt19937_64 stream seeded with `seed`, positions
 * and velocities into x and v (3n doubles each), bit for bit the reference's.
 * epsilon < 0 selects the reference's default softening 0.63 n^-0.22
 * (problem.hpp:290-292), which also truncates the mass fraction drawn; the
 * value used goes to *epsilon_used (may be NULL).  Host code: no device is
 * touched.  E_INVALID for n = 0, a NULL array or epsilon >= 1. */
int vipp_grav_problem_initialize(uint64_t n, uint64_t seed, double epsilon, double *x, double *v, double *epsilon_used);

/* Device pointers of the FP64 state (3n each) for callers that keep data
 * resident on the GPU; valid until destroy. */
int vipp_grav_device_state(vipp_grav_ctx *ctx, void **x, void **v);

#ifdef __cplusplus
}
#endif
#endif
