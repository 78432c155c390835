// gravity.hpp - device state of one gravity FMM context and the host-side
// entry points of each phase.  Cell numbering on the device is level-major:
// the cells of level d occupy ids [lvl_off[d], lvl_off[d+1]) and the eight
// children of the k-th internal cell of level d are the consecutive ids
// lvl_off[d+1] + 8k .. +7, in the reference's child-offset order
// (octree.hpp:164-176: offset bit 0/1/2 == 0 is the high x/y/z half).
#pragma once

#include "common.hpp"
#include "mlp.hpp"

namespace vg {

enum Phase { PH_TREE = 0, PH_LISTS, PH_UP, PH_M2L, PH_P2P, PH_DOWN, PH_RK4, PH_OPS, PH_COUNT };

struct TreeDev {
    // host-side shape of the last build
    int depth = 0;                 // deepest level holding cells
    int64_t ncells = 0;
    int64_t nleaves = 0;           // non-empty leaves
    std::vector<int64_t> lvl_off;  // depth + 2 entries
    std::vector<int64_t> int_off;  // internal cells of level d: int_ids[int_off[d] .. int_off[d+1])
    double box[6] = {0, 0, 0, 0, 0, 0};  // padded root box (min xyz, max xyz)

    // per cell
    DevBuf<uint32_t> start, count;   // particle range in sorted order
    DevBuf<int32_t> child, parent;   // first child (-1 for a leaf), parent (-1 for the root)
    DevBuf<uint8_t> octant;          // child offset within the parent (0..7)
    DevBuf<int8_t> level;
    DevBuf<double> bmin, bmax;       // 3 per cell: bounds exactly as the reference forms them
    DevBuf<int32_t> grid;            // 3 per cell: integer position at its level (low side = 0)
    DevBuf<int32_t> anc;             // (MAXD+1) per cell: ancestor id at every level <= own level
    DevBuf<int32_t> int_ids;         // internal cells, level by level
    DevBuf<int32_t> leaf_ids;        // non-empty leaves
    DevBuf<uint8_t> flags;           // scratch for compaction
    DevBuf<uint32_t> lv;             // level table of a device-driven build

    // interaction lists (reference order, device ids)
    DevBuf<int32_t> nb, ni, ap;              // 27, 27, 189 slots per cell
    DevBuf<int32_t> nb_cnt, ni_cnt, ap_cnt;  // per cell
    DevBuf<int32_t> nbr27;                   // 27 per cell: internal neighbour at offset (dx,dy,dz) or -1

    // reference numbering (built on demand for export)
    DevBuf<int32_t> ref_id;
    bool ref_valid = false;
};

// the expansion order and everything sized by it; the closure's fallback keeps
// a second set (order p_hi) and swaps it into the context while it runs
struct ExpSet {
    int p = 0, K = 0, ldk = 0;
    DevBuf<double> harm, op64;
    DevBuf<float> op32, M32, L32;
    DevBuf<uint16_t> opsplit, opsplit_ml, msplit;
};

// learned closure state (closure.hip)
struct Closure {
    bool on = false;
    double bound = 0;            // fallback where |MLP(x)| exceeds it
    int p_hi = 12;               // order of the fallback far field
    mlp::Net net;
    DevBuf<float> x, norm, y;    // per non-empty leaf: network input, |L~_25|, output
    DevBuf<float> l25;           // raw first 25 coefficients (export)
    DevBuf<int32_t> lev;         // leaf level (export)
    DevBuf<uint8_t> flag;        // per leaf: take the fallback
    DevBuf<uint32_t> scal;       // flagged-leaf counter
    int64_t nflag = 0;
    DevBuf<float> io;            // input shift (28), input scale (28), output scale (25)
    // fallback
    ExpSet hi;                   // order-p_hi expansions and operators
    DevBuf<uint8_t> need;        // per cell: on the ancestor chain of a flagged leaf
    DevBuf<int32_t> fb_leaves;   // flagged leaves
    DevBuf<int32_t> fb_int;      // needed internal cells, level by level
    DevBuf<uint32_t> fb_pos;     // compaction scratch
    std::vector<int64_t> fb_off; // fb_int offset of each level (depth + 2 entries)
};

struct Ctx {
    // configuration
    int64_t n = 0;
    int64_t s = 0;
    uint32_t thr = 0;  // floor(sqrt(8) s)
    int p = 4;
    int K = 25;        // (p+1)^2 real coefficients
    int ldk = 25;      // row stride of the expansion arrays (K, or K rounded up to 16 for the MFMA path)
    int prec = 1;
    double eps = 0, G = 1, mass = 0;
    uint32_t flags = 0;
    int device = 0;
    int ncu = 0;  // compute units of the device
    bool has_masses = false;
    bool uploaded = false;
    bool evaluated = false;
    std::string err;

    hipStream_t stream = nullptr;
    // split mode: the lists and P2P run on stream2 beside the upward pass and
    // M2L of the same evaluation (they share no buffer); M2L waits for the
    // lists (ev_lists), the scatter for P2P (join_p2p)
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_lists = nullptr;
    hipEvent_t ev_plan = nullptr;  // the P2P plan scalars are in the pinned words
    uint32_t p2p_chunk = 0;        // sources per P2P work item of the current plan
    bool p2p_pending = false;
    // pinned host words for the scalars read back during an evaluation (tree
    // level sizes, root box, P2P plan): a pinned copy returns sooner
    void *pinned = nullptr;
    static constexpr size_t PINNED_BYTES = 4096;
    bool timing = false;
    hipEvent_t ev[2 * PH_COUNT] = {};
    double phase_ms[PH_COUNT] = {};

    // FP64 state in the caller's particle order
    DevBuf<double> x, v, m, x0, v0, xo, vo;
    // force of the last evaluation in the caller's order (FP64 or FP32 by mode)
    DevBuf<double> acc64;
    DevBuf<float> acc32;
    DevBuf<double> pot64;  // potential (sum m/r without G) when requested

    // tree build scratch
    DevBuf<double> aabb_part;   // per-block partial boxes
    // partial boxes of the current x left by the RK4 stage that wrote it (0:
    // none); trusted only inside one step call, where nothing else writes x
    int aabb_parts = 0;
    DevBuf<double> box;         // 6 doubles + flag
    DevBuf<uint64_t> key[2];
    DevBuf<uint32_t> val[2];
    DevBuf<uint32_t> rs_hist;
    DevBuf<uint32_t> scal;      // small device scalars (counts, error flags)
    DevBuf<uint32_t> cmp_cnt;   // per-block counts of the multi-block compaction
    DevBuf<uint32_t> perm;      // sorted position -> caller index
    // particles in sorted order
    DevBuf<double> xs64;        // 4 per particle: x, y, z, m
    DevBuf<float> xs32;         // 4 per particle: x - root centre (FP32), m
    DevBuf<double> as64;        // 4 per particle: accel xyz, potential
    DevBuf<float> as32;

    TreeDev t;

    // expansions (scaled, real-packed (p+1)^2 per cell)
    DevBuf<double> M64, L64;
    DevBuf<float> M32, L32;
    // translation operators
    DevBuf<double> harm;        // harmonic tables of the operator offsets
    DevBuf<double> op64;        // M2M (8), L2L (8), M2L (343) blocks of K*K
    DevBuf<float> op32;
    DevBuf<uint16_t> opsplit;   // M2L operators as BF16 hi/lo MFMA fragments
    DevBuf<uint16_t> opsplit_ml;  // the 8 M2M and 8 L2L operators, the same way
    DevBuf<float> ml_part;      // split M2M: the 8 octant products of one level before their sum
    DevBuf<unsigned char> m2l_part;  // plain M2L cut into parts: the partial sums (bytes)
    DevBuf<uint16_t> msplit;    // multipoles as BF16 hi / lo halves (per cell ldk hi, then ldk lo)
    DevBuf<double> consts;      // A_n^m and recurrence constants (device copy)
    DevBuf<int16_t> ops_sel;    // the operator indices that are built (16 M2M / L2L, then octant 0's 189 M2L offsets)

    // ref-numbering scratch
    DevBuf<uint32_t> slist;
    DevBuf<uint64_t> rkey[2];
    DevBuf<uint32_t> rval[2];

    // P2P work items
    DevBuf<uint32_t> p2p_scal, p2p_nchunk, p2p_item_off, p2p_local, p2p_pre, p2p_btot;
    DevBuf<uint32_t> p2p_ticket;  // per leaf: work items of a split leaf that have finished
    DevBuf<int64_t> p2p_items;  // int2 (leaf index, chunk)
    DevBuf<float> p2p_part32;
    DevBuf<double> p2p_part64;

    double counts[8] = {};

    Closure cl;
};

// phases (each launches on ctx.stream)
void tree_build(Ctx &c);                 // AABB, codes, sort, cells, leaves, particle gather
void lists_build(Ctx &c);                // NB / NI / AP lists and the 27-neighbour table
void ref_numbering(Ctx &c);              // reference cell ids (for export only)
void operators_build(Ctx &c);            // M2M / L2L / M2L operators for the current root box
void fmm_upward(Ctx &c);                 // P2M + M2M
void fmm_m2l(Ctx &c);                    // M2L (plain VALU or split MFMA by mode)
void fmm_p2p(Ctx &c, bool potential);    // near field into the sorted accelerations (plan, then run)
void fmm_p2p_plan(Ctx &c);               // P2P work items; the item counts are read back without waiting
void fmm_p2p_run(Ctx &c, bool potential);  // waits for the item counts, then the P2P kernels
void fmm_downward(Ctx &c, bool potential);  // L2L + L2P
void join_p2p(Ctx &c);                       // ctx.stream waits for a P2P launched on ctx.stream2
void scatter_accel(Ctx &c, bool potential);  // sorted -> caller order
void rk4_init(Ctx &c);
void rk4_stage(Ctx &c, int stage, double dt);
void count_work(Ctx &c);
double kinetic_energy(Ctx &c);
double potential_energy(Ctx &c);
void closure_inputs(Ctx &c, bool with_l25);  // network inputs of every non-empty leaf from the current L32
void closure_apply(Ctx &c);                   // L~_25 += |L~_25| MLP(x) and fallback flags
bool closure_fallback_prepare(Ctx &c);        // order-p_hi locals of the flagged leaves; false if none is flagged
void closure_fallback_l2p(Ctx &c);            // order-p_hi L2P of the flagged leaves
void swap_expansions(Ctx &c, ExpSet &e);      // exchange the context's expansion set with e
void m2l_list(Ctx &c, const int32_t *parents, int np);      // M2L of the children of the given internal cells
void l2l_list(Ctx &c, const int32_t *parents, int64_t np);  // L2L into the children of the given internal cells (FP32)
void ml_split_level(Ctx &c, const int32_t *ids, int64_t n, bool m2m);  // M2M (m2m) or L2L of one level on the matrix cores
void l2p_list(Ctx &c, const int32_t *leaves, int64_t n, const uint8_t *skip);  // L2P of the given leaves (FP32)
void direct_subset(Ctx &c, const int64_t *idx, int64_t nidx, double *a, double *pot);  // FP64 O(N nidx) oracle forces
// the same with idx, a and pot in device memory, its partial sums in the
// library's static device scratch (no allocation); noalloc_init resolves that
// scratch's address on a device (once per device, with the device current)
void direct_subset_noalloc(Ctx &c, const int64_t *idx, int64_t nidx, double *a, double *pot);
void noalloc_init(int device);

void phase_begin(Ctx &c, Phase p);
void phase_end(Ctx &c, Phase p);

// one force evaluation at the current positions
void evaluate(Ctx &c, bool potential);

void radix_sort_pairs(hipStream_t st, uint64_t *&k, uint32_t *&v, uint64_t *&k_alt, uint32_t *&v_alt, int64_t n, int begin_bit,
                      int end_bit, DevBuf<uint32_t> &hist);

}  // namespace vg
