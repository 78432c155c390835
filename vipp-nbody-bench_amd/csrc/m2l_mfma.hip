// m2l_mfma.hip - M2L, and in split mode M2M and L2L, on the CDNA4 matrix
// cores with split BF16 operands.
//
// The M2L of one target is L~(t) = sum_o T_o M~(s(t, o)) over its (at most
// 189) source offsets o (fmm.hpp:277-304 in the scaled real basis of
// fmm.hip).  A block owns 32 NT targets of one child octant qt, so that every
// target sees the same 189 offsets in the same order (offsets follow from
// (parent-neighbour offset, source octant, qt) alone).  Per offset the block
// forms T_o (K x K) times the targets' source multipoles (K x 32 NT) on
// v_mfma_f32_32x32x16_bf16, two wavefronts per 32 output rows, accumulating
// in FP32 across all offsets.  Every block walks the offsets in the same
// order, so the operators stream through L2 shared by the blocks of an XCD.
//
// Precision: every M2L operand x is split as hi = bf16(x), lo = bf16(x - hi)
// (16 significant bits together) and a product is formed as
// hi*hi + hi*lo + lo*hi with FP32 accumulation; the dropped lo*lo term and the
// split residuals are ~2^-18 of the product.  T_o is split from its FP64
// value; the multipoles are split once per evaluation (k_split_m).  M2M and
// L2L use a three-way split (below).
//
// Operand layouts (checked on the device): for v_mfma_f32_32x32x16_bf16 lane l
// holds A[l%32][8(l/32)+e] and B[8(l/32)+e][l%32], e = 0..7, and
// C[8(r/4) + 4(l/32) + r%4][l%32], r = 0..15.  T_o is stored pre-swizzled in
// that A order, so a wavefront loads a fragment with one 16-byte read per lane.
#include "gravity.hpp"

namespace vg {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));


__host__ __device__ constexpr int mt_of(int K) { return (K + 31) / 32; }
// M2L row tiles: the odd counts 3 (p = 8) and 5 (p = 11) are padded to 4
// and 6, so the kernel's 2 MT wavefronts spread evenly over the 4 SIMDs of a
// CU (the padded rows are zero; p = 8: M2L 1.47 -> 1.38 ms and 29.4 -> 31.2
// steps/s on the 4M Plummer input, p = 11: 3.84 -> 2.56 ms)
__host__ __device__ constexpr int mt_m2l(int K) { return (mt_of(K) == 3 || mt_of(K) == 5) ? mt_of(K) + 1 : mt_of(K); }
__host__ __device__ constexpr int ks_of(int K) { return (K + 15) / 16; }

// FP64 operators (row-major K x K per offset) -> BF16 hi/lo A fragments of
// the M2L blocks: frag[o][mt][h][ks][lane][e], so that any run of consecutive
// m-tiles (hi and lo) is one contiguous range.  One block per (offset, m-tile):
// the tile's 32 rows are read row by row into LDS (coalesced), then each thread
// writes whole 8-coefficient fragment vectors.
__global__ __launch_bounds__(256) void k_split_ops(const int16_t *__restrict__ sel, const double *__restrict__ op64, int K, __bf16 *__restrict__ frag,
                                                   int MT, int NKS) {
    extern __shared__ double tile[];  // 32 rows x (16 NKS + 1)
    const int o = sel[16 + blockIdx.x / MT] - 16, mt = blockIdx.x % MT;  // the M2L offsets of the walk only
    const int kc = 16 * NKS, W = kc + 1;
    const double *src = op64 + static_cast<int64_t>(16 + o) * K * K;
    for (int i = threadIdx.x; i < 32 * kc; i += 256) {
        const int r = i / kc, col = i % kc, row = mt * 32 + r;
        tile[r * W + col] = (row < K && col < K) ? src[static_cast<int64_t>(row) * K + col] : 0.0;
    }
    __syncthreads();
    const int64_t base = (static_cast<int64_t>(o) * MT + mt) * 2 * NKS * 64;  // in 8-coefficient vectors
    for (int v = threadIdx.x; v < 2 * NKS * 64; v += 256) {
        const int h = v / (NKS * 64), ks = (v / 64) % NKS, l = v % 64;
        const int r = l & 31, c0 = ks * 16 + (l >> 5) * 8;
        bf16x8 out;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const double x = tile[r * W + c0 + e];
            const __bf16 hv = static_cast<__bf16>(static_cast<float>(x));
            out[e] = h ? static_cast<__bf16>(static_cast<float>(x - static_cast<double>(static_cast<float>(hv)))) : hv;
        }
        reinterpret_cast<bf16x8 *>(frag)[base + v] = out;
    }
}

// One block: 32*NT targets (children of octant qt of 32*NT parents) and NG
// wavefronts per 32-row m-tile of the output.  Without KS each of them owns
// NT/NG n-tiles of 32 targets over all k-steps; with KS (NG = 2) both own all
// NT n-tiles, one the even and one the odd k-steps, and their sums are added
// at the end, so every A fragment is loaded by one wavefront only.  Per offset
// step:
//   A (the wavefront's m-tile of T_o, hi and lo) streams from L2 straight
//     into registers, a few k-steps ahead;
//   B (the targets' source multipoles) sits in LDS as hi/lo MFMA fragments.
// The k-steps of an offset are cut into two halves, and B is double-buffered
// by half: while a half is multiplied from one buffer, the next half is
// gathered into registers and stored into the other, so the only pause is one
// barrier per half.  The B fragment of lane l is stored in slot
// l ^ (k8 & 15), k8 = 2 ks + l / 32 (ks counted within the half), so that
// both the row-wise stores and the fragment loads spread over the LDS banks.
template <int P, int NT, int NG, bool KS>
__global__ __launch_bounds__(64 * NG * mt_m2l((P + 1) * (P + 1))) void k_m2l_split(const int32_t *__restrict__ parents, int nparents,
                                                                             const int32_t *__restrict__ child, const uint32_t *__restrict__ count,
                                                                             const int32_t *__restrict__ nbr27, const uint4 *__restrict__ Ms,
                                                                             float *__restrict__ L, const bf16x8 *__restrict__ ops, int ldk) {
    constexpr int K = (P + 1) * (P + 1);
    constexpr int MT = mt_m2l(K), NKS = ks_of(K);
    constexpr int NTH = 64 * MT * NG;        // threads
    constexpr int NW = KS ? NT : NT / NG;    // n-tiles per wavefront
    static_assert(NT % NG == 0, "n-tiles must split evenly over the wavefronts of an m-tile");
    static_assert(!KS || (NG == 2 && NKS >= 3), "the k-step split pairs two wavefronts per m-tile");
    constexpr int TB = 32 * NT;              // targets per block
    constexpr int TILE = 2 * NKS * 64;       // bf16x8 vectors of one m-tile of an operator (hi and lo)
    constexpr int KH0 = KS ? 2 * ((NKS + 3) / 4) : (NKS + 1) / 2;  // k-steps of the first half (even with KS)
    constexpr int NH = NKS > KH0 ? 2 : 1;     // halves per offset step
    constexpr int CPT = (TB * 2 * KH0 + NTH - 1) / NTH;  // 8-coefficient chunks per thread and half
    __shared__ bf16x8 sB[2][KH0][NT][2][64];  // [buffer][k-step in half][n-tile][hi/lo][lane]
    __shared__ int16_t s_dn[189], s_qs[189], s_o[189];
    __shared__ int32_t s_sb[27][TB];         // first child of each parent neighbour (-1: none)
    __shared__ uint32_t s_live[7][TB];       // bit 8 dn + qs: that source cell holds particles
    __shared__ float s_sgn[16 * NKS];        // D_qt: sign of each coefficient under this octant's mirror
    __shared__ uint4 s_msk[2 * NKS];         // the same as BF16 sign-bit masks, 8 coefficients per entry
    __shared__ uint8_t s_nl[189];            // KS: bit n set when some target of n-tile n has a source at that step
    // XCD-aware placement: blocks go round-robin to the 8 XCDs, so XCD x takes
    // the contiguous run of work items [x per, (x+1) per); the 8 octant blocks
    // of a parent chunk, which read the same source multipoles and operators,
    // run side by side on one XCD and share its L2.
    const int nchunks = (nparents + TB - 1) / TB;
    const int w = (blockIdx.x & 7) * nchunks + (blockIdx.x >> 3);
    if (w >= 8 * nchunks) return;
    const int chunk = w >> 3, qt = w & 7;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform, so addresses built from it stay scalar
    const int mt = wid % MT, n0 = KS ? 0 : (wid / MT) * NW;
    const int kw = KS ? wid / MT : 0;  // KS: this wavefront takes the k-steps ks = kw mod 2
    // Every block walks the 189 offsets e of octant 0 in one fixed order.  For
    // octant qt the geometry is octant 0's mirrored in the axes where qt has a
    // bit set (parent neighbour d -> mirror(d), source octant q -> q ^ qt), and
    // the mirrored operator is D_qt T_e D_qt with D_qt diagonal +-1 (x -> -x:
    // (-1)^m on cosine, -(-1)^m on sine coefficients; y -> -y: -1 on sine;
    // z -> -z: (-1)^(n+m)).  So L = D_qt sum_e T_e (D_qt M): the signs go on the
    // gathered multipoles and on the result, and all blocks share 189 operators.
    if (tid == 0) m2l_walk(qt, s_o, s_dn, s_qs);
    for (int k = tid; k < 16 * NKS; k += NTH) s_sgn[k] = static_cast<float>(m2l_mirror_sign(qt, k));
    __syncthreads();
    for (int k8 = tid; k8 < 2 * NKS; k8 += NTH) {
        uint32_t w[4];
        for (int i = 0; i < 4; ++i)
            w[i] = (s_sgn[8 * k8 + 2 * i] < 0 ? 0x8000u : 0u) | (s_sgn[8 * k8 + 2 * i + 1] < 0 ? 0x80000000u : 0u);
        s_msk[k8] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    for (int t = tid; t < TB; t += NTH) {
        const int pi = chunk * TB + t;
        int32_t P0 = pi < nparents ? parents[pi] : -1;
        if (P0 >= 0 && count[child[P0] + qt] == 0) P0 = -1;
        uint32_t live[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int dn = 0; dn < 27; ++dn) {
            const int32_t sp = P0 >= 0 ? nbr27[27 * static_cast<int64_t>(P0) + dn] : -1;
            const int32_t sb = sp >= 0 ? child[sp] : -1;
            s_sb[dn][t] = sb;
            if (sb >= 0)
                for (int q = 0; q < 8; ++q)
                    if (count[sb + q] > 0) live[(8 * dn + q) >> 5] |= 1u << ((8 * dn + q) & 31);
        }
        for (int w = 0; w < 7; ++w) s_live[w][t] = live[w];
    }
    __syncthreads();
    if constexpr (KS) {  // n-tiles without any source at a step skip its products (their B is zero)
        for (int st = wid; st < 189; st += NTH / 64) {
            const int bit = 8 * s_dn[st] + s_qs[st];
            uint32_t m = 0;
            for (int h = 0; h < NT / 2; ++h) {
                const int t = 64 * h + lane;
                const uint64_t b = __ballot((s_live[bit >> 5][t] >> (bit & 31)) & 1u);
                m |= (b & 0xffffffffull ? 1u : 0u) << (2 * h);
                m |= (b >> 32 ? 1u : 0u) << (2 * h + 1);
            }
            if (lane == 0) s_nl[st] = static_cast<uint8_t>(m);
        }
        __syncthreads();
    }
    // source cell of target t at step st (-1: none or empty); st is uniform
    auto source_of = [&](int st, int t) -> int32_t {
        const int dn = __builtin_amdgcn_readfirstlane(s_dn[st]), qs = __builtin_amdgcn_readfirstlane(s_qs[st]);
        const int bit = 8 * dn + qs;
        return ((s_live[bit >> 5][t] >> (bit & 31)) & 1u) ? s_sb[dn][t] + qs : -1;
    };
    // gather half h of step st's B operand into registers: chunk c -> target
    // c / (2 kh), the 8 coefficients at k = 8 k8, k8 = 2 KH0 h + c % (2 kh), as
    // BF16 hi and lo halves (Ms: per cell ldk hi values then ldk lo values, 16
    // bytes per 8 coefficients)
    uint4 gh[CPT], gl[CPT];
    auto gather = [&](int st, int h) {
        const int kh = h ? NKS - KH0 : KH0;
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int c = tid + u * NTH;
            const int t = c / (2 * kh), k8 = 2 * KH0 * h + c % (2 * kh);
            const int32_t src_cell = (c < TB * 2 * kh) ? source_of(st, t) : -1;
            if (src_cell >= 0) {
                const uint4 *src = Ms + static_cast<int64_t>(src_cell) * (ldk / 4) + k8;  // ldk / 8 chunks of hi, then of lo
                gh[u] = src[0];
                gl[u] = src[ldk / 8];
            } else {
                gh[u] = make_uint4(0, 0, 0, 0);
                gl[u] = make_uint4(0, 0, 0, 0);
            }
        }
    };
    // store into buffer bf as MFMA B fragments with this octant's signs
    // (negation commutes with the split, so it is a flip of the sign bits):
    // lane (k8l % 2) * 32 + t % 32 of n-tile t / 32, k-step k8l / 2 of the half
    auto stash = [&](int bf, int h) {
        const int kh = h ? NKS - KH0 : KH0;
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int c = tid + u * NTH;
            if (c >= TB * 2 * kh) continue;
            const int t = c / (2 * kh), k8l = c % (2 * kh);
            const uint4 m = s_msk[2 * KH0 * h + k8l];
            const uint4 hv = make_uint4(gh[u].x ^ m.x, gh[u].y ^ m.y, gh[u].z ^ m.z, gh[u].w ^ m.w);
            const uint4 lv = make_uint4(gl[u].x ^ m.x, gl[u].y ^ m.y, gl[u].z ^ m.z, gl[u].w ^ m.w);
            const int ln = ((k8l & 1) * 32 + (t & 31)) ^ (k8l & 15);
            sB[bf][k8l >> 1][t >> 5][0][ln] = __builtin_bit_cast(bf16x8, hv);  // typed stores: the loads below are bf16x8
            sB[bf][k8l >> 1][t >> 5][1][ln] = __builtin_bit_cast(bf16x8, lv);
        }
    };
    f32x16 acc[NW];
#pragma unroll
    for (int n = 0; n < NW; ++n)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[n][r] = 0.0f;
    gather(0, 0);
    stash(0, 0);
    __syncthreads();
    // A fragments of this wavefront's m-tile for operator o (uniform), flat k-step q
    constexpr int TOT = 189 * NKS;
    constexpr int PF = 3;  // A prefetch distance in k-steps
    auto afrag = [&](int o, int ks, int h) -> bf16x8 {
        const bf16x8 *base = ops + (static_cast<int64_t>(o) * MT + mt) * TILE + h * NKS * 64 + ks * 64;
        return base[lane];
    };
    // B fragments of k-step ksl of buffer bf for this wavefront's n-tiles
    // (slot = lane ^ ((2 ksl + lane / 32) % 16) = (lane ^ lane / 32) ^ (2 ksl % 16):
    // one XOR with a constant on a per-lane byte offset)
    uint32_t lb = (lane ^ (lane >> 5)) * 16u;
    auto bfrag = [&](int bf, int ksl, bf16x8 *h, bf16x8 *l) {
        const uint32_t off = lb ^ (((2u * ksl) & 14u) << 4);
        const char *row = reinterpret_cast<const char *>(&sB[bf][ksl][n0][0][0]) + off;
#pragma unroll
        for (int n = 0; n < NW; ++n) {
            h[n] = *reinterpret_cast<const bf16x8 *>(row + (2 * n) * 1024);
            l[n] = *reinterpret_cast<const bf16x8 *>(row + (2 * n + 1) * 1024);
        }
    };
    if constexpr (KS) {
        // Wavefront position j of a step takes k-step ks = 2 j + kw (none when
        // ks >= NKS); A of positions p + 1 .. p + PF is in flight.  B is read
        // for the position itself.  Half h covers positions [h JH0, ...), the
        // same for both wavefronts of an m-tile, so every wavefront meets the
        // barriers at the same positions.
        constexpr int NJ = (NKS + 1) / 2, JH0 = KH0 / 2;
        constexpr int PFJ = 2;
        constexpr int NSO = (PFJ + NJ - 1) / NJ + 1;  // steps the prefetch of one step reaches
        // (positions without a k-step, and past the last step, load a valid
        // fragment that is not used: no branch, so the ring needs no copies)
        bf16x8 ah[PFJ], al[PFJ];
#pragma unroll
        for (int i = 0; i < PFJ; ++i) {
            const int ks = min(2 * (i % NJ) + kw, NKS - 1);
            const int o = __builtin_amdgcn_readfirstlane(s_o[i / NJ]);
            ah[i] = afrag(o, ks, 0);
            al[i] = afrag(o, ks, 1);
        }
        for (int st = 0; st < 189; ++st) {
            int so[NSO];  // operators of steps st .. st + NSO - 1, scalar
#pragma unroll
            for (int i = 0; i < NSO; ++i) so[i] = __builtin_amdgcn_readfirstlane(st + i < 189 ? s_o[st + i] : 0);
            asm volatile("" : "+v"(lb));  // recompute the B offsets per step rather than hold them all
            const uint32_t nl = __builtin_amdgcn_readfirstlane(s_nl[st]);
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const int h = j < JH0 ? 0 : 1, jl = j - h * JH0, jh = h ? NJ - JH0 : JH0;
                // with NKS odd the last k-step is shared: the even wavefront forms
                // hi*hi + hi*lo, the odd one lo*hi, so neither waits for the other
                constexpr bool SHARED_LAST = (NKS & 1) != 0;
                const bool last = SHARED_LAST && j == NJ - 1;
                const int ks = last ? NKS - 1 : 2 * j + kw, ksl = ks - h * KH0;
                const bool active = ks < NKS;
                const int bf = NH == 2 ? h : (st & 1);
                const int nst = (NH == 2 && h == 0) ? st : st + 1, nhf = (NH == 2 && h == 0) ? 1 : 0;
                const int nbf = NH == 2 ? nhf : ((st + 1) & 1);
                const bool has_next = nst < 189;
                const bf16x8 ch = ah[0], cl = al[0];
#pragma unroll
                for (int i = 0; i + 1 < PFJ; ++i) {
                    ah[i] = ah[i + 1];
                    al[i] = al[i + 1];
                }
                {
                    const int jp = j + PFJ, kp = min(2 * (jp % NJ) + kw, NKS - 1);
                    ah[PFJ - 1] = afrag(so[jp / NJ], kp, 0);
                    al[PFJ - 1] = afrag(so[jp / NJ], kp, 1);
                }
                if (jl == 0 && has_next) gather(nst, nhf);
                bf16x8 bh[NW], bl[NW];
                if (active) bfrag(bf, ksl, bh, bl);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int n = 0; n < NW; ++n) {
                    if (!((nl >> n) & 1u)) continue;
                    if (!last || kw == 0) {
                        acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ch, bh[n], acc[n], 0, 0, 0);
                        acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ch, bl[n], acc[n], 0, 0, 0);
                    }
                    if (!last || kw == 1) acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cl, bh[n], acc[n], 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
                if (jl + 1 == jh) {
                    if (has_next) stash(nbf, nhf);
                    __syncthreads();  // the next half is stored and this one is read
                }
            }
        }
        // add the odd-k-step sums to the even ones through LDS (B is no longer needed)
        static_assert(MT * NT * 16 * 64 * sizeof(float) <= sizeof(sB), "partial sums must fit in the B buffers");
        float *red = reinterpret_cast<float *>(&sB[0][0][0][0][0]);
        if (kw == 1)
#pragma unroll
            for (int n = 0; n < NW; ++n)
#pragma unroll
                for (int r = 0; r < 16; ++r) red[((mt * NT + n) * 16 + r) * 64 + lane] = acc[n][r];
        __syncthreads();
        if (kw == 1) return;  // no barrier follows
#pragma unroll
        for (int n = 0; n < NW; ++n)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[n][r] += red[((mt * NT + n) * 16 + r) * 64 + lane];
    } else {
        // Software pipeline, kept in place by scheduling barriers: during k-step q
        // the A fragments of q + 1 .. q + PF and the B fragments of q + 1 are in
        // flight.  The next half is gathered at the first k-step of a half, after
        // that k-step's A loads, and stored at the fourth (vector loads complete in
        // issue order, so the A loads issued after the gather already wait for it
        // there).
        // operator of flat k-step q: step q / NKS, k-step q % NKS
        auto op_of = [&](int q) -> int { return __builtin_amdgcn_readfirstlane(s_o[q / NKS]); };
        bf16x8 ah[PF], al[PF];  // A of k-steps q .. q + PF - 1
    #pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int o = op_of(i);
            ah[i] = afrag(o, i % NKS, 0);
            al[i] = afrag(o, i % NKS, 1);
        }
        bf16x8 bh[NW], bl[NW];
        bfrag(0, 0, bh, bl);
        for (int st = 0; st < 189; ++st) {
            asm volatile("" : "+v"(lb));  // recompute the B offsets per step rather than hold them all
    #pragma unroll
            for (int ks = 0; ks < NKS; ++ks) {
                const int h = ks < KH0 ? 0 : 1;             // half of this k-step
                const int ksl = ks - h * KH0, kh = h ? NKS - KH0 : KH0;
                const int bf = NH == 2 ? h : (st & 1);      // its buffer
                const int nst = (NH == 2 && h == 0) ? st : st + 1, nhf = (NH == 2 && h == 0) ? 1 : 0;  // the next half
                const int nbf = NH == 2 ? nhf : ((st + 1) & 1);
                const bool has_next = nst < 189;
                const int q = st * NKS + ks;
                const bf16x8 ch = ah[0], cl = al[0];
    #pragma unroll
                for (int i = 0; i + 1 < PF; ++i) {
                    ah[i] = ah[i + 1];
                    al[i] = al[i + 1];
                }
                if (q + PF < TOT) {
                    const int o = op_of(q + PF);
                    ah[PF - 1] = afrag(o, (ks + PF) % NKS, 0);
                    al[PF - 1] = afrag(o, (ks + PF) % NKS, 1);
                }
                if (ksl == 0 && has_next) gather(nst, nhf);
                bf16x8 nh[NW], nl[NW];
                if (ksl + 1 < kh) bfrag(bf, ksl + 1, nh, nl);
                if (ksl == (kh - 1 < 3 ? kh - 1 : 3) && has_next) stash(nbf, nhf);
                if (ksl + 1 == kh) {
                    // this half's B is in registers and the next half is stored:
                    // one barrier, then the first fragments of the next half
                    __syncthreads();
                    if (has_next) bfrag(nbf, 0, nh, nl);
                }
                __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
                for (int n = 0; n < NW; ++n) acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ch, bh[n], acc[n], 0, 0, 0);
    #pragma unroll
                for (int n = 0; n < NW; ++n) acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ch, bl[n], acc[n], 0, 0, 0);
    #pragma unroll
                for (int n = 0; n < NW; ++n) acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cl, bh[n], acc[n], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
                for (int n = 0; n < NW; ++n) {
                    bh[n] = nh[n];
                    bl[n] = nl[n];
                }
            }
        }
    }
    // C[8(r/4) + 4(l/32) + r%4][l%32] of n-tile n -> row of this m-tile, target 32 n + l%32
#pragma unroll
    for (int n = 0; n < NW; ++n) {
        const int t = 32 * (n0 + n) + (lane & 31);
        const int pi = chunk * TB + t;
        const int32_t P0 = pi < nparents ? parents[pi] : -1;
        if (P0 < 0) continue;
        const int32_t tgt = child[P0] + qt;
        if (count[tgt] == 0) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = mt * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (row < K) L[static_cast<int64_t>(tgt) * ldk + row] = acc[n][r] * s_sgn[row];
        }
    }
}

// M (FP32, row stride ldk, a multiple of 16) -> per cell ldk BF16 hi values
// then ldk lo values.  A block takes SPLIT_CELLS cells, a thread 4
// coefficients at a time (one 16-byte load, two 8-byte stores); the cell and
// the position in it come from a 32-bit division of the block-local index.
constexpr int SPLIT_CELLS = 32;
__global__ __launch_bounds__(256) void k_split_m(const float4 *__restrict__ M, int64_t n, int ldk, uint2 *__restrict__ Ms) {
    const uint32_t vpr = static_cast<uint32_t>(ldk) / 4;  // 4-coefficient vectors per cell
    const int64_t c0 = static_cast<int64_t>(blockIdx.x) * SPLIT_CELLS;
    const uint32_t nv = static_cast<uint32_t>(min<int64_t>(SPLIT_CELLS, n - c0)) * vpr;
    for (uint32_t t = threadIdx.x; t < nv; t += 256) {
        const uint32_t cl = t / vpr, k = t - cl * vpr;
        const int64_t c = c0 + cl;
        const float4 x = M[c * vpr + k];
        const float v[4] = {x.x, x.y, x.z, x.w};
        uint32_t h[4], l[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const __bf16 hb = static_cast<__bf16>(v[e]);
            h[e] = __builtin_bit_cast(uint16_t, hb);
            l[e] = __builtin_bit_cast(uint16_t, static_cast<__bf16>(v[e] - static_cast<float>(hb)));
        }
        Ms[c * 2 * vpr + k] = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
        Ms[c * 2 * vpr + vpr + k] = make_uint2(l[0] | (l[1] << 16), l[2] | (l[3] << 16));
    }
}

// ---------------------------------------------------------------------------
// M2M and L2L of one tree level on the matrix cores.  Their work is small
// (8 operators per level instead of 189 offsets), so they take a three-way
// split, x = h + m + l (BF16 each, 24 significant bits together), and six
// products, hh + hm + mh + hl + mm + lh, which keeps them at FP32 accuracy.
// The 8 M2M (child - parent) and 8 L2L (parent - child) operators are stored
// column-major for the FP32 kernels; their fragments are cut here in the A
// order of the M2L blocks, frag[blk][mt][part][ks][lane][8].
// ---------------------------------------------------------------------------
__device__ __forceinline__ void split3(float x, __bf16 &h, __bf16 &m, __bf16 &l) {
    h = static_cast<__bf16>(x);
    const float r = x - static_cast<float>(h);
    m = static_cast<__bf16>(r);
    l = static_cast<__bf16>(r - static_cast<float>(m));
}

__global__ __launch_bounds__(256) void k_split_ops_ml(const double *__restrict__ op64, int K, __bf16 *__restrict__ frag, int MT, int NKS) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    const int64_t per = static_cast<int64_t>(MT) * NKS * 64;
    if (i >= 16 * per) return;
    const int blk = static_cast<int>(i / per);
    const int rem = static_cast<int>(i % per);
    const int mt = rem / (NKS * 64), ks = (rem / 64) % NKS, l = rem % 64;
    const double *src = op64 + static_cast<int64_t>(blk) * K * K;
    const int row = mt * 32 + (l & 31);
    const int64_t vec = (static_cast<int64_t>(blk) * MT + mt) * 3 * NKS * 64 + ks * 64 + l;
    for (int e = 0; e < 8; ++e) {
        const int col = ks * 16 + (l >> 5) * 8 + e;
        const double x = (row < K && col < K) ? src[static_cast<int64_t>(col) * K + row] : 0.0;  // column-major
        // the first part from the FP64 value, the other two from the residual
        const __bf16 h = static_cast<__bf16>(static_cast<float>(x));
        __bf16 m, lo, rest;
        split3(static_cast<float>(x - static_cast<double>(static_cast<float>(h))), m, lo, rest);
        frag[vec * 8 + e] = h;
        frag[(vec + NKS * 64) * 8 + e] = m;
        frag[(vec + 2 * NKS * 64) * 8 + e] = lo;
    }
}

// Expansions of 32 cells (row stride ldk = 16 NKS; cells[t] < 0: zeros) ->
// LDS as h/m/l B fragments of one n-tile, the M2L kernel's slot order.  Rows
// K .. ldk-1 are read as zero whatever the buffer holds there (the locals'
// padding is never written).
template <int K, int NKS, int NTH>
__device__ __forceinline__ void stage_split(bf16x8 (*sB)[3][64], const float *__restrict__ src, int ldk, const int32_t *cells) {
    for (int c = threadIdx.x; c < 32 * 2 * NKS; c += NTH) {
        const int t = c / (2 * NKS), k8 = c % (2 * NKS);
        const int32_t cell = cells[t];
        float4 a = make_float4(0, 0, 0, 0), b = a;
        if (cell >= 0) {
            const float4 *q = reinterpret_cast<const float4 *>(src + static_cast<int64_t>(cell) * ldk + 8 * k8);
            a = q[0];
            b = q[1];
        }
        float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        bf16x8 h, m, l;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if (8 * k8 + e >= K) v[e] = 0.0f;
            __bf16 he, me, le;
            split3(v[e], he, me, le);
            h[e] = he;
            m[e] = me;
            l[e] = le;
        }
        const int ln = ((k8 & 1) * 32 + t) ^ (k8 & 15);
        sB[k8 >> 1][0][ln] = h;
        sB[k8 >> 1][1][ln] = m;
        sB[k8 >> 1][2][ln] = l;
    }
}

// acc += T (m-tile mt of fragment block fb) x B (LDS): the six products of
// the three-way split down to 2^-16 relative, smallest first
template <int NKS>
__device__ __forceinline__ void mfma_split_tile(const bf16x8 *__restrict__ fb, int mt, int lane, bf16x8 (*sB)[3][64], f32x16 &acc) {
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
        const bf16x8 ah = fb[(mt * 3) * NKS * 64 + ks * 64 + lane], am = fb[(mt * 3 + 1) * NKS * 64 + ks * 64 + lane],
                     al = fb[(mt * 3 + 2) * NKS * 64 + ks * 64 + lane];
        const int slot = lane ^ ((2 * ks + (lane >> 5)) & 15);
        const bf16x8 bh = sB[ks][0][slot], bm = sB[ks][1][slot], bl = sB[ks][2][slot];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bm, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bm, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
    }
}

// C of the wavefront's m-tile -> sO[t][row] (row stride OST, odd, so the 32
// targets of one row fall in different banks)
template <int OST>
__device__ __forceinline__ void park_tile(float *sO, int mt, int lane, const f32x16 &acc) {
#pragma unroll
    for (int r = 0; r < 16; ++r) sO[(lane & 31) * OST + mt * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)] = acc[r];
}

// M2M (solver_ref.hpp:196-197), first pass: block = (32 parents, child
// octant q) forms T_q M(child q) (zero for an empty child) into
// part[q][parent][row], rows K .. ldk-1 zero (zero rows of the operators).
template <int P>
__global__ __launch_bounds__(64 * mt_of((P + 1) * (P + 1))) void k_m2m_split(const int32_t *__restrict__ ids, int n, const int32_t *__restrict__ child,
                                                                          const uint32_t *__restrict__ count, const bf16x8 *__restrict__ frag,
                                                                          const float *__restrict__ M, int ldk, float *__restrict__ part) {
    constexpr int K = (P + 1) * (P + 1), MT = mt_of(K), NKS = ks_of(K), NTH = 64 * MT;
    constexpr int OST = MT * 32 + 1;
    __shared__ bf16x8 sB[NKS][3][64];
    __shared__ float sO[32 * OST];
    __shared__ int32_t s_cell[32];
    const int tid = threadIdx.x, lane = tid & 63;
    const int mt = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int q = blockIdx.x & 7, chunk = blockIdx.x >> 3;
    if (tid < 32) {
        const int32_t P0 = chunk * 32 + tid < n ? ids[chunk * 32 + tid] : -1;
        const int32_t cq = P0 >= 0 ? child[P0] + q : -1;
        s_cell[tid] = (cq >= 0 && count[cq] > 0) ? cq : -1;
    }
    __syncthreads();
    stage_split<K, NKS, NTH>(sB, M, ldk, s_cell);
    __syncthreads();
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    mfma_split_tile<NKS>(frag + static_cast<int64_t>(q) * MT * 3 * NKS * 64, mt, lane, sB, acc);
    park_tile<OST>(sO, mt, lane, acc);
    __syncthreads();
    const int nt = min(32, n - chunk * 32);
    for (int i = tid; i < nt * ldk; i += NTH) {
        const int t = i / ldk, row = i % ldk;
        part[(static_cast<int64_t>(q) * n + chunk * 32 + t) * ldk + row] = sO[t * OST + row];
    }
}

// M2M, second pass: M(parent) = sum over q = 0..7 of part[q][parent], in that order
__global__ __launch_bounds__(256) void k_m2m_sum(const int32_t *__restrict__ ids, int64_t n, int ldk, const float *__restrict__ part,
                                                 float *__restrict__ M) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= n * ldk) return;
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) s += part[q * n * ldk + i];
    M[static_cast<int64_t>(ids[i / ldk]) * ldk + i % ldk] = s;
}

// L2L (solver_ref.hpp:217-218): block = (32 parents, child octant q);
// L(child q) += T'_q L(parent) for every non-empty child.
template <int P>
__global__ __launch_bounds__(64 * mt_of((P + 1) * (P + 1))) void k_l2l_split(const int32_t *__restrict__ ids, int n, const int32_t *__restrict__ child,
                                                                          const uint32_t *__restrict__ count, const bf16x8 *__restrict__ frag,
                                                                          float *__restrict__ L, int ldk) {
    constexpr int K = (P + 1) * (P + 1), MT = mt_of(K), NKS = ks_of(K), NTH = 64 * MT;
    constexpr int OST = MT * 32 + 1;
    __shared__ bf16x8 sB[NKS][3][64];
    __shared__ float sO[32 * OST];
    __shared__ int32_t s_par[32], s_out[32];
    const int tid = threadIdx.x, lane = tid & 63;
    const int mt = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int q = blockIdx.x & 7, chunk = blockIdx.x >> 3;
    if (tid < 32) {
        const int32_t P0 = chunk * 32 + tid < n ? ids[chunk * 32 + tid] : -1;
        const int32_t cq = P0 >= 0 ? child[P0] + q : -1;
        s_par[tid] = P0;
        s_out[tid] = (cq >= 0 && count[cq] > 0) ? cq : -1;
    }
    __syncthreads();
    stage_split<K, NKS, NTH>(sB, L, ldk, s_par);
    __syncthreads();
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    mfma_split_tile<NKS>(frag + static_cast<int64_t>(8 + q) * MT * 3 * NKS * 64, mt, lane, sB, acc);
    park_tile<OST>(sO, mt, lane, acc);
    __syncthreads();
    for (int i = tid; i < 32 * K; i += NTH) {  // each child's row updated contiguously
        const int t = i / K, row = i % K;
        if (s_out[t] >= 0) L[static_cast<int64_t>(s_out[t]) * ldk + row] += sO[t * OST + row];
    }
}

template <int P>
void launch_ml(Ctx &c, const int32_t *ids, int64_t n, bool m2m) {
    constexpr int K = (P + 1) * (P + 1);
    const unsigned grid = static_cast<unsigned>(ceil_div(n, 32));
    const bf16x8 *frag = reinterpret_cast<const bf16x8 *>(c.opsplit_ml.p);
    if (m2m) {
        c.ml_part.ensure(static_cast<size_t>(8 * n * c.ldk));
        k_m2m_split<P><<<8 * grid, 64 * mt_of(K), 0, c.stream>>>(ids, static_cast<int>(n), c.t.child, c.t.count, frag, c.M32, c.ldk, c.ml_part);
        k_m2m_sum<<<ceil_div(n * c.ldk, 256), 256, 0, c.stream>>>(ids, n, c.ldk, c.ml_part, c.M32);
    } else
        k_l2l_split<P><<<8 * grid, 64 * mt_of(K), 0, c.stream>>>(ids, static_cast<int>(n), c.t.child, c.t.count, frag, c.L32, c.ldk);
    VG_LAUNCH_CHECK();
}

template <int P>
void launch_split(Ctx &c, const int32_t *parents, int np) {
    constexpr int K = (P + 1) * (P + 1);
    constexpr int NT = 4, NG = 2;  // 128 targets, 12 wavefronts: 3 per SIMD
    constexpr bool KS = ks_of(K) >= 3;
    const unsigned grid = 8 * ceil_div(np, 32 * NT);  // one block per (parent chunk, child octant)
    k_m2l_split<P, NT, NG, KS><<<grid, 64 * NG * mt_m2l(K), 0, c.stream>>>(parents, np, c.t.child, c.t.count, c.t.nbr27,
                                                                     reinterpret_cast<const uint4 *>(c.msplit.p), c.L32,
                                                                     reinterpret_cast<const bf16x8 *>(c.opsplit.p), c.ldk);
    VG_LAUNCH_CHECK();
}

}  // namespace

void m2l_split_operators(Ctx &c, int nvec) {
    const int MT = mt_of(c.K), NKS = ks_of(c.K), MTL = mt_m2l(c.K);
    const int64_t per = static_cast<int64_t>(MT) * NKS * 64;
    c.opsplit.ensure(static_cast<size_t>(343 * 2 * static_cast<int64_t>(MTL) * NKS * 64 * 8));
    // split mode runs M2M, M2L and L2L from the fragments: no FP32 copies
    (void)nvec;
    k_split_ops<<<189 * MTL, 256, 32 * (16 * NKS + 1) * sizeof(double), c.stream>>>(c.ops_sel, c.op64, c.K, reinterpret_cast<__bf16 *>(c.opsplit.p),
                                                                                  MTL, NKS);
    c.opsplit_ml.ensure(static_cast<size_t>(16 * 3 * per * 8));
    k_split_ops_ml<<<ceil_div(16 * per, 256), 256, 0, c.stream>>>(c.op64, c.K, reinterpret_cast<__bf16 *>(c.opsplit_ml.p), MT, NKS);
    VG_LAUNCH_CHECK();
}

void ml_split_level(Ctx &c, const int32_t *ids, int64_t n, bool m2m) {
    if (n <= 0) return;
    switch (c.p) {
        case 1: launch_ml<1>(c, ids, n, m2m); break;
        case 2: launch_ml<2>(c, ids, n, m2m); break;
        case 3: launch_ml<3>(c, ids, n, m2m); break;
        case 4: launch_ml<4>(c, ids, n, m2m); break;
        case 5: launch_ml<5>(c, ids, n, m2m); break;
        case 6: launch_ml<6>(c, ids, n, m2m); break;
        case 7: launch_ml<7>(c, ids, n, m2m); break;
        case 8: launch_ml<8>(c, ids, n, m2m); break;
        case 9: launch_ml<9>(c, ids, n, m2m); break;
        case 10: launch_ml<10>(c, ids, n, m2m); break;
        case 11: launch_ml<11>(c, ids, n, m2m); break;
        case 12: launch_ml<12>(c, ids, n, m2m); break;
        default: throw Error(1, "split M2M / L2L supports 1 <= p <= 12");
    }
}

void fmm_m2l_split_list(Ctx &c, const int32_t *parents, int np) {
    if (np <= 0) return;
    const int64_t nc = c.t.ncells;
    c.msplit.ensure(static_cast<size_t>(nc) * 2 * c.ldk);
    if (c.ldk % 16 != 0) throw Error(1, "split M2L: the coefficient stride must be a multiple of 16");
    k_split_m<<<ceil_div(nc, SPLIT_CELLS), 256, 0, c.stream>>>(reinterpret_cast<const float4 *>(c.M32.p), nc, c.ldk,
                                                             reinterpret_cast<uint2 *>(c.msplit.p));
    VG_LAUNCH_CHECK();
    switch (c.p) {
        case 1: launch_split<1>(c, parents, np); break;
        case 2: launch_split<2>(c, parents, np); break;
        case 3: launch_split<3>(c, parents, np); break;
        case 4: launch_split<4>(c, parents, np); break;
        case 5: launch_split<5>(c, parents, np); break;
        case 6: launch_split<6>(c, parents, np); break;
        case 7: launch_split<7>(c, parents, np); break;
        case 8: launch_split<8>(c, parents, np); break;
        case 9: launch_split<9>(c, parents, np); break;
        case 10: launch_split<10>(c, parents, np); break;
        case 11: launch_split<11>(c, parents, np); break;
        case 12: launch_split<12>(c, parents, np); break;
        default: throw Error(1, "split M2L supports 1 <= p <= 12");
    }
}

void fmm_m2l_split(Ctx &c) {
    TreeDev &t = c.t;
    if (t.depth < 1) return;
    const int64_t a = t.int_off[1], b = t.int_off.back();
    if (b <= a) return;
    fmm_m2l_split_list(c, t.int_ids.p + a, static_cast<int>(b - a));
}

}  // namespace vg
