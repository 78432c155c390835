// sort.hip - stable LSD radix sort of (u64 key, u32 value) pairs.
//
// Three kernels per digit of up to 9 bits (the passes split the key bits
// evenly: 36 bits take 4 passes of 9, 64 take 8 of 8): a tile histogram, a
// per-digit scan over the tiles (which also gives each digit's total), and a
// scatter that ranks equal digits with ballots, wavefront by wavefront in key
// order, so the order of equal keys is kept (the tree needs points of a cell
// in the caller's index order).  The scatter reorders each tile in LDS before
// writing, so its global writes are runs of consecutive addresses.
#include "gravity.hpp"

namespace vg {
namespace {

constexpr int RS_THREADS = 512;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_MAXBITS = 9;
constexpr int RS_BINS = 1 << RS_MAXBITS;  // bins held; a pass of b bits uses the first 2^b
constexpr int RS_BPT = RS_BINS / RS_THREADS;
constexpr int RS_WAVES = RS_THREADS / 64;
static_assert(RS_BINS % RS_THREADS == 0, "whole bins per thread");

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *__restrict__ key, int64_t n, int shift, unsigned mask, uint32_t *__restrict__ hist,
                                                        unsigned ntiles) {
    __shared__ uint32_t h[RS_BINS];
    const unsigned tile = blockIdx.x;
#pragma unroll
    for (int q = 0; q < RS_BPT; ++q) h[threadIdx.x + q * RS_THREADS] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(tile) * RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int64_t i = base + r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(key[i] >> shift) & mask], 1u);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RS_BPT; ++q) {
        const unsigned b = threadIdx.x + q * RS_THREADS;
        if (b > mask) break;
        hist[static_cast<size_t>(b) * ntiles + tile] = h[b];
    }
}

// one block per digit value: exclusive scan of hist[bin][0..ntiles) in
// place, and the digit's total into bin_total[bin]
__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t *__restrict__ hist, uint32_t *__restrict__ bin_total, unsigned ntiles) {
    __shared__ uint32_t warp_sum[16];
    __shared__ uint32_t carry;
    const unsigned bin = blockIdx.x;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    uint32_t *row = hist + static_cast<size_t>(bin) * ntiles;
    for (unsigned off = 0; off < ntiles; off += 1024) {
        const unsigned i = off + threadIdx.x;
        const uint32_t v = i < ntiles ? row[i] : 0u;
        // inclusive scan within the wavefront
        uint32_t s = v;
        const unsigned lane = threadIdx.x & 63;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(s, d, 64);
            if (lane >= static_cast<unsigned>(d)) s += o;
        }
        if (lane == 63) warp_sum[threadIdx.x >> 6] = s;
        __syncthreads();
        uint32_t wbase = 0;
        for (unsigned w = 0; w < (threadIdx.x >> 6); ++w) wbase += warp_sum[w];
        const uint32_t c0 = carry;
        if (i < ntiles) row[i] = c0 + wbase + s - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c0 + wbase + s;
        __syncthreads();
    }
    if (threadIdx.x == 0) bin_total[bin] = carry;
}

// exclusive scan of one value per thread over the block (RS_THREADS threads)
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *wsum, uint32_t &total) {
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(x, d, 64);
        if (lane >= static_cast<unsigned>(d)) x += o;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t wb = 0, t = 0;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
        wb += static_cast<unsigned>(w) < wave ? wsum[w] : 0u;
        t += wsum[w];
    }
    total = t;
    return wb + x - v;
}

// Scatter of one 4096-key tile (512 threads x 8 keys).  Wavefront w ranks the
// contiguous keys [512 w, 512 w + 512) of the tile in 8 rounds of 64, in
// order: the lanes with equal digits find each other with ballots, and a
// running count per (wavefront, digit) in LDS gives each key its rank among
// the wavefront's keys of that digit, so equal keys keep their order.  One
// block scan then turns the per-wavefront counts into tile-local starts; the
// keys are reordered in LDS and written out in that order, so consecutive
// threads write consecutive addresses of each digit's run.  hist holds the
// tile's start within each digit (exclusive over tiles), bin_total each
// digit's total, whose exclusive scan is the digit's start in the output.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *__restrict__ key, const uint32_t *__restrict__ val, uint64_t *__restrict__ key_out,
                                                           uint32_t *__restrict__ val_out, int64_t n, int shift, int bits,
                                                           const uint32_t *__restrict__ hist, const uint32_t *__restrict__ bin_total, unsigned ntiles) {
    static_assert(RS_BINS == RS_THREADS, "one digit per thread in the scans");
    static_assert(RS_TILE == RS_WAVES * RS_ITEMS * 64, "each wavefront ranks RS_ITEMS rounds of 64 keys");
    __shared__ uint64_t sk[RS_TILE];
    __shared__ uint32_t sv[RS_TILE];
    __shared__ uint32_t wc[RS_WAVES][RS_BINS];  // keys per (wavefront, digit); then the wavefront's tile-local start
    __shared__ uint32_t start[RS_BINS];         // tile-local start of each digit
    __shared__ uint32_t gbase[RS_BINS];         // output position of each digit's first key of this tile
    __shared__ uint32_t wsum[RS_WAVES];
    const unsigned mask = (1u << bits) - 1u;
    const unsigned tile = blockIdx.x;
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = static_cast<int64_t>(tile) * RS_TILE;
    const int nt = static_cast<int>(min<int64_t>(RS_TILE, n - base));
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) wc[w][threadIdx.x] = 0;
    __syncthreads();
    uint64_t k[RS_ITEMS];
    uint32_t v[RS_ITEMS], lp[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int j = static_cast<int>(wave) * (RS_ITEMS * 64) + r * 64 + static_cast<int>(lane);
        const bool ok = j < nt;
        k[r] = ok ? key[base + j] : 0ull;
        v[r] = ok ? val[base + j] : 0u;
        const unsigned d = static_cast<unsigned>((k[r] >> shift) & mask);
        uint64_t match = __ballot(ok);
        for (int b = 0; b < bits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            match &= bit ? bal : ~bal;
        }
        const unsigned rank = __popcll(match & lt);
        const unsigned cnt = __popcll(match);
        const uint32_t c0 = ok ? wc[wave][d] : 0u;  // read by every lane of the digit before its last lane updates it
        lp[r] = c0 + rank;
        if (ok && rank == cnt - 1) wc[wave][d] = c0 + cnt;
    }
    __syncthreads();
    {
        const unsigned b = threadIdx.x;  // digit
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) {
            const uint32_t c = wc[w][b];
            wc[w][b] = run;
            run += c;
        }
        uint32_t tot;
        const uint32_t st = block_excl_scan(run, wsum, tot);  // tile-local start of digit b
        __syncthreads();                                      // wsum is reused below
        const uint32_t bt = b <= mask ? bin_total[b] : 0u;
        const uint32_t dbase = block_excl_scan(bt, wsum, tot);  // output start of digit b
        start[b] = st;
        gbase[b] = b <= mask ? dbase + hist[static_cast<size_t>(b) * ntiles + tile] : 0u;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) wc[w][b] += st;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int j = static_cast<int>(wave) * (RS_ITEMS * 64) + r * 64 + static_cast<int>(lane);
        if (j < nt) {
            const unsigned d = static_cast<unsigned>((k[r] >> shift) & mask);
            const uint32_t pos = wc[wave][d] + lp[r];
            sk[pos] = k[r];
            sv[pos] = v[r];
        }
    }
    __syncthreads();
    // write out in LDS order: slot j of digit d goes to its output run
    for (int j = threadIdx.x; j < nt; j += RS_THREADS) {
        const uint64_t kk = sk[j];
        const unsigned d = static_cast<unsigned>((kk >> shift) & mask);
        const uint32_t pos = gbase[d] + (j - start[d]);
        key_out[pos] = kk;
        val_out[pos] = sv[j];
    }
}

}  // namespace

// Sorts (k, v) by bits [begin_bit, end_bit) of k; on return k/v point to the
// sorted data (the buffers may have been swapped with k_alt/v_alt).
void radix_sort_pairs(hipStream_t st, uint64_t *&k, uint32_t *&v, uint64_t *&k_alt, uint32_t *&v_alt, int64_t n, int begin_bit, int end_bit,
                      DevBuf<uint32_t> &hist) {
    if (n <= 1) return;
    const unsigned ntiles = ceil_div(static_cast<size_t>(n), RS_TILE);
    hist.ensure(static_cast<size_t>(RS_BINS) * ntiles + RS_BINS);
    uint32_t *bin_total = hist.p + static_cast<size_t>(RS_BINS) * ntiles;
    const int nbits = end_bit - begin_bit;
    const int passes = (nbits + RS_MAXBITS - 1) / RS_MAXBITS;
    const int width = (nbits + passes - 1) / passes;  // <= RS_MAXBITS
    for (int shift = begin_bit; shift < end_bit; shift += width) {
        const int bits = std::min(width, end_bit - shift);
        const unsigned nbins = 1u << bits;
        k_rs_hist<<<ntiles, RS_THREADS, 0, st>>>(k, n, shift, nbins - 1, hist.p, ntiles);
        k_rs_scan<<<nbins, 1024, 0, st>>>(hist.p, bin_total, ntiles);
        k_rs_scatter<<<ntiles, RS_THREADS, 0, st>>>(k, v, k_alt, v_alt, n, shift, bits, hist.p, bin_total, ntiles);
        VG_LAUNCH_CHECK();
        std::swap(k, k_alt);
        std::swap(v, v_alt);
    }
}

}  // namespace vg
