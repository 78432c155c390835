// sort.hip - stable LSD radix sort of (u64 key, u32 value) pairs.
//
// Three kernels per digit of up to 9 bits (the passes split the key bits
// evenly: 36 bits take 4 passes of 9, 64 take 8 of 8): a tile histogram, a
// per-digit scan over the tiles, and a scatter that ranks equal digits inside each 64-lane wavefront
// with ballots, so the order of equal keys is kept (the tree needs points of a
// cell in the caller's index order).  The scatter reorders each tile in LDS
// before writing, so its global writes are runs of consecutive addresses.
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
                                                        uint32_t *__restrict__ bin_total, unsigned ntiles) {
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
        const uint32_t c = h[b];
        hist[static_cast<size_t>(b) * ntiles + tile] = c;
        if (c) atomicAdd(&bin_total[b], c);
    }
}

// one block per digit value: exclusive scan of hist[bin][0..ntiles) plus the
// number of keys with a smaller digit.
__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t *__restrict__ hist, const uint32_t *__restrict__ bin_total, unsigned ntiles) {
    __shared__ uint32_t warp_sum[16];
    __shared__ uint32_t carry;
    const unsigned bin = blockIdx.x;
    {  // keys with a smaller digit
        uint32_t b = 0;
        for (unsigned q = threadIdx.x; q < bin; q += 1024) b += bin_total[q];
        for (int d = 32; d >= 1; d >>= 1) b += __shfl_xor(b, d, 64);
        if ((threadIdx.x & 63) == 0) warp_sum[threadIdx.x >> 6] = b;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int w = 0; w < 16; ++w) t += warp_sum[w];
            carry = t;
        }
        __syncthreads();
    }
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
}

// Scatter of one 4096-key tile (512 threads x 8 keys; measured against 256 x 16,
// 512 x 4 and 512 x 16, this shape ran the tree's sort fastest): the keys are first ranked stably by digit
// (ballots inside a wavefront, running counts across wavefronts and rounds)
// and reordered in LDS, then written out in that order, so consecutive threads
// write consecutive addresses of each digit's run.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *__restrict__ key, const uint32_t *__restrict__ val, uint64_t *__restrict__ key_out,
                                                           uint32_t *__restrict__ val_out, int64_t n, int shift, int bits,
                                                           const uint32_t *__restrict__ hist, unsigned ntiles) {
    __shared__ uint64_t sk[RS_TILE];
    __shared__ uint32_t sv[RS_TILE];
    __shared__ uint32_t run[RS_BINS];    // next tile-local slot of each digit
    __shared__ uint32_t start[RS_BINS];  // tile-local start of each digit
    __shared__ uint32_t wcnt[RS_WAVES][RS_BINS];
    __shared__ uint32_t wsum[RS_WAVES];
    __shared__ uint32_t qsum[RS_BPT];
    const unsigned mask = (1u << bits) - 1u;
    const unsigned tile = blockIdx.x;
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = static_cast<int64_t>(tile) * RS_TILE;
    const int nt = static_cast<int>(min<int64_t>(RS_TILE, n - base));
    // tile histogram and its exclusive scan (bins q RS_THREADS + threadIdx.x)
#pragma unroll
    for (int q = 0; q < RS_BPT; ++q) run[threadIdx.x + q * RS_THREADS] = 0;
    __syncthreads();
    for (int j = threadIdx.x; j < nt; j += RS_THREADS) atomicAdd(&run[(key[base + j] >> shift) & mask], 1u);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RS_BPT; ++q) {
        const unsigned b = threadIdx.x + q * RS_THREADS;
        const uint32_t c = run[b];
        uint32_t x = c;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(x, d, 64);
            if (lane >= static_cast<unsigned>(d)) x += o;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        uint32_t wb = 0;
        for (unsigned w = 0; w < wave; ++w) wb += wsum[w];
        start[b] = wb + x - c;  // within this group of RS_THREADS bins
        if (threadIdx.x == RS_THREADS - 1) qsum[q] = wb + x;
        __syncthreads();
    }
#pragma unroll
    for (int q = 1; q < RS_BPT; ++q) {
        uint32_t off = 0;
        for (int r = 0; r < q; ++r) off += qsum[r];
        start[threadIdx.x + q * RS_THREADS] += off;
    }
#pragma unroll
    for (int q = 0; q < RS_BPT; ++q) run[threadIdx.x + q * RS_THREADS] = start[threadIdx.x + q * RS_THREADS];
    __syncthreads();
    // stable local ranks, written to LDS in digit order
    for (int r = 0; r < RS_ITEMS; ++r) {
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w)
#pragma unroll
            for (int q = 0; q < RS_BPT; ++q) wcnt[w][threadIdx.x + q * RS_THREADS] = 0;
        __syncthreads();
        const int j = r * RS_THREADS + threadIdx.x;
        const bool ok = j < nt;
        uint64_t k = 0;
        uint32_t v = 0;
        unsigned d = 0;
        if (ok) {
            k = key[base + j];
            v = val[base + j];
            d = static_cast<unsigned>((k >> shift) & mask);
        }
        uint64_t match = __ballot(ok);
        for (int b = 0; b < bits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            match &= bit ? bal : ~bal;
        }
        const unsigned rank = __popcll(match & lt);
        const unsigned cnt = __popcll(match);
        if (ok && rank == cnt - 1) wcnt[wave][d] = cnt;
        __syncthreads();
        if (ok) {
            uint32_t pos = run[d] + rank;
            for (unsigned w = 0; w < wave; ++w) pos += wcnt[w][d];
            sk[pos] = k;
            sv[pos] = v;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RS_BPT; ++q) {
            const unsigned b = threadIdx.x + q * RS_THREADS;
            uint32_t add = 0;
#pragma unroll
            for (int w = 0; w < RS_WAVES; ++w) add += wcnt[w][b];
            run[b] += add;
        }
        __syncthreads();
    }
    // write out in LDS order: slot j of digit d goes to its global run
    for (int j = threadIdx.x; j < nt; j += RS_THREADS) {
        const uint64_t k = sk[j];
        const unsigned d = static_cast<unsigned>((k >> shift) & mask);
        const uint32_t pos = hist[static_cast<size_t>(d) * ntiles + tile] + (j - start[d]);
        key_out[pos] = k;
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
        VG_HIP(hipMemsetAsync(bin_total, 0, RS_BINS * sizeof(uint32_t), st));
        k_rs_hist<<<ntiles, RS_THREADS, 0, st>>>(k, n, shift, nbins - 1, hist.p, bin_total, ntiles);
        k_rs_scan<<<nbins, 1024, 0, st>>>(hist.p, bin_total, ntiles);
        k_rs_scatter<<<ntiles, RS_THREADS, 0, st>>>(k, v, k_alt, v_alt, n, shift, bits, hist.p, ntiles);
        VG_LAUNCH_CHECK();
        std::swap(k, k_alt);
        std::swap(v, v_alt);
    }
}

}  // namespace vg
