// sort.hip - stable LSD radix sort of (u64 key, u32 value) pairs.
//
// Three kernels per 8-bit digit: a tile histogram, a per-digit scan over the
// tiles, and a scatter that ranks equal digits inside each 64-lane wavefront
// with ballots, so the order of equal keys is kept (the tree needs points of a
// cell in the caller's index order).  The scatter reorders each tile in LDS
// before writing, so its global writes are runs of consecutive addresses.
#include "gravity.hpp"

namespace vg {
namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_BINS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *__restrict__ key, int64_t n, int shift, uint32_t *__restrict__ hist,
                                                        uint32_t *__restrict__ bin_total, unsigned ntiles) {
    __shared__ uint32_t h[RS_BINS];
    const unsigned tile = blockIdx.x;
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(tile) * RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int64_t i = base + r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(key[i] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    const uint32_t c = h[threadIdx.x];
    hist[static_cast<size_t>(threadIdx.x) * ntiles + tile] = c;
    if (c) atomicAdd(&bin_total[threadIdx.x], c);
}

// one block per digit value: exclusive scan of hist[bin][0..ntiles) plus the
// number of keys with a smaller digit.
__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t *__restrict__ hist, const uint32_t *__restrict__ bin_total, unsigned ntiles) {
    __shared__ uint32_t warp_sum[16];
    __shared__ uint32_t carry;
    const unsigned bin = blockIdx.x;
    if (threadIdx.x == 0) {
        uint32_t b = 0;
        for (unsigned q = 0; q < bin; ++q) b += bin_total[q];
        carry = b;
    }
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
}

// Scatter of one 4096-key tile: the keys are first ranked stably by digit
// (ballots inside a wavefront, running counts across wavefronts and rounds)
// and reordered in LDS, then written out in that order, so consecutive threads
// write consecutive addresses of each digit's run.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *__restrict__ key, const uint32_t *__restrict__ val, uint64_t *__restrict__ key_out,
                                                           uint32_t *__restrict__ val_out, int64_t n, int shift, const uint32_t *__restrict__ hist,
                                                           unsigned ntiles) {
    __shared__ uint64_t sk[RS_TILE];
    __shared__ uint32_t sv[RS_TILE];
    __shared__ uint32_t run[RS_BINS];    // next tile-local slot of each digit
    __shared__ uint32_t start[RS_BINS];  // tile-local start of each digit
    __shared__ uint32_t wcnt[RS_WAVES][RS_BINS];
    __shared__ uint32_t wsum[RS_WAVES];
    const unsigned tile = blockIdx.x;
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = static_cast<int64_t>(tile) * RS_TILE;
    const int nt = static_cast<int>(min<int64_t>(RS_TILE, n - base));
    // tile histogram and its exclusive scan
    run[threadIdx.x] = 0;
    __syncthreads();
    for (int j = threadIdx.x; j < nt; j += RS_THREADS) atomicAdd(&run[(key[base + j] >> shift) & 0xFF], 1u);
    __syncthreads();
    {
        const uint32_t c = run[threadIdx.x];
        uint32_t x = c;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(x, d, 64);
            if (lane >= static_cast<unsigned>(d)) x += o;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        uint32_t wb = 0;
        for (unsigned w = 0; w < wave; ++w) wb += wsum[w];
        start[threadIdx.x] = wb + x - c;
        run[threadIdx.x] = wb + x - c;
    }
    __syncthreads();
    // stable local ranks, written to LDS in digit order
    for (int r = 0; r < RS_ITEMS; ++r) {
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        const int j = r * RS_THREADS + threadIdx.x;
        const bool ok = j < nt;
        uint64_t k = 0;
        uint32_t v = 0;
        unsigned d = 0;
        if (ok) {
            k = key[base + j];
            v = val[base + j];
            d = static_cast<unsigned>((k >> shift) & 0xFF);
        }
        uint64_t match = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
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
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) add += wcnt[w][threadIdx.x];
        run[threadIdx.x] += add;
        __syncthreads();
    }
    // write out in LDS order: slot j of digit d goes to its global run
    for (int j = threadIdx.x; j < nt; j += RS_THREADS) {
        const uint64_t k = sk[j];
        const unsigned d = static_cast<unsigned>((k >> shift) & 0xFF);
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
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        VG_HIP(hipMemsetAsync(bin_total, 0, RS_BINS * sizeof(uint32_t), st));
        k_rs_hist<<<ntiles, RS_THREADS, 0, st>>>(k, n, shift, hist.p, bin_total, ntiles);
        k_rs_scan<<<RS_BINS, 1024, 0, st>>>(hist.p, bin_total, ntiles);
        k_rs_scatter<<<ntiles, RS_THREADS, 0, st>>>(k, v, k_alt, v_alt, n, shift, hist.p, ntiles);
        VG_LAUNCH_CHECK();
        std::swap(k, k_alt);
        std::swap(v, v_alt);
    }
}

}  // namespace vg
