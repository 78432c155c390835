// mlp.hip - the learned closure's network: 28 -> 128 -> 128 -> 25, GELU on
// the hidden layers, FP16 weights and activations with FP32 accumulation in
// the forward pass (v_mfma_f32_16x16x16_f16), FP32 backward pass and Adam.
//
// Parameters are one FP32 vector (the master copy), row-major [out][in]:
//   W1 (128 x 28), b1 (128), W2 (128 x 128), b2 (128), W3 (25 x 128), b3 (25).
// The forward pass reads FP16 copies of W1..W3 prepacked in MFMA A-operand
// order; the backward pass uses the same FP16-rounded weights (as FP32), so
// the gradient is that of the function the forward pass evaluates (the FP16
// rounding of activations is passed straight through).
//
// Forward mapping: a wavefront evaluates 16 samples.  Layer l computes
// Z = W X^T: A = W (out x in, 16 x 16 tiles), B = X^T (in x samples).  For
// v_mfma_f32_16x16x16_f16, lane l holds A[l%16][4(l/16)+e], B[4(l/16)+e][l%16]
// and D[4(l/16)+r][l%16] (e, r = 0..3): the output tile t of one layer sits in
// the lanes exactly as the k-step t of the next layer's B operand, so hidden
// activations never leave registers.
#include "gravity.hpp"
#include "mlp.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>

namespace vg {
namespace mlp {
namespace {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

// A fragments: layer 1: 8 out-tiles x 2 k-steps, layer 2: 8 x 8, layer 3: 2 x 8
constexpr int F1 = 0, F2 = 16, F3 = 80, NFRAG = 96;
__host__ __device__ constexpr int frag_base(int layer) { return layer == 0 ? F1 : layer == 1 ? F2 : F3; }
__host__ __device__ constexpr int ksteps(int layer) { return layer == 0 ? 2 : 8; }

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// FP32 params -> FP16 A fragments [frag][lane][4] and FP16-rounded FP32 weights
__global__ __launch_bounds__(256) void k_pack(const float *__restrict__ par, _Float16 *__restrict__ frag, float *__restrict__ wr) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < NPAR) wr[i] = static_cast<float>(static_cast<_Float16>(par[i]));
    if (i >= NFRAG * 64) return;
    const int f = i / 64, l = i % 64;
    int layer, t, s;
    if (f < F2) layer = 0, t = f / 2, s = f % 2;
    else if (f < F3) layer = 1, t = (f - F2) / 8, s = (f - F2) % 8;
    else layer = 2, t = (f - F3) / 8, s = (f - F3) % 8;
    const int nin = layer == 0 ? IN : H, nout = layer == 2 ? OUT : H;
    const int woff = layer == 0 ? OW1 : layer == 1 ? OW2 : OW3;
    const int row = 16 * t + (l & 15);
    for (int e = 0; e < 4; ++e) {
        const int col = 16 * s + 4 * (l >> 4) + e;
        const float w = (row < nout && col < nin) ? par[woff + row * nin + col] : 0.0f;
        frag[(static_cast<int64_t>(f) * 64 + l) * 4 + e] = static_cast<_Float16>(w);
    }
}

// Forward pass of 16 samples per wavefront, 4 wavefronts per block.  With
// TRAIN the FP16-rounded input, pre-activations and FP16-rounded activations
// are stored for the backward pass (n x 32, n x 128 each).
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_forward(const half4 *__restrict__ frag, const float *__restrict__ par, const float *__restrict__ x, int64_t n,
                                                 float *__restrict__ y, float *__restrict__ x16, float *__restrict__ z1, float *__restrict__ a1,
                                                 float *__restrict__ z2, float *__restrict__ a2) {
    const int lane = threadIdx.x & 63;
    const int64_t j = (blockIdx.x * 4ll + (threadIdx.x >> 6)) * 16 + (lane & 15);  // this lane's sample
    const bool live = j < n;
    const int q = lane >> 4;                                                        // row group: rows 4q .. 4q+3 of a tile
    half4 b[8];
    // layer 1 input: k-step s holds inputs 16 s + 4 q + e
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int k = 16 * s + 4 * q + e;
            const float v = (live && k < IN) ? x[j * IN + k] : 0.0f;
            b[s][e] = static_cast<_Float16>(v);
            if (TRAIN && live) x16[j * 32 + k] = static_cast<float>(b[s][e]);
        }
    half4 h[8];
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        const int ks = ksteps(layer);
        const float *bias = par + (layer == 0 ? OB1 : OB2);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < ks; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x16f16(frag[(frag_base(layer) + t * ks + s) * 64 + lane], b[s], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 16 * t + 4 * q + r;
                const float z = acc[r] + bias[o];
                const _Float16 a = static_cast<_Float16>(gelu(z));
                h[t][r] = a;
                if (TRAIN && live) {
                    (layer == 0 ? z1 : z2)[j * H + o] = z;
                    (layer == 0 ? a1 : a2)[j * H + o] = static_cast<float>(a);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) b[t] = h[t];
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 8; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x16f16(frag[(F3 + t * 8 + s) * 64 + lane], b[s], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int o = 16 * t + 4 * q + r;
            if (live && o < OUT) y[j * OUT + o] = acc[r] + par[OB3 + o];
        }
    }
}

// Backward through the activations, one block of 128 threads per 8 samples:
// dY = 2 w (y - t) / (W 25) with W the sum of the sample weights w (1 when
// none are given), dZ2 = (dY W3) * gelu'(Z2), dZ1 = (dZ2 W2) * gelu'(Z1).
// Also the per-block weighted sum of squared errors.
__global__ __launch_bounds__(128) void k_backward_act(const float *__restrict__ y, const float *__restrict__ tgt, const float *__restrict__ wt,
                                                      int64_t n, float inv_count,
                                                      const float *__restrict__ wr, const float *__restrict__ z1, const float *__restrict__ z2,
                                                      float *__restrict__ dy, float *__restrict__ dz1, float *__restrict__ dz2,
                                                      float *__restrict__ sse) {
    __shared__ float s_dy[OUT], s_dz2[H];
    __shared__ float s_red[128];
    const int h = threadIdx.x;
    float err = 0.0f;
    for (int u = 0; u < 8; ++u) {
        const int64_t j = blockIdx.x * 8ll + u;
        if (j >= n) break;  // uniform across the block
        const float wj = wt ? wt[j] : 1.0f;
        if (h < OUT) {
            const float d = y[j * OUT + h] - tgt[j * OUT + h];
            err += wj * d * d;
            s_dy[h] = 2.0f * wj * d * inv_count;
            dy[j * OUT + h] = s_dy[h];
        }
        __syncthreads();
        float g = 0.0f;
        for (int o = 0; o < OUT; ++o) g += s_dy[o] * wr[OW3 + o * H + h];
        const float d2 = g * gelu_grad(z2[j * H + h]);
        s_dz2[h] = d2;
        dz2[j * H + h] = d2;
        __syncthreads();
        float g1 = 0.0f;
        for (int o = 0; o < H; ++o) g1 += s_dz2[o] * wr[OW2 + o * H + h];
        dz1[j * H + h] = g1 * gelu_grad(z1[j * H + h]);
        __syncthreads();
    }
    s_red[h] = err;
    __syncthreads();
    for (int w = 64; w > 0; w >>= 1) {
        if (h < w) s_red[h] += s_red[h + w];
        __syncthreads();
    }
    if (h == 0) sse[blockIdx.x] = s_red[0];
}

// dW[o][i] = sum_j dZ[j][o] A[j][i] (A row stride lda), db[o] = sum_j dZ[j][o]:
// one thread per weight, samples in order (deterministic).
__global__ __launch_bounds__(256) void k_weight_grad(const float *__restrict__ dz, int nout, const float *__restrict__ a, int nin, int lda, int64_t n,
                                                     float *__restrict__ gw, float *__restrict__ gb) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < nout * nin) {
        const int o = i / nin, c = i % nin;
        float s = 0.0f;
        for (int64_t j = 0; j < n; ++j) s += dz[j * nout + o] * a[j * lda + c];
        gw[i] = s;
    } else if (i < nout * nin + nout) {
        const int o = i - nout * nin;
        float s = 0.0f;
        for (int64_t j = 0; j < n; ++j) s += dz[j * nout + o];
        gb[o] = s;
    }
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ par, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= NPAR) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    par[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void k_gather_rows(const float *__restrict__ src, int w, const uint32_t *__restrict__ idx, int64_t n,
                                                     float *__restrict__ dst) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= n * w) return;
    dst[i] = src[static_cast<int64_t>(idx[i / w]) * w + i % w];
}

}  // namespace

void Net::set_params(hipStream_t st, const float *par_dev) {
    frag.ensure(NFRAG * 64 * 4);
    wr.ensure(NPAR);
    if (par_dev != par.p) {
        par.ensure(NPAR);
        VG_HIP(hipMemcpyAsync(par.p, par_dev, NPAR * sizeof(float), hipMemcpyDeviceToDevice, st));
    }
    k_pack<<<ceil_div(std::max(NPAR, NFRAG * 64), 256), 256, 0, st>>>(par, frag, wr);
    VG_LAUNCH_CHECK();
}

void Net::forward(hipStream_t st, const float *x, int64_t n, float *y) const {
    if (n <= 0) return;
    k_forward<false><<<ceil_div(n, 64), 256, 0, st>>>(reinterpret_cast<const half4 *>(frag.p), par, x, n, y, nullptr, nullptr, nullptr, nullptr,
                                                      nullptr);
    VG_LAUNCH_CHECK();
}

// loss and gradient of the (weighted) mean squared error over n samples (device
// arrays; w may be null, wsum is the sum of the weights)
float Net::loss_grad(hipStream_t st, const float *x, const float *t, const float *w, double wsum, int64_t n, float *grad) {
    y.ensure(n * OUT);
    x16.ensure(n * 32);
    z1.ensure(n * H), a1.ensure(n * H), z2.ensure(n * H), a2.ensure(n * H);
    dy.ensure(n * OUT), dz1.ensure(n * H), dz2.ensure(n * H);
    const int64_t nb = ceil_div(n, 8);
    sse.ensure(nb);
    k_forward<true><<<ceil_div(n, 64), 256, 0, st>>>(reinterpret_cast<const half4 *>(frag.p), par, x, n, y, x16, z1, a1, z2, a2);
    VG_LAUNCH_CHECK();
    const double wtot = w ? wsum : static_cast<double>(n);
    k_backward_act<<<static_cast<unsigned>(nb), 128, 0, st>>>(y, t, w, n, static_cast<float>(1.0 / (wtot * OUT)), wr, z1, z2, dy, dz1, dz2, sse);
    VG_LAUNCH_CHECK();
    k_weight_grad<<<ceil_div(OUT * H + OUT, 256), 256, 0, st>>>(dy, OUT, a2, H, H, n, grad + OW3, grad + OB3);
    k_weight_grad<<<ceil_div(H * H + H, 256), 256, 0, st>>>(dz2, H, a1, H, H, n, grad + OW2, grad + OB2);
    k_weight_grad<<<ceil_div(H * IN + H, 256), 256, 0, st>>>(dz1, H, x16, IN, 32, n, grad + OW1, grad + OB1);
    VG_LAUNCH_CHECK();
    std::vector<float> h(nb);
    VG_HIP(hipMemcpyAsync(h.data(), sse.p, nb * sizeof(float), hipMemcpyDeviceToHost, st));
    VG_HIP(hipStreamSynchronize(st));
    double s = 0;
    for (float v : h) s += v;
    return static_cast<float>(s / (wtot * OUT));
}

void Net::adam_step(hipStream_t st, const float *grad, float lr, int step) {
    if (m.cap < static_cast<size_t>(NPAR)) {
        m.ensure(NPAR), v.ensure(NPAR);
        VG_HIP(hipMemsetAsync(m.p, 0, NPAR * sizeof(float), st));
        VG_HIP(hipMemsetAsync(v.p, 0, NPAR * sizeof(float), st));
    }
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, static_cast<float>(step)), c2 = 1.0f - std::pow(b2, static_cast<float>(step));
    k_adam<<<ceil_div(NPAR, 256), 256, 0, st>>>(par, grad, m, v, lr, b1, b2, 1e-8f, c1, c2);
    VG_LAUNCH_CHECK();
    set_params(st, par);
}

void gather_rows(hipStream_t st, const float *src, int w, const uint32_t *idx, int64_t n, float *dst) {
    if (n <= 0) return;
    k_gather_rows<<<ceil_div(n * w, 256), 256, 0, st>>>(src, w, idx, n, dst);
    VG_LAUNCH_CHECK();
}

}  // namespace mlp
}  // namespace vg
