// mlp.hpp - the learned closure's network (28 -> 128 -> 128 -> 25, GELU).
#pragma once

#include "common.hpp"

namespace vg {
namespace mlp {

constexpr int IN = 28, H = 128, OUT = 25;
// offsets of the FP32 parameter vector: W1 (H x IN), b1, W2 (H x H), b2, W3 (OUT x H), b3
constexpr int OW1 = 0, OB1 = OW1 + H * IN, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + OUT * H;
constexpr int NPAR = OB3 + OUT;  // 23,449

struct Net {
    DevBuf<float> par;        // FP32 master parameters
    DevBuf<float> wr;         // parameters rounded to FP16 (as FP32): the backward pass's weights
    DevBuf<_Float16> frag;    // FP16 weights in MFMA A-operand order
    // training state
    DevBuf<float> m, v;       // Adam moments
    DevBuf<float> y, x16, z1, a1, z2, a2, dy, dz1, dz2, sse;

    void set_params(hipStream_t st, const float *par_dev);  // copy (device) + pack
    void forward(hipStream_t st, const float *x, int64_t n, float *y) const;
    // (weighted) mean squared error and its gradient; w may be null (all 1), wsum = sum of w
    float loss_grad(hipStream_t st, const float *x, const float *t, const float *w, double wsum, int64_t n, float *grad);
    void adam_step(hipStream_t st, const float *grad, float lr, int step);
};

void gather_rows(hipStream_t st, const float *src, int w, const uint32_t *idx, int64_t n, float *dst);

}  // namespace mlp
}  // namespace vg
