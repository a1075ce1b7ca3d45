// Launchers of the pressure-network kernels (nn_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nn_sizes.h"

namespace stm {
namespace nn {

typedef __bf16 bf16;

// How a convolution reads its input: channels [0, ca) from `a` (directly,
// as the 2x2 max of an image twice the size, or as the nearest upsampling of
// an image half the size), channels [ca, ca + cb) from `b` directly.
enum : int { IN_DIRECT = 0, IN_POOL = 1, IN_UP = 2 };
struct ConvIn {
    const bf16* a;
    int ca, mode;
    const bf16* b;
    int cb;
    const int* live = nullptr;   // device count: images at or beyond it are skipped
    float* acc32 = nullptr;      // FP32 scratch for the partial sums of split K ...
    int64_t acc_cap = 0;         // ... holding this many floats (split s writes [s*n, s*n + n))
    // the 1x1 head fused into a 32-channel convolution (no split K):
    // head_y[pix] = head_b[0] + sum_c out[pix, c] * head_w[c], channels in
    // order, over the bf16-rounded outputs (the operations of nn::head)
    const float* head_w = nullptr;
    const float* head_b = nullptr;
    float* head_y = nullptr;
};

// 3x3 convolution, padding 1, NHWC bf16, fp32 accumulation on the matrix
// cores.  w is [cout][9][cin] bf16; flags 1 = add bias (fp32), 2 = ReLU.
// Writes bf16 `out` and/or fp32 `out32` (either may be null).
// live (device int, may be null): only images [0, *live) are computed
void conv3x3(hipStream_t, const bf16* in, int cin, const bf16* w, const float* bias, bf16* out,
             float* out32, int cout, int nb, int H, int W, int flags, const int* live = nullptr);
// The staged-tile convolution takes every image whose sides are multiples
// of 8; conv3x3_in is that convolution with a described input.
bool conv_tiles_fit(int H, int W);
void conv3x3_in(hipStream_t, const ConvIn& src, const bf16* w, const float* bias, bf16* out,
                float* out32, int cout, int nb, int H, int W, int flags);
// w'[ci][8 - tap][co] = w[co][tap][ci]: the backward-input weights
void weight_flip(hipStream_t, const bf16* w, bf16* wt, int cin, int cout);
// The gradient and loss reductions leave per-block partial sums in `part`
// (device scratch, reused launch after launch on one stream) and add them in
// a fixed order, so a gradient is the same from run to run.  Sizes in floats:
// grad_partials and loss_partials (nn_sizes.h).
// dw32[co][tap][ci] += sum_pix gz[pix][co] * in[pix + off(tap)][ci]
void conv3x3_dw(hipStream_t, const bf16* gz, int cout, const bf16* in, int cin, float* dw32, int nb,
                int H, int W, float* part);
void relu_bwd(hipStream_t, int64_t n, const bf16* y, bf16* g);
// db[c] += sum_pix g[pix][c]
void bias_grad(hipStream_t, const bf16* g, int c, int64_t npix, float* db, float* part);
void maxpool(hipStream_t, const bf16* x, bf16* y, int nb, int H, int W, int c,
             const int* live = nullptr);
// accumulates into gx
void maxpool_bwd(hipStream_t, const bf16* x, const bf16* y, const bf16* gy, bf16* gx, int nb,
                 int H, int W, int c);
// x (nb, H/2, W/2, c) -> channels [c0, c0 + c) of y (nb, H, W, cy)
void upsample(hipStream_t, const bf16* x, bf16* y, int nb, int H, int W, int c, int cy, int c0,
              const int* live = nullptr);
void upsample_bwd(hipStream_t, const bf16* gy, bf16* gx, int nb, int H, int W, int c, int cy,
                  int c0);
void channel_copy(hipStream_t, const bf16* src, int cs, int s0, bf16* dst, int cd, int d0, int c,
                  int64_t npix, int add, const int* live = nullptr, int64_t hw = 1);
void head(hipStream_t, const bf16* x, int c, const float* w, const float* b, float* y,
          int64_t npix, const int* live = nullptr, int64_t hw = 1);
// gx = gy w; dw += sum gy x; db += sum gy
void head_bwd(hipStream_t, const bf16* x, int c, const float* w, const float* gy, bf16* gx,
              float* dw, float* db, int64_t npix, float* part);
// denom[s] = sum m t^2 of sample s; loss_out += the masked, normalised MSE; gy its gradient
void loss(hipStream_t, const float* y, const float* t, const float* m, float* denom, float* gy,
          float* loss_out, int64_t npix_per, int nb, float* part);
void adam(hipStream_t, int64_t n, float* p, const float* g, float* m, float* v, bf16* pb, float lr,
          float b1, float b2, float eps, int step, float clip);
void to_bf16(hipStream_t, int64_t n, const float* x, bf16* y);

}  // namespace nn
}  // namespace stm
