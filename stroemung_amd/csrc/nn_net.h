// The pressure network (host side): a 4-level U-Net with 32 base channels,
// 3x3 convolutions and ReLU, BF16 activations and weights, FP32 accumulation
// and FP32 master weights trained with Adam.
//
//   level 0 (H x W):        conv 32pad->32, conv 32->32      ... conv 96->32, conv 32->32, head 1x1 32->1
//   level 1 (H/2 x W/2):    conv 32->64, conv 64->64         ... conv 192->64, conv 64->64
//   level 2 (H/4 x W/4):    conv 64->128, conv 128->128      ... conv 384->128, conv 128->128
//   level 3 (H/8 x W/8):    conv 128->256, conv 256->256
// down: 2x2 max pooling; up: nearest 2x upsampling concatenated with the skip.
// Input channels: 0 = scaled residual, 1 = fluid flag, 2..31 zero.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <string>
#include <vector>

#include "nn_kernels.h"

struct stm_ctx;

namespace stm {

hipStream_t ctx_stream(stm_ctx* ctx);
int ctx_device(stm_ctx* ctx);

struct Conv {
    int cin, cout;
    int64_t woff, boff;          // offsets into the parameter buffers
    int level;                   // spatial level 0..3
    int x, y;                    // activation indices of input and output
};

struct Act {
    int level, c;
    nn::bf16* v = nullptr;       // value, nb x H_l x W_l x c
    nn::bf16* g = nullptr;       // gradient
};

struct Net {
    stm_ctx* ctx = nullptr;
    uint64_t serial = 0;         // unique per net in the process (graph keys)
    hipStream_t stream = nullptr;
    int H = 0, W = 0, cap = 0;   // spatial size and batch capacity
    std::vector<Conv> convs;
    std::vector<Act> acts;
    int64_t nparam = 0, head_w = 0, head_b = 0;
    float *p32 = nullptr, *g32 = nullptr, *m32 = nullptr, *v32 = nullptr;
    nn::bf16* pb = nullptr;      // bf16 copy of the parameters
    nn::bf16* wflip = nullptr;   // scratch for the backward-input weights
    // FP32 partial sums: one slice per split of a split-K convolution in the
    // forward pass, the per-block sums of the gradient and loss reductions in
    // training (scratch_floats in all, at least splitk_cap)
    float* splitk = nullptr;
    int64_t scratch_floats = 0;
    // the split-K share in floats: two batches of the first level's 32-channel
    // output, four splits of any deeper level's (cout doubles as the sides halve)
    int64_t splitk_cap() const { return (int64_t)2 * cap * H * W * 32; }
    float* y = nullptr;          // head output, cap x H x W (fp32)
    float* gy = nullptr;
    float* target = nullptr;     // training batch: targets and fluid mask
    float* mask = nullptr;
    float* denom = nullptr;
    float* loss = nullptr;       // device scalar
    int adam_step = 0;
    // device-resident dataset of (input, target) pairs
    nn::bf16* dx = nullptr;      // n x H x W x 2
    float* dt = nullptr;         // n x H x W
    int dcap = 0, dsize = 0, dwrite = 0;   // capacity, samples held, next slot (ring)
    int32_t* d_idx = nullptr;
    // timing
    double fwd_ms = 0, train_ms = 0;
    uint64_t fwd_calls = 0, train_steps = 0, fwd_samples = 0, train_samples = 0;
    std::string err;

    int h(int level) const { return H >> level; }
    int w(int level) const { return W >> level; }
    nn::bf16* input() const { return acts[0].v; }
    double flops_forward() const;   // per sample

    int init(stm_ctx* c, int H, int W, int cap, uint64_t seed);
    void release();
    void forward(int nb);          // materialises every activation (training)
    // pooling / upsampling / concatenation fused into staging; live (device
    // int, may be null) limits the work to batch entries [0, *live)
    void forward_infer(int nb, const int* live = nullptr);
    void backward(int nb);
    int train_step(int nb, float lr);
};

}  // namespace stm

struct stm_net {
    stm::Net net;
};

namespace stm {
void ctx_attach_net(stm_ctx* ctx, stm_net* net);
void ctx_detach_net(stm_ctx* ctx, stm_net* net);
}  // namespace stm
