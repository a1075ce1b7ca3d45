// The pressure network: layer plan, parameters, forward / backward / Adam,
// the device-resident training set and the C ABI of the network.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/stroemung_hip.h"
#include "nn_kernels.h"
#include "nn_net.h"

using namespace stm;
using nn::bf16;

namespace {

#define NCHK(call)                                                                     \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            err = std::string(#call ": ") + hipGetErrorString(e_);                     \
            return STM_ERR_HIP;                                                        \
        }                                                                              \
    } while (0)

uint64_t splitmix(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double unif(uint64_t& s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }
double normal(uint64_t& s) {
    const double u1 = unif(s) + 1e-300, u2 = unif(s);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

thread_local std::string g_net_error;

}  // namespace

namespace stm {

double Net::flops_forward() const {
    double f = 0;
    for (const Conv& c : convs) f += 2.0 * 9.0 * c.cin * c.cout * h(c.level) * (double)w(c.level);
    return f + 2.0 * 32.0 * H * (double)W;
}

int Net::init(stm_ctx* c, int H_, int W_, int cap_, uint64_t seed) {
    ctx = c;
    stream = ctx_stream(c);
    H = H_;
    W = W_;
    cap = cap_;
    if (H % 8 || W % 8 || H < 8 || W < 8 || cap < 1) {
        err = "network size must be a multiple of 8 in both directions";
        return STM_ERR_INVALID_ARG;
    }
    // activations: (level, channels), in the order of the plan in nn_net.h
    const int spec[][2] = {{0, 32},  {0, 32}, {0, 32}, {1, 32},  {1, 64},  {1, 64},  {2, 64},
                           {2, 128}, {2, 128}, {3, 128}, {3, 256}, {3, 256}, {2, 384}, {2, 128},
                           {2, 128}, {1, 192}, {1, 64}, {1, 64}, {0, 96},  {0, 32},  {0, 32}};
    for (auto& s : spec) acts.push_back(Act{s[0], s[1]});
    // convolutions: input activation -> output activation
    const int plan[][2] = {{0, 1},   {1, 2},   {3, 4},   {4, 5},   {6, 7},   {7, 8},   {9, 10},
                           {10, 11}, {12, 13}, {13, 14}, {15, 16}, {16, 17}, {18, 19}, {19, 20}};
    int64_t off = 0;
    for (auto& pl : plan) {
        Conv cv;
        cv.x = pl[0];
        cv.y = pl[1];
        cv.cin = acts[cv.x].c;
        cv.cout = acts[cv.y].c;
        cv.level = acts[cv.y].level;
        cv.woff = off;
        off += (int64_t)cv.cout * 9 * cv.cin;
        cv.boff = off;
        off += cv.cout;
        off = (off + 7) / 8 * 8;   // keep every weight block 16-byte aligned in bf16
        convs.push_back(cv);
    }
    head_w = off;
    off += 32;
    head_b = off;
    off += 8;
    nparam = off;

    NCHK(hipSetDevice(ctx_device(c)));
    const size_t pbytes = sizeof(float) * nparam;
    NCHK(hipMalloc(&p32, pbytes));
    NCHK(hipMalloc(&g32, pbytes));
    NCHK(hipMalloc(&m32, pbytes));
    NCHK(hipMalloc(&v32, pbytes));
    NCHK(hipMalloc(&pb, sizeof(bf16) * nparam));
    NCHK(hipMemset(m32, 0, pbytes));
    NCHK(hipMemset(v32, 0, pbytes));
    NCHK(hipMemset(g32, 0, pbytes));
    int64_t maxw = 0;
    for (auto& cv : convs) maxw = std::max(maxw, (int64_t)cv.cout * 9 * cv.cin);
    NCHK(hipMalloc(&wflip, sizeof(bf16) * maxw));
    // the largest conv output is level 0's: cap x H x W x 32 (levels below have
    // a quarter of the pixels and twice the channels)
    scratch_floats = std::max(splitk_cap(), nn::loss_partials((int64_t)H * W, cap, 32));
    for (auto& cv : convs)
        scratch_floats = std::max(
            scratch_floats,
            nn::grad_partials(cv.cin, cv.cout, (int64_t)cap * h(cv.level) * w(cv.level)));
    NCHK(hipMalloc(&splitk, sizeof(float) * (size_t)scratch_floats));
    NCHK(hipMemset(splitk, 0, sizeof(float) * (size_t)scratch_floats));
    for (auto& a : acts) {
        const size_t n = (size_t)cap * h(a.level) * w(a.level) * a.c;
        NCHK(hipMalloc(&a.v, sizeof(bf16) * n));
        NCHK(hipMalloc(&a.g, sizeof(bf16) * n));
        NCHK(hipMemset(a.v, 0, sizeof(bf16) * n));
    }
    const size_t npx = (size_t)cap * H * W;
    NCHK(hipMalloc(&y, sizeof(float) * npx));
    NCHK(hipMalloc(&gy, sizeof(float) * npx));
    NCHK(hipMalloc(&target, sizeof(float) * npx));
    NCHK(hipMalloc(&mask, sizeof(float) * npx));
    NCHK(hipMalloc(&denom, sizeof(float) * cap));
    NCHK(hipMalloc(&loss, sizeof(float)));
    NCHK(hipMalloc(&d_idx, sizeof(int32_t) * cap));

    // He initialisation from a SplitMix64 stream; the first layer sees two
    // real input channels.
    std::vector<float> p(nparam, 0.f);
    uint64_t s = seed;
    for (size_t k = 0; k < convs.size(); k++) {
        const Conv& cv = convs[k];
        const int fan = 9 * (k == 0 ? 2 : cv.cin);
        const double sd = std::sqrt(2.0 / fan);
        for (int64_t i = 0; i < (int64_t)cv.cout * 9 * cv.cin; i++) {
            const int ci = (int)(i % cv.cin);
            p[cv.woff + i] = (k == 0 && ci >= 2) ? 0.f : (float)(sd * normal(s));
        }
    }
    for (int i = 0; i < 32; i++) p[head_w + i] = (float)(0.1 * std::sqrt(1.0 / 32) * normal(s));
    NCHK(hipMemcpy(p32, p.data(), pbytes, hipMemcpyHostToDevice));
    nn::to_bf16(stream, nparam, p32, pb);
    NCHK(hipStreamSynchronize(stream));
    return STM_OK;
}

void Net::release() {
    void* bufs[] = {p32, g32, m32, v32, pb, wflip, splitk, y, gy, target, mask, denom, loss,
                    d_idx, dx, dt};
    for (void* b : bufs)
        if (b) hipFree(b);
    for (auto& a : acts) {
        if (a.v) hipFree(a.v);
        if (a.g) hipFree(a.g);
    }
    acts.clear();
}

void Net::forward(int nb) {
    auto conv = [&](int k) {
        const Conv& c = convs[k];
        nn::conv3x3(stream, acts[c.x].v, c.cin, pb + c.woff, p32 + c.boff, acts[c.y].v, nullptr,
                    c.cout, nb, h(c.level), w(c.level), 3);
    };
    auto npix = [&](int level) { return (int64_t)nb * h(level) * w(level); };
    conv(0);
    conv(1);
    nn::maxpool(stream, acts[2].v, acts[3].v, nb, h(0), w(0), 32);
    conv(2);
    conv(3);
    nn::maxpool(stream, acts[5].v, acts[6].v, nb, h(1), w(1), 64);
    conv(4);
    conv(5);
    nn::maxpool(stream, acts[8].v, acts[9].v, nb, h(2), w(2), 128);
    conv(6);
    conv(7);
    nn::upsample(stream, acts[11].v, acts[12].v, nb, h(2), w(2), 256, 384, 0);
    nn::channel_copy(stream, acts[8].v, 128, 0, acts[12].v, 384, 256, 128, npix(2), 0);
    conv(8);
    conv(9);
    nn::upsample(stream, acts[14].v, acts[15].v, nb, h(1), w(1), 128, 192, 0);
    nn::channel_copy(stream, acts[5].v, 64, 0, acts[15].v, 192, 128, 64, npix(1), 0);
    conv(10);
    conv(11);
    nn::upsample(stream, acts[17].v, acts[18].v, nb, h(0), w(0), 64, 96, 0);
    nn::channel_copy(stream, acts[2].v, 32, 0, acts[18].v, 96, 64, 32, npix(0), 0);
    conv(12);
    conv(13);
    nn::head(stream, acts[20].v, 32, p32 + head_w, p32 + head_b, y, npix(0));
}

// Inference: the same network with pooling, upsampling and concatenation
// done while each convolution stages its input, on every level whose sides
// are multiples of 8; other levels materialise those inputs as in forward().
void Net::forward_infer(int nb, const int* live) {
    using nn::ConvIn;
    using nn::conv_tiles_fit;
    auto npix = [&](int level) { return (int64_t)nb * h(level) * w(level); };
    auto run = [&](int k, const ConvIn& in) {
        const Conv& c = convs[k];
        ConvIn src = in;
        src.live = live;
        src.acc32 = splitk;
        src.acc_cap = splitk_cap();
        nn::conv3x3_in(stream, src, pb + c.woff, p32 + c.boff, acts[c.y].v, nullptr, c.cout, nb,
                   h(c.level), w(c.level), 3);
    };
    auto plain = [&](int k) {
        const Conv& c = convs[k];
        if (conv_tiles_fit(h(c.level), w(c.level)))
            run(k, ConvIn{acts[c.x].v, c.cin, nn::IN_DIRECT, nullptr, 0});
        else
            nn::conv3x3(stream, acts[c.x].v, c.cin, pb + c.woff, p32 + c.boff, acts[c.y].v,
                        nullptr, c.cout, nb, h(c.level), w(c.level), 3, live);
    };
    // conv k whose input is the pooled activation `src` (level - 1)
    auto pooled = [&](int k, int src, int pooled_act) {
        const Conv& c = convs[k];
        if (conv_tiles_fit(h(c.level), w(c.level))) {
            run(k, ConvIn{acts[src].v, acts[src].c, nn::IN_POOL, nullptr, 0});
        } else {
            nn::maxpool(stream, acts[src].v, acts[pooled_act].v, nb, h(c.level - 1),
                        w(c.level - 1), acts[src].c, live);
            plain(k);
        }
    };
    // conv k whose input is [up(low), skip] materialised in `cat`
    auto joined = [&](int k, int low, int skip, int cat) {
        const Conv& c = convs[k];
        const int cl = acts[low].c, cs = acts[skip].c;
        if (conv_tiles_fit(h(c.level), w(c.level))) {
            run(k, ConvIn{acts[low].v, cl, nn::IN_UP, acts[skip].v, cs});
        } else {
            nn::upsample(stream, acts[low].v, acts[cat].v, nb, h(c.level), w(c.level), cl, cl + cs,
                         0, live);
            nn::channel_copy(stream, acts[skip].v, cs, 0, acts[cat].v, cl + cs, cl, cs,
                             npix(c.level), 0, live, (int64_t)h(c.level) * w(c.level));
            plain(k);
        }
    };
    plain(0);
    plain(1);
    pooled(2, 2, 3);
    plain(3);
    pooled(4, 5, 6);
    plain(5);
    pooled(6, 8, 9);
    plain(7);
    joined(8, 11, 8, 12);
    plain(9);
    joined(10, 14, 5, 15);
    plain(11);
    joined(12, 17, 2, 18);
    if (conv_tiles_fit(h(0), w(0))) {   // the head in the last convolution's epilogue
        ConvIn in{acts[convs[13].x].v, convs[13].cin, nn::IN_DIRECT, nullptr, 0};
        in.head_w = p32 + head_w;
        in.head_b = p32 + head_b;
        in.head_y = y;
        run(13, in);
    } else {
        plain(13);
        nn::head(stream, acts[20].v, 32, p32 + head_w, p32 + head_b, y, npix(0), live,
                 (int64_t)h(0) * w(0));
    }
}

void Net::backward(int nb) {
    auto npix = [&](int level) { return (int64_t)nb * h(level) * w(level); };
    // gradient through conv k (ReLU included); the gradient of its output is in acts[y].g
    auto conv_bwd = [&](int k, bool need_gx) {
        const Conv& c = convs[k];
        const int hh = h(c.level), ww = w(c.level);
        const int64_t np = npix(c.level);
        nn::relu_bwd(stream, np * c.cout, acts[c.y].v, acts[c.y].g);
        nn::conv3x3_dw(stream, acts[c.y].g, c.cout, acts[c.x].v, c.cin, g32 + c.woff, nb, hh, ww,
                       splitk);
        nn::bias_grad(stream, acts[c.y].g, c.cout, np, g32 + c.boff, splitk);
        if (need_gx) {
            nn::weight_flip(stream, pb + c.woff, wflip, c.cin, c.cout);
            nn::conv3x3(stream, acts[c.y].g, c.cout, wflip, nullptr, acts[c.x].g, nullptr, c.cin,
                        nb, hh, ww, 0);
        }
    };
    nn::head_bwd(stream, acts[20].v, 32, p32 + head_w, gy, acts[20].g, g32 + head_w,
                 g32 + head_b, npix(0), splitk);
    conv_bwd(13, true);
    conv_bwd(12, true);
    nn::upsample_bwd(stream, acts[18].g, acts[17].g, nb, h(0), w(0), 64, 96, 0);
    nn::channel_copy(stream, acts[18].g, 96, 64, acts[2].g, 32, 0, 32, npix(0), 0);
    conv_bwd(11, true);
    conv_bwd(10, true);
    nn::upsample_bwd(stream, acts[15].g, acts[14].g, nb, h(1), w(1), 128, 192, 0);
    nn::channel_copy(stream, acts[15].g, 192, 128, acts[5].g, 64, 0, 64, npix(1), 0);
    conv_bwd(9, true);
    conv_bwd(8, true);
    nn::upsample_bwd(stream, acts[12].g, acts[11].g, nb, h(2), w(2), 256, 384, 0);
    nn::channel_copy(stream, acts[12].g, 384, 256, acts[8].g, 128, 0, 128, npix(2), 0);
    conv_bwd(7, true);
    conv_bwd(6, true);
    nn::maxpool_bwd(stream, acts[8].v, acts[9].v, acts[9].g, acts[8].g, nb, h(2), w(2), 128);
    conv_bwd(5, true);
    conv_bwd(4, true);
    nn::maxpool_bwd(stream, acts[5].v, acts[6].v, acts[6].g, acts[5].g, nb, h(1), w(1), 64);
    conv_bwd(3, true);
    conv_bwd(2, true);
    nn::maxpool_bwd(stream, acts[2].v, acts[3].v, acts[3].g, acts[2].g, nb, h(0), w(0), 32);
    conv_bwd(1, true);
    conv_bwd(0, false);
}

// One Adam step on the batch already in acts[0].v / target / mask.
int Net::train_step(int nb, float lr) {
    NCHK(hipMemsetAsync(g32, 0, sizeof(float) * nparam, stream));
    forward(nb);
    nn::loss(stream, y, target, mask, denom, gy, loss, (int64_t)H * W, nb, splitk);
    backward(nb);
    adam_step++;
    nn::adam(stream, nparam, p32, g32, m32, v32, pb, lr, 0.9f, 0.999f, 1e-8f, adam_step, 1.0f);
    return STM_OK;
}

}  // namespace stm

// ------------------------------------------------------------ kernels for host batches

namespace {

// x (nb, H, W, 2) fp32 -> network input (nb, H, W, 32) bf16; t, mask from x[..., 1]
__global__ void k_pack_host(const float* __restrict__ x, const float* __restrict__ t, int64_t npix,
                            bf16* __restrict__ a0, float* __restrict__ tt, float* __restrict__ m)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    for (int c = 0; c < 32; c++) a0[p * 32 + c] = (bf16)(c < 2 ? x[p * 2 + c] : 0.f);
    if (t) {
        tt[p] = t[p];
        m[p] = x[p * 2 + 1] > 0.f ? 1.f : 0.f;
    }
}

// dataset batch: samples idx[0..nb) of (dx, dt) -> network input, target, mask
__global__ void k_pack_batch(const bf16* __restrict__ dx, const float* __restrict__ dt,
                             const int32_t* __restrict__ idx, int64_t hw, int nb,
                             bf16* __restrict__ a0, float* __restrict__ tt, float* __restrict__ m)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= hw * nb) return;
    const int s = (int)(k / hw);
    const int64_t p = k - (int64_t)s * hw, src = (int64_t)idx[s] * hw + p;
    const bf16 x0 = dx[src * 2], x1 = dx[src * 2 + 1];
    bf16* o = a0 + k * 32;
    o[0] = x0;
    o[1] = x1;
    for (int c = 2; c < 32; c++) o[c] = (bf16)0.f;
    tt[k] = dt[src];
    m[k] = (float)x1 > 0.f ? 1.f : 0.f;
}

int fail_net(stm_net* n, int code, const char* msg) {
    g_net_error = msg;
    if (n) n->net.err = msg;
    return code;
}

}  // namespace

// =================================================================== C ABI

extern "C" {

int stm_net_create(stm_ctx* ctx, uint32_t height, uint32_t width, uint32_t batch, uint64_t seed,
                   stm_net** out) {
    if (!ctx || !out) return fail_net(nullptr, STM_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    stm_net* n = new stm_net();
    const int rc = n->net.init(ctx, (int)height, (int)width, (int)batch, seed);
    if (rc) {
        g_net_error = n->net.err;
        n->net.release();
        delete n;
        return rc;
    }
    static std::atomic<uint64_t> next_serial{1};
    n->net.serial = next_serial++;
    stm::ctx_attach_net(ctx, n);
    *out = n;
    return STM_OK;
}

void stm_net_destroy(stm_net* n) {
    if (!n) return;
    if (n->net.ctx) stm::ctx_detach_net(n->net.ctx, n);
    hipStreamSynchronize(n->net.stream);
    n->net.release();
    delete n;
}

const char* stm_net_last_error(const stm_net* n) {
    return n ? n->net.err.c_str() : g_net_error.c_str();
}

int stm_net_param_count(stm_net* n, uint64_t* count) {
    if (!n || !count) return fail_net(n, STM_ERR_INVALID_ARG, "null argument");
    *count = (uint64_t)n->net.nparam;
    return STM_OK;
}

int stm_net_get_params(stm_net* n, float* out) {
    if (!n || !out) return fail_net(n, STM_ERR_INVALID_ARG, "null argument");
    std::string& err = n->net.err;
    NCHK(hipMemcpyAsync(out, n->net.p32, sizeof(float) * n->net.nparam, hipMemcpyDeviceToHost,
                        n->net.stream));
    NCHK(hipStreamSynchronize(n->net.stream));
    return STM_OK;
}

int stm_net_set_params(stm_net* n, const float* in) {
    if (!n || !in) return fail_net(n, STM_ERR_INVALID_ARG, "null argument");
    std::string& err = n->net.err;
    Net& net = n->net;
    NCHK(hipMemcpyAsync(net.p32, in, sizeof(float) * net.nparam, hipMemcpyHostToDevice, net.stream));
    nn::to_bf16(net.stream, net.nparam, net.p32, net.pb);
    NCHK(hipStreamSynchronize(net.stream));
    return STM_OK;
}

int stm_net_forward(stm_net* n, uint32_t nb, const float* x, float* y) {
    if (!n || !x || !y || nb == 0) return fail_net(n, STM_ERR_INVALID_ARG, "invalid argument");
    Net& net = n->net;
    std::string& err = net.err;
    if ((int)nb > net.cap) return fail_net(n, STM_ERR_INVALID_ARG, "batch larger than capacity");
    const int64_t npix = (int64_t)nb * net.H * net.W;
    float* dxh = nullptr;
    NCHK(hipMalloc(&dxh, sizeof(float) * npix * 2));
    hipMemcpyAsync(dxh, x, sizeof(float) * npix * 2, hipMemcpyHostToDevice, net.stream);
    k_pack_host<<<(unsigned)((npix + 255) / 256), 256, 0, net.stream>>>(dxh, nullptr, npix,
                                                                       net.input(), nullptr,
                                                                       nullptr);
    net.forward_infer((int)nb);
    hipMemcpyAsync(y, net.y, sizeof(float) * npix, hipMemcpyDeviceToHost, net.stream);
    hipError_t e = hipStreamSynchronize(net.stream);
    hipFree(dxh);
    if (e != hipSuccess) return fail_net(n, STM_ERR_HIP, hipGetErrorString(e));
    return STM_OK;
}

int stm_net_gradients(stm_net* n, uint32_t nb, const float* x, const float* t, float* loss,
                      float* grads) {
    if (!n || !x || !t || !loss || !grads || nb == 0)
        return fail_net(n, STM_ERR_INVALID_ARG, "invalid argument");
    Net& net = n->net;
    std::string& err = net.err;
    if ((int)nb > net.cap) return fail_net(n, STM_ERR_INVALID_ARG, "batch larger than capacity");
    const int64_t npix = (int64_t)nb * net.H * net.W;
    float *dxh = nullptr, *dth = nullptr;
    NCHK(hipMalloc(&dxh, sizeof(float) * npix * 2));
    NCHK(hipMalloc(&dth, sizeof(float) * npix));
    hipMemcpyAsync(dxh, x, sizeof(float) * npix * 2, hipMemcpyHostToDevice, net.stream);
    hipMemcpyAsync(dth, t, sizeof(float) * npix, hipMemcpyHostToDevice, net.stream);
    k_pack_host<<<(unsigned)((npix + 255) / 256), 256, 0, net.stream>>>(dxh, dth, npix,
                                                                       net.input(), net.target,
                                                                       net.mask);
    hipMemsetAsync(net.g32, 0, sizeof(float) * net.nparam, net.stream);
    hipMemsetAsync(net.loss, 0, sizeof(float), net.stream);
    net.forward((int)nb);
    nn::loss(net.stream, net.y, net.target, net.mask, net.denom, net.gy, net.loss,
             (int64_t)net.H * net.W, (int)nb, net.splitk);
    net.backward((int)nb);
    hipMemcpyAsync(loss, net.loss, sizeof(float), hipMemcpyDeviceToHost, net.stream);
    hipMemcpyAsync(grads, net.g32, sizeof(float) * net.nparam, hipMemcpyDeviceToHost, net.stream);
    hipError_t e = hipStreamSynchronize(net.stream);
    hipFree(dxh);
    hipFree(dth);
    if (e != hipSuccess) return fail_net(n, STM_ERR_HIP, hipGetErrorString(e));
    return STM_OK;
}

int stm_net_dataset_reserve(stm_net* n, uint32_t capacity) {
    if (!n) return fail_net(n, STM_ERR_INVALID_ARG, "null net");
    Net& net = n->net;
    std::string& err = net.err;
    if (net.dx) hipFree(net.dx);
    if (net.dt) hipFree(net.dt);
    net.dx = nullptr;
    net.dt = nullptr;
    const size_t hw = (size_t)net.H * net.W;
    NCHK(hipMalloc(&net.dx, sizeof(bf16) * hw * 2 * capacity));
    NCHK(hipMalloc(&net.dt, sizeof(float) * hw * capacity));
    net.dcap = (int)capacity;
    net.dsize = 0;
    net.dwrite = 0;
    return STM_OK;
}

int stm_net_dataset_size(stm_net* n, uint32_t* size) {
    if (!n || !size) return fail_net(n, STM_ERR_INVALID_ARG, "null argument");
    *size = (uint32_t)n->net.dsize;
    return STM_OK;
}

int stm_net_dataset_get(stm_net* n, uint32_t index, float* x, float* t) {
    if (!n || !x || !t) return fail_net(n, STM_ERR_INVALID_ARG, "null argument");
    Net& net = n->net;
    std::string& err = net.err;
    if ((int)index >= net.dsize) return fail_net(n, STM_ERR_INVALID_ARG, "sample out of range");
    const size_t hw = (size_t)net.H * net.W;
    std::vector<bf16> xb(hw * 2);
    NCHK(hipMemcpyAsync(xb.data(), net.dx + hw * 2 * index, sizeof(bf16) * hw * 2,
                        hipMemcpyDeviceToHost, net.stream));
    NCHK(hipMemcpyAsync(t, net.dt + hw * index, sizeof(float) * hw, hipMemcpyDeviceToHost,
                        net.stream));
    NCHK(hipStreamSynchronize(net.stream));
    for (size_t k = 0; k < hw * 2; k++) x[k] = (float)xb[k];
    return STM_OK;
}

int stm_net_train(stm_net* n, uint32_t steps, uint32_t batch, float lr, uint64_t seed,
                  float* last_loss) {
    if (!n || steps == 0 || batch == 0) return fail_net(n, STM_ERR_INVALID_ARG, "invalid argument");
    Net& net = n->net;
    std::string& err = net.err;
    if ((int)batch > net.cap) return fail_net(n, STM_ERR_INVALID_ARG, "batch larger than capacity");
    if (net.dsize < (int)batch) return fail_net(n, STM_ERR_STATE, "dataset smaller than a batch");
    const int64_t hw = (int64_t)net.H * net.W;
    uint64_t s = seed;
    std::vector<int32_t> idx(batch);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, net.stream);
    float lsum = 0.f;
    int lcount = 0;
    for (uint32_t it = 0; it < steps; it++) {
        for (uint32_t b = 0; b < batch; b++) idx[b] = (int32_t)(splitmix(s) % (uint64_t)net.dsize);
        hipMemcpyAsync(net.d_idx, idx.data(), sizeof(int32_t) * batch, hipMemcpyHostToDevice,
                       net.stream);
        k_pack_batch<<<(unsigned)((hw * batch + 255) / 256), 256, 0, net.stream>>>(
            net.dx, net.dt, net.d_idx, hw, (int)batch, net.input(), net.target, net.mask);
        hipMemsetAsync(net.loss, 0, sizeof(float), net.stream);
        net.train_step((int)batch, lr);
        if (it + 8 >= steps) {   // average the loss over the last steps
            float l = 0.f;
            hipMemcpyAsync(&l, net.loss, sizeof(float), hipMemcpyDeviceToHost, net.stream);
            hipStreamSynchronize(net.stream);
            lsum += l;
            lcount++;
        }
    }
    hipEventRecord(e1, net.stream);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (e != hipSuccess) return fail_net(n, STM_ERR_HIP, hipGetErrorString(e));
    net.train_ms += ms;
    net.train_steps += steps;
    net.train_samples += (uint64_t)steps * batch;
    if (last_loss) *last_loss = lcount ? lsum / lcount : 0.f;
    return STM_OK;
}

int stm_net_benchmark(stm_net* n, uint32_t nb, uint32_t reps, double* fwd_ms, double* train_ms) {
    if (!n || nb == 0 || reps == 0) return fail_net(n, STM_ERR_INVALID_ARG, "invalid argument");
    Net& net = n->net;
    std::string& err = net.err;
    if ((int)nb > net.cap) return fail_net(n, STM_ERR_INVALID_ARG, "batch larger than capacity");
    hipEvent_t e0, e1, e2;
    NCHK(hipEventCreate(&e0));
    NCHK(hipEventCreate(&e1));
    NCHK(hipEventCreate(&e2));
    // the input is whatever the buffers hold: the matrix-core time does not
    // depend on the values
    net.forward_infer((int)nb);   // warm-up
    NCHK(hipEventRecord(e0, net.stream));
    for (uint32_t r = 0; r < reps; r++) net.forward_infer((int)nb);
    NCHK(hipEventRecord(e1, net.stream));
    if (train_ms) {
        NCHK(hipMemsetAsync(net.target, 0, sizeof(float) * (size_t)nb * net.H * net.W, net.stream));
        NCHK(hipMemsetAsync(net.mask, 0, sizeof(float) * (size_t)nb * net.H * net.W, net.stream));
        // lr = 0 keeps the parameters; Adam's moments change, which timing ignores
        for (uint32_t r = 0; r < reps; r++) net.train_step((int)nb, 0.f);
    }
    NCHK(hipEventRecord(e2, net.stream));
    NCHK(hipEventSynchronize(e2));
    float a = 0.f, b = 0.f;
    NCHK(hipEventElapsedTime(&a, e0, e1));
    NCHK(hipEventElapsedTime(&b, e1, e2));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    if (fwd_ms) *fwd_ms = a / reps;
    if (train_ms) *train_ms = b / reps;
    return STM_OK;
}

int stm_net_stats(stm_net* n, stm_net_stats_t* out) {
    if (!n || !out) return fail_net(n, STM_ERR_INVALID_ARG, "null argument");
    const Net& net = n->net;
    stm_net_stats_t o{};
    o.forward_ms = net.fwd_ms;
    o.forward_calls = net.fwd_calls;
    o.forward_samples = net.fwd_samples;
    o.train_ms = net.train_ms;
    o.train_steps = net.train_steps;
    o.train_samples = net.train_samples;
    o.flops_forward_per_sample = net.flops_forward();
    o.params = (uint64_t)net.nparam;
    o.dataset_size = (uint32_t)net.dsize;
    *out = o;
    return STM_OK;
}

}  // extern "C"
