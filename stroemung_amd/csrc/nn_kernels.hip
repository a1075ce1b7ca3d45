// gfx950 kernels of the pressure network: a U-Net of 3x3 convolutions in
// BF16 with FP32 accumulation on the matrix cores.
//
// Tensors are NHWC (channels contiguous), bf16, with channel counts that are
// multiples of 32.  A 3x3 convolution with padding 1 is an implicit GEMM:
//   forward      out[pix, co]  = sum_{tap, ci} in[pix + off(tap), ci] * w[co, tap, ci]
//   backward-in  gin[pix, ci]  = sum_{tap, co} gz[pix - off(tap), co] * w[co, tap, ci]
//                              = forward with w'[ci, 8 - tap, co] (rotated, transposed)
//   backward-w   dw[co, tap, ci] = sum_pix gz[pix, co] * in[pix + off(tap), ci]
// v_mfma_f32_16x16x32_bf16 operands (checked on the device): lane l holds row
// l % 16 of A and column l % 16 of B, each with k = 8 * (l / 16) + 0..7, and
// column l % 16, rows 4 * (l / 16) + 0..3 of the f32 result.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "nn_kernels.h"

namespace stm {
namespace nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 load8(const bf16* p) {
    return *reinterpret_cast<const bf16x8*>(p);
}
__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
    for (int i = 0; i < 8; i++) z[i] = (bf16)0.0f;
    return z;
}

// ------------------------------------------------------------ convolution

// Block: 256 threads = 4 waves in a 2 x 2 grid; block tile BM pixels x 64
// output channels, each wave (BM/2) x 32 as (BM/32) x 2 MFMA tiles.
// flags: 1 = add bias, 2 = ReLU.
template <int BM>
__global__ void __launch_bounds__(256)
k_conv3x3(const bf16* __restrict__ in, int cin, const bf16* __restrict__ w,
          const float* __restrict__ bias, bf16* __restrict__ out, float* __restrict__ out32,
          int cout, int nb, int H, int W, int flags, const int* __restrict__ live)
{
    constexpr int MT = BM / 32;          // M tiles per wave
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int npix = (live ? min(nb, *live) : nb) * H * W;
    if ((int)(blockIdx.x * BM) >= npix) return;   // the whole block is beyond the live images
    const int m_base = blockIdx.x * BM + wm * (BM / 2);
    const int n_base = blockIdx.y * 64 + wn * 32;
    if (n_base >= cout) return;          // 32-channel layers use half the tile (no barriers here)
    const int row = lane & 15, kq = (lane >> 4) * 8;

    // the pixels this lane loads for each M tile
    int pb[MT], ph[MT], pw[MT];
    bool pv[MT];
#pragma unroll
    for (int t = 0; t < MT; t++) {
        const int pix = m_base + t * 16 + row;
        pv[t] = pix < npix;
        const int pp = pv[t] ? pix : 0;
        pb[t] = pp / (H * W);
        const int r = pp - pb[t] * H * W;
        ph[t] = r / W;
        pw[t] = r - ph[t] * W;
    }
    f32x4 acc[MT][2];
#pragma unroll
    for (int t = 0; t < MT; t++)
        for (int u = 0; u < 2; u++) acc[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int tap = 0; tap < 9; tap++) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const bf16* arow[MT];
        bool aok[MT];
#pragma unroll
        for (int t = 0; t < MT; t++) {
            const int hh = ph[t] + dy, ww = pw[t] + dx;
            aok[t] = pv[t] && hh >= 0 && hh < H && ww >= 0 && ww < W;
            arow[t] = in + ((int64_t)(pb[t] * H + (aok[t] ? hh : 0)) * W + (aok[t] ? ww : 0)) * cin;
        }
        const bf16* brow[2];
#pragma unroll
        for (int u = 0; u < 2; u++) brow[u] = w + ((int64_t)(n_base + u * 16 + row) * 9 + tap) * cin;
        for (int c0 = 0; c0 < cin; c0 += 32) {
            bf16x8 a[MT], b[2];
#pragma unroll
            for (int t = 0; t < MT; t++) a[t] = aok[t] ? load8(arow[t] + c0 + kq) : zero8();
#pragma unroll
            for (int u = 0; u < 2; u++) b[u] = load8(brow[u] + c0 + kq);
#pragma unroll
            for (int t = 0; t < MT; t++)
#pragma unroll
                for (int u = 0; u < 2; u++)
                    acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], b[u], acc[t][u], 0, 0, 0);
        }
    }
    // epilogue: lane holds column n = l % 16, rows 4 * (l / 16) + i
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int n = n_base + u * 16 + row;
        const float bv = (flags & 1) ? bias[n] : 0.f;
#pragma unroll
        for (int t = 0; t < MT; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int pix = m_base + t * 16 + (lane >> 4) * 4 + i;
                if (pix >= npix) continue;
                float v = acc[t][u][i] + bv;
                if (flags & 2) v = v > 0.f ? v : 0.f;
                if (out) out[(int64_t)pix * cout + n] = (bf16)v;
                if (out32) out32[(int64_t)pix * cout + n] = v;
            }
    }
}

// LDS-tiled variant.  A workgroup owns a TH x TW pixel tile of one image
// (TH * TW = 64 or 128) and NW = 32 * WN output channels; for every
// 32-channel slice of the input it stages in LDS the (TH + 2) x (TW + 2)
// patch and the slice's weights of its NW channels, and serves all nine taps
// and all four waves from them: each input value is read from memory once per
// slice and channel block, each weight once per workgroup.  Waves: (4 / WN) in
// M x WN in N; a wave owns MT M tiles of 16 pixels (RM = 16 / TW image rows
// each) on consecutive rows, and 32 output channels (two N tiles).  The
// MFMAs take the weights as their first operand, so each result holds
// channels as rows and pixels as columns, and a lane ends with four
// consecutive channels of one pixel to store.
//
// The MFMA operand of tap (dy, dx) for M tile i is the patch block that
// starts RM * i + dy rows below the wave's first row, shifted by dx: the
// wave walks the distinct starts s once per dx and feeds each fragment to
// every (i, dy) with RM * i + dy = s, so a fragment is read from LDS once for
// up to three taps.  Each output's taps are added in the order dx-major,
// dy ascending (0, 3, 6, 1, 4, 7, 2, 5, 8), slice after slice.
//
// The input is described by ConvIn: channels [0, ca) come from `a` read
// directly, as the 2x2 max of an image twice the size (the U-Net's pooling)
// or as the nearest upsampling of an image half the size; channels
// [ca, ca + cb) come from `b` read directly (the skip half of a
// concatenation).  Pooling, upsampling and concatenation thus happen while
// staging, without materialising their outputs.
//
// With PIPE (chosen for workgroups with at least three input slices, in
// launches of at most two workgroups per CU) patch and weights are
// double-buffered in LDS and the loads run two slices ahead in registers, so
// one barrier per slice suffices and the load latency hides behind two
// slices of matrix work.  The other launches (few slices, or enough
// workgroups that occupancy hides the latency) stage each slice and then
// multiply, in half the LDS.
template <int TW, int TH, int WN, bool POOL, bool PIPE>
__global__ void __launch_bounds__(256)
k_conv3x3_tile(ConvIn src, const bf16* __restrict__ w, const float* __restrict__ bias,
               bf16* __restrict__ out, float* __restrict__ out32, int cout, int nb, int H, int W,
               int flags)
{
    constexpr int PH = TH + 2, PW = TW + 2;
    constexpr int CS = 40;                       // padded channel stride of a staged pixel
    constexpr int CSW = 40;                      // ... of a staged (channel, tap) weight row
    constexpr int NW = 32 * WN;                  // output channels of the workgroup
    constexpr int WM = 4 / WN;                   // waves along M
    constexpr int RM = 16 / TW;                  // image rows per M tile
    constexpr int MT = TH * TW / WM / 16;        // M tiles per wave
    constexpr int NS = RM * (MT - 1) + 3;        // distinct fragment starts per dx
    constexpr int NPV = PH * PW * 4;             // 16-byte vectors of one staged slice
    constexpr int NV = (NPV + 255) / 256;        // of them per thread
    constexpr int NWVEC = NW * 9 * 4;            // 16-byte vectors of one slice's weights
    constexpr int NWV = (NWVEC + 255) / 256;
    constexpr int NB = PIPE ? 2 : 1;
    static_assert(TW == 8 || TW == 16, "tile width");
    static_assert(MT >= 1 && TH * TW % (WM * 16) == 0, "tile shape");
    __shared__ __attribute__((aligned(16))) bf16 patch[NB][PH * PW * CS];
    __shared__ __attribute__((aligned(16))) bf16 wl[NB][NW * 9 * CSW];

    const int cin = src.ca + src.cb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int tiles_w = (W + TW - 1) / TW, tiles_h = (H + TH - 1) / TH;
    int t = blockIdx.x;
    const int tw0 = (t % tiles_w) * TW;
    t /= tiles_w;
    const int th0 = (t % tiles_h) * TH;
    const int b = t / tiles_h;
    if (src.live && b >= *src.live) return;      // uniform over the block, before any barrier
    const int nb0 = blockIdx.y * NW;
    const int n_base = nb0 + wn * 32;
    const int row = lane & 15, kq = (lane >> 4) * 8;

    // this lane's LDS offsets: its pixel of the wave's first M tile at tap
    // (0, 0), and its weight row (channel n_base + row) at tap 0
    const int aoff = ((wm * MT * RM + row / TW) * PW + row % TW) * CS + kq;
    const int boff = (wn * 32 + row) * 9 * CSW + kq;
    f32x4 acc[MT][2];
#pragma unroll
    for (int i = 0; i < MT; i++)
        for (int u = 0; u < 2; u++) acc[i][u] = f32x4{0.f, 0.f, 0.f, 0.f};

    // split K: block z of gridDim.z takes its share of the 32-channel slices
    const int nsl = cin / 32, z = blockIdx.z, nz = gridDim.z;
    const int sl0 = z * nsl / nz, sl1 = (z + 1) * nsl / nz;

    // Staging: the loads of a slice are issued unconditionally (addresses
    // clamped into the image, out-of-image vectors zeroed when stored), with
    // no branch between them, so the compiler can count them exactly; a
    // pooled input (POOL, src.mode == IN_POOL, never with a skip part) keeps
    // its four window values in registers and takes their max when storing.
    constexpr bool pool = POOL;
    constexpr int NR = POOL ? 4 : 1;
    struct Stage {               // one slice's loads in flight
        bf16x8 raw[NV][NR];
        bf16x8 w[NWV];
        uint32_t valid;
    };
    auto fetch = [&](Stage& g, int c0) {
        const bool first = c0 < src.ca;
        const bf16* base = first ? src.a : src.b;
        const int cs = first ? src.ca : src.cb, cc = first ? c0 : c0 - src.ca;
        const int up = (first && src.mode == IN_UP) ? 1 : 0, dn = pool ? 1 : 0;
        const int Hs = (H << dn) >> up, Ws = (W << dn) >> up;     // source image
        g.valid = 0;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const int v = threadIdx.x + k * 256;
            const int vv = v < NPV ? v : 0;
            const int pix = vv >> 2, part = vv & 3;
            const int ph = pix / PW, pw = pix - ph * PW;
            const int gh = th0 + ph - 1, gw = tw0 + pw - 1;
            if (v < NPV && gh >= 0 && gh < H && gw >= 0 && gw < W) g.valid |= 1u << k;
            const int ghc = min(max(gh, 0), H - 1), gwc = min(max(gw, 0), W - 1);
            const int64_t r0 = (((int64_t)b * Hs + ((ghc << dn) >> up)) * Ws + ((gwc << dn) >> up))
                               * cs + cc + part * 8;
            g.raw[k][0] = load8(base + r0);
            if constexpr (POOL) {
                const int64_t rs = (int64_t)Ws * cs;
                g.raw[k][1] = load8(base + r0 + cs);
                g.raw[k][2] = load8(base + r0 + rs);
                g.raw[k][3] = load8(base + r0 + rs + cs);
            }
        }
        // the weights w[nb0 + n][tap][c0 .. c0 + 31] of the slice: vector
        // v = (n * 9 + tap) * 4 + part
#pragma unroll
        for (int k = 0; k < NWV; k++) {
            const int v = threadIdx.x + k * 256;
            const int vv = v < NWVEC ? v : 0;
            g.w[k] = load8(w + (int64_t)(nb0 + (vv >> 2) / 9) * 9 * cin +
                           (int64_t)((vv >> 2) % 9) * cin + c0 + (vv & 3) * 8);
        }
    };
    auto put = [&](const Stage& g, int buf) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const int v = threadIdx.x + k * 256;
            if (v >= NPV) continue;
            bf16x8 x = g.raw[k][0];
            if constexpr (POOL) {
#pragma unroll
                for (int e = 0; e < 8; e++)
                    x[e] = (bf16)fmaxf(fmaxf((float)g.raw[k][0][e], (float)g.raw[k][1][e]),
                                       fmaxf((float)g.raw[k][2][e], (float)g.raw[k][3][e]));
            }
            if (!((g.valid >> k) & 1u)) x = zero8();
            *reinterpret_cast<bf16x8*>(&patch[buf][(v >> 2) * CS + (v & 3) * 8]) = x;
        }
#pragma unroll
        for (int k = 0; k < NWV; k++) {
            const int v = threadIdx.x + k * 256;
            if (v >= NWVEC) continue;
            *reinterpret_cast<bf16x8*>(&wl[buf][(v >> 2) * CSW + (v & 3) * 8]) = g.w[k];
        }
    };
    auto mfmas = [&](int buf) {
        const bf16* pa = &patch[buf][aoff];
        const bf16* pb = &wl[buf][boff];
#pragma unroll
        for (int dx = 0; dx < 3; dx++) {
            bf16x8 bq[3][2];
#pragma unroll
            for (int dy = 0; dy < 3; dy++)
#pragma unroll
                for (int u = 0; u < 2; u++)
                    bq[dy][u] = *reinterpret_cast<const bf16x8*>(pb + (u * 16 * 9 + dy * 3 + dx) * CSW);
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(pa + (s * PW + dx) * CS);
#pragma unroll
                for (int dy = 0; dy < 3; dy++) {
                    const int d = s - dy;
                    if (d < 0 || d % RM != 0 || d / RM >= MT) continue;
#pragma unroll
                    for (int u = 0; u < 2; u++)   // channels as rows, pixels as columns
                        acc[d / RM][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            bq[dy][u], a, acc[d / RM][u], 0, 0, 0);
                }
            }
        }
    };

    if constexpr (PIPE) {
        // Two slices in flight: slice s is multiplied out of one LDS buffer
        // while slice s + 1 (loaded during the previous slice) is stored to
        // the other and slice s + 2 is being loaded, so each load has two
        // slices' time to arrive.  Slice s of the split sits in buffer
        // (s - sl0) & 1 and in stage g0 / g1 by the same parity.  The
        // workgroup has at least three slices (the launcher's condition);
        // the loads in the main loop are unconditional, so the compiler's
        // wait counts stay exact.
        Stage g0, g1;
        fetch(g0, sl0 * 32);
        fetch(g1, min(sl0 + 1, sl1 - 1) * 32);
        put(g0, 0);
        __syncthreads();
        int s = sl0;
        for (; s + 3 < sl1; s += 2) {
            fetch(g0, (s + 2) * 32);
            __builtin_amdgcn_sched_barrier(0);   // the loads issue before the MFMAs
            mfmas(0);
            put(g1, 1);          // buffer 1 was last read before the previous barrier
            __syncthreads();
            fetch(g1, (s + 3) * 32);
            __builtin_amdgcn_sched_barrier(0);
            mfmas(1);
            put(g0, 0);
            __syncthreads();
        }
        // the last one to three slices: s in buffer 0, s + 1 in g1
        if (s + 2 < sl1) fetch(g0, (s + 2) * 32);
        mfmas(0);
        if (s + 1 < sl1) {
            put(g1, 1);
            __syncthreads();
            mfmas(1);
            if (s + 2 < sl1) {
                put(g0, 0);
                __syncthreads();
                mfmas(0);
            }
        }
    } else {
        Stage g;
        for (int s = sl0; s < sl1; s++) {
            if (s > sl0) __syncthreads();        // the previous slice's MFMAs are done
            fetch(g, s * 32);
            put(g, 0);
            __syncthreads();
            mfmas(0);
        }
    }
    // epilogue: the result tile is (channels x pixels), so the lane holds
    // pixel row of each M tile and its four consecutive channels
    // n_base + 16 u + 4 (l / 16) + 0..3: one 8-byte store of bf16 (16 bytes
    // of fp32) per tile and channel half
    const int q4 = (lane >> 4) * 4;
    // with the fused head (32 channels, WN == 1, no split) each wave also
    // stages its bf16 outputs, one row of CS per pixel, where the patch was
    constexpr int HROWS = MT * 16;
    static_assert(WN != 1 || 4 * HROWS <= PH * PW, "head staging fits the patch buffer");
    const bool head = WN == 1 && src.head_y != nullptr;      // uniform
    bf16* hs = &patch[0][0] + wave * HROWS * CS;
    if (head) __syncthreads();       // every wave's MFMAs are done with the patch
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int n = n_base + u * 16 + q4;
        float4 bv = {0.f, 0.f, 0.f, 0.f};
        if ((flags & 1) && nz == 1) bv = *reinterpret_cast<const float4*>(bias + n);
#pragma unroll
        for (int i = 0; i < MT; i++) {
            const int m = (wm * MT + i) * 16 + row;
            const int gh = th0 + m / TW, gw = tw0 + m % TW;
            if (gh >= H || gw >= W) continue;
            const int64_t o = (((int64_t)b * H + gh) * W + gw) * cout + n;
            const f32x4 r = acc[i][u];
            if (nz > 1) {   // split z's partial sum in its own slice; k_conv_finish adds them
                *reinterpret_cast<float4*>(src.acc32 + (int64_t)z * ((int64_t)nb * H * W * cout) + o) =
                    float4{r[0], r[1], r[2], r[3]};
                continue;
            }
            float v[4] = {r[0] + bv.x, r[1] + bv.y, r[2] + bv.z, r[3] + bv.w};
            if (flags & 2)
#pragma unroll
                for (int j = 0; j < 4; j++) v[j] = v[j] > 0.f ? v[j] : 0.f;
            typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
            const bf16x4 vb = {(bf16)v[0], (bf16)v[1], (bf16)v[2], (bf16)v[3]};
            if (out) *reinterpret_cast<bf16x4*>(out + o) = vb;
            if (out32) *reinterpret_cast<float4*>(out32 + o) = float4{v[0], v[1], v[2], v[3]};
            if (head) *reinterpret_cast<bf16x4*>(hs + (i * 16 + row) * CS + u * 16 + q4) = vb;
        }
    }
    if (head) {
        __syncthreads();
        if (lane < HROWS) {
            const int m = wm * HROWS + lane;
            const int gh = th0 + m / TW, gw = tw0 + m % TW;
            if (gh < H && gw < W) {
                float sum = src.head_b[0];
#pragma unroll
                for (int c = 0; c < 32; c += 8) {
                    const bf16x8 x = *reinterpret_cast<const bf16x8*>(hs + lane * CS + c);
#pragma unroll
                    for (int e = 0; e < 8; e++) sum += (float)x[e] * src.head_w[c + e];
                }
                src.head_y[((int64_t)b * H + gh) * W + gw] = sum;
            }
        }
    }
}

// Completes a split-K convolution: the ks partial sums added in split order
// (the same bits on every run), then bias, ReLU and BF16; four consecutive
// channels per thread (n and cout are multiples of 4).
__global__ void __launch_bounds__(256)
k_conv_finish(const float* __restrict__ acc32, int ks, const float* __restrict__ bias,
              bf16* __restrict__ out, float* __restrict__ out32, int64_t n, int cout, int flags,
              const int* __restrict__ live, int64_t per_image)
{
    const int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (k >= n) return;
    if (live && k >= (int64_t)*live * per_image) return;   // images the splits skipped
    float4 v = *reinterpret_cast<const float4*>(acc32 + k);
    for (int z = 1; z < ks; z++) {
        const float4 a = *reinterpret_cast<const float4*>(acc32 + (int64_t)z * n + k);
        v.x = v.x + a.x;
        v.y = v.y + a.y;
        v.z = v.z + a.z;
        v.w = v.w + a.w;
    }
    if (flags & 1) {
        const float4 bv = *reinterpret_cast<const float4*>(bias + (int)(k % cout));
        v.x = v.x + bv.x;
        v.y = v.y + bv.y;
        v.z = v.z + bv.z;
        v.w = v.w + bv.w;
    }
    if (flags & 2) {
        v.x = v.x > 0.f ? v.x : 0.f;
        v.y = v.y > 0.f ? v.y : 0.f;
        v.z = v.z > 0.f ? v.z : 0.f;
        v.w = v.w > 0.f ? v.w : 0.f;
    }
    if (out) {
        typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
        *reinterpret_cast<bf16x4*>(out + k) = bf16x4{(bf16)v.x, (bf16)v.y, (bf16)v.z, (bf16)v.w};
    }
    if (out32) *reinterpret_cast<float4*>(out32 + k) = v;
}

// w'[ci][8 - tap][co] = w[co][tap][ci]
__global__ void k_weight_flip(const bf16* __restrict__ w, bf16* __restrict__ wt, int cin, int cout)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)cin * 9 * cout;
    if (k >= n) return;
    const int ci = (int)(k % cin);
    const int tap = (int)((k / cin) % 9);
    const int co = (int)(k / ((int64_t)cin * 9));
    wt[((int64_t)ci * 9 + (8 - tap)) * cout + co] = w[k];
}

// dw[co][tap][ci] over a pixel chunk: sum of gz[pix][co] * in[pix+off][ci].
// Block: 256 threads, a tile of up to 64 co x 64 ci for one tap; chunks of
// 128 pixels are staged transposed in LDS ([channel][pixel]) so each lane's 8
// consecutive pixels (the MFMA's k) are one 16-byte read.  The four waves
// cover the tile in 32 x 32 pieces and split the chunk's four 32-pixel k
// steps among themselves when the tile is narrower than 64 x 64, so no wave
// multiplies padding; those waves then add their sums in LDS in wave order.
// With one pixel chunk (chunk_stride 0) the block adds its sums into dw; with
// several, chunk z writes its own slice out[z * chunk_stride + ...] and
// k_chunk_sum adds the slices in chunk order, so the result does not depend
// on the order in which blocks finish.
__global__ void __launch_bounds__(256)
k_conv3x3_dw(const bf16* __restrict__ gz, int cout, const bf16* __restrict__ in, int cin,
             float* __restrict__ out, int64_t chunk_stride, int nb, int H, int W,
             int pix_per_block)
{
    constexpr int KP = 128, LS = KP + 8;
    __shared__ __attribute__((aligned(16))) bf16 smem[2][64][LS];
    bf16 (*sa)[LS] = smem[0];   // [co][pix]
    bf16 (*sb)[LS] = smem[1];   // [ci][pix]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tap = blockIdx.z % 9;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int co0 = blockIdx.y * 64;
    const int ci0 = blockIdx.x * 64;
    const int npix = nb * H * W;
    const int p_begin = (blockIdx.z / 9) * pix_per_block;
    const int p_end = min(npix, p_begin + pix_per_block);
    const int ci_n = cin - ci0 < 64 ? cin - ci0 : 64;    // 32 or 64
    const int co_n = cout - co0 < 64 ? cout - co0 : 64;
    const int wo = co_n / 32, wi = ci_n / 32, ks = 4 / (wo * wi);
    const int mo = wave % wo, ni = (wave / wo) % wi, kg = wave / (wo * wi);
    f32x4 acc[2][2];
    for (int t = 0; t < 2; t++)
        for (int u = 0; u < 2; u++) acc[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int row = lane & 15, kq = (lane >> 4) * 8;
    // staging: thread -> pixel j = tid / 2, channels [32 h, 32 h + 32), h = tid % 2
    const int sj = threadIdx.x >> 1, sh = (threadIdx.x & 1) * 32;
    for (int p0 = p_begin; p0 < p_end; p0 += KP) {
        const int pix = p0 + sj;
        bf16x8 va[4], vb[4];
        for (int q = 0; q < 4; q++) {
            va[q] = zero8();
            vb[q] = zero8();
        }
        if (pix < p_end) {
            const int b = pix / (H * W), r = pix - b * H * W;
            const int hh = r / W + dy, ww = r % W + dx;
            const bool inside = hh >= 0 && hh < H && ww >= 0 && ww < W;
            for (int q = 0; q < 4; q++) {
                const int c = sh + q * 8;
                if (c < co_n) va[q] = load8(gz + (int64_t)pix * cout + co0 + c);
                if (inside && c < ci_n)
                    vb[q] = load8(in + ((int64_t)(b * H + hh) * W + ww) * cin + ci0 + c);
            }
        }
        __syncthreads();
        for (int q = 0; q < 4; q++)
            for (int i = 0; i < 8; i++) {
                sa[sh + q * 8 + i][sj] = va[q][i];
                sb[sh + q * 8 + i][sj] = vb[q][i];
            }
        __syncthreads();
        for (int k = kg; k < 4; k += ks) {
            bf16x8 a[2], bb[2];
            for (int t = 0; t < 2; t++) a[t] = load8(&sa[mo * 32 + t * 16 + row][k * 32 + kq]);
            for (int u = 0; u < 2; u++) bb[u] = load8(&sb[ni * 32 + u * 16 + row][k * 32 + kq]);
            for (int t = 0; t < 2; t++)
                for (int u = 0; u < 2; u++)
                    acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], bb[u], acc[t][u], 0, 0, 0);
        }
    }
    if (ks > 1) {
        // wave w + g * wo * wi holds k group g of wave w's piece: 16 floats a
        // lane, 4 KB a wave, in the staging memory (34 KB) once every wave is
        // past its last MFMA read
        float* red = reinterpret_cast<float*>(&smem[0][0][0]);
        __syncthreads();
        if (kg > 0)
            for (int t = 0; t < 2; t++)
                for (int u = 0; u < 2; u++)
                    for (int i = 0; i < 4; i++)
                        red[(wave * 16 + (t * 2 + u) * 4 + i) * 64 + lane] = acc[t][u][i];
        __syncthreads();
        if (kg == 0)
            for (int g = 1; g < ks; g++) {
                const int w2 = wave + g * wo * wi;
                for (int t = 0; t < 2; t++)
                    for (int u = 0; u < 2; u++)
                        for (int i = 0; i < 4; i++)
                            acc[t][u][i] += red[(w2 * 16 + (t * 2 + u) * 4 + i) * 64 + lane];
            }
    }
    if (kg != 0) return;
    float* dst = out + (int64_t)(blockIdx.z / 9) * chunk_stride;
    for (int t = 0; t < 2; t++)
        for (int u = 0; u < 2; u++) {
            const int ci = ci0 + ni * 32 + u * 16 + row;
            for (int i = 0; i < 4; i++) {
                const int co = co0 + mo * 32 + t * 16 + (lane >> 4) * 4 + i;
                const int64_t o = ((int64_t)co * 9 + tap) * cin + ci;
                dst[o] = chunk_stride ? acc[t][u][i] : dst[o] + acc[t][u][i];
            }
        }
}

// dst[i] += part[i] + part[n + i] + ... + part[(chunks - 1) n + i], in that order
__global__ void k_chunk_sum(const float* __restrict__ part, int chunks, int64_t n,
                            float* __restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = part[i];
#pragma unroll 8
    for (int c = 1; c < chunks; c++) s += part[(int64_t)c * n + i];
    dst[i] += s;
}

// Block j: the sum over k of part[k * stride + col0 + j], k in [0, nparts):
// thread t adds k = t, t + 256, ... in order, then a fixed LDS tree.  dst[j]
// gets it added (add) or assigned.
__global__ void __launch_bounds__(256)
k_block_sum(const float* __restrict__ part, int nparts, int stride, int col0,
            float* __restrict__ dst, int add)
{
    __shared__ float red[256];
    const int j = blockIdx.x;
    float s = 0.f;
    for (int k = threadIdx.x; k < nparts; k += 256) s += part[(int64_t)k * stride + col0 + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) dst[j] = add ? dst[j] + red[0] : red[0];
}

// ------------------------------------------------------------ elementwise

// g = g * (y > 0): gradient through a ReLU whose output is y
__global__ void k_relu_bwd(int64_t n, const bf16* __restrict__ y, bf16* __restrict__ g)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && !((float)y[k] > 0.f)) g[k] = (bf16)0.f;
}

// part[block][c] = sum over the block's pixels of g[pix][c].  c is 32, 64,
// 128 or 256: thread t takes the 8 channels of vector v = t % (c / 8) (one
// 16-byte load per pixel) for pixels j, j + lanes, ... of the block's chunk
// (j = t / (c / 8), lanes = 256 / (c / 8)), each channel added in pixel
// order; then the lanes' partials of each channel are added in LDS in lane
// order.  The loads of several pixels are in flight at once.
__global__ void __launch_bounds__(256)
k_bias_grad(const bf16* __restrict__ g, int c, int npix, int per, float* __restrict__ part)
{
    __shared__ float red[256 * 8];
    const int nv = c / 8, v = threadIdx.x % nv, j = threadIdx.x / nv, lanes = 256 / nv;
    const int p0 = blockIdx.x * per, p1 = min(npix, p0 + per);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
    for (int p = p0 + j; p < p1; p += lanes) {
        const bf16x8 x = load8(g + (int64_t)p * c + v * 8);
#pragma unroll
        for (int e = 0; e < 8; e++) s[e] += (float)x[e];
    }
#pragma unroll
    for (int e = 0; e < 8; e++) red[threadIdx.x * 8 + e] = s[e];
    __syncthreads();
    if (threadIdx.x < c) {
        const int cv = threadIdx.x / 8, ce = threadIdx.x % 8;
        float t = 0.f;
        for (int k = 0; k < lanes; k++) t += red[(k * nv + cv) * 8 + ce];
        part[(int64_t)blockIdx.x * c + threadIdx.x] = t;
    }
}

// 2x2 max pooling, H and W even
__global__ void k_maxpool(const bf16* __restrict__ x, bf16* __restrict__ y, int nb, int H, int W, int c,
                          const int* __restrict__ live)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n = (int64_t)(live ? min(nb, *live) : nb) * Ho * Wo * c;
    if (k >= n) return;
    const int ch = (int)(k % c);
    const int64_t q = k / c;
    const int wo = (int)(q % Wo), ho = (int)((q / Wo) % Ho), b = (int)(q / ((int64_t)Wo * Ho));
    const int64_t base = ((int64_t)(b * H + 2 * ho) * W + 2 * wo) * c + ch;
    float m = (float)x[base];
    m = fmaxf(m, (float)x[base + c]);
    m = fmaxf(m, (float)x[base + (int64_t)W * c]);
    m = fmaxf(m, (float)x[base + (int64_t)W * c + c]);
    y[k] = (bf16)m;
}

// gradient of max pooling: to the first maximal input of each window
__global__ void k_maxpool_bwd(const bf16* __restrict__ x, const bf16* __restrict__ y,
                              const bf16* __restrict__ gy, bf16* __restrict__ gx, int nb, int H,
                              int W, int c)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n = (int64_t)nb * Ho * Wo * c;
    if (k >= n) return;
    const int ch = (int)(k % c);
    const int64_t q = k / c;
    const int wo = (int)(q % Wo), ho = (int)((q / Wo) % Ho), b = (int)(q / ((int64_t)Wo * Ho));
    const int64_t base = ((int64_t)(b * H + 2 * ho) * W + 2 * wo) * c + ch;
    const int64_t off[4] = {0, c, (int64_t)W * c, (int64_t)W * c + c};
    const float m = (float)y[k];
    int hit = 0;
    for (int i = 0; i < 4; i++) {
        const bool take = !hit && (float)x[base + off[i]] == m;
        if (take) gx[base + off[i]] = (bf16)((float)gx[base + off[i]] + (float)gy[k]);
        hit |= take;
    }
}

// nearest 2x upsampling of x (nb, H/2, W/2, c) into channels [c0, c0 + c) of
// y (nb, H, W, cy) -- the concatenation with the skip happens in place
__global__ void k_upsample(const bf16* __restrict__ x, bf16* __restrict__ y, int nb, int H, int W,
                           int c, int cy, int c0, const int* __restrict__ live)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)(live ? min(nb, *live) : nb) * H * W * c;
    if (k >= n) return;
    const int ch = (int)(k % c);
    const int64_t q = k / c;
    const int w = (int)(q % W), h = (int)((q / W) % H), b = (int)(q / ((int64_t)W * H));
    y[q * cy + c0 + ch] = x[((int64_t)(b * (H / 2) + h / 2) * (W / 2) + w / 2) * c + ch];
}

__global__ void k_upsample_bwd(const bf16* __restrict__ gy, bf16* __restrict__ gx, int nb, int H,
                               int W, int c, int cy, int c0)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n = (int64_t)nb * Ho * Wo * c;
    if (k >= n) return;
    const int ch = (int)(k % c);
    const int64_t q = k / c;
    const int wo = (int)(q % Wo), ho = (int)((q / Wo) % Ho), b = (int)(q / ((int64_t)Wo * Ho));
    float s = 0.f;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            s += (float)gy[((int64_t)(b * H + 2 * ho + i) * W + 2 * wo + j) * cy + c0 + ch];
    gx[k] = (bf16)s;
}

// copy channels: y[:, c0:c0+c] = x (forward of the skip half of a concat) or
// x = y[:, c0:c0+c] (its backward), add = accumulate into x
__global__ void k_channel_copy(const bf16* __restrict__ src, int cs, int s0, bf16* __restrict__ dst,
                               int cd, int d0, int c, int64_t npix, int add,
                               const int* __restrict__ live, int64_t hw)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npix * c) return;
    if (live && k / c >= (int64_t)*live * hw) return;
    const int ch = (int)(k % c);
    const int64_t p = k / c;
    const bf16 v = src[p * cs + s0 + ch];
    bf16& d = dst[p * cd + d0 + ch];
    d = add ? (bf16)((float)d + (float)v) : v;
}

// 1x1 head: y[pix] = b + sum_c x[pix, c] * w[c]   (fp32 out)
__global__ void k_head(const bf16* __restrict__ x, int c, const float* __restrict__ w,
                       const float* __restrict__ b, float* __restrict__ y, int64_t npix,
                       const int* __restrict__ live, int64_t hw)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    if (live && p >= (int64_t)*live * hw) return;
    float s = b[0];
    const bf16* xp = x + p * c;              // c is a multiple of 8: 16-byte loads
    for (int i = 0; i < c; i += 8) {
        const bf16x8 q = load8(xp + i);
#pragma unroll
        for (int e = 0; e < 8; e++) s += (float)q[e] * w[i + e];
    }
    y[p] = s;
}

// head backward: gx[pix, c] = gy[pix] * w[c]; per block, part[block][0] = sum gy
// and part[block][1 + c] = sum gy * x[c] over the block's pixels
__global__ void k_head_bwd(const bf16* __restrict__ x, int c, const float* __restrict__ w,
                           const float* __restrict__ gy, bf16* __restrict__ gx,
                           float* __restrict__ part, int64_t npix)
{
    __shared__ float red[4][33];
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float g = p < npix ? gy[p] : 0.f;
    if (p < npix)
        for (int i = 0; i < c; i++) gx[p * c + i] = (bf16)(g * w[i]);
    // per wave: shuffle sums of g and of g * x[i], then the 4 waves in LDS
    for (int i = -1; i < c; i++) {
        float v = i < 0 ? g : (p < npix ? g * (float)x[p * c + i] : 0.f);
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[wave][i + 1] = v;
    }
    __syncthreads();
    if (threadIdx.x <= c) {
        const float t = (red[0][threadIdx.x] + red[1][threadIdx.x]) +
                        (red[2][threadIdx.x] + red[3][threadIdx.x]);
        part[(int64_t)blockIdx.x * (c + 1) + threadIdx.x] = t;
    }
}

// masked MSE, normalised per sample: loss_s = sum m (y-t)^2 / sum m t^2;
// gy = 2 m (y - t) / (denominator * nb).  A sample whose target is zero (a
// solve that made no change) has weight zero.
__global__ void k_loss(const float* __restrict__ y, const float* __restrict__ t,
                       const float* __restrict__ m, const float* __restrict__ denom,
                       float* __restrict__ gy, float* __restrict__ part, int64_t npix_per,
                       int nb)
{
    __shared__ float red[256];
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = npix_per * nb;
    float l = 0.f;
    if (k < n) {
        const int s = (int)(k / npix_per);
        const float d = y[k] - t[k];
        const float mm = denom[s] > 1e-12f ? m[k] : 0.f;
        l = mm * d * d / denom[s];
        gy[k] = 2.f * mm * d / (denom[s] * nb);
    }
    red[threadIdx.x] = l / nb;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// per-sample sum of m * t^2 (+1e-30), in parts: part[block x][sample]
__global__ void k_target_norm(const float* __restrict__ t, const float* __restrict__ m,
                              float* __restrict__ part, int64_t npix_per)
{
    __shared__ float red[256];
    const int s = blockIdx.y;
    float acc = 0.f;
    for (int64_t k = blockIdx.x * 256 + threadIdx.x; k < npix_per; k += (int64_t)gridDim.x * 256) {
        const float v = t[s * npix_per + k];
        acc += m[s * npix_per + k] * v * v;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[(int64_t)blockIdx.x * gridDim.y + s] = red[0] + (blockIdx.x == 0 ? 1e-30f : 0.f);
}

// Adam on fp32 master weights; refreshes the bf16 copy used by the kernels
__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, bf16* __restrict__ pb,
                       float lr, float b1, float b2, float eps, float c1, float c2, float clip)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float gk = g[k] * clip;
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float upd = lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    p[k] -= upd;
    if (pb) pb[k] = (bf16)p[k];
}

__global__ void k_to_bf16(int64_t n, const float* __restrict__ x, bf16* __restrict__ y)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] = (bf16)x[k];
}

// ------------------------------------------------------------ launchers

static inline unsigned nblk(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

template <int TW, int TH, int WN>
static void conv_tile(hipStream_t st, const ConvIn& src, const bf16* w, const float* bias,
                      bf16* out, float* out32, int cout, int nb, int H, int W, int flags) {
    const unsigned tiles = (unsigned)(((W + TW - 1) / TW) * ((H + TH - 1) / TH) * nb);
    const unsigned nblocks = tiles * (unsigned)(cout / (32 * WN));
    // small layers (fewer workgroups than CUs, e.g. the deep levels of a small
    // batch): split the input channels over up to 4 blocks per tile, each with
    // at least two 32-channel slices, writing its partial sums to its own
    // slice of the scratch.  From one workgroup per CU up, the finishing pass
    // costs more than the split gains.
    const int nsl = (src.ca + src.cb) / 32;
    const int64_t nout = (int64_t)nb * H * W * cout;
    int ks = 1;
    if (src.acc32 && nblocks < 256 && !src.head_y)
        while (ks < 4 && nblocks * ks < 1024 && nsl / (2 * ks) >= 2 &&
               2 * ks * nout <= src.acc_cap)
            ks *= 2;
    dim3 grid(tiles, cout / (32 * WN), ks);
    const bool pool = src.mode == IN_POOL, pipe = nsl / ks >= 3 && nblocks * ks <= 512;
#define STM_CONV_TILE(P, Q)                                                                    \
    k_conv3x3_tile<TW, TH, WN, P, Q><<<grid, 256, 0, st>>>(src, w, bias, out, out32, cout, nb, H, \
                                                           W, flags)
    if (pool && pipe) STM_CONV_TILE(true, true);
    else if (pool) STM_CONV_TILE(true, false);
    else if (pipe) STM_CONV_TILE(false, true);
    else STM_CONV_TILE(false, false);
#undef STM_CONV_TILE
    if (ks > 1)
        k_conv_finish<<<nblk(nout / 4), 256, 0, st>>>(src.acc32, ks, bias, out, out32, nout, cout,
                                                  flags, src.live, (int64_t)H * W * cout);
}

bool conv_tiles_fit(int H, int W) { return W % 8 == 0 && H % 8 == 0; }

void conv3x3_in(hipStream_t st, const ConvIn& src, const bf16* w, const float* bias, bf16* out,
                float* out32, int cout, int nb, int H, int W, int flags) {
    // two N blocks per workgroup only where that still leaves >= 512 workgroups
    const int64_t pix = (int64_t)nb * H * W;
    const bool n64 = cout % 64 == 0 && (pix / 128) * (cout / 64) >= 512;
    if (W % 16 == 0) {
        if (n64) conv_tile<16, 8, 2>(st, src, w, bias, out, out32, cout, nb, H, W, flags);
        else conv_tile<16, 8, 1>(st, src, w, bias, out, out32, cout, nb, H, W, flags);
    } else if (H % 16 == 0) {
        if (n64) conv_tile<8, 16, 2>(st, src, w, bias, out, out32, cout, nb, H, W, flags);
        else conv_tile<8, 16, 1>(st, src, w, bias, out, out32, cout, nb, H, W, flags);
    } else {   // 8 x 8 tiles of 64 pixels
        conv_tile<8, 8, 1>(st, src, w, bias, out, out32, cout, nb, H, W, flags);
    }
}

void conv3x3(hipStream_t st, const bf16* in, int cin, const bf16* w, const float* bias, bf16* out,
             float* out32, int cout, int nb, int H, int W, int flags, const int* live) {
    const int64_t npix = (int64_t)nb * H * W;
    if (conv_tiles_fit(H, W)) {
        const ConvIn src{in, cin, IN_DIRECT, nullptr, 0, live};
        conv3x3_in(st, src, w, bias, out, out32, cout, nb, H, W, flags);
        return;
    }
    if (npix >= 64 * 256) {
        dim3 grid(nblk(npix, 128), (cout + 63) / 64);
        k_conv3x3<128><<<grid, 256, 0, st>>>(in, cin, w, bias, out, out32, cout, nb, H, W, flags,
                                             live);
    } else {
        dim3 grid(nblk(npix, 64), (cout + 63) / 64);
        k_conv3x3<64><<<grid, 256, 0, st>>>(in, cin, w, bias, out, out32, cout, nb, H, W, flags,
                                            live);
    }
}
void weight_flip(hipStream_t st, const bf16* w, bf16* wt, int cin, int cout) {
    const int64_t n = (int64_t)cin * 9 * cout;
    k_weight_flip<<<nblk(n), 256, 0, st>>>(w, wt, cin, cout);
}
void conv3x3_dw(hipStream_t st, const bf16* gz, int cout, const bf16* in, int cin, float* dw32,
                int nb, int H, int W, float* part) {
    const int64_t npix = (int64_t)nb * H * W;
    const int ci_blocks = (cin + 63) / 64;
    const int per = dw_pixels_per_chunk(npix);
    const int chunks = (int)dw_chunks(npix);   // partials: grad_partials (nn_sizes.h)
    const int64_t nw = (int64_t)cout * 9 * cin;
    dim3 grid(ci_blocks, (cout + 63) / 64, 9 * chunks);
    if (chunks == 1) {
        k_conv3x3_dw<<<grid, 256, 0, st>>>(gz, cout, in, cin, dw32, 0, nb, H, W, per);
        return;
    }
    k_conv3x3_dw<<<grid, 256, 0, st>>>(gz, cout, in, cin, part, nw, nb, H, W, per);
    k_chunk_sum<<<nblk(nw), 256, 0, st>>>(part, chunks, nw, dw32);
}
void relu_bwd(hipStream_t st, int64_t n, const bf16* y, bf16* g) {
    k_relu_bwd<<<nblk(n), 256, 0, st>>>(n, y, g);
}
void bias_grad(hipStream_t st, const bf16* g, int c, int64_t npix, float* db, float* part) {
    // bias channels are 32, 64, 128 or 256; each block leaves c partial sums
    // and at most about one block per CU runs, so the block sums stay short
    const int per = bias_pixels_per_block(npix);
    const int nblocks = (int)bias_blocks(npix);   // partials: grad_partials (nn_sizes.h)
    k_bias_grad<<<nblocks, 256, 0, st>>>(g, c, (int)npix, per, part);
    k_block_sum<<<c, 256, 0, st>>>(part, nblocks, c, 0, db, 1);
}
void maxpool(hipStream_t st, const bf16* x, bf16* y, int nb, int H, int W, int c, const int* live) {
    k_maxpool<<<nblk((int64_t)nb * H * W * c / 4), 256, 0, st>>>(x, y, nb, H, W, c, live);
}
void maxpool_bwd(hipStream_t st, const bf16* x, const bf16* y, const bf16* gy, bf16* gx, int nb,
                 int H, int W, int c) {
    k_maxpool_bwd<<<nblk((int64_t)nb * H * W * c / 4), 256, 0, st>>>(x, y, gy, gx, nb, H, W, c);
}
void upsample(hipStream_t st, const bf16* x, bf16* y, int nb, int H, int W, int c, int cy, int c0,
              const int* live) {
    k_upsample<<<nblk((int64_t)nb * H * W * c), 256, 0, st>>>(x, y, nb, H, W, c, cy, c0, live);
}
void upsample_bwd(hipStream_t st, const bf16* gy, bf16* gx, int nb, int H, int W, int c, int cy,
                  int c0) {
    k_upsample_bwd<<<nblk((int64_t)nb * H * W * c / 4), 256, 0, st>>>(gy, gx, nb, H, W, c, cy, c0);
}
void channel_copy(hipStream_t st, const bf16* src, int cs, int s0, bf16* dst, int cd, int d0, int c,
                  int64_t npix, int add, const int* live, int64_t hw) {
    k_channel_copy<<<nblk(npix * c), 256, 0, st>>>(src, cs, s0, dst, cd, d0, c, npix, add, live,
                                                   hw);
}
void head(hipStream_t st, const bf16* x, int c, const float* w, const float* b, float* y,
          int64_t npix, const int* live, int64_t hw) {
    k_head<<<nblk(npix), 256, 0, st>>>(x, c, w, b, y, npix, live, hw);
}
void head_bwd(hipStream_t st, const bf16* x, int c, const float* w, const float* gy, bf16* gx,
              float* dw, float* db, int64_t npix, float* part) {
    const int nblocks = (int)nblk(npix);
    k_head_bwd<<<nblocks, 256, 0, st>>>(x, c, w, gy, gx, part, npix);
    k_block_sum<<<1, 256, 0, st>>>(part, nblocks, c + 1, 0, db, 1);
    k_block_sum<<<c, 256, 0, st>>>(part, nblocks, c + 1, 1, dw, 1);
}
void loss(hipStream_t st, const float* y, const float* t, const float* m, float* denom, float* gy,
          float* loss_out, int64_t npix_per, int nb, float* part) {
    k_target_norm<<<dim3(64, nb), 256, 0, st>>>(t, m, part, npix_per);
    k_block_sum<<<nb, 256, 0, st>>>(part, 64, nb, 0, denom, 0);
    const int nblocks = (int)nblk(npix_per * nb);
    k_loss<<<nblocks, 256, 0, st>>>(y, t, m, denom, gy, part, npix_per, nb);
    k_block_sum<<<1, 256, 0, st>>>(part, nblocks, 1, 0, loss_out, 1);
}
void adam(hipStream_t st, int64_t n, float* p, const float* g, float* m, float* v, bf16* pb,
          float lr, float b1, float b2, float eps, int step, float clip) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    k_adam<<<nblk(n), 256, 0, st>>>(n, p, g, m, v, pb, lr, b1, b2, eps, c1, c2, clip);
}
void to_bf16(hipStream_t st, int64_t n, const float* x, bf16* y) {
    k_to_bf16<<<nblk(n), 256, 0, st>>>(n, x, y);
}

}  // namespace nn
}  // namespace stm
