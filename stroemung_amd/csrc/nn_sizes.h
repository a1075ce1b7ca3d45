// Launch sizes of the network's gradient and loss reductions and the bounds
// on their per-block partial sums (plain C++, shared by the library and its
// CPU tests).  The partials live in one scratch buffer sized by these bounds
// for the largest batch, so every smaller batch must fit as well.
#pragma once

#include <stdint.h>

#include <algorithm>

namespace stm {
namespace nn {

// 256-thread blocks over n items
inline int64_t blocks256(int64_t n) { return (n + 255) / 256; }

// weight gradient: pixel chunks of 1024 << k pixels, at most 256 full
// chunks until a chunk reaches 1 << 20 pixels (several workgroups per CU on
// every layer of the small frames, so their load latencies overlap)
inline int dw_pixels_per_chunk(int64_t npix) {
    int per = 1024;
    while (npix / per > 256 && per < (1 << 20)) per *= 2;
    return per;
}
inline int64_t dw_chunks(int64_t npix) {
    const int per = dw_pixels_per_chunk(npix);
    return (npix + per - 1) / per;
}

// bias gradient: blocks of 128 << k pixels, at most 256 full blocks
inline int bias_pixels_per_block(int64_t npix) {
    int per = 128;
    while (npix / per > 256) per *= 2;
    return per;
}
inline int64_t bias_blocks(int64_t npix) {
    const int per = bias_pixels_per_block(npix);
    return (npix + per - 1) / per;
}

// Partials (floats) of conv3x3_dw and bias_grad for any pixel count up to
// npix.  The chunk counts are not monotone in the pixel count: at most 257
// chunks until the chunk size stops growing, at most 257 bias blocks.
inline int64_t grad_partials(int cin, int cout, int64_t npix) {
    const int64_t chunks =
        std::min((npix + 1023) / 1024, std::max<int64_t>(257, (npix + (1 << 20) - 1) >> 20));
    const int64_t dw = chunks > 1 ? chunks * cout * 9 * cin : 0;
    const int64_t db = std::min<int64_t>((npix + 127) / 128, 257) * cout;
    return std::max(dw, db);
}

// Partials of loss (per-sample sums over 64 blocks, then one per 256-pixel
// block) and of head_bwd with head_c input channels (head_c + 1 per block)
// for up to nb samples of npix_per pixels.
inline int64_t loss_partials(int64_t npix_per, int nb, int head_c) {
    const int64_t npix = npix_per * nb;
    return std::max<int64_t>({(int64_t)64 * nb, blocks256(npix), blocks256(npix) * (head_c + 1)});
}

}  // namespace nn
}  // namespace stm
