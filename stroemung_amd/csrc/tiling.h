// Tiling of one axis for the blocked red-black solver (plain C++, shared by
// the library and its CPU tests).
#pragma once

#include <vector>

namespace stm {

// Output parts of one axis of n cells for the blocked solver: region E
// cells, halo H, part starts and region origins multiples of q (1 in x, 2 in
// y).  A middle part has H cells of halo on both sides (at most E - 2H
// cells).  The first part's region starts q cells before the channel and the
// last part's region reaches past its end, where nothing feeds the tile, so
// those parts hold up to E - q - H and E - 1 - H cells.
struct AxisPart {
    int start, len, origin;
};
inline std::vector<AxisPart> plan_axis(int n, int E, int H, int q) {
    std::vector<AxisPart> out;
    if (n <= E - q - 1) {
        out.push_back(AxisPart{0, n, -q});
        return out;
    }
    const int o = E - 2 * H, e0 = E - q - H, e1 = E - 1 - H;
    int P = 2;
    while (e0 + (P - 2) * o + e1 < n) P++;
    // every part but the last at its largest; the last gets the rest, which
    // is at most e1 (P is enough) and more than H - 1 (P is the least: with
    // one part fewer, n > e0 + (P - 3) o + e1, so the rest exceeds e1 - o)
    int start = 0;
    for (int k = 0; k < P - 1; k++) {
        const int len = k == 0 ? e0 : o;
        out.push_back(AxisPart{start, len, k == 0 ? -q : start - H});
        start += len;
    }
    out.push_back(AxisPart{start, n - start, start - H});
    return out;
}

}  // namespace stm
