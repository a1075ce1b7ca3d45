// Temporally blocked red-black SOR for gfx950.
//
// One launch advances every active channel by up to K full sweeps (ghost
// refresh, red half, black half).  Each workgroup owns an output tile of at
// most OXM x OYM cells; it loads the tile plus a halo of H = 2K + 3 cells
// into LDS, performs the K sweeps there, and writes the output tile back.
// A dependency reaches one cell per half-sweep plus one through a ghost
// refresh, so after K sweeps every cell at distance >= 2K + 1 from the loaded
// region is exact; the two extra cells cover the final ghost refresh and
// the residual.  The result is therefore bit-identical to K separate global
// sweeps (same operations, same operands), while HBM / cache traffic drops
// by roughly K times.
//
// LDS layout: p split by absolute colour (x + y) & 1, each colour with one
// slot per (x, y-pair).  A thread owns pairs (lx, 2m), (lx, 2m+1); in a
// half-sweep it updates the one of colour c, reading the four neighbours of
// the other colour from consecutive slots -- no divergence, no bank
// conflicts.  rhs and per-cell codes stay in registers.
//
// p ping-pongs between two buffers (tiles read neighbours' halos, so a
// launch may not write what another workgroup of the same launch reads).
// The last workgroup of a channel (atomic ticket) folds the per-tile
// residual partials in tile order, updates the sweep count and stops the
// channel when sum r^2 <= tol2 or at itermax.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "stm_internal.h"

namespace stm {

namespace {

enum : uint8_t { C_UPDATE = 1, C_OUT = 2, C_EDGE_SHIFT = 2 };   // edge kind in bits 2..5

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
    return v;
}

}  // namespace

template <int OXM, int OYM, int K, int NT>
__global__ void __launch_bounds__(NT)
k_rb_tb(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
        const TbTile* __restrict__ tiles, const uint8_t* __restrict__ flags,
        const double* __restrict__ rhs, double* __restrict__ pa, double* __restrict__ pb,
        double* __restrict__ part, int32_t* __restrict__ counters,
        int32_t* __restrict__ active_out, int check)
{
    constexpr int H = 2 * K + 3;
    constexpr int EX = OXM + 2 * H;
    constexpr int EYR = OYM + 2 * H;
    constexpr int EY = EYR + (EYR & 1);
    constexpr int HY = EY / 2;
    constexpr int NPAIR = EX * HY;
    constexpr int PPT = (NPAIR + NT - 1) / NT;
    constexpr int NW = NT / 64;

    __shared__ double P[2][NPAIR];
    __shared__ double wsum[NW];
    __shared__ int last;

    const TbTile t = tiles[blockIdx.x];
    const int c = t.chan;
    const ChanDesc& d = cd[c];
    ChanState& s = cs[c];
    if (!s.active) return;                       // uniform per workgroup
    const int left = d.itermax - (int)s.sor_iters;
    const int nsw = left < K ? left : K;
    const int par = s.pbuf;
    const double* __restrict__ pin = par ? pb : pa;
    double* __restrict__ pout = par ? pa : pb;
    const int ex0 = t.x0 - H, ey0 = t.y0 - H;
    const int nx = d.nx, ny = d.ny;
    const int64_t pitch = d.pitch, base = d.off;
    const int tid = threadIdx.x;

    // ---- load: p into LDS, rhs and cell codes into registers
    double rr[PPT][2];
    uint8_t code[PPT][2];
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        const int k = tid + j * NT;
        const int lx = k / HY, m = k - (k / HY) * HY;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int ly = 2 * m + q;
            const int gx = ex0 + lx, gy = ey0 + ly;
            const bool inside = k < NPAIR && ly < EYR && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
            double pv = 0.0, rv = 0.0;
            uint8_t cc = 0;
            if (inside) {
                const int64_t g = base + gx * pitch + gy;
                pv = pin[g];
                rv = rhs[g];
                const uint8_t fl = flags[g];
                const bool inner = lx >= 1 && lx <= EX - 2 && ly >= 1 && ly <= EYR - 2;
                if (inner && flag_fluid(fl) && gx >= 1 && gx <= nx - 2 && gy >= 1 && gy <= ny - 2)
                    cc |= C_UPDATE;
                if (inner && !flag_fluid(fl)) cc |= (uint8_t)(flag_edge(fl) << C_EDGE_SHIFT);
                if (lx >= H && lx < H + t.xn && ly >= H && ly < H + t.yn) cc |= C_OUT;
            }
            rr[j][q] = rv;
            code[j][q] = cc;
            if (k < NPAIR) P[(ex0 + lx + ey0 + ly) & 1][k] = pv;
        }
    }
    __syncthreads();

    const double omw = d.omw, mid = d.mid, rdx2 = d.rdx2, rdy2 = d.rdy2;

    // ghost pressure of every refreshable boundary cell (src/grid/mod.rs:343-412)
    auto refresh = [&]() {
#pragma unroll
        for (int j = 0; j < PPT; j++) {
            const int k = tid + j * NT;
            const int lx = k / HY, m = k - (k / HY) * HY;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int e = code[j][q] >> C_EDGE_SHIFT;
                if (!e) continue;
                const int col = (ex0 + lx + ey0 + 2 * m + q) & 1, o = col ^ 1;
                const double* Q = P[o];
                const int r = lx * HY + m;
                double v;
                switch (e) {
                case E_N: v = Q[r - 1 + q]; break;
                case E_NE: v = (Q[r - 1 + q] + Q[r + HY]) * 0.5; break;
                case E_E: v = Q[r + HY]; break;
                case E_SE: v = (Q[r + q] + Q[r + HY]) * 0.5; break;
                case E_S: v = Q[r + q]; break;
                case E_SW: v = (Q[r + q] + Q[r - HY]) * 0.5; break;
                case E_W: v = Q[r - HY]; break;
                default: v = (Q[r - 1 + q] + Q[r - HY]) * 0.5; break;   // E_NW
                }
                P[col][k] = v;
            }
        }
        __syncthreads();
    };

    for (int sw = 0; sw < nsw; sw++) {
        refresh();
#pragma unroll
        for (int colour = 0; colour < 2; colour++) {
            const int o = colour ^ 1;
#pragma unroll
            for (int j = 0; j < PPT; j++) {
                const int k = tid + j * NT;
                const int lx = k / HY, m = k - (k / HY) * HY;
                const int q = colour ^ ((ex0 + lx + ey0) & 1);
                const uint8_t cc = q ? code[j][1] : code[j][0];
                if (!(cc & C_UPDATE)) continue;
                const double* Q = P[o];
                const int r = lx * HY + m;
                const double pe = Q[r + HY], pw = Q[r - HY];
                const double pn = Q[r - 1 + q], ps = Q[r + q];
                const double rh = q ? rr[j][1] : rr[j][0];
                const double tt = fma(pe + pw, rdx2, (ps + pn) * rdy2);
                P[colour][k] = fma(omw, P[colour][k], mid * (tt - rh));
            }
            __syncthreads();
        }
    }

    // ---- residual of the new state on the output cells
    double acc = 0.0;
    if (check) {
        refresh();
#pragma unroll
        for (int j = 0; j < PPT; j++) {
            const int k = tid + j * NT;
            const int lx = k / HY, m = k - (k / HY) * HY;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const uint8_t cc = code[j][q];
                if ((cc & (C_UPDATE | C_OUT)) != (C_UPDATE | C_OUT)) continue;
                const int col = (ex0 + lx + ey0 + 2 * m + q) & 1, o = col ^ 1;
                const double* Q = P[o];
                const int r = lx * HY + m;
                const double pc = P[col][k];
                const double a = (Q[r + HY] - pc) - (pc - Q[r - HY]);
                const double b = (Q[r + q] - pc) - (pc - Q[r - 1 + q]);
                const double res = fma(a, rdx2, b * rdy2) - rr[j][q];
                acc = acc + res * res;
            }
        }
    }

    // ---- write the output tile
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        const int k = tid + j * NT;
        const int lx = k / HY, m = k - (k / HY) * HY;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            if (!(code[j][q] & C_OUT)) continue;
            const int ly = 2 * m + q;
            const int64_t g = base + (int64_t)(ex0 + lx) * pitch + (ey0 + ly);
            pout[g] = P[(ex0 + lx + ey0 + ly) & 1][k];
        }
    }

    // ---- per-tile partial, then the channel's last workgroup folds them
    acc = wave_sum(acc);
    if ((tid & 63) == 0) wsum[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
        double b = 0.0;
        for (int w = 0; w < NW; w++) b = b + wsum[w];
        part[(int64_t)blockIdx.x * NPART + 2] = b;
        __threadfence();
        last = atomicAdd(&counters[c], 1) == d.ntb - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double sum = 0.0;
    for (int i = d.tb0 + tid; i < d.tb0 + d.ntb; i += NT)
        sum = sum + __builtin_nontemporal_load(part + (int64_t)i * NPART + 2);
    sum = wave_sum(sum);
    if ((tid & 63) == 0) wsum[tid >> 6] = sum;
    __syncthreads();
    if (tid == 0) {
        double res2 = 0.0;
        for (int w = 0; w < NW; w++) res2 = res2 + wsum[w];
        s.sor_iters += (uint32_t)nsw;
        s.pbuf = par ^ 1;
        counters[c] = 0;
        if (check) {
            s.res2 = res2;
            s.norm2 = res2 / d.fluid_cells;
            if (res2 <= s.tol2) {
                s.active = 0;
                s.exhausted = 0;
            }
        }
        if (s.active && (int)s.sor_iters >= d.itermax) {
            s.active = 0;
            s.exhausted = check ? 1 : 0;
        }
        active_out[c] = s.active;
    }
}

// After a blocked solve: channels whose latest p is in the second buffer
// copy it back.
__global__ void __launch_bounds__(TILE_THREADS)
k_pbuf_fix(const ChanDesc* __restrict__ cd, const ChanState* __restrict__ cs,
           const Tile* __restrict__ tiles, const double* __restrict__ pb, double* __restrict__ pa)
{
    const Tile t = tiles[blockIdx.x];
    if (!cs[t.chan].pbuf) return;
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    if (x >= d.nx || y >= d.ny) return;
    const int64_t i = d.off + x * (int64_t)d.pitch + y;
    pa[i] = pb[i];
}

__global__ void k_pbuf_reset(int nchan, ChanState* __restrict__ cs) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nchan) cs[c].pbuf = 0;
}

// ---------------------------------------------------------------- variants

template <int OXM, int OYM, int K>
static void launch_variant(hipStream_t st, int ntb, const ChanDesc* cd, ChanState* cs,
                           const TbTile* tiles, const uint8_t* flags, const double* rhs,
                           double* pa, double* pb, double* part, int32_t* counters,
                           int32_t* active, int check) {
    k_rb_tb<OXM, OYM, K, 256><<<ntb, 256, 0, st>>>(cd, cs, tiles, flags, rhs, pa, pb, part,
                                                    counters, active, check);
}

// Instantiated variants: sweeps per launch K and the largest output tile.
const TbVariant TB_VARIANTS[] = {
    {1, 64, 64}, {2, 64, 64}, {4, 64, 64}, {8, 48, 48},
    {2, 32, 64}, {4, 32, 64}, {4, 48, 96}, {8, 32, 64},
};
const int N_TB_VARIANTS = sizeof(TB_VARIANTS) / sizeof(TB_VARIANTS[0]);

void launch_rb_tb(hipStream_t st, int variant, int ntb, const ChanDesc* cd, ChanState* cs,
                  const TbTile* tiles, const uint8_t* flags, const double* rhs, double* pa,
                  double* pb, double* part, int32_t* counters, int32_t* active, int check) {
#define V(i, K, X, Y) \
    case i: launch_variant<X, Y, K>(st, ntb, cd, cs, tiles, flags, rhs, pa, pb, part, counters, active, check); break;
    switch (variant) {
        V(0, 1, 64, 64)
        V(1, 2, 64, 64)
        V(2, 4, 64, 64)
        V(3, 8, 48, 48)
        V(4, 2, 32, 64)
        V(5, 4, 32, 64)
        V(6, 4, 48, 96)
        V(7, 8, 32, 64)
    default: break;
    }
#undef V
}

void launch_pbuf_fix(hipStream_t st, int ntiles, int nchan, const ChanDesc* cd, ChanState* cs,
                     const Tile* tiles, const double* pb, double* pa) {
    k_pbuf_fix<<<ntiles, dim3(TILE_Y, TILE_X), 0, st>>>(cd, cs, tiles, pb, pa);
    k_pbuf_reset<<<(nchan + 63) / 64, 64, 0, st>>>(nchan, cs);
}

}  // namespace stm
