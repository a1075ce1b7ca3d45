// Temporally blocked red-black SOR for gfx950.
//
// Launch L of a solve advances every running channel by up to K full sweeps.
// Each workgroup owns an output tile; it loads the tile plus a halo of
// H = 2K + 4 cells, performs the K sweeps on chip and writes the output tile
// back.  A dependency reaches one cell per half-sweep plus one through a
// ghost refresh, so after K sweeps every cell at distance >= 2K + 1 from the
// loaded region is exact; the remaining halo covers the final ghost refresh
// and the residual (H is even so that loaded rows start on a 16-byte pair).
// Nothing beyond a channel's edge feeds a tile, so on an edge side the region
// needs no halo: it starts one cell outside the channel (two in y, keeping
// pairs aligned), which leaves the edge tiles H - 1 more output cells (each
// tile carries its region origin).
// The result is bit-identical to K separate global sweeps of
//     refresh all ghosts; red half-sweep; black half-sweep
// (same operations on the same operands) while the cache / HBM traffic drops
// by about K times.
//
// Schedule inside a launch: ghosts take the colour of their cell and read
// only fluid cells of the other colour, and a cell of colour c reads only
// cells of the other colour.  So the red ghosts (sources black) can be
// refreshed in the red phase and the black ghosts (sources red) in the black
// phase, each from values that the phase does not change -- the same values
// the start-of-sweep refresh sees.  Two barriers per sweep.
//
// Registers and LDS: the loaded region is EX x EY cells, EY = 2*HY.  Thread
// (strip s, pair m) owns rows [s*R, s*R + R) of the y-pair (2m, 2m+1) -- one
// red and one black cell per row.  E / W neighbours come from its own
// registers except at the strip ends; of N / S one is its own other cell.
// Every new value is published to LDS (split by absolute colour, one slot
// per (x, pair)) for threads that read it as N / S or across a strip end.
// R is even, so a tile's colour pattern is fixed by the parity of its x
// origin; each parity has its own compiled body without colour selects.
//
// No fences, no second kernel: launch L writes one residual partial per tile;
// every workgroup of launch L + 1 folds the previous launch's partials of its
// channel in tile order, so all take the same stop decision; the channel's
// first tile records it.  p ping-pongs between two buffers (workgroups read
// each other's halos).
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

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* wsum) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) s = s + wsum[w];
    __syncthreads();
    return s;
}

// Everything after the load, for a tile whose (even-x, even-y) cells have
// colour RP0.  Own cell (row r, slot q) has colour RP0 ^ (r & 1) ^ q.
template <int K, int R, int HY, int NS, int RP0>
struct Blk {
    static constexpr int NT = NS * HY;
    static constexpr int EX = NS * R;
    // each colour's LDS plane has HY guard slots before and after the EX x HY
    // region, so the neighbour reads of cells on the region's border stay in
    // bounds (their results are discarded); slot (lx, m) is at OFF + lx*HY + m
    static constexpr int OFF = HY;
    static constexpr int PS = EX * HY + 2 * HY;

    __device__ static __forceinline__ int col(int r, int q) { return RP0 ^ (r & 1) ^ q; }
    // own slot of colour c in row r
    __device__ static __forceinline__ int slot(int r, int c) { return c ^ RP0 ^ (r & 1); }

    // the four neighbours (colour col(r,q) ^ 1) of own cell (r, q)
    template <int r, int q>
    __device__ static __forceinline__ void nbrs(const double (*P)[PS],
                                                const double (&pv)[R][2], int lx0, int m,
                                                double& pn, double& ps, double& pe, double& pw) {
        constexpr int o = RP0 ^ (r & 1) ^ q ^ 1;
        const int lx = lx0 + r;
        const double x = P[o][OFF + lx * HY + m - 1 + 2 * q];
        pn = q ? pv[r][0] : x;
        ps = q ? x : pv[r][1];
        if constexpr (r + 1 < R) pe = pv[r + 1][q]; else pe = P[o][OFF + (lx + 1) * HY + m];
        if constexpr (r >= 1) pw = pv[r - 1][q]; else pw = P[o][OFF + (lx - 1) * HY + m];
    }

    // ghost pressure (src/grid/mod.rs:343-412) of own ghosts of colour g
    template <int r, int g>
    __device__ static __forceinline__ void refresh_row(double (*P)[PS], double (&pv)[R][2],
                                                       const uint8_t (&cc)[R][2], int lx0, int m) {
        constexpr int q = g ^ RP0 ^ (r & 1);
        const int e = cc[r][q] >> C_EDGE_SHIFT;
        if (!e) return;
        double pn, ps, pe, pw;
        nbrs<r, q>(P, pv, lx0, m, pn, ps, pe, pw);
        double v;
        switch (e) {
        case E_N: v = pn; break;
        case E_NE: v = (pn + pe) * 0.5; break;
        case E_E: v = pe; break;
        case E_SE: v = (ps + pe) * 0.5; break;
        case E_S: v = ps; break;
        case E_SW: v = (ps + pw) * 0.5; break;
        case E_W: v = pw; break;
        default: v = (pn + pw) * 0.5; break;   // E_NW
        }
        pv[r][q] = v;
        P[g][OFF + (lx0 + r) * HY + m] = v;
    }

    template <int g, int r = 0>
    __device__ static __forceinline__ void refresh(double (*P)[PS], double (&pv)[R][2],
                                                   const uint8_t (&cc)[R][2], int lx0, int m) {
        if constexpr (r < R) {
            refresh_row<r, g>(P, pv, cc, lx0, m);
            refresh<g, r + 1>(P, pv, cc, lx0, m);
        }
    }

    // Red-black update of own fluid cells of colour c.  Every LDS read of the
    // phase is of the other colour, which the phase does not change, so they
    // are issued first as one batch (x[r]: N or S of row r; xe / xw: E of the
    // last row and W of the first, across the strip ends) and the wave waits
    // once.  Branch-free: every own cell of the colour is computed and the
    // cells that are not updated (ghosts, solids, the region's border) keep
    // their value; bit 2r+q of `upd` says which are updated.
    template <int c>
    __device__ static __forceinline__ void update(double (*P)[PS], double (&pv)[R][2],
                                                  const double (&rr)[R][2], uint32_t upd,
                                                  int lx0, int m, double omw, double mid,
                                                  double rdx2, double rdy2) {
        constexpr int o = c ^ 1;
        double x[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int q = c ^ RP0 ^ (r & 1);
            x[r] = P[o][OFF + (lx0 + r) * HY + m - 1 + 2 * q];
        }
        const double xe = P[o][OFF + (lx0 + R) * HY + m];
        const double xw = P[o][OFF + (lx0 - 1) * HY + m];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int q = c ^ RP0 ^ (r & 1);
            // cells (r, q ^ 1), (r +- 1, q) have colour c ^ 1: unchanged here
            const double pn = q ? pv[r][0] : x[r];
            const double ps = q ? x[r] : pv[r][1];
            const double pe = r + 1 < R ? pv[r + 1 < R ? r + 1 : r][q] : xe;
            const double pw = r >= 1 ? pv[r >= 1 ? r - 1 : r][q] : xw;
            const double tt = fma(pe + pw, rdx2, (ps + pn) * rdy2);
            const double nv = fma(omw, pv[r][q], mid * (tt - rr[r][q]));
            pv[r][q] = ((upd >> (2 * r + q)) & 1u) ? nv : pv[r][q];
            P[c][OFF + (lx0 + r) * HY + m] = pv[r][q];
        }
    }

    // sum of r^2 over own fluid output cells (ghosts fresh)
    template <int r = 0, int q = 0>
    __device__ static __forceinline__ double resid(double (*P)[PS], const double (&pv)[R][2],
                                                   const double (&rr)[R][2],
                                                   const uint8_t (&cc)[R][2], int lx0, int m,
                                                   double rdx2, double rdy2, double acc) {
        if constexpr (r < R) {
            if ((cc[r][q] & (C_UPDATE | C_OUT)) == (C_UPDATE | C_OUT)) {
                double pn, ps, pe, pw;
                nbrs<r, q>(P, pv, lx0, m, pn, ps, pe, pw);
                const double pc = pv[r][q];
                const double a = (pe - pc) - (pc - pw);
                const double b = (ps - pc) - (pc - pn);
                const double res = fma(a, rdx2, b * rdy2) - rr[r][q];
                acc = acc + res * res;
            }
            if constexpr (q == 0) return resid<r, 1>(P, pv, rr, cc, lx0, m, rdx2, rdy2, acc);
            else return resid<r + 1, 0>(P, pv, rr, cc, lx0, m, rdx2, rdy2, acc);
        } else {
            return acc;
        }
    }

    __device__ static __forceinline__ double run(double (*P)[PS], double (&pv)[R][2],
                                                 const double (&rr)[R][2],
                                                 const uint8_t (&cc)[R][2], int lx0, int m,
                                                 int nsw, bool check, const ChanDesc& d) {
        const double omw = d.omw, mid = d.mid, rdx2 = d.rdx2, rdy2 = d.rdy2;
        uint32_t upd = 0;
        int ghost = 0;
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int q = 0; q < 2; q++) {
                P[col(r, q)][OFF + (lx0 + r) * HY + m] = pv[r][q];
                if (cc[r][q] & C_UPDATE) upd |= 1u << (2 * r + q);
                if (cc[r][q] >> C_EDGE_SHIFT) ghost = 1;
            }
        // ghost refreshes only in waves that own a ghost cell (wave-uniform)
        ghost = __ballot(ghost) != 0;
        __syncthreads();
        if (ghost) refresh<1>(P, pv, cc, lx0, m);   // black ghosts, for the first red half
        __syncthreads();
        for (int sw = 0; sw < nsw; sw++) {
            update<0>(P, pv, rr, upd, lx0, m, omw, mid, rdx2, rdy2);
            if (ghost) refresh<0>(P, pv, cc, lx0, m);   // red ghosts, for the black half
            __syncthreads();
            update<1>(P, pv, rr, upd, lx0, m, omw, mid, rdx2, rdy2);
            if (ghost && (check || sw + 1 < nsw)) refresh<1>(P, pv, cc, lx0, m);
            __syncthreads();
        }
        double acc = 0.0;
        if (check) {
            if (ghost) refresh<0>(P, pv, cc, lx0, m);   // red ghosts from the final black
            __syncthreads();
            acc = resid(P, pv, rr, cc, lx0, m, rdx2, rdy2, 0.0);
        }
        return acc;
    }
};

}  // namespace

template <int K, int R, int HY, int NS>
__global__ void __launch_bounds__(NS * HY)
k_rb_strip(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
           const TbTile* __restrict__ tiles, const uint8_t* __restrict__ flags,
           const double* __restrict__ rhs, double* __restrict__ pa, double* __restrict__ pb,
           const double* __restrict__ part_prev, double* __restrict__ part_cur,
           int32_t* __restrict__ active_out, int L, int mode)
{
    constexpr int NT = NS * HY;
    constexpr int H = 2 * K + 4;
    constexpr int EX = NS * R;
    constexpr int EY = 2 * HY;
    static_assert(NT % 64 == 0 && NT <= 1024, "workgroup size");
    static_assert(R % 2 == 0, "rows per thread must be even");
    static_assert(EX > 2 * H && EY > 2 * H, "tile smaller than its halo");

    __shared__ double P[2][EX * HY + 2 * HY];   // plus guard slots, see Blk::OFF
    __shared__ double wsum[NT / 64];

    const TbTile t = tiles[blockIdx.x];
    const int c = t.chan;
    const ChanDesc& d = cd[c];
    ChanState& s = cs[c];
    if (s.stop_launch >= 0) return;              // uniform: the channel is done
    const int tid = threadIdx.x;
    const bool check = mode & TB_CHECK;
    const double* __restrict__ pin = (L & 1) ? pb : pa;
    double* __restrict__ pout = (L & 1) ? pa : pb;

    const int ex0 = t.rx0, ey0 = t.ry0;          // region origin; ey0 is even
    const int nx = d.nx, ny = d.ny;
    const int64_t pitch = d.pitch, base = d.off;
    const int strip = tid / HY, m = tid - (tid / HY) * HY;
    const int lx0 = strip * R;
    const int gy = ey0 + 2 * m;                  // even: (gy, gy+1) is one aligned pair

    // ---- load (issued before the stop decision so the latencies overlap);
    // the solve is for rhs - mean (the all-Neumann problem needs zero mean)
    const double mean = s.rhs_mean;
    double pv[R][2], rr[R][2];
    uint8_t cc[R][2];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int lx = lx0 + r, gx = ex0 + lx;
        const bool row_in = gx >= 0 && gx < nx;
        const int64_t g = base + gx * pitch + gy;
        double2 pp = make_double2(0.0, 0.0), rh = make_double2(0.0, 0.0);
        uint8_t f0 = 0xff, f1 = 0xff;             // 0xff: outside the channel
        if (row_in && gy >= 0 && gy + 1 < ny) {
            pp = *reinterpret_cast<const double2*>(pin + g);
            rh = *reinterpret_cast<const double2*>(rhs + g);
            f0 = flags[g];
            f1 = flags[g + 1];
        } else if (row_in && gy >= 0 && gy < ny) {
            pp.x = pin[g];
            rh.x = rhs[g];
            f0 = flags[g];
        }
        pv[r][0] = pp.x;
        pv[r][1] = pp.y;
        rr[r][0] = rh.x - mean;
        rr[r][1] = rh.y - mean;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const uint8_t fl = q ? f1 : f0;
            const int ly = 2 * m + q, gyq = gy + q;
            uint8_t code = 0;
            if (fl != 0xff) {
                const bool inner = lx >= 1 && lx <= EX - 2 && ly >= 1 && ly <= EY - 2;
                if (inner && flag_fluid(fl) && gx >= 1 && gx <= nx - 2 && gyq >= 1 && gyq <= ny - 2)
                    code |= C_UPDATE;
                if (inner && !flag_fluid(fl)) code |= (uint8_t)(flag_edge(fl) << C_EDGE_SHIFT);
                if (gx >= t.x0 && gx < t.x0 + t.xn && gyq >= t.y0 && gyq < t.y0 + t.yn)
                    code |= C_OUT;
            }
            cc[r][q] = code;
        }
    }

    // ---- stop decision on the previous launch (identical in every tile)
    const int itermax = (s.budget > 0 && s.budget < d.itermax) ? s.budget : d.itermax;
    const int64_t sweeps = (int64_t)L * K;
    const int done = sweeps < itermax ? (int)sweeps : itermax;
    if (L > 0) {
        double acc = 0.0;
        if (check)
            for (int i = d.tb0 + tid; i < d.tb0 + d.ntb; i += NT) acc = acc + part_prev[i];
        const double res2 = block_sum<NT>(acc, wsum);
        const bool conv = check && res2 <= s.tol2;
        // a residual that is not a number never meets the test: stop there
        // (the tick then reports it) instead of sweeping to itermax
        const bool nan = check && res2 != res2;
        if (conv || nan || done >= itermax) {
            if (t.idx == d.tb0 && tid == 0) {
                s.stop_launch = L - 1;
                s.sor_iters = (uint32_t)done;
                s.pbuf = L & 1;
                s.active = 0;
                s.exhausted = check && !conv;
                if (check) {
                    s.res2 = res2;
                    s.norm2 = res2 / d.fluid_cells;
                }
                active_out[c] = 0;
            }
            return;
        }
    }
    const int nsw = itermax - done < K ? itermax - done : K;

    const double acc = (ex0 & 1) ? Blk<K, R, HY, NS, 1>::run(P, pv, rr, cc, lx0, m, nsw, check, d)
                                 : Blk<K, R, HY, NS, 0>::run(P, pv, rr, cc, lx0, m, nsw, check, d);

    // ---- write the output tile
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t g = base + (int64_t)(ex0 + lx0 + r) * pitch + gy;
        const bool o0 = cc[r][0] & C_OUT, o1 = cc[r][1] & C_OUT;
        if (o0 && o1) *reinterpret_cast<double2*>(pout + g) = make_double2(pv[r][0], pv[r][1]);
        else if (o0) pout[g] = pv[r][0];
        else if (o1) pout[g + 1] = pv[r][1];
    }
    const double b = block_sum<NT>(acc, wsum);
    if (tid == 0) part_cur[t.idx] = b;
}

// After a blocked solve: channels whose latest p is in the second buffer
// copy it back.
__global__ void __launch_bounds__(TILE_THREADS)
k_pbuf_fix(const ChanDesc* __restrict__ cd, const ChanState* __restrict__ cs,
           const stm::Tile* __restrict__ tiles, const double* __restrict__ pb,
           double* __restrict__ pa)
{
    const stm::Tile t = tiles[blockIdx.x];
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

// (index, K sweeps per launch, R rows per thread, HY pairs per row, NS strips)
// loaded region (NS*R) x (2*HY); output tile (NS*R - 2H) x (2*HY - 2H) with
// H = 2K + 4; LDS NS*R*2*HY*8 bytes; workgroup NS*HY threads.
#define STM_TB_VARIANTS(V)   \
    V(0, 1, 4, 32, 16)       \
    V(1, 2, 4, 32, 16)       \
    V(2, 2, 4, 64, 16)       \
    V(3, 4, 4, 32, 16)       \
    V(4, 4, 4, 64, 16)       \
    V(5, 4, 6, 64, 16)       \
    V(6, 8, 6, 64, 16)       \
    V(7, 2, 2, 64, 16)       \
    V(8, 8, 12, 64, 8)       \
    V(9, 6, 6, 48, 16)

#define V_ENTRY(i, K, R, HY, NS) \
    {K, NS * R - 2 * (2 * K + 4), 2 * HY - 2 * (2 * K + 4), NS * R, 2 * HY},
const TbVariant TB_VARIANTS[] = {STM_TB_VARIANTS(V_ENTRY)};
#undef V_ENTRY
const int N_TB_VARIANTS = sizeof(TB_VARIANTS) / sizeof(TB_VARIANTS[0]);

void launch_rb_tb(hipStream_t st, int variant, int ntb, const ChanDesc* cd, ChanState* cs,
                  const TbTile* tiles, const uint8_t* flags, const double* rhs, double* pa,
                  double* pb, const double* part_prev, double* part_cur, int32_t* active,
                  int L, int mode) {
#define V_CASE(i, K, R, HY, NS)                                                              \
    case i:                                                                                  \
        k_rb_strip<K, R, HY, NS><<<ntb, NS * HY, 0, st>>>(cd, cs, tiles, flags, rhs, pa, pb, \
                                                          part_prev, part_cur, active, L,    \
                                                          mode);                             \
        break;
    switch (variant) {
        STM_TB_VARIANTS(V_CASE)
    default: break;
    }
#undef V_CASE
}

void launch_pbuf_fix(hipStream_t st, int ntiles, int nchan, const ChanDesc* cd, ChanState* cs,
                     const stm::Tile* tiles, const double* pb, double* pa) {
    k_pbuf_fix<<<ntiles, dim3(TILE_Y, TILE_X), 0, st>>>(cd, cs, tiles, pb, pa);
    k_pbuf_reset<<<(nchan + 63) / 64, 64, 0, st>>>(nchan, cs);
}

}  // namespace stm
