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
// cells of the other colour.  So the red ghosts (sources black) are
// refreshed in the red phase and the black ghosts (sources red) in the black
// phase, each from values that the phase does not change -- the same values
// the start-of-sweep refresh sees -- and from the same neighbour reads as the
// phase's updates.  Two barriers per sweep.
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

// For a ghost cell, the fluid neighbours its pressure copies
// (src/grid/mod.rs:343-412): N, S, E, W; one bit for a straight edge, a
// vertical and a horizontal one for a corner, whose ghost is their mean.
enum : uint32_t { C_GN = 1, C_GS = 2, C_GE = 4, C_GW = 8 };

__device__ __forceinline__ uint32_t ghost_bits(uint8_t edge) {
    switch (edge) {
    case E_N: return C_GN;
    case E_NE: return C_GN | C_GE;
    case E_E: return C_GE;
    case E_SE: return C_GS | C_GE;
    case E_S: return C_GS;
    case E_SW: return C_GS | C_GW;
    case E_W: return C_GW;
    case E_NW: return C_GN | C_GW;
    default: return 0;
    }
}

// What a thread knows of its 2R own cells, cell i = 2r + q, packed so the
// codes take a few registers: bit i of `upd` = updated by SOR, of `out` = in
// the output tile; 4 ghost bits per cell.
template <int R>
struct Codes {
    static constexpr int NG = (2 * R + 7) / 8;
    uint32_t upd, out;
    uint32_t g[NG];
    __device__ __forceinline__ uint32_t ghost(int i) const {
        return (g[i >> 3] >> (4 * (i & 7))) & 15u;
    }
};

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

    // The neighbour values of colour c ^ 1 that own cells of colour c read:
    // N / S from LDS (x[r]; one of N / S is the thread's own other cell),
    // E / W from the thread's registers except across the strip ends (xe,
    // xw).  The LDS reads are issued as one batch, so the wave waits once.
    template <int c>
    __device__ static __forceinline__ void reads(double (*P)[PS], int lx0, int m, double (&x)[R],
                                                 double& xe, double& xw) {
        constexpr int o = c ^ 1;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int q = c ^ RP0 ^ (r & 1);
            x[r] = P[o][OFF + (lx0 + r) * HY + m - 1 + 2 * q];
        }
        xe = P[o][OFF + (lx0 + R) * HY + m];
        xw = P[o][OFF + (lx0 - 1) * HY + m];
    }

    // Red-black update of own fluid cells of colour c (bit 2r+q of cc.upd);
    // branch-free: every own cell of the colour is computed and the cells
    // that are not updated (ghosts, solids, the region's border) keep their
    // value.  Every value read has colour c ^ 1, which the phase does not
    // change.
    template <int c>
    __device__ static __forceinline__ void update(double (*P)[PS], double (&pv)[R][2],
                                                  const double (&rr)[R][2], const Codes<R>& cc,
                                                  int lx0, int m, double omw, double mid,
                                                  double rdx2, double rdy2) {
        double x[R], xe, xw;
        reads<c>(P, lx0, m, x, xe, xw);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int q = c ^ RP0 ^ (r & 1);
            const double pn = q ? pv[r][0] : x[r];
            const double ps = q ? x[r] : pv[r][1];
            const double pe = r + 1 < R ? pv[r + 1 < R ? r + 1 : r][q] : xe;
            const double pw = r >= 1 ? pv[r >= 1 ? r - 1 : r][q] : xw;
            const double tt = fma(pe + pw, rdx2, (ps + pn) * rdy2);
            const double nv = fma(omw, pv[r][q], mid * (tt - rr[r][q]));
            pv[r][q] = ((cc.upd >> (2 * r + q)) & 1u) ? nv : pv[r][q];
            P[c][OFF + (lx0 + r) * HY + m] = pv[r][q];
        }
    }

    // Ghost cells of colour c take their fluid neighbours' pressure
    // (src/grid/mod.rs:343-412: one neighbour, or (a + b) * 0.5 of a
    // vertical and a horizontal one) -- the reference's start-of-sweep copy:
    // a ghost of colour c reads only fluid cells of colour c ^ 1 and is read
    // only by them, so refreshing it in colour c's phase gives the values
    // that copy gives.  Branch-free over the thread's cells, batched reads;
    // run only by waves that hold a ghost cell (the choice is wave-uniform).
    template <int c>
    __device__ static __forceinline__ void ghosts(double (*P)[PS], double (&pv)[R][2],
                                                  const Codes<R>& cc, int lx0, int m) {
        double x[R], xe, xw;
        reads<c>(P, lx0, m, x, xe, xw);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int q = c ^ RP0 ^ (r & 1);
            const uint32_t g = cc.ghost(2 * r + q);
            const double pn = q ? pv[r][0] : x[r];
            const double ps = q ? x[r] : pv[r][1];
            const double pe = r + 1 < R ? pv[r + 1 < R ? r + 1 : r][q] : xe;
            const double pw = r >= 1 ? pv[r >= 1 ? r - 1 : r][q] : xw;
            const double a = (g & C_GN) ? pn : ps;
            const double b = (g & C_GE) ? pe : pw;
            const bool vert = g & (C_GN | C_GS), horz = g & (C_GE | C_GW);
            const double gv = vert ? (horz ? (a + b) * 0.5 : a) : b;
            pv[r][q] = g ? gv : pv[r][q];
            P[c][OFF + (lx0 + r) * HY + m] = pv[r][q];
        }
    }

    // sum of r^2 over own fluid output cells (ghosts fresh); the LDS reads
    // (both colours' N / S neighbours and the strip-end E / W ones) first
    __device__ static __forceinline__ double resid(double (*P)[PS], const double (&pv)[R][2],
                                                   const double (&rr)[R][2],
                                                   const Codes<R>& cc, int lx0, int m,
                                                   double rdx2, double rdy2) {
        double xn[R], xs[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            xn[r] = P[col(r, 0) ^ 1][OFF + (lx0 + r) * HY + m - 1];   // N of (r, 0)
            xs[r] = P[col(r, 1) ^ 1][OFF + (lx0 + r) * HY + m + 1];   // S of (r, 1)
        }
        double xe[2], xw[2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            xe[q] = P[col(R - 1, q) ^ 1][OFF + (lx0 + R) * HY + m];
            xw[q] = P[col(0, q) ^ 1][OFF + (lx0 - 1) * HY + m];
        }
        double acc = 0.0;
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const double pn = q ? pv[r][0] : xn[r];
                const double ps = q ? xs[r] : pv[r][1];
                const double pe = r + 1 < R ? pv[r + 1 < R ? r + 1 : r][q] : xe[q];
                const double pw = r >= 1 ? pv[r >= 1 ? r - 1 : r][q] : xw[q];
                const double pc = pv[r][q];
                const double a = (pe - pc) - (pc - pw);
                const double b = (ps - pc) - (pc - pn);
                const double res = fma(a, rdx2, b * rdy2) - rr[r][q];
                const bool in = ((cc.upd & cc.out) >> (2 * r + q)) & 1u;
                acc = acc + (in ? res * res : 0.0);
            }
        return acc;
    }

    __device__ static __forceinline__ double run(double (*P)[PS], double (&pv)[R][2],
                                                 const double (&rr)[R][2],
                                                 const Codes<R>& cc, int lx0, int m,
                                                 int nsw, bool check, const ChanDesc& d) {
        const double omw = d.omw, mid = d.mid, rdx2 = d.rdx2, rdy2 = d.rdy2;
        int ghost = 0;
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int q = 0; q < 2; q++) P[col(r, q)][OFF + (lx0 + r) * HY + m] = pv[r][q];
#pragma unroll
        for (int k = 0; k < Codes<R>::NG; k++) ghost |= cc.g[k] != 0;
        // the ghost refresh only in waves that own a ghost cell (wave-uniform)
        ghost = __ballot(ghost) != 0;
        __syncthreads();
        if (ghost) ghosts<1>(P, pv, cc, lx0, m);   // black ghosts, for the first red half
        __syncthreads();
        Codes<R> ck = cc;
        for (int sw = 0; sw < nsw; sw++) {
            // the lane masks the codes give are cheap to form: form them in each
            // sweep rather than keep them all live in scalar registers
            asm volatile("" : "+v"(ck.upd));
#pragma unroll
            for (int k = 0; k < Codes<R>::NG; k++) asm volatile("" : "+v"(ck.g[k]));
            update<0>(P, pv, rr, ck, lx0, m, omw, mid, rdx2, rdy2);
            if (ghost) ghosts<0>(P, pv, ck, lx0, m);   // red ghosts, for the black half
            __syncthreads();
            update<1>(P, pv, rr, ck, lx0, m, omw, mid, rdx2, rdy2);
            if (ghost && (check || sw + 1 < nsw)) ghosts<1>(P, pv, ck, lx0, m);
            __syncthreads();
        }
        double acc = 0.0;
        if (check) {
            if (ghost) ghosts<0>(P, pv, cc, lx0, m);   // red ghosts from the final black
            __syncthreads();
            acc = resid(P, pv, rr, cc, lx0, m, rdx2, rdy2);
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
    Codes<R> cc;
    cc.upd = cc.out = 0;
#pragma unroll
    for (int k = 0; k < Codes<R>::NG; k++) cc.g[k] = 0;
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
            const uint16_t ff = *reinterpret_cast<const uint16_t*>(flags + g);   // g is even
            f0 = (uint8_t)(ff & 0xff);
            f1 = (uint8_t)(ff >> 8);
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
            const int ly = 2 * m + q, gyq = gy + q, i = 2 * r + q;
            if (fl != 0xff) {
                const bool inner = lx >= 1 && lx <= EX - 2 && ly >= 1 && ly <= EY - 2;
                if (inner && flag_fluid(fl) && gx >= 1 && gx <= nx - 2 && gyq >= 1 && gyq <= ny - 2)
                    cc.upd |= 1u << i;
                if (inner && !flag_fluid(fl)) cc.g[i >> 3] |= ghost_bits(flag_edge(fl)) << (4 * (i & 7));
                if (gx >= t.x0 && gx < t.x0 + t.xn && gyq >= t.y0 && gyq < t.y0 + t.yn)
                    cc.out |= 1u << i;
            }
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
        const bool o0 = (cc.out >> (2 * r)) & 1u, o1 = (cc.out >> (2 * r + 1)) & 1u;
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
