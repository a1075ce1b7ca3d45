// gfx950 kernels for the pressure Poisson solve (POISSON of NaSt2D):
//  * the reference-compatible lexicographic SOR, run as a diagonal wavefront
//    inside one workgroup per channel -- bit-exact with the reference's
//    x-outer / y-inner in-place sweep, including its unmasked sequential
//    residual fold and stopping rule (src/simulation.rs:216-285);
//  * the set-up of the red-black solve (its sweeps live in kernels_rbtb.hip)
//    and the ghost refresh used by the stage entry points.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "stm_internal.h"

namespace stm {

__device__ __forceinline__ double ghost_value(const double* __restrict__ p, int64_t i,
                                              int64_t P, uint8_t edge) {
    // src/grid/mod.rs:343-412; (a + b) / 2.0 equals (a + b) * 0.5 exactly.
    switch (edge) {
    case E_N: return p[i - 1];
    case E_NE: return (p[i - 1] + p[i + P]) * 0.5;
    case E_E: return p[i + P];
    case E_SE: return (p[i + 1] + p[i + P]) * 0.5;
    case E_S: return p[i + 1];
    case E_SW: return (p[i + 1] + p[i - P]) * 0.5;
    case E_W: return p[i - P];
    case E_NW: return (p[i - 1] + p[i - P]) * 0.5;
    default: return p[i];
    }
}

// Ghost pressure of boundary cells (all colours, or one).  Reads fluid cells
// only and writes boundary cells only, so it is order-free.
__global__ void __launch_bounds__(256)
k_ghost(int nbtot, const BEntry* __restrict__ be, const ChanState* __restrict__ cs,
        double* __restrict__ p, int colour, int only_active)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbtot) return;
    const BEntry e = be[k];
    if (e.edge == E_NONE) return;
    if (colour >= 0 && e.colour != colour) return;
    if (only_active && !cs[e.chan].active) return;
    p[e.gidx] = ghost_value(p, e.gidx, e.pitch, e.edge);
}

// Sum slot `slot` of the partials of one channel's tiles in a fixed order.
__device__ double chan_sum(const ChanDesc& d, const double* __restrict__ part, int slot,
                           double* sh) {
    const int tid = threadIdx.x;
    double acc = 0.0;
    for (int k = d.tile0 + tid; k < d.tile0 + d.ntiles; k += blockDim.x)
        acc = acc + part[(int64_t)k * NPART + slot];
    sh[tid] = acc;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (tid < st) sh[tid] = sh[tid] + sh[tid + st];
        __syncthreads();
    }
    return sh[0];
}

// Per-tile sums over interior fluid cells: pass 0 sum of rhs (slot 0),
// pass 1 sum of (rhs - mean)^2 (slot 1).
__global__ void __launch_bounds__(TILE_THREADS)
k_rhs_sums(const ChanDesc* __restrict__ cd, const ChanState* __restrict__ cs,
           const Tile* __restrict__ tiles, const uint8_t* __restrict__ flags,
           const double* __restrict__ rhs, double* __restrict__ part, int pass)
{
    __shared__ double sh[TILE_THREADS];
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (x >= 1 && x <= d.nx - 2 && y >= 1 && y <= d.ny - 2) {
        const int64_t i = d.off + x * (int64_t)d.pitch + y;
        if (flag_fluid(flags[i])) {
            const double r = pass ? rhs[i] - cs[t.chan].rhs_mean : rhs[i];
            v = pass ? r * r : r;
        }
    }
    sh[tid] = v;
    __syncthreads();
    for (int st = TILE_THREADS / 2; st > 0; st >>= 1) {
        if (tid < st) sh[tid] = sh[tid] + sh[tid + st];
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x * NPART + pass] = sh[0];
}

// The mean of rhs over interior fluid cells.  The pressure problem has
// Neumann conditions all round (every boundary's ghost copies its fluid
// neighbour), so it is solvable only for rhs of zero mean; with inflow and
// outflow the mean is not zero and no iteration can push the residual below
// it.  The red-black solver solves for rhs - mean.
__global__ void __launch_bounds__(256)
k_rhs_mean(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
           const double* __restrict__ part)
{
    __shared__ double sh[256];
    const ChanDesc& d = cd[blockIdx.x];
    const double sum = chan_sum(d, part, 0, sh);
    if (threadIdx.x == 0) cs[blockIdx.x].rhs_mean = d.n_inner > 0.0 ? sum / d.n_inner : 0.0;
}

// Start of a red-black solve: tol2 = eps^2 * sum((rhs - mean)^2) over
// interior fluid cells.
__global__ void __launch_bounds__(256)
k_solve_begin(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
              const double* __restrict__ part, int32_t* __restrict__ active_out)
{
    __shared__ double sh[256];
    const ChanDesc& d = cd[blockIdx.x];
    const double rsum2 = chan_sum(d, part, 1, sh);
    if (threadIdx.x == 0) {
        ChanState& s = cs[blockIdx.x];
        const bool rb = d.solver == SOLVER_RB;
        s.rhs2 = rsum2;
        s.tol2 = d.eps2 * rsum2;
        s.sor_iters = 0;
        s.exhausted = 0;
        s.active = (rb && d.itermax > 0) ? 1 : 0;
        s.stop_launch = s.active ? -1 : 0;
        s.pbuf = 0;
        s.budget = 0;
        active_out[blockIdx.x] = s.active;
    }
}

// ------------------------------------------------ lexicographic (reference)

// src/math.rs:176-186, squared -- one term of the reference's norm fold.
__device__ __forceinline__ double ref_residual_sq(const double* __restrict__ p,
                                                  const double* __restrict__ rhs, int64_t i,
                                                  int64_t P, double dx2, double dy2) {
    const double pc = p[i];
    const double part1 = ((p[i + P] - pc) - (pc - p[i - P])) / dx2;
    const double part2 = ((p[i + 1] - pc) - (pc - p[i - 1])) / dy2;
    const double r = (part1 + part2) - rhs[i];
    return r * r;
}

// acc + s[0] + s[1] + ... + s[n - 1], added left to right, from LDS: every
// lane of the wave reads the same address (a broadcast, no bank conflict),
// sixteen reads in flight, then their additions, one dependent add per term.
__device__ __forceinline__ double lds_ordered_sum(double acc, const double* s, int n) {
    int k = 0;
    for (; k + 16 <= n; k += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = s[k + u];
#pragma unroll
        for (int u = 0; u < 16; u++) acc = acc + v[u];
    }
    for (; k < n; k++) acc = acc + s[k];
    return acc;
}

// The sequential sum from 0.0 of the interior terms t[(x, y)] >= 0, x-major
// (x outer, y inner), by the first wave of the workgroup: the wave loads 64
// consecutive terms of a row at once, eight such chunks per batch, puts the
// batch in `stage` (512 doubles of LDS) in that order and adds it term by
// term, so the additions are the sequential ones in the same order, bit for
// bit.  Lanes past the end of a row add +0.0, which leaves the sum unchanged:
// it starts at +0.0 and adds squares, so it is never -0.0 (a NaN or infinity
// stays what it is).  A wave's LDS accesses complete in order, so the batch
// it stored is what it reads back (the wavefront-scope fences keep the
// compiler from moving them).
__device__ __forceinline__ double ordered_row_sum(const double* __restrict__ t, int64_t base,
                                                  int64_t P, int nxi, int nyi, double* stage) {
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
    if (nxi <= 0 || nyi <= 0) return acc;
    const int64_t chunks = (int64_t)nxi * ((nyi + 63) / 64);
    int x = 1, y0 = 0;   // the next chunk to load: row x, terms 1 + y0 ..
    // that chunk: every lane loads (an address clamped into the channel, so
    // the loads are unconditional and all of a batch are in flight together),
    // and `in` says whether the lane's term is part of the sum (not past the
    // row or the last row)
    auto take = [&](bool& in) -> double {
        in = x <= nxi && y0 + lane < nyi;
        const double v = t[base + (int64_t)min(x, nxi) * P + 1 + min(y0 + lane, nyi - 1)];
        y0 += 64;
        if (y0 >= nyi) {
            y0 = 0;
            x++;
        }
        return v;
    };
    constexpr int U = 8;
    for (int64_t k = 0; k < chunks; k += U) {
        double v[U];
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = take(in[u]);
#pragma unroll
        for (int u = 0; u < U; u++) stage[u * 64 + lane] = in[u] ? v[u] : 0.0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        acc = lds_ordered_sum(acc, stage, U * 64);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    return acc;
}

// src/simulation.rs:216-227: sum r^2 over every interior cell, x-major,
// sequentially from 0.0, divided by the fluid count.  The squares are formed
// in parallel; the first wave then folds them in the reference's order.
__device__ double ref_norm_squared(const ChanDesc& d, const double* __restrict__ p,
                                   const double* __restrict__ rhs, double* __restrict__ sq,
                                   double* sh_out) {
    const int nxi = d.nx - 2, nyi = d.ny - 2;
    const int64_t P = d.pitch;
    for (int k = threadIdx.x; k < nxi * nyi; k += blockDim.x) {
        const int x = 1 + k / nyi, y = 1 + k % nyi;
        const int64_t i = d.off + x * P + y;
        sq[i] = ref_residual_sq(p, rhs, i, P, d.delx2, d.dely2);
    }
    __syncthreads();
    __shared__ double stage[512];
    if (threadIdx.x < 64) {
        const double acc = ordered_row_sum(sq, d.off, P, nxi, nyi, stage);
        if (threadIdx.x == 0) *sh_out = acc / d.fluid_cells;
    }
    __syncthreads();
    return *sh_out;
}

// Construction (src/simulation.rs:94-96 -> 229-237): initial_norm_squared
// when the state carries none.
__global__ void __launch_bounds__(1024)
k_initial_norm(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
               const double* __restrict__ rhs, const double* __restrict__ p,
               double* __restrict__ scratch)
{
    __shared__ double sh;
    const ChanDesc d = cd[blockIdx.x];
    ChanState& s = cs[blockIdx.x];
    if (s.has_init) return;
    const double n = ref_norm_squared(d, p, rhs, scratch, &sh);
    if (threadIdx.x == 0) {
        s.init_norm2 = n;
        s.has_init = 1;
    }
}

// src/simulation.rs:266-270: the SOR update of one fluid cell from its x + 1,
// x - 1, y + 1 and y - 1 neighbours, in the reference's association.
__device__ __forceinline__ double lex_update(const ChanDesc& d, double pc, double pxp, double pxm,
                                             double pyp, double pym, double r) {
    return (d.omw * pc) + (d.mid * ((((pxp + pxm) / d.delx2) + ((pyp + pym) / d.dely2)) - r));
}

// Channels of at most this many cells (ring included) solve with p, rhs, the
// flags and the norm's terms in LDS (115 KB).
constexpr int LEX_LDS_CELLS = 4608;

// Lane i of the wave gets lane i - 1's v (lane 0 gets 0.0): a DPP wave shift.
__device__ __forceinline__ double lane_shift_up(double v) {
    const int64_t b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((int64_t)hi << 32) | (int64_t)(uint32_t)lo);
}

// src/simulation.rs:239-285 for channels with solver == SOLVER_LEX.
// Lexicographic Gauss-Seidel order (x outer, y inner, in place) makes cell
// (x, y) read the new W and N values and the old E and S values; sweeping the
// anti-diagonals x + y = const in increasing order, with a barrier between
// them, reproduces exactly those reads, so the result is bit-identical.
//
// A channel that fits LEX_LDS_CELLS is staged in LDS once, iterated there and
// written back: the same operations on the same values in the same order.
// When no anti-diagonal is longer than 64 cells (the shorter side is at most
// 64), the first wave sweeps alone: a wave's LDS accesses complete in order,
// so it needs no workgroup barrier between diagonals (the wavefront-scope
// fence keeps the compiler from reordering).  With at most 64 interior rows
// each lane owns a row y and takes its cells x = 1, 2, .. on successive
// diagonals: the W neighbour is the lane's own previous result, the N
// neighbour the previous result of the lane below (a DPP shift), and E, S,
// rhs and the flag -- values no earlier diagonal of the sweep writes -- are
// read from LDS one diagonal ahead.  Otherwise (at most 64 interior columns)
// a lane takes one cell of each diagonal, all reads from LDS.
__global__ void __launch_bounds__(1024)
k_lex_solve(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
            const BEntry* __restrict__ be, const uint8_t* __restrict__ flags,
            const double* __restrict__ rhs, double* __restrict__ p, double* __restrict__ scratch)
{
    __shared__ double sh_norm;
    __shared__ double sp[LEX_LDS_CELLS], srhs[LEX_LDS_CELLS], ssq[LEX_LDS_CELLS];
    __shared__ uint8_t sfl[LEX_LDS_CELLS];
    const ChanDesc d = cd[blockIdx.x];
    if (d.solver != SOLVER_LEX) return;
    ChanState& s = cs[blockIdx.x];
    const int nxi = d.nx - 2, nyi = d.ny - 2;
    const int64_t P = d.pitch;
    const double init = s.init_norm2;
    const int tid = threadIdx.x;
    double norm2 = 0.0;
    int iters = d.itermax, exhausted = 1;
    const int ny = d.ny, ncell = d.nx * d.ny;
    if (ncell <= LEX_LDS_CELLS) {
        for (int l = tid; l < ncell; l += blockDim.x) {
            const int x = l / ny, y = l - (l / ny) * ny;
            const int64_t g = d.off + x * P + y;
            sp[l] = p[g];
            srhs[l] = rhs[g];
            sfl[l] = flags[g];
        }
        __syncthreads();
        const bool one_wave = (nxi < nyi ? nxi : nyi) <= 64;
        const bool row_lanes = nyi <= 64;
        for (int it = 0; it < d.itermax; it++) {
            // copy_pressure_to_boundaries, src/grid/mod.rs:343-412
            for (int k = tid; k < d.nb; k += blockDim.x) {
                const BEntry e = be[d.boff + k];
                if (e.edge != E_NONE) {
                    const int64_t o = e.gidx - d.off;
                    const int l = (int)(o / P) * ny + (int)(o % P);
                    sp[l] = ghost_value(sp, l, ny, e.edge);
                }
            }
            __syncthreads();
            if (row_lanes) {
                if (tid < 64) {
                    // lane tid owns row y (rows past nyi idle, their addresses
                    // clamped to row nyi); cur = the cell x the lane takes next,
                    // with its old value pc, E = (x + 1, y), S = (x, y + 1), rhs,
                    // flag, and for lane 0 the ghost N = (x, 0)
                    const int y = tid + 1;
                    const bool row = y <= nyi;
                    const int yc = row ? y : nyi;
                    double w = sp[yc];                                       // (0, y)
                    double pc = sp[ny + yc];                                 // (1, y)
                    double e = sp[min(2, d.nx - 1) * ny + yc];               // (2, y)
                    double sv = sp[ny + yc + 1];                             // (1, y + 1)
                    double r = srhs[ny + yc];
                    uint8_t f = sfl[ny + yc];
                    double ng = sp[ny];                                      // (1, 0)
                    double mine = 0.0;   // the lane's latest result
                    for (int dd = 2; dd <= nxi + nyi; dd++) {
                        const double below = lane_shift_up(mine);
                        const int x = dd - y;
                        if (row && x >= 1 && x <= nxi) {
                            // the next cell's values, read ahead (clamped past nxi)
                            const int xn = x + 1;
                            const int ln = xn * ny + yc;
                            const double e_n = sp[min(xn + 1, d.nx - 1) * ny + yc];
                            const double s_n = sp[ln + 1];
                            const double r_n = srhs[ln];
                            const uint8_t f_n = sfl[ln];
                            const double g_n = sp[xn * ny];
                            const double nb = tid == 0 ? ng : below;
                            const double nv = lex_update(d, pc, e, w, sv, nb, r);
                            const double res = flag_fluid(f) ? nv : pc;
                            if (flag_fluid(f)) sp[x * ny + yc] = nv;
                            w = res;
                            mine = res;
                            pc = e;
                            e = e_n;
                            sv = s_n;
                            r = r_n;
                            f = f_n;
                            ng = g_n;
                        }
                        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    }
                }
                __syncthreads();
            } else if (one_wave) {
                if (tid < 64) {
                    for (int dd = 2; dd <= nxi + nyi; dd++) {
                        const int xlo = dd - nyi > 1 ? dd - nyi : 1;
                        const int xhi = dd - 1 < nxi ? dd - 1 : nxi;
                        const int x = xlo + tid;
                        if (x <= xhi) {
                            // the flag and the values are read together (an
                            // interior cell's neighbours are all in the
                            // channel); only a fluid cell stores its update
                            const int l = x * ny + (dd - x);
                            const uint8_t f = sfl[l];
                            const double nv = lex_update(d, sp[l], sp[l + ny], sp[l - ny],
                                                         sp[l + 1], sp[l - 1], srhs[l]);
                            if (flag_fluid(f)) sp[l] = nv;
                        }
                        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    }
                }
                __syncthreads();
            } else {
                for (int dd = 2; dd <= nxi + nyi; dd++) {
                    const int xlo = dd - nyi > 1 ? dd - nyi : 1;
                    const int xhi = dd - 1 < nxi ? dd - 1 : nxi;
                    for (int x = xlo + tid; x <= xhi; x += blockDim.x) {
                        const int l = x * ny + (dd - x);
                        const uint8_t f = sfl[l];
                        const double nv = lex_update(d, sp[l], sp[l + ny], sp[l - ny], sp[l + 1],
                                                     sp[l - 1], srhs[l]);
                        if (flag_fluid(f)) sp[l] = nv;
                    }
                    __syncthreads();
                }
            }
            // src/simulation.rs:216-227, as ref_norm_squared, on the LDS copy:
            // the terms in the reference's order (x-major over the interior),
            // then added in that order
            for (int k = tid; k < nxi * nyi; k += blockDim.x) {
                const int l = (1 + k / nyi) * ny + 1 + k % nyi;
                ssq[k] = ref_residual_sq(sp, srhs, l, ny, d.delx2, d.dely2);
            }
            __syncthreads();
            if (tid < 64) {
                const double acc = lds_ordered_sum(0.0, ssq, nxi * nyi);
                if (tid == 0) sh_norm = acc / d.fluid_cells;
            }
            __syncthreads();
            norm2 = sh_norm;
            if ((norm2 < init) || (norm2 < d.eps2)) {
                iters = it + 1;
                exhausted = 0;
                break;
            }
        }
        for (int l = tid; l < ncell; l += blockDim.x) {
            const int x = l / ny, y = l - (l / ny) * ny;
            p[d.off + x * P + y] = sp[l];
        }
        if (tid == 0) {
            s.sor_iters = (uint32_t)iters;
            s.norm2 = norm2;
            s.exhausted = exhausted;
            s.active = 0;
        }
        return;
    }
    for (int it = 0; it < d.itermax; it++) {
        // copy_pressure_to_boundaries, src/grid/mod.rs:343-412
        for (int k = threadIdx.x; k < d.nb; k += blockDim.x) {
            const BEntry e = be[d.boff + k];
            if (e.edge != E_NONE) p[e.gidx] = ghost_value(p, e.gidx, e.pitch, e.edge);
        }
        __syncthreads();
        for (int dd = 2; dd <= nxi + nyi; dd++) {
            const int xlo = dd - nyi > 1 ? dd - nyi : 1;
            const int xhi = dd - 1 < nxi ? dd - 1 : nxi;
            for (int x = xlo + (int)threadIdx.x; x <= xhi; x += blockDim.x) {
                const int y = dd - x;
                const int64_t i = d.off + x * P + y;
                // the flag and the values are read together; only a fluid
                // cell stores its update
                const uint8_t f = flags[i];
                const double nv = lex_update(d, p[i], p[i + P], p[i - P], p[i + 1], p[i - 1],
                                             rhs[i]);
                if (flag_fluid(f)) p[i] = nv;
            }
            __syncthreads();
        }
        norm2 = ref_norm_squared(d, p, rhs, scratch, &sh_norm);
        if ((norm2 < init) || (norm2 < d.eps2)) {
            iters = it + 1;
            exhausted = 0;
            break;
        }
    }
    if (threadIdx.x == 0) {
        s.sor_iters = (uint32_t)iters;
        s.norm2 = norm2;
        s.exhausted = exhausted;
        s.active = 0;
    }
}

// ---------------------------------------------------------------- launchers

#define TB dim3(TILE_Y, TILE_X)

void launch_ghost(hipStream_t st, int nbtot, const BEntry* be, const ChanState* cs, double* p,
                  int colour, int only_active) {
    if (nbtot > 0)
        k_ghost<<<(nbtot + 255) / 256, 256, 0, st>>>(nbtot, be, cs, p, colour, only_active);
}
void launch_rhs_stats(hipStream_t st, int ntiles, int nchan, const ChanDesc* cd, ChanState* cs,
                      const Tile* tiles, const uint8_t* flags, const double* rhs, double* part) {
    k_rhs_sums<<<ntiles, dim3(TILE_Y, TILE_X), 0, st>>>(cd, cs, tiles, flags, rhs, part, 0);
    k_rhs_mean<<<nchan, 256, 0, st>>>(cd, cs, part);
    k_rhs_sums<<<ntiles, dim3(TILE_Y, TILE_X), 0, st>>>(cd, cs, tiles, flags, rhs, part, 1);
}
void launch_solve_begin(hipStream_t st, int nchan, const ChanDesc* cd, ChanState* cs,
                        const double* part, int32_t* active) {
    k_solve_begin<<<nchan, 256, 0, st>>>(cd, cs, part, active);
}
void launch_initial_norm(hipStream_t st, int nchan, const ChanDesc* cd, ChanState* cs,
                         const double* rhs, const double* p, double* scratch) {
    k_initial_norm<<<nchan, 1024, 0, st>>>(cd, cs, rhs, p, scratch);
}
void launch_lex_solve(hipStream_t st, int nchan, const ChanDesc* cd, ChanState* cs,
                      const BEntry* be, const uint8_t* flags, const double* rhs, double* p,
                      double* scratch) {
    k_lex_solve<<<nchan, 1024, 0, st>>>(cd, cs, be, flags, rhs, p, scratch);
}

}  // namespace stm
