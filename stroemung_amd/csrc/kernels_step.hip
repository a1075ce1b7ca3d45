// gfx950 kernels for the non-pressure stages of one NaSt2D tick:
// SETBCOND, COMP_FG, COMP_RHS, ADAP_UV with its restore list, the
// max|u|,|v| / range reductions and the adaptive time step.
//
// Built with -ffp-contract=off: every expression keeps the reference's
// association order and no a*b+c is fused, so these stages are bit-exact
// against the Rust reference (which never fuses).  Plain '/' is the IEEE
// correctly rounded divide on gfx950.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "stm_internal.h"

namespace stm {

__device__ __forceinline__ double sq(double x) { return x * x; }

// Wall rule for NoSlip / Inflow / moving-wall cells: the ghost tangential
// velocity.  NoSlip and Inflow negate (src/grid/mod.rs:443, 463, ...); the
// moving wall reflects about its own velocity.
__device__ __forceinline__ double reflect(bool moving, double b, double fluid) {
    return moving ? (2.0 * b) - fluid : -fluid;
}

// ---------------------------------------------------------------- SETBCOND
//
// src/grid/mod.rs:414-651.  The reference walks the (x, y)-sorted boundary
// list once.  Its reads all hit fluid cells, its writes to fluid cells are
// constants (v[N] = bv by North-type wall cells, u[W] = bu by West-type wall
// cells) with a single writer each, and the only read that follows its
// writer in list order is v[W] of a West / NorthWest cell, written by the wall
// cell at (x-1, y+1).  So: phase 1 computes every boundary cell's own values
// and its restore values from the pre-pass fields (substituting that writer's
// constant for v[W]); phase 2, after a barrier, stores the fluid-face
// constants.  Each phase is a kernel over every boundary cell of every
// channel (one thread per cell); the kernel boundary is the barrier.
__global__ void __launch_bounds__(256)
k_bc_cells(int nbtot, const BEntry* __restrict__ be, double2* __restrict__ restore,
           double* __restrict__ u, double* __restrict__ v)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nbtot) {
        const BEntry e = be[k];
        const int64_t i = e.gidx, P = e.pitch;
        const int64_t N = i - 1, S = i + 1, E = i + P, W = i - P;
        const double ub = u[i], vb = v[i];
        double nu = ub, nv = vb, ru, rv;
        const uint8_t ed = e.edge;
        if (ed == E_NONE) {
            ru = ub;
            rv = vb;
        } else if (e.type == T_OUTFLOW) {  // src/grid/mod.rs:489-536
            const double vw = e.vw_rule ? e.vw_const : (ed == E_W || ed == E_NW ? v[W] : 0.0);
            switch (ed) {
            case E_N:  nu = u[N]; nv = v[N]; break;
            case E_NE: nu = u[N]; nv = v[E]; break;
            case E_E:  nu = u[E]; nv = v[E]; break;
            case E_SE: nu = u[E]; nv = v[S]; break;
            case E_S:  nu = u[S]; nv = v[S]; break;
            case E_SW: nu = u[W]; nv = v[S]; break;
            case E_W:  nu = u[W]; nv = vw; break;
            case E_NW: nu = u[N]; nv = vw; break;
            }
            // src/grid/mod.rs:595-648; an outflow cell writes no fluid face
            const bool wt = ed == E_SW || ed == E_W || ed == E_NW;
            const bool nt = ed == E_N || ed == E_NE || ed == E_NW;
            ru = wt ? u[W] : nu;
            rv = nt ? v[N] : nv;
        } else {  // NoSlip / Inflow (src/grid/mod.rs:437-488, 537-586), moving wall
            const bool mv = e.type == T_MOVING;
            const double bu = e.bu, bv = e.bv;
            const double vw = e.vw_rule ? e.vw_const : (ed == E_W || ed == E_NW ? v[W] : 0.0);
            switch (ed) {
            case E_N:  nu = reflect(mv, bu, u[N]); break;
            case E_NE: nu = bu; nv = reflect(mv, bv, v[E]); break;
            case E_E:  nu = bu; nv = reflect(mv, bv, v[E]); break;
            case E_SE: nu = bu; nv = bv; break;
            case E_S:  nu = reflect(mv, bu, u[S]); nv = bv; break;
            case E_SW: nu = reflect(mv, bu, u[S]); nv = bv; break;
            case E_W:  nv = reflect(mv, bv, vw); break;
            case E_NW: nu = reflect(mv, bu, u[N]); nv = reflect(mv, bv, vw); break;
            }
            const bool wt = ed == E_SW || ed == E_W || ed == E_NW;
            const bool nt = ed == E_N || ed == E_NE || ed == E_NW;
            ru = wt ? bu : nu;
            rv = nt ? bv : nv;
        }
        u[i] = nu;
        v[i] = nv;
        restore[k] = make_double2(ru, rv);
    }
}

__global__ void __launch_bounds__(256)
k_bc_faces(int nbtot, const BEntry* __restrict__ be, double* __restrict__ u,
           double* __restrict__ v)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbtot) return;
    const BEntry e = be[k];
    const uint8_t ed = e.edge;
    if (ed == E_NONE || e.type == T_OUTFLOW) return;
    if (ed == E_N || ed == E_NE || ed == E_NW) v[e.gidx - 1] = e.bv;
    if (ed == E_SW || ed == E_W || ed == E_NW) u[e.gidx - e.pitch] = e.bu;
}

// ---------------------------------------------------------------- COMP_FG

// src/math.rs:162-174
__device__ __forceinline__ double laplacian(double c, double n, double s, double e, double w,
                                            double delx, double dely) {
    const double d2edx2 = ((e - (2. * c)) + w) / sq(delx);
    const double d2edy2 = ((s - (2. * c)) + n) / sq(dely);
    return d2edx2 + d2edy2;
}

// F and G of one interior cell (src/simulation.rs:349-392 with
// src/math.rs:19-150).  ne = (x+1, y-1), sw = (x-1, y+1).
__device__ __forceinline__ void fg_cell(
    double uc, double un, double us, double ue, double uw, double usw,
    double vc, double vn, double vs, double ve, double vw, double vne,
    double delx, double dely, double delt, double gamma, double re, double& F, double& G)
{
    // du2dx, src/math.rs:19-33
    double du2dx;
    {
        const double l1 = sq(uc + ue), r1 = sq(uw + uc);
        const double left = l1 - r1;
        const double l2 = fabs(uc + ue) * (uc - ue);
        const double r2 = fabs(uw + uc) * (uw - uc);
        du2dx = (left + (gamma * (l2 - r2))) / (4.0 * delx);
    }
    // duvdy, src/math.rs:97-120
    double duvdy;
    {
        const double l1 = (vc + ve) * (uc + us);
        const double r1 = (vn + vne) * (un + uc);
        const double left = l1 - r1;
        const double l2 = fabs(vc + ve) * (uc - us);
        const double r2 = fabs(vn + vne) * (un - uc);
        duvdy = (left + (gamma * (l2 - r2))) / (4.0 * dely);
    }
    // duvdx, src/math.rs:53-77
    double duvdx;
    {
        const double l1 = (uc + us) * (vc + ve);
        const double r1 = (uw + usw) * (vw + vc);
        const double left = l1 - r1;
        const double l2 = fabs(uc + us) * (vc - ve);
        const double r2 = fabs(uw + usw) * (vw - vc);
        duvdx = (left + (gamma * (l2 - r2))) / (4.0 * delx);
    }
    // dv2dy, src/math.rs:136-150
    double dv2dy;
    {
        const double l1 = sq(vc + vs), r1 = sq(vn + vc);
        const double left = l1 - r1;
        const double l2 = fabs(vc + vs) * (vc - vs);
        const double r2 = fabs(vn + vc) * (vn - vc);
        dv2dy = (left + (gamma * (l2 - r2))) / (4.0 * dely);
    }
    const double lu = laplacian(uc, un, us, ue, uw, delx, dely);
    const double lv = laplacian(vc, vn, vs, ve, vw, delx, dely);
    F = uc + (delt * (((lu / re) - du2dx) - duvdy));
    G = vc + (delt * (((lv / re) - duvdx) - dv2dy));
}

// src/simulation.rs:122-202 at cell (x, y) of a channel: F, G on interior
// cells; then f = u, g = v on boundary cells, and on fluid cells whose east
// (south) neighbour is a boundary, f = u (g = v) -- the West-type
// (North-type) restore of src/simulation.rs:167-201.  wf / wg: f / g is
// written; fluid cells on the outer ring keep their f, g.
__device__ __forceinline__ void fg_at(const ChanDesc& d, double delt,
                                      const uint8_t* __restrict__ flags,
                                      const double* __restrict__ u, const double* __restrict__ v,
                                      int x, int y, int64_t i, double& F, double& G, bool& wf,
                                      bool& wg) {
    const int64_t P = d.pitch;
    if (!flag_fluid(flags[i])) {
        F = u[i];
        G = v[i];
        wf = wg = true;
        return;
    }
    F = G = 0.0;
    wf = wg = false;
    if (x >= 1 && x <= d.nx - 2 && y >= 1 && y <= d.ny - 2) {
        fg_cell(u[i], u[i - 1], u[i + 1], u[i + P], u[i - P], u[i - P + 1],
                v[i], v[i - 1], v[i + 1], v[i + P], v[i - P], v[i + P - 1],
                d.delx, d.dely, delt, d.gamma, d.reynolds, F, G);
        wf = wg = true;
    }
    if (x + 1 < d.nx && !flag_fluid(flags[i + P])) { F = u[i]; wf = true; }
    if (y + 1 < d.ny && !flag_fluid(flags[i + 1])) { G = v[i]; wg = true; }
}

__global__ void __launch_bounds__(TILE_THREADS)
k_fg(const ChanDesc* __restrict__ cd, const ChanState* __restrict__ cs,
     const Tile* __restrict__ tiles, const uint8_t* __restrict__ flags,
     const double* __restrict__ u, const double* __restrict__ v,
     double* __restrict__ f, double* __restrict__ g)
{
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    if (x >= d.nx || y >= d.ny) return;
    const int64_t i = d.off + x * (int64_t)d.pitch + y;
    double F, G;
    bool wf, wg;
    fg_at(d, cs[t.chan].delt, flags, u, v, x, y, i, F, G, wf, wg);
    if (wf) f[i] = F;
    if (wg) g[i] = G;
}

// COMP_FG and COMP_RHS (src/simulation.rs:204-214) in one pass over a tile
// of TILE_X rows of TILE_Y cells: a fifth row of threads computes f of the
// row before the tile, and the first lane of each row g of the cell before
// it, so the rhs of every tile cell takes its f(x-1, y) and g(x, y-1) from
// LDS -- the values k_fg would have stored -- and f, g are not read back.
// The same expressions in the same order: bit-identical to k_fg + k_rhs.
__global__ void __launch_bounds__(TILE_THREADS + TILE_Y)
k_fg_rhs(const ChanDesc* __restrict__ cd, const ChanState* __restrict__ cs,
         const Tile* __restrict__ tiles, const uint8_t* __restrict__ flags,
         const double* __restrict__ u, const double* __restrict__ v,
         double* __restrict__ f, double* __restrict__ g, double* __restrict__ rhs)
{
    __shared__ double sf[TILE_X + 1][TILE_Y];
    __shared__ double sg[TILE_X + 1][TILE_Y + 1];
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const double delt = cs[t.chan].delt;
    const int r = threadIdx.y, l = threadIdx.x;       // row r - 1 of the tile
    const int x = t.x0 - 1 + r, y = t.y0 + l;
    const int64_t P = d.pitch;
    if (x >= 0 && x < d.nx && y < d.ny) {
        const int64_t i = d.off + x * P + y;
        double F, G;
        bool wf, wg;
        fg_at(d, delt, flags, u, v, x, y, i, F, G, wf, wg);
        sf[r][l] = wf ? F : f[i];
        sg[r][l + 1] = wg ? G : g[i];
        if (r > 0) {
            if (wf) f[i] = F;
            if (wg) g[i] = G;
        }
        if (r > 0 && l == 0 && y >= 1) {             // g of the cell before the tile's column
            double F2, G2;
            bool wf2, wg2;
            fg_at(d, delt, flags, u, v, x, y - 1, i - 1, F2, G2, wf2, wg2);
            sg[r][0] = wg2 ? G2 : g[i - 1];
        }
    }
    __syncthreads();
    if (r > 0 && x >= 1 && x < d.nx && y >= 1 && y < d.ny) {
        const int64_t i = d.off + x * P + y;
        rhs[i] = (((sf[r][l] - sf[r - 1][l]) / d.delx) + ((sg[r][l + 1] - sg[r][l]) / d.dely))
                 / delt;
    }
}

// Deterministic block sum: fixed LDS tree over TILE_THREADS values.
__device__ __forceinline__ double block_sum(double val, double* sh) {
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    sh[tid] = val;
    __syncthreads();
    for (int s = TILE_THREADS / 2; s > 0; s >>= 1) {
        if (tid < s) sh[tid] = sh[tid] + sh[tid + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// src/simulation.rs:204-214: rhs on x in [1, nx-1], y in [1, ny-1].
__global__ void __launch_bounds__(TILE_THREADS)
k_rhs(const ChanDesc* __restrict__ cd, const ChanState* __restrict__ cs,
      const Tile* __restrict__ tiles, const double* __restrict__ f,
      const double* __restrict__ g, double* __restrict__ rhs)
{
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    if (x >= 1 && x < d.nx && y >= 1 && y < d.ny) {
        const int64_t P = d.pitch, i = d.off + x * P + y;
        rhs[i] = (((f[i] - f[i - P]) / d.delx) + ((g[i] - g[i - 1]) / d.dely))
                 / cs[t.chan].delt;
    }
}

// ---------------------------------------------------------------- ADAP_UV

// src/simulation.rs:313-320: the net effect of the restore list, per cell.
__global__ void __launch_bounds__(256)
k_restore(int nbtot, const BEntry* __restrict__ be, const double2* __restrict__ restore,
          double* __restrict__ u, double* __restrict__ v)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbtot) return;
    const double2 r = restore[k];
    const int64_t i = be[k].gidx;
    u[i] = r.x;
    v[i] = r.y;
}

// ---------------------------------------------------------------- reductions

// Per-tile partials over fluid cells: max|u|, max|v| over fluid faces, the
// min / max of u^2+v^2 (src/grid/mod.rs:253-268) and of p (237-251).  min and
// max ignore a NaN operand (fmin / fmax, as Real::min / max and the oracle's
// comparisons do) and are order-free, so the fold order is free too: a
// shuffle tree inside each wave, then the four waves' results.
__device__ __forceinline__ double wave_fmax(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_fmin(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__global__ void __launch_bounds__(TILE_THREADS)
k_ranges(const ChanDesc* __restrict__ cd, const Tile* __restrict__ tiles,
         const uint8_t* __restrict__ flags, const double* __restrict__ u,
         const double* __restrict__ v, const double* __restrict__ p, double* __restrict__ part)
{
    constexpr int NW = TILE_THREADS / 64;
    __shared__ double sh[6][NW];
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    double um = 0.0, vm = 0.0, smin = __DBL_MAX__, smax = 0.0, pmin = __DBL_MAX__, pmax = 0.0;
    if (x < d.nx && y < d.ny) {
        const int64_t P = d.pitch, i = d.off + x * P + y;
        if (flag_fluid(flags[i])) {
            const double uc = u[i], vc = v[i];
            um = fmax(0.0, fabs(uc));
            vm = fmax(0.0, fabs(vc));
            if (x > 0 && !flag_fluid(flags[i - P])) um = fmax(um, fabs(u[i - P]));
            if (y > 0 && !flag_fluid(flags[i - 1])) vm = fmax(vm, fabs(v[i - 1]));
            const double sp = sq(uc) + sq(vc);
            smin = fmin(__DBL_MAX__, sp);
            smax = fmax(0.0, sp);
            pmin = fmin(__DBL_MAX__, p[i]);
            pmax = fmax(0.0, p[i]);
        }
    }
    um = wave_fmax(um);
    vm = wave_fmax(vm);
    smin = wave_fmin(smin);
    smax = wave_fmax(smax);
    pmin = wave_fmin(pmin);
    pmax = wave_fmax(pmax);
    if ((tid & 63) == 0) {
        const int w = tid >> 6;
        sh[0][w] = um; sh[1][w] = vm; sh[2][w] = smin;
        sh[3][w] = smax; sh[4][w] = pmin; sh[5][w] = pmax;
    }
    __syncthreads();
    if (tid < 6) {
        const bool lo = tid == 2 || tid == 4;
        double r = sh[tid][0];
#pragma unroll
        for (int w = 1; w < NW; w++) r = lo ? fmin(r, sh[tid][w]) : fmax(r, sh[tid][w]);
        part[blockIdx.x * NPART + tid] = r;
    }
}

// ADAP_UV (src/simulation.rs:287-311) fused with the range partials of
// k_ranges.  Runs after the restore list has been applied: every boundary
// cell is in that list, so the list's values are the boundary cells' final
// u, v whatever ADAP_UV computes for them, and this kernel writes only the
// fluid cells of [0, nx-2] x [0, ny-2] (same expression, same bits); the
// boundary faces the max |u|, |v| reads are then final too.
__global__ void __launch_bounds__(TILE_THREADS)
k_adap_ranges(const ChanDesc* __restrict__ cd, const ChanState* __restrict__ cs,
              const Tile* __restrict__ tiles, const uint8_t* __restrict__ flags,
              const double* __restrict__ f, const double* __restrict__ g,
              const double* __restrict__ p, double* __restrict__ u, double* __restrict__ v,
              double* __restrict__ part)
{
    constexpr int NW = TILE_THREADS / 64;
    __shared__ double sh[6][NW];
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    double um = 0.0, vm = 0.0, smin = __DBL_MAX__, smax = 0.0, pmin = __DBL_MAX__, pmax = 0.0;
    if (x < d.nx && y < d.ny) {
        const int64_t P = d.pitch, i = d.off + x * P + y;
        if (flag_fluid(flags[i])) {
            const double pc = p[i];
            double uc, vc;
            if (x <= d.nx - 2 && y <= d.ny - 2) {
                const double delt = cs[t.chan].delt;
                uc = f[i] - ((delt / d.delx) * (p[i + P] - pc));
                vc = g[i] - ((delt / d.dely) * (p[i + 1] - pc));
                u[i] = uc;
                v[i] = vc;
            } else {
                uc = u[i];
                vc = v[i];
            }
            um = fmax(0.0, fabs(uc));
            vm = fmax(0.0, fabs(vc));
            if (x > 0 && !flag_fluid(flags[i - P])) um = fmax(um, fabs(u[i - P]));
            if (y > 0 && !flag_fluid(flags[i - 1])) vm = fmax(vm, fabs(v[i - 1]));
            const double sp = sq(uc) + sq(vc);
            smin = fmin(__DBL_MAX__, sp);
            smax = fmax(0.0, sp);
            pmin = fmin(__DBL_MAX__, pc);
            pmax = fmax(0.0, pc);
        }
    }
    um = wave_fmax(um);
    vm = wave_fmax(vm);
    smin = wave_fmin(smin);
    smax = wave_fmax(smax);
    pmin = wave_fmin(pmin);
    pmax = wave_fmax(pmax);
    if ((tid & 63) == 0) {
        const int w = tid >> 6;
        sh[0][w] = um; sh[1][w] = vm; sh[2][w] = smin;
        sh[3][w] = smax; sh[4][w] = pmin; sh[5][w] = pmax;
    }
    __syncthreads();
    if (tid < 6) {
        const bool lo = tid == 2 || tid == 4;
        double r = sh[tid][0];
#pragma unroll
        for (int w = 1; w < NW; w++) r = lo ? fmin(r, sh[tid][w]) : fmax(r, sh[tid][w]);
        part[blockIdx.x * NPART + tid] = r;
    }
}

// One workgroup per channel: fold the range partials of its tiles (min and
// max are order-free, so the result is exact).
// want_pressure: 1 = always (construction), 2 = only if the solve hit
// itermax (src/simulation.rs:283), 0 = never.
__global__ void __launch_bounds__(256)
k_ranges_reduce(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
                const double* __restrict__ part, int want_pressure)
{
    __shared__ double sh[6][256];
    const int c = blockIdx.x, tid = threadIdx.x;
    const ChanDesc& d = cd[c];
    double um = 0.0, vm = 0.0, smin = __DBL_MAX__, smax = 0.0, pmin = __DBL_MAX__, pmax = 0.0;
    for (int k = d.tile0 + tid; k < d.tile0 + d.ntiles; k += 256) {
        const double* q = part + (int64_t)k * NPART;
        um = fmax(um, q[0]);
        vm = fmax(vm, q[1]);
        smin = fmin(smin, q[2]);
        smax = fmax(smax, q[3]);
        pmin = fmin(pmin, q[4]);
        pmax = fmax(pmax, q[5]);
    }
    sh[0][tid] = um; sh[1][tid] = vm; sh[2][tid] = smin;
    sh[3][tid] = smax; sh[4][tid] = pmin; sh[5][tid] = pmax;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st) {
            sh[0][tid] = fmax(sh[0][tid], sh[0][tid + st]);
            sh[1][tid] = fmax(sh[1][tid], sh[1][tid + st]);
            sh[2][tid] = fmin(sh[2][tid], sh[2][tid + st]);
            sh[3][tid] = fmax(sh[3][tid], sh[3][tid + st]);
            sh[4][tid] = fmin(sh[4][tid], sh[4][tid + st]);
            sh[5][tid] = fmax(sh[5][tid], sh[5][tid + st]);
        }
        __syncthreads();
    }
    if (tid) return;
    um = sh[0][0]; vm = sh[1][0]; smin = sh[2][0]; smax = sh[3][0]; pmin = sh[4][0]; pmax = sh[5][0];
    ChanState& s = cs[c];
    s.umax = um;
    s.vmax = vm;
    s.speed_range[0] = sqrt(smin);
    s.speed_range[1] = sqrt(smax);
    if (want_pressure == 1 || (want_pressure == 2 && s.exhausted)) {
        s.pressure_range[0] = pmin;
        s.pressure_range[1] = pmax;
    }
}

// Adaptive time step (Griebel, Dornseifer, Neunhoeffer: NaSt2D COMP_DELT):
// delt = tau * min(Re/2 / (1/dx^2 + 1/dy^2), dx/|u|max, dy/|v|max).
__global__ void k_delt(int nchan, const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchan) return;
    const ChanDesc& d = cd[c];
    if (!(d.tau > 0.0)) return;
    ChanState& s = cs[c];
    double dt = (d.reynolds / 2.0) / ((1.0 / sq(d.delx)) + (1.0 / sq(d.dely)));
    if (s.umax > 0.0) { const double t = d.delx / s.umax; if (t < dt) dt = t; }
    if (s.vmax > 0.0) { const double t = d.dely / s.vmax; if (t < dt) dt = t; }
    s.delt = d.tau * dt;
}

// src/simulation.rs:330-331
__global__ void k_advance_time(int nchan, ChanState* __restrict__ cs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchan) return;
    cs[c].time = cs[c].time + cs[c].delt;
    cs[c].iterations += 1;
}

// Plain streaming copy (16 B per lane per step): the bandwidth probe and the
// known-size reference for hardware byte counters.
__global__ void __launch_bounds__(256)
k_copy(int64_t n2, const double2* __restrict__ src, double2* __restrict__ dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n2; k += stride)
        dst[k] = src[k];
}

void launch_copy(hipStream_t st, int64_t bytes, const double* src, double* dst) {
    const int64_t n2 = bytes / 16;
    k_copy<<<256 * 16, 256, 0, st>>>(n2, reinterpret_cast<const double2*>(src),
                                     reinterpret_cast<double2*>(dst));
}

// u = v = p = 0 at the listed cells (a grid edit's changed cells).
__global__ void __launch_bounds__(256)
k_zero_cells(int n, const int64_t* __restrict__ idx, double* __restrict__ u,
             double* __restrict__ v, double* __restrict__ p)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t i = idx[k];
    u[i] = 0.0;
    v[i] = 0.0;
    p[i] = 0.0;
}

__global__ void __launch_bounds__(256)
k_fill(int64_t n, double* __restrict__ a, double val)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = val;
}

// ---------------------------------------------------------------- launchers

#define TB dim3(TILE_Y, TILE_X)

void launch_bc_velocity(hipStream_t st, int nbtot, const BEntry* be, double2* restore, double* u,
                        double* v) {
    if (nbtot <= 0) return;
    k_bc_cells<<<(nbtot + 255) / 256, 256, 0, st>>>(nbtot, be, restore, u, v);
    k_bc_faces<<<(nbtot + 255) / 256, 256, 0, st>>>(nbtot, be, u, v);
}
void launch_fg(hipStream_t st, int ntiles, const ChanDesc* cd, const ChanState* cs,
               const Tile* tiles, const uint8_t* flags, const double* u, const double* v,
               double* f, double* g) {
    k_fg<<<ntiles, TB, 0, st>>>(cd, cs, tiles, flags, u, v, f, g);
}
void launch_fg_rhs(hipStream_t st, int ntiles, const ChanDesc* cd, const ChanState* cs,
                   const Tile* tiles, const uint8_t* flags, const double* u, const double* v,
                   double* f, double* g, double* rhs) {
    k_fg_rhs<<<ntiles, dim3(TILE_Y, TILE_X + 1), 0, st>>>(cd, cs, tiles, flags, u, v, f, g, rhs);
}
void launch_rhs(hipStream_t st, int ntiles, const ChanDesc* cd, const ChanState* cs,
                const Tile* tiles, const double* f, const double* g, double* rhs) {
    k_rhs<<<ntiles, TB, 0, st>>>(cd, cs, tiles, f, g, rhs);
}
void launch_restore(hipStream_t st, int nbtot, const BEntry* be, const double2* restore,
                    double* u, double* v) {
    if (nbtot > 0) k_restore<<<(nbtot + 255) / 256, 256, 0, st>>>(nbtot, be, restore, u, v);
}
void launch_ranges(hipStream_t st, int ntiles, int nchan, const ChanDesc* cd, ChanState* cs,
                   const Tile* tiles, const uint8_t* flags, const double* u, const double* v,
                   const double* p, double* part, int want_pressure) {
    k_ranges<<<ntiles, TB, 0, st>>>(cd, tiles, flags, u, v, p, part);
    k_ranges_reduce<<<nchan, 256, 0, st>>>(cd, cs, part, want_pressure);
}
void launch_adap_ranges(hipStream_t st, int ntiles, int nchan, const ChanDesc* cd, ChanState* cs,
                        const Tile* tiles, const uint8_t* flags, const double* f,
                        const double* g, const double* p, double* u, double* v, double* part,
                        int want_pressure) {
    k_adap_ranges<<<ntiles, TB, 0, st>>>(cd, cs, tiles, flags, f, g, p, u, v, part);
    k_ranges_reduce<<<nchan, 256, 0, st>>>(cd, cs, part, want_pressure);
}
void launch_delt(hipStream_t st, int nchan, const ChanDesc* cd, ChanState* cs) {
    k_delt<<<(nchan + 63) / 64, 64, 0, st>>>(nchan, cd, cs);
}
void launch_advance_time(hipStream_t st, int nchan, ChanState* cs) {
    k_advance_time<<<(nchan + 63) / 64, 64, 0, st>>>(nchan, cs);
}
void launch_zero_cells(hipStream_t st, int n, const int64_t* idx, double* u, double* v,
                       double* p) {
    if (n > 0) k_zero_cells<<<(n + 255) / 256, 256, 0, st>>>(n, idx, u, v, p);
}
void launch_fill(hipStream_t st, int64_t n, double* a, double val) {
    if (n > 0) k_fill<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, a, val);
}

}  // namespace stm
