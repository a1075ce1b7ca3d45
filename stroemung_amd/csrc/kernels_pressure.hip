// gfx950 kernels for the pressure Poisson solve (POISSON of NaSt2D):
//  * the reference-compatible lexicographic SOR, run as a diagonal wavefront
//    inside one workgroup per channel -- bit-exact with the reference's
//    x-outer / y-inner in-place sweep, including its unmasked sequential
//    residual fold and stopping rule (src/simulation.rs:216-285);
//  * red-black SOR with a ghost refresh per sweep and a fluid-masked
//    residual, reduced deterministically, checked every `check_every` sweeps.
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

// ------------------------------------------------------------ red-black SOR

// One half-sweep over interior fluid cells with (x + y) & 1 == colour:
//   p = fma(omw, p, mid * (fma(pE + pW, 1/dx^2, (pS + pN) * (1/dy^2)) - rhs))
// `sweep` is the index of this sweep inside the current check interval; a
// channel whose solve has stopped, or that would exceed itermax, skips it.
__global__ void __launch_bounds__(TILE_THREADS)
k_rb_half(const ChanDesc* __restrict__ cd, const ChanState* __restrict__ cs,
          const Tile* __restrict__ tiles, const uint8_t* __restrict__ flags,
          const double* __restrict__ rhs, double* __restrict__ p, int colour, int sweep)
{
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const ChanState& s = cs[t.chan];
    if (!s.active || (int)s.sor_iters + sweep >= d.itermax) return;
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    if (x < 1 || x > d.nx - 2 || y < 1 || y > d.ny - 2 || ((x + y) & 1) != colour) return;
    const int64_t P = d.pitch, i = d.off + x * P + y;
    if (!flag_fluid(flags[i])) return;
    const double tt = fma(p[i + P] + p[i - P], d.rdx2, (p[i + 1] + p[i - 1]) * d.rdy2);
    p[i] = fma(d.omw, p[i], d.mid * (tt - rhs[i]));
}

// Per-tile sum of the masked residual^2 (ghosts must be fresh):
//   r = fma((pE-p)-(p-pW), 1/dx^2, ((pS-p)-(p-pN)) * (1/dy^2)) - rhs
__global__ void __launch_bounds__(TILE_THREADS)
k_resid(const ChanDesc* __restrict__ cd, const ChanState* __restrict__ cs,
        const Tile* __restrict__ tiles, const uint8_t* __restrict__ flags,
        const double* __restrict__ rhs, const double* __restrict__ p, double* __restrict__ part)
{
    __shared__ double sh[TILE_THREADS];
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    if (!cs[t.chan].active) return;
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    double r2 = 0.0;
    if (x >= 1 && x <= d.nx - 2 && y >= 1 && y <= d.ny - 2) {
        const int64_t P = d.pitch, i = d.off + x * P + y;
        if (flag_fluid(flags[i])) {
            const double pc = p[i];
            const double a = (p[i + P] - pc) - (pc - p[i - P]);
            const double b = (p[i + 1] - pc) - (pc - p[i - 1]);
            const double r = fma(a, d.rdx2, b * d.rdy2) - rhs[i];
            r2 = r * r;
        }
    }
    sh[tid] = r2;
    __syncthreads();
    for (int st = TILE_THREADS / 2; st > 0; st >>= 1) {
        if (tid < st) sh[tid] = sh[tid] + sh[tid + st];
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x * NPART + 1] = sh[0];
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

// Start of a red-black solve: tol2 = eps^2 * sum(rhs^2) over interior fluid.
__global__ void __launch_bounds__(256)
k_solve_begin(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
              const double* __restrict__ part, int32_t* __restrict__ active_out)
{
    __shared__ double sh[256];
    const ChanDesc& d = cd[blockIdx.x];
    const double rsum2 = chan_sum(d, part, 0, sh);
    if (threadIdx.x == 0) {
        ChanState& s = cs[blockIdx.x];
        const bool rb = d.solver == SOLVER_RB;
        s.rhs2 = rsum2;
        s.tol2 = d.eps2 * rsum2;
        s.sor_iters = 0;
        s.exhausted = 0;
        s.active = (rb && d.itermax > 0) ? 1 : 0;
        active_out[blockIdx.x] = s.active;
    }
}

// After `nsweeps` sweeps: fold the residual, count the sweeps, stop the
// channel when sum r^2 <= tol2 or at itermax.
__global__ void __launch_bounds__(256)
k_solve_check(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
              const double* __restrict__ part, int nsweeps, int32_t* __restrict__ active_out)
{
    __shared__ double sh[256];
    const ChanDesc& d = cd[blockIdx.x];
    ChanState& s = cs[blockIdx.x];
    if (!s.active) return;   // uniform across the block
    const double res2 = chan_sum(d, part, 1, sh);
    if (threadIdx.x == 0) {
        const int left = d.itermax - (int)s.sor_iters;
        s.sor_iters += (uint32_t)(nsweeps < left ? nsweeps : left);
        s.res2 = res2;
        s.norm2 = res2 / d.fluid_cells;
        if (res2 <= s.tol2) {
            s.active = 0;
            s.exhausted = 0;
        } else if ((int)s.sor_iters >= d.itermax) {
            s.active = 0;
            s.exhausted = 1;
        }
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

// src/simulation.rs:216-227: sum r^2 over every interior cell, x-major,
// sequentially from 0.0, divided by the fluid count.  The squares are formed
// in parallel; one thread then folds them in the reference's order.
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
    if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int x = 1; x <= nxi; x++) {
            const double* row = sq + d.off + x * P;
            for (int y = 1; y <= nyi; y++) acc = acc + row[y];
        }
        *sh_out = acc / d.fluid_cells;
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

// src/simulation.rs:239-285 for channels with solver == SOLVER_LEX.
// Lexicographic Gauss-Seidel order (x outer, y inner, in place) makes cell
// (x, y) read the new W and N values and the old E and S values; sweeping the
// anti-diagonals x + y = const in increasing order, with a barrier between
// them, reproduces exactly those reads, so the result is bit-identical.
__global__ void __launch_bounds__(1024)
k_lex_solve(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
            const BEntry* __restrict__ be, const uint8_t* __restrict__ flags,
            const double* __restrict__ rhs, double* __restrict__ p, double* __restrict__ scratch)
{
    __shared__ double sh_norm;
    const ChanDesc d = cd[blockIdx.x];
    if (d.solver != SOLVER_LEX) return;
    ChanState& s = cs[blockIdx.x];
    const int nxi = d.nx - 2, nyi = d.ny - 2;
    const int64_t P = d.pitch;
    const double init = s.init_norm2;
    double norm2 = 0.0;
    int iters = d.itermax, exhausted = 1;
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
                if (!flag_fluid(flags[i])) continue;
                const double pc = p[i];
                p[i] = (d.omw * pc)
                     + (d.mid * ((((p[i + P] + p[i - P]) / d.delx2)
                                  + ((p[i + 1] + p[i - 1]) / d.dely2)) - rhs[i]));
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
void launch_rb_half(hipStream_t st, int ntiles, const ChanDesc* cd, const ChanState* cs,
                    const Tile* tiles, const uint8_t* flags, const double* rhs, double* p,
                    int colour, int sweep) {
    k_rb_half<<<ntiles, TB, 0, st>>>(cd, cs, tiles, flags, rhs, p, colour, sweep);
}
void launch_resid(hipStream_t st, int ntiles, const ChanDesc* cd, const ChanState* cs,
                  const Tile* tiles, const uint8_t* flags, const double* rhs, const double* p,
                  double* part) {
    k_resid<<<ntiles, TB, 0, st>>>(cd, cs, tiles, flags, rhs, p, part);
}
void launch_solve_begin(hipStream_t st, int nchan, const ChanDesc* cd, ChanState* cs,
                        const double* part, int32_t* active) {
    k_solve_begin<<<nchan, 256, 0, st>>>(cd, cs, part, active);
}
void launch_solve_check(hipStream_t st, int nchan, const ChanDesc* cd, ChanState* cs,
                        const double* part, int nsweeps, int32_t* active) {
    k_solve_check<<<nchan, 256, 0, st>>>(cd, cs, part, nsweeps, active);
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
