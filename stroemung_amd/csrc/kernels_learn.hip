// gfx950 kernels joining the FP64 pressure solve and the BF16 pressure network.
//
// A learned cycle on channel c:
//   r  = Lap(p) - (rhs - mean)              interior fluid cells, ghosts refreshed first
//   x  = -r / rms(r)                        network input, channel 0 (channel 1: fluid mask)
//   y  = net(x)                             fp32, one value per cell
//   q  = Lap(y)                             y with the same Neumann ghosts as p
//   a  = -<r, q> / <q, q>                   minimises |r + a q|, so the step never raises
//   p += a y                                the residual norm
// after which budgeted red-black sweeps smooth what the network left.  The
// same kernels build training pairs from plain solves: the residual at the
// start of a solve (and after a few sweeps) with the normalised pressure
// change that the solve went on to make.
//
// Lap is the five-point operator of the red-black solver with the same
// operation order, so a residual computed here equals the solver's own.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "stm_internal.h"

namespace stm {

namespace {

__device__ __forceinline__ bool interior(const ChanDesc& d, int x, int y) {
    return x >= 1 && x <= d.nx - 2 && y >= 1 && y <= d.ny - 2;
}

__device__ __forceinline__ double lap(const double* __restrict__ a, int64_t i, int64_t P,
                                      double rdx2, double rdy2) {
    const double c = a[i];
    return fma((a[i + P] - c) - (c - a[i - P]), rdx2, ((a[i + 1] - c) - (c - a[i - 1])) * rdy2);
}

// Sum of one block's values into part[blockIdx.x * NPART + slot].
__device__ void tile_sum(double v, double* sh, double* __restrict__ part, int slot) {
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int st = TILE_THREADS / 2; st > 0; st >>= 1) {
        if (tid < st) sh[tid] = sh[tid] + sh[tid + st];
        __syncthreads();
    }
    if (tid == 0) part[(int64_t)blockIdx.x * NPART + slot] = sh[0];
    __syncthreads();
}

// One channel's partials of `slot` summed in a fixed order (256 threads).
__device__ double chan_total(const ChanDesc& d, const double* __restrict__ part, int slot,
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
    const double r = sh[0];
    __syncthreads();
    return r;
}

constexpr int S_RES2 = 2, S_RQ = 3, S_QQ = 4, S_DSUM = 5, S_DSQ = 6;

}  // namespace

// r = Lap(p) - (rhs - mean) at interior fluid cells, 0 elsewhere; slot 2 = sum r^2.
// p's ghosts must be current.
__global__ void __launch_bounds__(TILE_THREADS)
k_learn_residual(const ChanDesc* __restrict__ cd, const ChanState* __restrict__ cs,
                 const Tile* __restrict__ tiles, const uint8_t* __restrict__ flags,
                 const double* __restrict__ rhs, const double* __restrict__ p,
                 double* __restrict__ res, double* __restrict__ part)
{
    __shared__ double sh[TILE_THREADS];
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    double r2 = 0.0;
    if (x < d.nx && y < d.ny) {
        const int64_t i = d.off + x * (int64_t)d.pitch + y;
        double r = 0.0;
        if (interior(d, x, y) && flag_fluid(flags[i])) {
            r = lap(p, i, d.pitch, d.rdx2, d.rdy2) - (rhs[i] - cs[t.chan].rhs_mean);
            r2 = r * r;
        }
        res[i] = r;
    }
    tile_sum(r2, sh, part, S_RES2);
}

// Per channel: has the solve converged, and the input scale.  first = 1 at
// the start of a solve.
__global__ void __launch_bounds__(256)
k_learn_scale(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
              LearnChan* __restrict__ lc, const double* __restrict__ part, int first)
{
    __shared__ double sh[256];
    const int c = blockIdx.x;
    const ChanDesc& d = cd[c];
    const double res2 = chan_total(d, part, S_RES2, sh);
    if (threadIdx.x) return;
    ChanState& s = cs[c];
    LearnChan& l = lc[c];
    if (first) {
        l.sweeps = 0;
        l.run = (d.solver == SOLVER_RB && d.itermax > 0) ? 1 : 0;
    }
    if (l.run && res2 <= s.tol2) {
        l.run = 0;
        s.res2 = res2;
        s.norm2 = res2 / d.fluid_cells;
        s.exhausted = 0;
    }
    l.scale = (l.run && res2 > 0.0) ? sqrt(d.n_inner / res2) : 0.0;
    l.alpha = 0.0;
}

// Network input of channel b (batch entry b): channel 0 = -r * scale,
// channel 1 = interior fluid mask, channels 2..31 zero.  Interior cell
// (x, y) is pixel (x - 1, y - 1); pixels beyond the interior are zero.
// Optionally also a compact 2-channel dataset sample in slot (slot0 + b) % dcap.
__global__ void __launch_bounds__(256)
k_learn_pack(const ChanDesc* __restrict__ cd, const LearnChan* __restrict__ lc,
             const uint8_t* __restrict__ flags, const double* __restrict__ res, int nch, int H,
             int W, __bf16* __restrict__ a0, __bf16* __restrict__ dx, int slot0, int dcap)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t hw = (int64_t)H * W;
    if (k >= hw * nch) return;
    const int b = (int)(k / hw);
    const int64_t q = k - (int64_t)b * hw;
    const int h = (int)(q / W), w = (int)(q - (int64_t)h * W);
    const ChanDesc& d = cd[b];
    float v0 = 0.f, v1 = 0.f;
    if (h <= d.nx - 3 && w <= d.ny - 3) {
        const int64_t i = d.off + (h + 1) * (int64_t)d.pitch + (w + 1);
        if (flag_fluid(flags[i])) {
            v0 = (float)(-res[i] * lc[b].scale);
            v1 = 1.f;
        }
    }
    if (a0) {
        typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
        bf16x8 z;
        for (int j = 0; j < 8; j++) z[j] = (__bf16)0.f;
        bf16x8* o = reinterpret_cast<bf16x8*>(a0 + k * 32);
        bf16x8 first = z;
        first[0] = (__bf16)v0;
        first[1] = (__bf16)v1;
        o[0] = first;
        o[1] = z;
        o[2] = z;
        o[3] = z;
    }
    if (dx) {
        const int64_t s = (slot0 + b) % dcap;
        dx[(s * hw + q) * 2] = (__bf16)v0;
        dx[(s * hw + q) * 2 + 1] = (__bf16)v1;
    }
}

// Correction field: y at interior fluid cells of running channels, 0 at the
// other cells (boundary ghosts are set afterwards).
__global__ void __launch_bounds__(TILE_THREADS)
k_learn_unpack(const ChanDesc* __restrict__ cd, const Tile* __restrict__ tiles,
               const LearnChan* __restrict__ lc, const uint8_t* __restrict__ flags,
               const float* __restrict__ y, int H, int W, double* __restrict__ cf)
{
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, yy = t.y0 + threadIdx.x;
    if (x >= d.nx || yy >= d.ny) return;
    const int64_t i = d.off + x * (int64_t)d.pitch + yy;
    double v = 0.0;
    if (lc[t.chan].run && interior(d, x, yy) && flag_fluid(flags[i]))
        v = (double)y[((int64_t)t.chan * H + (x - 1)) * W + (yy - 1)];
    cf[i] = v;
}

// q = Lap(cf) at interior fluid cells: slot 3 = sum r q, slot 4 = sum q q.
__global__ void __launch_bounds__(TILE_THREADS)
k_learn_dots(const ChanDesc* __restrict__ cd, const Tile* __restrict__ tiles,
             const uint8_t* __restrict__ flags, const double* __restrict__ res,
             const double* __restrict__ cf, double* __restrict__ part)
{
    __shared__ double sh[TILE_THREADS];
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    double rq = 0.0, qq = 0.0;
    if (x < d.nx && y < d.ny && interior(d, x, y)) {
        const int64_t i = d.off + x * (int64_t)d.pitch + y;
        if (flag_fluid(flags[i])) {
            const double q = lap(cf, i, d.pitch, d.rdx2, d.rdy2);
            rq = res[i] * q;
            qq = q * q;
        }
    }
    tile_sum(rq, sh, part, S_RQ);
    tile_sum(qq, sh, part, S_QQ);
}

__global__ void __launch_bounds__(256)
k_learn_alpha(const ChanDesc* __restrict__ cd, LearnChan* __restrict__ lc,
              const double* __restrict__ part)
{
    __shared__ double sh[256];
    const int c = blockIdx.x;
    const ChanDesc& d = cd[c];
    const double rq = chan_total(d, part, S_RQ, sh);
    const double qq = chan_total(d, part, S_QQ, sh);
    if (threadIdx.x) return;
    LearnChan& l = lc[c];
    l.alpha = (l.run && qq > 0.0) ? -rq / qq : 0.0;
}

// p += alpha * cf at interior fluid cells.
__global__ void __launch_bounds__(TILE_THREADS)
k_learn_apply(const ChanDesc* __restrict__ cd, const Tile* __restrict__ tiles,
              const LearnChan* __restrict__ lc, const uint8_t* __restrict__ flags,
              const double* __restrict__ cf, double* __restrict__ p)
{
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const double a = lc[t.chan].alpha;
    if (a == 0.0) return;
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    if (x >= d.nx || y >= d.ny || !interior(d, x, y)) return;
    const int64_t i = d.off + x * (int64_t)d.pitch + y;
    if (flag_fluid(flags[i])) p[i] = fma(a, cf[i], p[i]);
}

// Start of a blocked phase: running channels get `budget` sweeps (0 = the
// rest of itermax), the others are marked done.
__global__ void k_learn_arm(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
                            const LearnChan* __restrict__ lc, int32_t* __restrict__ active_out,
                            int nch, int budget)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nch) return;
    ChanState& s = cs[c];
    const LearnChan& l = lc[c];
    s.pbuf = 0;
    s.sor_iters = 0;
    if (l.run) {
        const int rest = cd[c].itermax - l.sweeps;
        s.budget = budget > 0 ? budget : (rest > 1 ? rest : 1);
        s.active = 1;
        s.stop_launch = -1;
        s.exhausted = 0;
    } else {
        s.budget = 0;
        s.active = 0;
        s.stop_launch = 0;
    }
    active_out[c] = s.active;
}

// End of a blocked phase: count its sweeps; a channel that stopped inside
// its budget has converged.  final = 1 also publishes the solve's totals.
__global__ void k_learn_collect(ChanState* __restrict__ cs, LearnChan* __restrict__ lc, int nch,
                                int final_phase)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nch) return;
    ChanState& s = cs[c];
    LearnChan& l = lc[c];
    if (l.run) {
        l.sweeps += (int)s.sor_iters;
        if (!s.exhausted || final_phase) l.run = 0;
    }
    if (final_phase) {
        s.sor_iters = (uint32_t)l.sweeps;
        s.budget = 0;
        s.active = 0;
    }
}

// Capture: pass 0 sums d = p - p0 (slot 5), pass 1 sums (d - mean)^2 (slot 6),
// over interior fluid cells.
__global__ void __launch_bounds__(TILE_THREADS)
k_learn_delta_sums(const ChanDesc* __restrict__ cd, const Tile* __restrict__ tiles,
                   const LearnChan* __restrict__ lc, const uint8_t* __restrict__ flags,
                   const double* __restrict__ p, const double* __restrict__ p0,
                   double* __restrict__ part, int pass)
{
    __shared__ double sh[TILE_THREADS];
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    double v = 0.0;
    if (x < d.nx && y < d.ny && interior(d, x, y)) {
        const int64_t i = d.off + x * (int64_t)d.pitch + y;
        if (flag_fluid(flags[i])) {
            const double dl = p[i] - p0[i];
            v = pass ? (dl - lc[t.chan].dmean) * (dl - lc[t.chan].dmean) : dl;
        }
    }
    tile_sum(v, sh, part, pass ? S_DSQ : S_DSUM);
}

__global__ void __launch_bounds__(256)
k_learn_delta_stats(const ChanDesc* __restrict__ cd, LearnChan* __restrict__ lc,
                    const double* __restrict__ part, int pass)
{
    __shared__ double sh[256];
    const int c = blockIdx.x;
    const ChanDesc& d = cd[c];
    const double sum = chan_total(d, part, pass ? S_DSQ : S_DSUM, sh);
    if (threadIdx.x) return;
    LearnChan& l = lc[c];
    if (!pass) l.dmean = d.n_inner > 0.0 ? sum / d.n_inner : 0.0;
    else l.dscale = sum > 0.0 ? sqrt(d.n_inner / sum) : 0.0;
}

// Capture: the normalised pressure change as the training target of the
// sample in slot (slot0 + b) % dcap; zero outside interior fluid cells.
__global__ void __launch_bounds__(256)
k_learn_target(const ChanDesc* __restrict__ cd, const LearnChan* __restrict__ lc,
               const uint8_t* __restrict__ flags, const double* __restrict__ p,
               const double* __restrict__ p0, int nch, int H, int W, float* __restrict__ dt,
               int slot0, int dcap)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t hw = (int64_t)H * W;
    if (k >= hw * nch) return;
    const int b = (int)(k / hw);
    const int64_t q = k - (int64_t)b * hw;
    const int h = (int)(q / W), w = (int)(q - (int64_t)h * W);
    const ChanDesc& d = cd[b];
    float v = 0.f;
    if (h <= d.nx - 3 && w <= d.ny - 3) {
        const int64_t i = d.off + (h + 1) * (int64_t)d.pitch + (w + 1);
        if (flag_fluid(flags[i])) v = (float)((p[i] - p0[i] - lc[b].dmean) * lc[b].dscale);
    }
    dt[((slot0 + b) % dcap) * hw + q] = v;
}

// ------------------------------------------------------------ launchers

static inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }
static const dim3 TB(TILE_Y, TILE_X);

void launch_learn_residual(hipStream_t st, int ntiles, const ChanDesc* cd, const ChanState* cs,
                           const Tile* tiles, const uint8_t* flags, const double* rhs,
                           const double* p, double* res, double* part) {
    k_learn_residual<<<ntiles, TB, 0, st>>>(cd, cs, tiles, flags, rhs, p, res, part);
}
void launch_learn_scale(hipStream_t st, int nch, const ChanDesc* cd, ChanState* cs, LearnChan* lc,
                        const double* part, int first) {
    k_learn_scale<<<nch, 256, 0, st>>>(cd, cs, lc, part, first);
}
void launch_learn_pack(hipStream_t st, int nch, const ChanDesc* cd, const LearnChan* lc,
                       const uint8_t* flags, const double* res, int H, int W, __bf16* a0,
                       __bf16* dx, int slot0, int dcap) {
    k_learn_pack<<<nb256((int64_t)H * W * nch), 256, 0, st>>>(cd, lc, flags, res, nch, H, W, a0,
                                                              dx, slot0, dcap);
}
void launch_learn_unpack(hipStream_t st, int ntiles, const ChanDesc* cd, const Tile* tiles,
                         const LearnChan* lc, const uint8_t* flags, const float* y, int H, int W,
                         double* cf) {
    k_learn_unpack<<<ntiles, TB, 0, st>>>(cd, tiles, lc, flags, y, H, W, cf);
}
void launch_learn_dots(hipStream_t st, int ntiles, const ChanDesc* cd, const Tile* tiles,
                       const uint8_t* flags, const double* res, const double* cf, double* part) {
    k_learn_dots<<<ntiles, TB, 0, st>>>(cd, tiles, flags, res, cf, part);
}
void launch_learn_alpha(hipStream_t st, int nch, const ChanDesc* cd, LearnChan* lc,
                        const double* part) {
    k_learn_alpha<<<nch, 256, 0, st>>>(cd, lc, part);
}
void launch_learn_apply(hipStream_t st, int ntiles, const ChanDesc* cd, const Tile* tiles,
                        const LearnChan* lc, const uint8_t* flags, const double* cf, double* p) {
    k_learn_apply<<<ntiles, TB, 0, st>>>(cd, tiles, lc, flags, cf, p);
}
void launch_learn_arm(hipStream_t st, int nch, const ChanDesc* cd, ChanState* cs,
                      const LearnChan* lc, int32_t* active, int budget) {
    k_learn_arm<<<nb256(nch), 256, 0, st>>>(cd, cs, lc, active, nch, budget);
}
void launch_learn_collect(hipStream_t st, int nch, ChanState* cs, LearnChan* lc, int final_phase) {
    k_learn_collect<<<nb256(nch), 256, 0, st>>>(cs, lc, nch, final_phase);
}
void launch_learn_target(hipStream_t st, int ntiles, int nch, const ChanDesc* cd,
                         const Tile* tiles, LearnChan* lc, const uint8_t* flags, const double* p,
                         const double* p0, double* part, int H, int W, float* dt, int slot0,
                         int dcap) {
    k_learn_delta_sums<<<ntiles, TB, 0, st>>>(cd, tiles, lc, flags, p, p0, part, 0);
    k_learn_delta_stats<<<nch, 256, 0, st>>>(cd, lc, part, 0);
    k_learn_delta_sums<<<ntiles, TB, 0, st>>>(cd, tiles, lc, flags, p, p0, part, 1);
    k_learn_delta_stats<<<nch, 256, 0, st>>>(cd, lc, part, 1);
    k_learn_target<<<nb256((int64_t)H * W * nch), 256, 0, st>>>(cd, lc, flags, p, p0, nch, H, W,
                                                                dt, slot0, dcap);
}

}  // namespace stm
