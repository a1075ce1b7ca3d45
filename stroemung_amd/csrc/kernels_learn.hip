// gfx950 kernels joining the FP64 pressure solve and the BF16 pressure network.
//
// With A = -Lap (symmetric positive definite on mean-free fields, Neumann
// ghosts) the pressure equation is A p = b, b = -(rhs - mean), and the error
// e = p* - p solves A e = r with r = b - A p = Lap(p) - (rhs - mean).
//
// The network works on a grid coarsened by f (1, 2, 4, 8 or 16) and acts as the
// coarse solver of a two-grid cycle whose smoother is red-black SOR:
//   r   = Lap(p) - (rhs - mean)             fine interior fluid cells, ghosts current
//   rc  = R r                               coarse pixel = sum of r over its f x f
//                                           fine cells / f^2 (solid cells add 0)
//   x   = -rc / rms(rc)                     network input channel 0; channel 1 is the
//                                           fluid fraction of the pixel
//   y   = net(x)                            trained to approximate e_c rms(rc) / h_c^2,
//                                           h_c^2 = f^2 dx dy: the grid-unit inverse
//                                           Laplacian of x on the coarse grid
//   z   = P y                               bilinear between coarse pixel centres,
//                                           weighted by fluid fraction
//   p  += a z                               a = h_c^2 / rms(rc) (the network's own
//                                           magnitude), or the a minimising the energy
//                                           norm of the new error,
//                                           a = <r, z> / <z, A z> = -<r, z> / <z, Lap z>
// after which budgeted red-black sweeps remove what the coarse grid cannot
// carry.  SOR with omega near 2 damps every mode shorter than a few dozen
// cells at a rate near omega - 1 per sweep and only the long modes slowly;
// those are what a coarse grid resolves and the network's receptive field
// covers.  f = 1 is the network on the fine grid itself.
//
// The same kernels build training pairs from plain solves: the coarse
// residual at the start of a solve (and after a few sweeps) with R of the
// mean-free pressure change the solve went on to make, in the units of y.
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

// Lap(a - b) at i, differences taken cell by cell.
__device__ __forceinline__ double lap_diff(const double* __restrict__ a,
                                           const double* __restrict__ b, int64_t i, int64_t P,
                                           double rdx2, double rdy2) {
    const double c = a[i] - b[i];
    const double e = a[i + P] - b[i + P], w = a[i - P] - b[i - P];
    const double n = a[i - 1] - b[i - 1], so = a[i + 1] - b[i + 1];
    return fma((e - c) - (c - w), rdx2, ((so - c) - (c - n)) * rdy2);
}

// Sums of N values over one tile block into part[blockIdx.x * NPART +
// slot[j]]: a fixed-order shuffle reduction inside each wave, then the
// TILE_THREADS / 64 wave sums added in wave order.  sh holds N * that many
// doubles; a call ends with a barrier, so sh can be reused at once.
template <int N>
__device__ void tile_sums(const double (&v)[N], double* sh, double* __restrict__ part,
                          const int (&slot)[N]) {
    constexpr int NW = TILE_THREADS / 64;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    double r[N];
#pragma unroll
    for (int j = 0; j < N; j++) r[j] = v[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int j = 0; j < N; j++) r[j] = r[j] + __shfl_down(r[j], off, 64);
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < N; j++) sh[j * NW + wave] = r[j];
    __syncthreads();
    if (tid < N) {
        double acc = 0.0;
#pragma unroll
        for (int w = 0; w < NW; w++) acc = acc + sh[tid * NW + w];
        part[(int64_t)blockIdx.x * NPART + slot[tid]] = acc;
    }
    __syncthreads();
}

__device__ void tile_sum(double v, double* sh, double* __restrict__ part, int slot) {
    const double vv[1] = {v};
    const int ss[1] = {slot};
    tile_sums<1>(vv, sh, part, ss);
}

// Block sum (256 threads, 1-D) returned to every thread.
__device__ double block_total(double v, double* sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st) sh[tid] = sh[tid] + sh[tid + st];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// One channel's partials of `slot` summed in a fixed order (256 threads).
__device__ double chan_total(const ChanDesc& d, const double* __restrict__ part, int slot,
                             double* sh) {
    double acc = 0.0;
    for (int k = d.tile0 + threadIdx.x; k < d.tile0 + d.ntiles; k += blockDim.x)
        acc = acc + part[(int64_t)k * NPART + slot];
    return block_total(acc, sh);
}

constexpr int S_RES2 = 2, S_RZ = 3, S_ZQ = 4, S_DSUM = 5;
// two-direction step (slots 0 and 1 are free once the solve has begun)
constexpr int S_RS = 0, S_SQ = 1, S_ZS = 6;
// three-direction step: s2 = prev - prev2
constexpr int S_RS2 = 7, S_S2Q = 8, S_ZS2 = 9, S_S1S2 = 10;

// Sum over the interior fluid cells of coarse pixel (hc, wc) of a[g], or of
// (a[g] - a0[g]) - shift when a0 is given; *nf = the number of those cells.
__device__ __forceinline__ double block_fine_sum(const ChanDesc& d, const uint8_t* __restrict__ flags,
                                                 const double* __restrict__ a,
                                                 const double* __restrict__ a0, double shift,
                                                 int hc, int wc, int f, int* nf) {
    double sum = 0.0;
    int n = 0;
    for (int i = 0; i < f; i++) {
        const int x = hc * f + i + 1;
        if (x > d.nx - 2) break;
        for (int j = 0; j < f; j++) {
            const int y = wc * f + j + 1;
            if (y > d.ny - 2) break;
            const int64_t g = d.off + x * (int64_t)d.pitch + y;
            if (flag_fluid(flags[g])) {
                sum += a0 ? (a[g] - a0[g]) - shift : a[g];
                n++;
            }
        }
    }
    *nf = n;
    return sum;
}

}  // namespace

// r = Lap(p) - (rhs - mean) at interior fluid cells, 0 elsewhere; slot 2 = sum r^2.
// p's ghosts must be current.  After the first phase of a solve, tiles of
// channels that are no longer running are skipped (their r and slot 2 are
// read only while a channel runs).
__global__ void __launch_bounds__(TILE_THREADS)
k_learn_residual(const ChanDesc* __restrict__ cd, const ChanState* __restrict__ cs,
                 const LearnChan* __restrict__ lc, int first,
                 const Tile* __restrict__ tiles, const uint8_t* __restrict__ flags,
                 const double* __restrict__ rhs, const double* __restrict__ p,
                 double* __restrict__ res, double* __restrict__ part)
{
    __shared__ double sh[TILE_THREADS];
    const Tile t = tiles[blockIdx.x];
    if (!first && !lc[t.chan].run) return;       // uniform over the block, before any barrier
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

// Per channel: has the solve converged?  first = 1 at the start of a solve.
__global__ void __launch_bounds__(256)
k_learn_check(const ChanDesc* __restrict__ cd, ChanState* __restrict__ cs,
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
        l.hist = 0;
        l.run = (d.solver == SOLVER_RB && d.itermax > 0) ? 1 : 0;
    }
    if (l.run && res2 <= s.tol2) {
        l.run = 0;
        s.res2 = res2;
        s.norm2 = res2 / d.fluid_cells;
        s.exhausted = 0;
    } else if (l.run && l.sweeps >= d.itermax) {
        // the budgeted phases used up itermax: the solve ends here, exhausted,
        // with this residual (the plain solve's result at itermax)
        l.run = 0;
        s.res2 = res2;
        s.norm2 = res2 / d.fluid_cells;
        s.exhausted = 1;
    }
    l.alpha = 0.0;
    l.beta = 0.0;
    l.gamma = 0.0;
}

// Coarse residual rc and fluid fraction mc of every pixel of an H x W frame
// (grid: pixel blocks x channels); per-block (sum rc^2, fluid pixels) into
// cpart[c * gridDim.x + blockIdx.x].
__global__ void __launch_bounds__(256)
k_learn_restrict(const ChanDesc* __restrict__ cd, const LearnChan* __restrict__ lc,
                 const uint8_t* __restrict__ flags, const double* __restrict__ res, int H, int W,
                 int f, float* __restrict__ rc, float* __restrict__ mc,
                 double2* __restrict__ cpart)
{
    __shared__ double sh[256];
    const int c = blockIdx.y;
    const int64_t hw = (int64_t)H * W;
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const ChanDesc& d = cd[c];
    double v = 0.0, fr = 0.0;
    if (q < hw) {
        if (lc[c].run) {
            const int hc = (int)(q / W), wc = (int)(q - (int64_t)hc * W);
            int nf = 0;
            const double sum = block_fine_sum(d, flags, res, nullptr, 0.0, hc, wc, f, &nf);
            v = sum / (double)(f * f);
            fr = (double)nf / (double)(f * f);
        }
        rc[c * hw + q] = (float)v;
        mc[c * hw + q] = (float)fr;
    }
    const double s2 = block_total(v * v, sh);
    const double n = block_total(fr > 0.0 ? 1.0 : 0.0, sh);
    if (threadIdx.x == 0) cpart[(int64_t)c * gridDim.x + blockIdx.x] = make_double2(s2, n);
}

// Network input scale 1 / rms(rc) over the channel's fluid pixels; save >= 0
// keeps it for captured sample `save`.
__global__ void __launch_bounds__(256)
k_learn_nscale(LearnChan* __restrict__ lc, const double2* __restrict__ cpart, int nblk, int save)
{
    __shared__ double sh[256];
    const int c = blockIdx.x;
    double s2 = 0.0, n = 0.0;
    for (int k = threadIdx.x; k < nblk; k += 256) {
        const double2 v = cpart[(int64_t)c * nblk + k];
        s2 = s2 + v.x;
        n = n + v.y;
    }
    s2 = block_total(s2, sh);
    n = block_total(n, sh);
    if (threadIdx.x) return;
    LearnChan& l = lc[c];
    l.scale = (l.run && s2 > 0.0) ? sqrt(n / s2) : 0.0;
    if (save >= 0) l.cscale[save] = l.scale;
}

// The running channels in order: map[s] = channel of batch entry s,
// inv[c] = batch entry of channel c (-1 if not running), *live = count.
__global__ void __launch_bounds__(64)
k_learn_compact(const LearnChan* __restrict__ lc, int nch, int32_t* __restrict__ map,
                int32_t* __restrict__ inv, int32_t* __restrict__ live)
{
    // one wave, 64 channels per step: a ballot of the running flags and the
    // count of set bits below each lane give the packed positions in order
    const int lane = threadIdx.x;
    int base = 0;
    for (int c0 = 0; c0 < nch; c0 += 64) {
        const int c = c0 + lane;
        const bool run = c < nch && lc[c].run;
        const unsigned long long m = __ballot(run);
        const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
        if (c < nch) inv[c] = run ? pos : -1;
        if (run) map[pos] = c;
        base += __popcll(m);
    }
    if (lane == 0) *live = base;
}

// Network input (H x W x 32 bf16 per batch entry: -rc * scale, fluid
// fraction, zeros) and/or a compact 2-channel dataset sample in slot
// (slot0 + b) % dcap.  With a map, batch entry s (s < *live) holds channel
// map[s] and the dataset is not written; without, entry b holds channel b.
__global__ void __launch_bounds__(256)
k_learn_pack(const LearnChan* __restrict__ lc, const float* __restrict__ rc,
             const float* __restrict__ mc, int nch, int H, int W, __bf16* __restrict__ a0,
             __bf16* __restrict__ dx, int slot0, int dcap, const int32_t* __restrict__ map,
             const int32_t* __restrict__ live)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t hw = (int64_t)H * W;
    if (k >= hw * nch) return;
    const int b = (int)(k / hw);
    const int64_t q = k - (int64_t)b * hw;
    if (map && b >= *live) return;
    const int c = map ? map[b] : b;
    const int64_t src = (int64_t)c * hw + q;
    const float v0 = (float)(-(double)rc[src] * lc[c].scale), v1 = mc[src];
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
    if (dx && !map) {
        const int64_t s = (slot0 + b) % dcap;
        dx[(s * hw + q) * 2] = (__bf16)v0;
        dx[(s * hw + q) * 2 + 1] = (__bf16)v1;
    }
}

// Correction z = P y at interior fluid cells of running channels, 0 at the
// other cells (boundary ghosts are set afterwards).  Fine interior index u
// lies at coarse coordinate (u + 1/2) / f - 1/2 (pixel centres at integers);
// the four surrounding pixels are weighted bilinearly and by fluid fraction.
__global__ void __launch_bounds__(TILE_THREADS)
k_learn_prolong(const ChanDesc* __restrict__ cd, const Tile* __restrict__ tiles,
                const LearnChan* __restrict__ lc, const uint8_t* __restrict__ flags,
                const float* __restrict__ y, const float* __restrict__ mc, int H, int W, int f,
                const int32_t* __restrict__ inv, double* __restrict__ cf)
{
    const Tile t = tiles[blockIdx.x];
    if (!lc[t.chan].run) return;                 // z of a stopped channel is never used
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, yy = t.y0 + threadIdx.x;
    if (x >= d.nx || yy >= d.ny) return;
    const int64_t i = d.off + x * (int64_t)d.pitch + yy;
    double v = 0.0;
    if (interior(d, x, yy) && flag_fluid(flags[i])) {
        const int hc_n = (d.nx - 2 + f - 1) / f, wc_n = (d.ny - 2 + f - 1) / f;
        const double xi = fmin(fmax(((x - 1) + 0.5) / f - 0.5, 0.0), (double)(hc_n - 1));
        const double eta = fmin(fmax(((yy - 1) + 0.5) / f - 0.5, 0.0), (double)(wc_n - 1));
        const int i0 = (int)xi, j0 = (int)eta;
        const int i1 = i0 + 1 < hc_n ? i0 + 1 : i0, j1 = j0 + 1 < wc_n ? j0 + 1 : j0;
        const double tx = xi - i0, ty = eta - j0;
        const int64_t base = (int64_t)t.chan * H * W;     // channel's frame in mc
        const int64_t ybase = (int64_t)(inv ? inv[t.chan] : t.chan) * H * W;  // its batch entry
        const int64_t k00 = (int64_t)i0 * W + j0, k01 = (int64_t)i0 * W + j1;
        const int64_t k10 = (int64_t)i1 * W + j0, k11 = (int64_t)i1 * W + j1;
        const double w00 = (1.0 - tx) * (1.0 - ty) * mc[base + k00];
        const double w01 = (1.0 - tx) * ty * mc[base + k01];
        const double w10 = tx * (1.0 - ty) * mc[base + k10], w11 = tx * ty * mc[base + k11];
        const double ws = (w00 + w01) + (w10 + w11);
        if (ws > 0.0)
            v = ((w00 * y[ybase + k00] + w01 * y[ybase + k01]) +
                 (w10 * y[ybase + k10] + w11 * y[ybase + k11])) / ws;
    }
    cf[i] = v;
}

// q = Lap(z) at interior fluid cells: slot 3 = sum r z, slot 4 = sum z q
// (running channels only: the step length of the others is 0).  With
// prev != nullptr also s = p - prev, the change of p over the previous cycle
// (p and prev both hold ghosts consistent with their interiors, so s does):
// slot 0 = sum r s, slot 1 = sum s Lap(s), slot 6 = sum z Lap(s).
__global__ void __launch_bounds__(TILE_THREADS)
k_learn_dots(const ChanDesc* __restrict__ cd, const Tile* __restrict__ tiles,
             const LearnChan* __restrict__ lc, const uint8_t* __restrict__ flags,
             const double* __restrict__ res, const double* __restrict__ cf,
             const double* __restrict__ p, const double* __restrict__ prev,
             const double* __restrict__ prev2, double* __restrict__ part)
{
    __shared__ double sh[TILE_THREADS];
    const Tile t = tiles[blockIdx.x];
    if (!lc[t.chan].run) return;                 // uniform over the block, before any barrier
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    double rz = 0.0, zq = 0.0, rs = 0.0, sq = 0.0, zs = 0.0;
    double rs2 = 0.0, s2q = 0.0, zs2 = 0.0, s1s2 = 0.0;
    if (x < d.nx && y < d.ny && interior(d, x, y)) {
        const int64_t i = d.off + x * (int64_t)d.pitch + y;
        if (flag_fluid(flags[i])) {
            const double q = lap(cf, i, d.pitch, d.rdx2, d.rdy2);
            rz = res[i] * cf[i];
            zq = cf[i] * q;
            if (prev) {
                const double sv = p[i] - prev[i];
                const double ls = lap_diff(p, prev, i, d.pitch, d.rdx2, d.rdy2);
                rs = res[i] * sv;
                sq = sv * ls;
                zs = cf[i] * ls;
                if (prev2) {
                    const double s2 = prev[i] - prev2[i];
                    const double l2 = lap_diff(prev, prev2, i, d.pitch, d.rdx2, d.rdy2);
                    rs2 = res[i] * s2;
                    s2q = s2 * l2;
                    zs2 = cf[i] * l2;
                    s1s2 = sv * l2;
                }
            }
        }
    }
    if (prev2) {
        const double v[9] = {rz, zq, rs, sq, zs, rs2, s2q, zs2, s1s2};
        const int slots[9] = {S_RZ, S_ZQ, S_RS, S_SQ, S_ZS, S_RS2, S_S2Q, S_ZS2, S_S1S2};
        tile_sums<9>(v, sh, part, slots);
    } else if (prev) {
        const double v[5] = {rz, zq, rs, sq, zs};
        const int slots[5] = {S_RZ, S_ZQ, S_RS, S_SQ, S_ZS};
        tile_sums<5>(v, sh, part, slots);
    } else {
        const double v[2] = {rz, zq};
        const int slots[2] = {S_RZ, S_ZQ};
        tile_sums<2>(v, sh, part, slots);
    }
}

// step 0: energy-norm line search; step 1: the network's own magnitude;
// step 2: p + a z + b s with (a, b) minimising the energy norm of the error
// over span{z, s} (s = change of p over the previous cycle; first = 1 on the
// first cycle of a solve, which has no s).  With A = -Lap, symmetric positive
// definite on the mean-free fluid functions, and r = b - A p the residual:
//   [<z,Az> <z,As>] [a]   [<r,z>]
//   [<z,As> <s,As>] [b] = [<r,s>]
// falling back to the line search along z when z and s are nearly parallel.
__global__ void __launch_bounds__(256)
k_learn_alpha(const ChanDesc* __restrict__ cd, LearnChan* __restrict__ lc,
              const double* __restrict__ part, int step, int f, int first)
{
    __shared__ double sh[256];
    const int c = blockIdx.x;
    const ChanDesc& d = cd[c];
    const double rz = chan_total(d, part, S_RZ, sh);
    const double zq = chan_total(d, part, S_ZQ, sh);
    // kept previous p's: step 2 has one after the first cycle, step 3 counts
    // them per channel (the graph-replayed cycles cannot be told apart)
    const int hist = step == 3 ? lc[c].hist : ((step == 2 && !first) ? 1 : 0);
    double rs = 0.0, sq = 0.0, zs = 0.0, rs2 = 0.0, s2q = 0.0, zs2 = 0.0, s1s2 = 0.0;
    if (hist >= 1) {
        rs = chan_total(d, part, S_RS, sh);
        sq = chan_total(d, part, S_SQ, sh);
        zs = chan_total(d, part, S_ZS, sh);
    }
    if (hist >= 2) {
        rs2 = chan_total(d, part, S_RS2, sh);
        s2q = chan_total(d, part, S_S2Q, sh);
        zs2 = chan_total(d, part, S_ZS2, sh);
        s1s2 = chan_total(d, part, S_S1S2, sh);
    }
    if (threadIdx.x) return;
    LearnChan& l = lc[c];
    double a = 0.0, b = 0.0, g = 0.0;
    if (l.run && l.scale > 0.0) {
        if (step == 1) {
            a = (double)(f * f) * d.delx * d.dely / l.scale;
        } else if (zq < 0.0) {               // <z, Lap z> < 0 for any z not constant
            a = -rz / zq;
            if (hist >= 1 && sq < 0.0) {
                const double zaz = -zq, sas = -sq, zas = -zs;
                const double det = zaz * sas - zas * zas;
                if (det > 1e-10 * zaz * sas) {
                    a = (rz * sas - zas * rs) / det;
                    b = (zaz * rs - zas * rz) / det;
                }
            }
            if (hist >= 2 && sq < 0.0 && s2q < 0.0) {
                // Cholesky of the 3 x 3 energy matrix over (z, s1, s2); any
                // pivot below 1e-10 of its diagonal keeps the 2-direction step
                const double m00 = -zq, m01 = -zs, m02 = -zs2;
                const double m11 = -sq, m12 = -s1s2, m22 = -s2q;
                const double l00 = sqrt(m00), l10 = m01 / l00, l20 = m02 / l00;
                const double d11 = m11 - l10 * l10;
                if (d11 > 1e-10 * m11) {
                    const double l11 = sqrt(d11), l21 = (m12 - l20 * l10) / l11;
                    const double d22 = m22 - l20 * l20 - l21 * l21;
                    if (d22 > 1e-10 * m22) {
                        const double l22 = sqrt(d22);
                        const double y0 = rz / l00, y1 = (rs - l10 * y0) / l11;
                        const double y2 = (rs2 - l20 * y0 - l21 * y1) / l22;
                        g = y2 / l22;
                        b = (y1 - l21 * g) / l11;
                        a = (y0 - l10 * b - l20 * g) / l00;
                    }
                }
            }
        }
    }
    l.alpha = a;
    l.beta = b;
    l.gamma = g;
    if (step == 3 && l.run) l.hist = min(l.hist + 1, 2);   // this cycle's p is kept
}

// p += alpha * z at interior fluid cells; with prev != nullptr (the
// two-direction step) p += alpha * z + beta * (p - prev), and prev takes the
// old p.
__global__ void __launch_bounds__(TILE_THREADS)
k_learn_apply(const ChanDesc* __restrict__ cd, const Tile* __restrict__ tiles,
              const LearnChan* __restrict__ lc, const uint8_t* __restrict__ flags,
              const double* __restrict__ cf, double* __restrict__ p, double* __restrict__ prev,
              double* __restrict__ prev2)
{
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const double a = lc[t.chan].alpha;
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    if (prev) {
        // two-direction step: every cell of a running channel also keeps
        // the p it had (ghosts included) for the next cycle's s
        if (!lc[t.chan].run || x >= d.nx || y >= d.ny) return;
        const double b = lc[t.chan].beta;
        const int64_t i = d.off + x * (int64_t)d.pitch + y;
        const double po = p[i], pp = prev[i];
        if (interior(d, x, y) && flag_fluid(flags[i])) {
            double v = fma(b, po - pp, fma(a, cf[i], po));
            if (prev2) v = fma(lc[t.chan].gamma, pp - prev2[i], v);
            p[i] = v;
        }
        if (prev2) prev2[i] = pp;
        prev[i] = po;
        return;
    }
    if (a == 0.0) return;
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
    // the sweeps left below itermax bound every phase (k_learn_check ends a
    // solve whose phases used them all before it is armed again)
    const int rest = cd[c].itermax - l.sweeps;
    if (l.run && rest > 0) {
        s.budget = (budget > 0 && budget < rest) ? budget : rest;
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

// End of a budgeted phase run as `launches` unchecked blocked launches: each
// running channel swept its armed budget (the phase's, or less where itermax
// is near; counted here).  A channel whose budget ended inside the phase was
// stopped by the launch that took that decision, which recorded where its
// latest p is; otherwise it is in the second buffer when `launches` is odd.
// They stay running: the next residual test decides whether they have
// converged (or used up itermax).
__global__ void k_learn_phase_end(ChanState* __restrict__ cs, LearnChan* __restrict__ lc,
                                  int nch, int launches)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nch || !lc[c].run) return;
    ChanState& s = cs[c];
    lc[c].sweeps += s.budget;
    if (s.stop_launch < 0) {
        s.pbuf = launches & 1;
        s.stop_launch = launches - 1;
    }
    s.sor_iters = (uint32_t)s.budget;
    s.exhausted = 1;
    s.active = 0;
}

// Capture: slot 5 = sum of p - p0 over interior fluid cells.
__global__ void __launch_bounds__(TILE_THREADS)
k_learn_delta_sum(const ChanDesc* __restrict__ cd, const Tile* __restrict__ tiles,
                  const uint8_t* __restrict__ flags, const double* __restrict__ p,
                  const double* __restrict__ p0, double* __restrict__ part)
{
    __shared__ double sh[TILE_THREADS];
    const Tile t = tiles[blockIdx.x];
    const ChanDesc& d = cd[t.chan];
    const int x = t.x0 + threadIdx.y, y = t.y0 + threadIdx.x;
    double v = 0.0;
    if (x < d.nx && y < d.ny && interior(d, x, y)) {
        const int64_t i = d.off + x * (int64_t)d.pitch + y;
        if (flag_fluid(flags[i])) v = p[i] - p0[i];
    }
    tile_sum(v, sh, part, S_DSUM);
}

__global__ void __launch_bounds__(256)
k_learn_delta_mean(const ChanDesc* __restrict__ cd, LearnChan* __restrict__ lc,
                   const double* __restrict__ part)
{
    __shared__ double sh[256];
    const int c = blockIdx.x;
    const ChanDesc& d = cd[c];
    const double sum = chan_total(d, part, S_DSUM, sh);
    if (threadIdx.x == 0) lc[c].dmean = d.n_inner > 0.0 ? sum / d.n_inner : 0.0;
}

// Capture: the training target of the sample in slot (slot0 + b) % dcap --
// R of the mean-free pressure change times scale / h_c^2 with the scale of
// the stored residual, i.e. in the units of the network's output.
__global__ void __launch_bounds__(256)
k_learn_target(const ChanDesc* __restrict__ cd, const LearnChan* __restrict__ lc,
               const uint8_t* __restrict__ flags, const double* __restrict__ p,
               const double* __restrict__ p0, int nch, int H, int W, int f,
               float* __restrict__ dt, int slot0, int dcap, int which)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t hw = (int64_t)H * W;
    if (k >= hw * nch) return;
    const int b = (int)(k / hw);
    const int64_t q = k - (int64_t)b * hw;
    const int hc = (int)(q / W), wc = (int)(q - (int64_t)hc * W);
    const ChanDesc& d = cd[b];
    int nf = 0;
    const double sum = block_fine_sum(d, flags, p, p0, lc[b].dmean, hc, wc, f, &nf);
    const double h2 = (double)(f * f) * d.delx * d.dely;
    dt[((slot0 + b) % dcap) * hw + q] = (float)(sum / (double)(f * f) * (lc[b].cscale[which] / h2));
}

// ------------------------------------------------------------ launchers

static inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }
static const dim3 TB(TILE_Y, TILE_X);

int learn_coarse_blocks(int H, int W) { return (int)nb256((int64_t)H * W); }

void launch_learn_residual(hipStream_t st, int ntiles, const ChanDesc* cd, const ChanState* cs,
                           const LearnChan* lc, int first, const Tile* tiles,
                           const uint8_t* flags, const double* rhs, const double* p, double* res,
                           double* part) {
    k_learn_residual<<<ntiles, TB, 0, st>>>(cd, cs, lc, first, tiles, flags, rhs, p, res, part);
}
void launch_learn_check(hipStream_t st, int nch, const ChanDesc* cd, ChanState* cs, LearnChan* lc,
                        const double* part, int first) {
    k_learn_check<<<nch, 256, 0, st>>>(cd, cs, lc, part, first);
}
void launch_learn_restrict(hipStream_t st, int nch, const ChanDesc* cd, LearnChan* lc,
                           const uint8_t* flags, const double* res, int H, int W, int f,
                           float* rc, float* mc, double2* cpart, int save) {
    const int nblk = learn_coarse_blocks(H, W);
    k_learn_restrict<<<dim3(nblk, nch), 256, 0, st>>>(cd, lc, flags, res, H, W, f, rc, mc, cpart);
    k_learn_nscale<<<nch, 256, 0, st>>>(lc, cpart, nblk, save);
}
void launch_learn_compact(hipStream_t st, int nch, const LearnChan* lc, int32_t* map, int32_t* inv,
                          int32_t* live) {
    k_learn_compact<<<1, 64, 0, st>>>(lc, nch, map, inv, live);
}
void launch_learn_pack(hipStream_t st, int nch, const LearnChan* lc, const float* rc,
                       const float* mc, int H, int W, __bf16* a0, __bf16* dx, int slot0,
                       int dcap, const int32_t* map, const int32_t* live) {
    k_learn_pack<<<nb256((int64_t)H * W * nch), 256, 0, st>>>(lc, rc, mc, nch, H, W, a0, dx,
                                                              slot0, dcap, map, live);
}
void launch_learn_prolong(hipStream_t st, int ntiles, const ChanDesc* cd, const Tile* tiles,
                          const LearnChan* lc, const uint8_t* flags, const float* y,
                          const float* mc, int H, int W, int f, const int32_t* inv, double* cf) {
    k_learn_prolong<<<ntiles, TB, 0, st>>>(cd, tiles, lc, flags, y, mc, H, W, f, inv, cf);
}
void launch_learn_dots(hipStream_t st, int ntiles, const ChanDesc* cd, const Tile* tiles,
                       const LearnChan* lc, const uint8_t* flags, const double* res,
                       const double* cf, const double* p, const double* prev,
                       const double* prev2, double* part) {
    k_learn_dots<<<ntiles, TB, 0, st>>>(cd, tiles, lc, flags, res, cf, p, prev, prev2, part);
}
void launch_learn_alpha(hipStream_t st, int nch, const ChanDesc* cd, LearnChan* lc,
                        const double* part, int step, int f, int first) {
    k_learn_alpha<<<nch, 256, 0, st>>>(cd, lc, part, step, f, first);
}
void launch_learn_apply(hipStream_t st, int ntiles, const ChanDesc* cd, const Tile* tiles,
                        const LearnChan* lc, const uint8_t* flags, const double* cf, double* p,
                        double* prev, double* prev2) {
    k_learn_apply<<<ntiles, TB, 0, st>>>(cd, tiles, lc, flags, cf, p, prev, prev2);
}
void launch_learn_arm(hipStream_t st, int nch, const ChanDesc* cd, ChanState* cs,
                      const LearnChan* lc, int32_t* active, int budget) {
    k_learn_arm<<<nb256(nch), 256, 0, st>>>(cd, cs, lc, active, nch, budget);
}
void launch_learn_phase_end(hipStream_t st, int nch, ChanState* cs, LearnChan* lc,
                            int launches) {
    k_learn_phase_end<<<nb256(nch), 256, 0, st>>>(cs, lc, nch, launches);
}
void launch_learn_collect(hipStream_t st, int nch, ChanState* cs, LearnChan* lc, int final_phase) {
    k_learn_collect<<<nb256(nch), 256, 0, st>>>(cs, lc, nch, final_phase);
}
void launch_learn_target(hipStream_t st, int ntiles, int nch, const ChanDesc* cd,
                         const Tile* tiles, LearnChan* lc, const uint8_t* flags, const double* p,
                         const double* p0, double* part, int H, int W, int f, float* dt,
                         int slot0, int dcap, int which) {
    k_learn_delta_sum<<<ntiles, TB, 0, st>>>(cd, tiles, flags, p, p0, part);
    k_learn_delta_mean<<<nch, 256, 0, st>>>(cd, lc, part);
    k_learn_target<<<nb256((int64_t)H * W * nch), 256, 0, st>>>(cd, lc, flags, p, p0, nch, H, W,
                                                                f, dt, slot0, dcap, which);
}

}  // namespace stm
