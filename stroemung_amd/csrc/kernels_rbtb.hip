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
// cells of the other colour.  A straight ghost (one fluid neighbour) is read
// only by that neighbour, which takes its own value in its place -- the value
// the start-of-sweep refresh would copy.  A corner ghost (two fluid
// neighbours) of colour c is refreshed in the phase of colour c, from values
// that the phase does not change.  Every ghost is refreshed after the last
// sweep.  Two barriers per sweep.  The update has a few forms (with or
// without the substitutions, see Blk::update), chosen per wave.
//
// Launch order: workgroups go to the XCDs round-robin; the host orders each
// channel's tiles so that every XCD gets one compact block of them (shared
// halos then hit in that XCD's L2).  The tile carries its channel's geometry,
// so the loads issue after a single scalar load.
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

#include <type_traits>

#include "kernels.h"
#include "stm_internal.h"

namespace stm {

namespace {

// For a ghost cell, the fluid neighbours its pressure copies
// (src/grid/mod.rs:343-412): N, S, E, W; one bit for a straight edge, a
// vertical and a horizontal one for a corner, whose ghost is their mean.
enum : uint32_t { C_GN = 1, C_GS = 2, C_GE = 4, C_GW = 8 };

// ghost_bits(e), 4 bits per edge kind E_N .. E_NW (0 for E_NONE and above
// E_NW), so the codes are formed without branches
constexpr uint64_t GHOST_TAB =
    (uint64_t)C_GN << (4 * E_N) | (uint64_t)(C_GN | C_GE) << (4 * E_NE) |
    (uint64_t)C_GE << (4 * E_E) | (uint64_t)(C_GS | C_GE) << (4 * E_SE) |
    (uint64_t)C_GS << (4 * E_S) | (uint64_t)(C_GS | C_GW) << (4 * E_SW) |
    (uint64_t)C_GW << (4 * E_W) | (uint64_t)(C_GN | C_GW) << (4 * E_NW);

__device__ __forceinline__ uint32_t ghost_bits(uint32_t edge) {
    return (uint32_t)(GHOST_TAB >> (4 * (edge & 15u))) & 15u;
}

// The forms a wave runs (wave-uniform bits): it holds a ghost cell, a corner
// ghost, a fluid cell that is not updated (SEL); its vertical substitutions
// (2 bits at W_VSHIFT: none, per lane, per cell); per-cell E ones (HX) and
// W ones (HW: a wall on the low x side has only W, one on the high side
// only E); every ghost of the wave is straight and its fluid neighbour is a
// cell of the same thread (OWN).
enum : int { W_GHOST = 1, W_DIAG = 2, W_SEL = 4, W_VSHIFT = 3, W_HX = 32, W_OWN = 64,
             W_HW = 128 };

// What a thread knows of its 2R own cells, cell i = 2r + q, packed so the
// codes take a few registers: bit i of `upd` = updated by SOR, of `out` = in
// the output tile, of `diag` = a corner ghost (two fluid neighbours); the
// edge kind of each ghost cell (4 bits per cell, 0 for any other cell); and
// for an updated cell, bit i of `vsub` / `hsub` (N / E) and bit SH + i (S /
// W): that neighbour is a straight ghost whose one fluid neighbour is this
// cell.
template <int R>
struct Codes {
    static_assert(2 * R <= 32, "one bit per cell in a word");
    static constexpr int NG = (2 * R + 7) / 8;
    static constexpr int SH = 2 * R <= 16 ? 16 : 32;
    using Sub = std::conditional_t<(SH + 2 * R > 32), uint64_t, uint32_t>;
    uint32_t upd, out, diag;
    Sub vsub, hsub;
    uint32_t g[NG];
    __device__ __forceinline__ uint32_t edge(int i) const {
        return (g[i >> 3] >> (4 * (i & 7))) & 15u;
    }
};

// SWAR helpers on words of four cells' flag bytes (cell 4k + j in byte j of
// word k).  zero7: bit 7 of each byte set where the byte is 0 (bytes below
// 0x80); ones: bit 7 of each byte to the full byte; nib7: bits 7, 15, 23, 31
// to bits 0 .. 3.
__device__ __forceinline__ uint32_t zero7(uint32_t t) {
    const uint32_t nz = (((t & 0x7f7f7f7fu) + 0x7f7f7f7fu) | t) & 0x80808080u;
    return nz ^ 0x80808080u;
}
__device__ __forceinline__ uint32_t ones(uint32_t b7) { return (b7 << 1) - (b7 >> 7); }
__device__ __forceinline__ uint32_t nib7(uint32_t m) {
    const uint32_t x = (m >> 7) & 0x01010101u;
    return (x | (x >> 7) | (x >> 14) | (x >> 21)) & 15u;
}

// Straight ghosts of kind `kind` (one fluid neighbour, on that side) among
// four flag bytes: bit 7 of each matching byte.  Not fluid and that edge
// kind; the 0xff of a cell outside the channel has the fluid bit, so it is
// none.
__device__ __forceinline__ uint32_t straight7(uint32_t w, uint32_t kind) {
    return zero7((w & 0x1f1f1f1fu) ^ (kind * 0x01010101u));
}

// One DPP move of a double (both halves); lanes of rows outside ROWS keep
// `old`'s value.
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_move(double v, double old) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), CTRL,
                                               ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), CTRL,
                                               ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

// The sum of v over the wave, in VALU (DPP) steps rather than LDS round
// trips, in a fixed order: pairs, quads, the quads of a row, then rows
// (r0 + r1) and (r2 + r3), then (r2 + r3) + (r0 + r1).  Uniform result.
__device__ __forceinline__ double wave_sum(double v) {
    v = v + dpp_move<0xb1>(v, 0.0);          // quad_perm [1, 0, 3, 2]
    v = v + dpp_move<0x4e>(v, 0.0);          // quad_perm [2, 3, 0, 1]
    v = v + dpp_move<0x124>(v, 0.0);         // row_ror 4
    v = v + dpp_move<0x128>(v, 0.0);         // row_ror 8
    v = v + dpp_move<0x142, 0xa>(v, 0.0);    // row_bcast 15 into rows 1, 3
    v = v + dpp_move<0x143, 0xc>(v, 0.0);    // row_bcast 31 into rows 2, 3
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// The workgroup's sum of v, in wave order, for thread 0 (the last use of
// wsum in the kernel: no barrier after it).
template <int NT>
__device__ __forceinline__ double block_sum_to_first(double v, double* wsum) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
#pragma unroll
        for (int w = 0; w < NT / 64; w++) s = s + wsum[w];
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

    // Ghost value (src/grid/mod.rs:343-412): one fluid neighbour, or
    // (a + b) * 0.5 of a vertical and a horizontal one; ghost bits g != 0.
    __device__ static __forceinline__ double ghost_value(uint32_t g, double pn, double ps,
                                                         double pe, double pw) {
        const double a = (g & C_GN) ? pn : ps;
        const double b = (g & C_GE) ? pe : pw;
        const bool vert = g & (C_GN | C_GS), horz = g & (C_GE | C_GW);
        return vert ? (horz ? (a + b) * 0.5 : a) : b;
    }

    // Red-black update of the own cells of colour c.  Every value read has
    // colour c ^ 1, which the phase does not change.
    //
    // What the cells that are not updated hold during the sweeps: a
    // straight ghost (src/grid/mod.rs:343-412: one fluid neighbour) copies
    // its one fluid neighbour at the start of each sweep, and that neighbour
    // is its only reader, so the reader takes its own value in its place --
    // the same operand -- and the ghost's stored value is not read; a
    // corner ghost is refreshed after the update of its colour, before its
    // readers run; every ghost is refreshed after the last sweep; a cell
    // with no fluid neighbour (solid) is read by no update, and an output
    // cell of that kind is taken from the input at write-back; cells on the
    // region's border or outside the channel are never output, and what
    // they hold reaches the output no sooner than their staleness would.
    // So the update may store into all of them (no select) unless the wave
    // holds a fluid cell that is not updated (a fluid cell on the channel's
    // frame: SEL, the select on cc.upd).
    //
    // The substitutions: VM 1 = the reader's straight ghost is its own other
    // cell of the pair, the same way for every updated cell of that slot in
    // the lane (the channel's walls along x): one lane flag per slot (vown
    // bit 0: q = 1 reads N, bit 1: q = 0 reads S); VM 2 = per-cell bits
    // (cc.vsub); HX = per-cell E / W bits (cc.hsub).  The form is chosen
    // per wave.
    template <int c, bool SEL, int VM, int HX>
    __device__ static __forceinline__ void update(double (*P)[PS], double (&pv)[R][2],
                                                  const double (&rr)[R][2], const Codes<R>& cc,
                                                  uint32_t vown, int lx0, int m, double omw,
                                                  double mid, double rdx2, double rdy2) {
        double x[R], xe, xw;
        reads<c>(P, lx0, m, x, xe, xw);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int q = c ^ RP0 ^ (r & 1), i = 2 * r + q;
            const double pc = pv[r][q];
            double pn = q ? pv[r][0] : x[r];
            double ps = q ? x[r] : pv[r][1];
            double pe = r + 1 < R ? pv[r + 1 < R ? r + 1 : r][q] : xe;
            double pw = r >= 1 ? pv[r >= 1 ? r - 1 : r][q] : xw;
            if constexpr (VM == 1) {
                if (q) pn = (vown & 1u) ? pc : pn;
                else ps = (vown & 2u) ? pc : ps;
            }
            if constexpr (VM == 2) {
                pn = ((cc.vsub >> i) & 1u) ? pc : pn;
                ps = ((cc.vsub >> (Codes<R>::SH + i)) & 1u) ? pc : ps;
            }
            if constexpr ((HX & 1) != 0) pe = ((cc.hsub >> i) & 1u) ? pc : pe;
            if constexpr ((HX & 2) != 0) pw = ((cc.hsub >> (Codes<R>::SH + i)) & 1u) ? pc : pw;
            const double tt = fma(pe + pw, rdx2, (ps + pn) * rdy2);
            const double nv = fma(omw, pc, mid * (tt - rr[r][q]));
            pv[r][q] = SEL ? (((cc.upd >> i) & 1u) ? nv : pc) : nv;
            P[c][OFF + (lx0 + r) * HY + m] = pv[r][q];
        }
    }

    // Ghost cells of colour c take their fluid neighbours' pressure
    // (src/grid/mod.rs:343-412: one neighbour, or (a + b) * 0.5 of a
    // vertical and a horizontal one) -- the reference's start-of-sweep copy:
    // a ghost of colour c reads only fluid cells of colour c ^ 1 and is read
    // only by them, so refreshing it in colour c's phase gives the values
    // that copy gives.  Inside the sweeps only the corner ghosts (DIAG: two
    // readers) need it; after the last sweep every ghost takes the final
    // values.  Branch-free over the thread's cells, batched reads; run only
    // by waves that hold such a cell (the choice is wave-uniform).
    template <int c, bool DIAG>
    __device__ static __forceinline__ void ghosts(double (*P)[PS], double (&pv)[R][2],
                                                  const Codes<R>& cc, int lx0, int m) {
        double x[R], xe, xw;
        reads<c>(P, lx0, m, x, xe, xw);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int q = c ^ RP0 ^ (r & 1);
            const uint32_t g = ghost_bits(cc.edge(2 * r + q));
            const double pn = q ? pv[r][0] : x[r];
            const double ps = q ? x[r] : pv[r][1];
            const double pe = r + 1 < R ? pv[r + 1 < R ? r + 1 : r][q] : xe;
            const double pw = r >= 1 ? pv[r >= 1 ? r - 1 : r][q] : xw;
            const bool on = DIAG ? ((cc.diag >> (2 * r + q)) & 1u) != 0 : g != 0;
            pv[r][q] = on ? ghost_value(g, pn, ps, pe, pw) : pv[r][q];
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

    // The loaded cells and (FLAGS) their flags (FL, two per (row, pair)) into
    // LDS; with FLAGS returns W_GHOST / W_DIAG: the wave holds a ghost / a
    // corner ghost (wave-uniform), else 0.  Ends with a barrier.
    template <bool FLAGS, int NF>
    __device__ static __forceinline__ int init(double (*P)[PS], uint16_t* FL,
                                               const double (&pv)[R][2],
                                               const uint32_t (&fo)[NF], const Codes<R>& cc,
                                               int lx0, int m) {
        int any = 0;
#pragma unroll
        for (int r = 0; r < R; r++) {
#pragma unroll
            for (int q = 0; q < 2; q++) P[col(r, q)][OFF + (lx0 + r) * HY + m] = pv[r][q];
            if (FLAGS) FL[(lx0 + r) * HY + m] = (uint16_t)(fo[r >> 1] >> (16 * (r & 1)));
        }
        int w = 0;
        if (FLAGS) {
#pragma unroll
            for (int k = 0; k < Codes<R>::NG; k++) any |= cc.g[k] != 0;
            w = __ballot(any) != 0 ? W_GHOST : 0;
            w |= __ballot(cc.diag != 0) != 0 ? W_DIAG : 0;
        }
        __syncthreads();
        return w;
    }

    // The final refresh of a wave whose ghosts are all straight with their
    // fluid neighbour in the thread (W_OWN): each takes that neighbour's
    // value from the registers -- the other cell of its pair (bit i of cpv),
    // the next row (bit i of cph) or the previous row (bit SH + i of cph).
    // Only that neighbour reads such a ghost, from the registers too, so
    // nothing goes to LDS.  The sources are fluid cells, which no copy
    // changes.
    __device__ static __forceinline__ void own_refresh(double (*pv)[2], uint32_t cpv,
                                                       typename Codes<R>::Sub cph) {
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int i = 2 * r + q;
                double v = pv[r][q];
                v = ((cpv >> i) & 1u) ? pv[r][q ^ 1] : v;
                if (r + 1 < R) v = ((cph >> i) & 1u) ? pv[r + 1 < R ? r + 1 : r][q] : v;
                if (r >= 1) v = ((cph >> (Codes<R>::SH + i)) & 1u) ? pv[r >= 1 ? r - 1 : r][q] : v;
                pv[r][q] = v;
            }
    }

    // Before the sweeps (a barrier follows): the black corner ghosts, for
    // the first red half.
    __device__ static __forceinline__ void start(double (*P)[PS], double (&pv)[R][2],
                                                 const Codes<R>& cc, int lx0, int m, int wm) {
        if (wm & W_DIAG) ghosts<1, true>(P, pv, cc, lx0, m);
    }

    // The sweeps, from a barrier after start(); the final refresh of every
    // ghost, and the residual when checked.
    // The sweeps in one form of the update (each form has its own loop, so
    // no form's registers meet another's inside it; the forms without corner
    // ghosts have no branch in the loop).
    template <bool SEL, int VM, int HX, bool DIAG>
    __device__ static __forceinline__ void sweeps(double (*P)[PS], double (&pv)[R][2],
                                                  const double (&rr)[R][2], Codes<R>& cc,
                                                  uint32_t vown, int lx0, int m, int nsw,
                                                  double omw, double mid, double rdx2,
                                                  double rdy2) {
        for (int sw = 0; sw < nsw; sw++) {
            // the lane masks the codes give are cheap to form: form them in each
            // sweep rather than keep them all live in scalar registers (the
            // empty asm "changes" the codes, so no mask is hoisted; the codes
            // themselves stay one copy)
            asm volatile("" : "+v"(cc.upd), "+v"(cc.diag), "+v"(cc.vsub), "+v"(cc.hsub));
#pragma unroll
            for (int k = 0; k < Codes<R>::NG; k++) asm volatile("" : "+v"(cc.g[k]));
            // red cells, and the red corner ghosts for the black half
            update<0, SEL, VM, HX>(P, pv, rr, cc, vown, lx0, m, omw, mid, rdx2, rdy2);
            if (DIAG) ghosts<0, true>(P, pv, cc, lx0, m);
            __syncthreads();
            // black cells, and the black corner ghosts for the next red half
            update<1, SEL, VM, HX>(P, pv, rr, cc, vown, lx0, m, omw, mid, rdx2, rdy2);
            if (DIAG && sw + 1 < nsw) ghosts<1, true>(P, pv, cc, lx0, m);
            __syncthreads();
        }
    }

    // The sweeps, from a barrier after start(), in the wave's form; the final
    // refresh of every ghost, and the residual when checked.
    __device__ static __forceinline__ double run(double (*P)[PS], double (&pv)[R][2],
                                                 const double (&rr)[R][2],
                                                 Codes<R>& cc, uint32_t vown, uint32_t cpv,
                                                 typename Codes<R>::Sub cph, int lx0, int m,
                                                 int nsw, bool check, int wm,
                                                 double omw, double mid, double rdx2,
                                                 double rdy2, bool lds_refresh) {
        const bool ghost = wm & W_GHOST;
        const int vm = (wm >> W_VSHIFT) & 3;
        const int hx = ((wm & W_HX) ? 1 : 0) | ((wm & W_HW) ? 2 : 0);
#define STM_SW(S, V, X, D) sweeps<S, V, X, D>(P, pv, rr, cc, vown, lx0, m, nsw, omw, mid, rdx2, rdy2)
#define STM_SWX(V)                                                           \
    do {                                                                     \
        if (hx == 0) STM_SW(false, V, 0, false);                             \
        else if (hx == 1) STM_SW(false, V, 1, false);                        \
        else if (hx == 2) STM_SW(false, V, 2, false);                        \
        else STM_SW(false, V, 3, false);                                     \
    } while (0)
        // a wave with a corner ghost or a fluid cell that is not updated runs
        // the general form: the select, per-cell substitutions, the corner
        // refresh
        if (wm & (W_SEL | W_DIAG)) STM_SW(true, 2, 3, true);
        else if (vm == 0) STM_SWX(0);
        else if (vm == 1) STM_SWX(1);
        else STM_SWX(2);
#undef STM_SWX
#undef STM_SW
        // every ghost from the final values, for the output and the residual
        if (ghost) {
            if (wm & W_OWN) {
                own_refresh(pv, cpv, cph);
            } else {
                ghosts<0, false>(P, pv, cc, lx0, m);
                ghosts<1, false>(P, pv, cc, lx0, m);
            }
        }
        double acc = 0.0;
        if (check) {
            // the residual reads neighbours through LDS: a workgroup none of
            // whose waves refreshes ghosts there finds them ordered by the
            // last sweep's barrier
            if (lds_refresh) __syncthreads();
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
           int32_t* __restrict__ active_out, uint32_t* __restrict__ codes, int L, int mode)
{
    constexpr int NT = NS * HY;
    constexpr int H = 2 * K + 4;
    constexpr int EX = NS * R;
    constexpr int EY = 2 * HY;
    static_assert(NT % 64 == 0 && NT <= 1024, "workgroup size");
    static_assert(R % 2 == 0, "rows per thread must be even");
    static_assert(EX > 2 * H && EY > 2 * H, "tile smaller than its halo");

    __shared__ double P[2][EX * HY + 2 * HY];   // plus guard slots, see Blk::OFF
    __shared__ uint16_t FL[EX * HY];            // flags of (row, pair), for the neighbours
    __shared__ double wsum[NT / 64];

    // The tile carries its channel's geometry: the region's loads issue after
    // one scalar load, together with the channel's state.
    const TbTile t = tiles[blockIdx.x];
    const int c = t.chan;
    ChanState& s = cs[c];
    // read here, tested only after the loads: no load waits for it
    const int stopped = s.stop_launch;
    const double mean = s.rhs_mean;
    // the channel's constants for the stop decision and the sweeps, read
    // with its state: each would otherwise be a memory round trip after a
    // barrier (the compiler does not move a load above the stop test)
    const ChanDesc& d = cd[c];
    const int budget = s.budget, d_itermax = d.itermax;
    const double tol2 = s.tol2;
    const double omw = d.omw, mid = d.mid, rdx2 = d.rdx2, rdy2 = d.rdy2;
    const int tid = threadIdx.x;
    const bool check = mode & TB_CHECK;
    const double* __restrict__ pin = (L & 1) ? pb : pa;
    double* __restrict__ pout = (L & 1) ? pa : pb;

    const int ex0 = t.rx0, ey0 = t.ry0;          // region origin; ey0 is even
    const int nx = t.nx, ny = t.ny;
    const int64_t pitch = t.pitch, base = t.off;
    const int strip = tid / HY, m = tid - (tid / HY) * HY;
    const int lx0 = strip * R;
    const int gy = ey0 + 2 * m;                  // even: (gy, gy+1) is one aligned pair

    // The previous launch's residual partials of this channel: the first
    // share is loaded with the region below, unconditionally (clamped index).
    const bool sum_parts = L > 0 && check;
    const double part0 = part_prev[t.tb0 + min(tid, t.ntb - 1)];

    // ---- load; the solve is for rhs - mean (the all-Neumann problem needs
    // zero mean).  Every address is clamped into the channel (rows of pitch
    // >= ny + 1 for odd ny, so an aligned pair load at the last even y is in
    // bounds) and values outside the channel are masked afterwards: the
    // loads carry no branches and issue back to back.
    double pv[R][2], rr[R][2];
    // own cells' flags, 0xff outside the channel, packed: row r in bits
    // 16r .. 16r + 15 of fo (q = 0 in the low byte)
    uint32_t fo[(2 * R + 3) / 4] = {};
    const int gyc = min(max(gy, 0), (ny - 1) & ~1);   // even
    const bool y0in = gy >= 0 && gy < ny, y1in = gy >= 0 && gy + 1 < ny;
    double2 pp[R], rh[R];
    uint32_t ff[R];
    // a thread's codes (CACHE: 2R <= 16 cells, two 16-byte words: upd | out,
    // diag | cpv, vsub, hsub; cph, g[0], g[1], vown | wm | rest | the
    // workgroup's LDS-refresh and corner bits), loaded
    // unconditionally with the region and used when mode has TB_CODES_GET
    constexpr bool CACHE = 2 * R <= 16;
    uint4 k0 = {}, k1 = {};
    if constexpr (CACHE) {
        const uint4* kp = reinterpret_cast<const uint4*>(codes) + ((int64_t)t.idx * NT + tid) * 2;
        k0 = kp[0];
        k1 = kp[1];
    }
    const bool get = CACHE && (mode & TB_CODES_GET);   // uniform: the codes come from the table
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t row = base + (int64_t)min(max(ex0 + lx0 + r, 0), nx - 1) * pitch;
        pp[r] = *reinterpret_cast<const double2*>(pin + row + gyc);
        rh[r] = *reinterpret_cast<const double2*>(rhs + row + gyc);
    }
    // the flags only feed the codes
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t row = base + (int64_t)min(max(ex0 + lx0 + r, 0), nx - 1) * pitch;
        ff[r] = get ? 0u : (uint32_t)*reinterpret_cast<const uint16_t*>(flags + row + gyc);
    }
    // every load is issued before any of them is used: the wave waits once
    // for the whole region, row by row, instead of once per batch the
    // scheduler would otherwise form
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int gx = ex0 + lx0 + r;
        const bool row_in = gx >= 0 && gx < nx;
        const bool in0 = row_in && y0in, in1 = row_in && y1in;
        pv[r][0] = in0 ? pp[r].x : 0.0;
        pv[r][1] = in1 ? pp[r].y : 0.0;
        rr[r][0] = (in0 ? rh[r].x : 0.0) - mean;
        rr[r][1] = (in1 ? rh[r].y : 0.0) - mean;
        fo[r >> 1] |= ((in0 ? ff[r] & 0xffu : 0xffu) | (in1 ? ff[r] & 0xff00u : 0xff00u)) << (16 * (r & 1));
        // formed here, not after the stop test: otherwise the raw rhs loads
        // stay live (and spill) through the code section
        asm volatile("" : "+v"(rr[r][0]), "+v"(rr[r][1]), "+v"(pv[r][0]), "+v"(pv[r][1]));
    }
#pragma unroll
    for (int k = 0; k < (2 * R + 3) / 4; k++) asm volatile("" : "+v"(fo[k]));
    // this thread's first share of the partials, taken with the region
    double pacc = sum_parts && tid < t.ntb ? part0 : 0.0;
    asm volatile("" : "+v"(pacc));
    // a channel that has stopped: nothing to do (uniform; the loads above
    // are in bounds and their results unused)
    if (stopped >= 0) return;

    // ---- every thread adds its share of the partials in a fixed order,
    // waves add theirs, and after the barrier below every thread adds the wave sums
    // in wave order: the same sum, bit for bit, in every tile of the channel
    {
        if (sum_parts)
            for (int i = t.tb0 + tid + NT; i < t.tb0 + t.ntb; i += NT) pacc = pacc + part_prev[i];
        const double w = wave_sum(pacc);
        if ((tid & 63) == 0) wsum[tid >> 6] = w;
    }
    // ---- the cell codes: computed from the flags (and, with TB_CODES_PUT,
    // stored for the later launches of the sequence, which load them with
    // TB_CODES_GET: the flags and the tiles do not change inside a sequence
    // of launches, and a tile whose channel has stopped stays stopped)
    Codes<R> cc;
    int wm;
    uint32_t vown = 0, cpv = 0;
    typename Codes<R>::Sub cph = 0;
    uint32_t rest = 0;   // output cells that keep their value (see the write-back)
    bool bdiag = true;   // some wave of the workgroup holds a corner ghost
    bool blds = true;    // some wave refreshes its ghosts through LDS after the sweeps
    if (get) {
        cc.upd = k0.x & 0xffffu;
        cc.out = k0.x >> 16;
        cc.diag = k0.y & 0xffffu;
        cpv = k0.y >> 16;
        cc.vsub = k0.z;
        cc.hsub = k0.w;
        cph = k1.x;
        cc.g[0] = k1.y;
        if constexpr (Codes<R>::NG > 1) cc.g[Codes<R>::NG > 1 ? 1 : 0] = k1.z;
        vown = k1.w & 0xffu;
        wm = (int)((k1.w >> 8) & 0xffu);
        rest = (k1.w >> 16) & 0x3fffu;
        bdiag = (k1.w >> 31) != 0;
        blds = ((k1.w >> 30) & 1u) != 0;
        if (ex0 & 1) Blk<K, R, HY, NS, 1>::template init<false>(P, FL, pv, fo, cc, lx0, m);
        else Blk<K, R, HY, NS, 0>::template init<false>(P, FL, pv, fo, cc, lx0, m);
    } else {
        // ---- the codes of the own cells, four cells at a time (the flag bytes
        // of word k of fo are cells 4k .. 4k + 3; byte masks, then cell bits):
        // updated = an interior fluid cell inside the region's border; a ghost =
        // a non-fluid cell inside the border (its edge kind kept); the output
        // tile; 0xff (outside the channel) is none of them
        constexpr int NW = (2 * R + 3) / 4;
        static_assert(2 * NW == R, "two rows per word");
        auto rows = [](bool a, bool b) -> uint32_t {   // byte mask of rows 2k, 2k + 1
            return (a ? 0x0000ffffu : 0u) | (b ? 0xffff0000u : 0u);
        };
        auto slots = [](bool a, bool b) -> uint32_t {  // byte mask of slots q = 0, 1
            return (a ? 0x00ff00ffu : 0u) | (b ? 0xff00ff00u : 0u);
        };
        const uint32_t yin = slots(2 * m >= 1 && 2 * m <= EY - 2, 2 * m + 1 <= EY - 2);
        const uint32_t yupd = slots(gy >= 1 && gy <= ny - 2, gy + 1 >= 1 && gy + 1 <= ny - 2);
        const uint32_t yout = slots(gy >= t.y0 && gy < t.y0 + t.yn,
                                    gy + 1 >= t.y0 && gy + 1 < t.y0 + t.yn);
        auto xin = [&](int r) { const int lx = lx0 + r; return lx >= 1 && lx <= EX - 2; };
        auto xupd = [&](int r) { const int gx = ex0 + lx0 + r; return gx >= 1 && gx <= nx - 2; };
        auto xout = [&](int r) { const int gx = ex0 + lx0 + r; return gx >= t.x0 && gx < t.x0 + t.xn; };
        cc.upd = cc.out = cc.diag = 0;
        cc.vsub = cc.hsub = 0;
        uint32_t keep = 0;   // fluid cells inside the border that are not updated
        uint32_t ghm = 0;    // ghost cells
        uint32_t upd7[NW];   // bit 7 of each updated cell's byte
    #pragma unroll
        for (int k = 0; k < Codes<R>::NG; k++) cc.g[k] = 0;
    #pragma unroll
        for (int k = 0; k < NW; k++) {
            const uint32_t w = fo[k];
            const uint32_t valid = ones(zero7(~w) ^ 0x80808080u);
            const uint32_t fluid = ones((w << 3) & 0x80808080u);
            const uint32_t rin = rows(xin(2 * k), xin(2 * k + 1)) & yin & valid;
            const uint32_t ru = rows(xupd(2 * k), xupd(2 * k + 1)) & yupd;
            const uint32_t u = rin & fluid & ru;
            const uint32_t gh = rin & ~fluid;
            const uint32_t o = valid & rows(xout(2 * k), xout(2 * k + 1)) & yout;
            const uint32_t e4 = w & 0x0f0f0f0fu & gh;                       // ghosts' edge kinds
            const uint32_t corner = (zero7(e4) ^ 0x80808080u) & ~(e4 << 7);   // kind 2, 4, 6 or 8
            upd7[k] = u & 0x80808080u;
            cc.upd |= nib7(u) << (4 * k);
            cc.out |= nib7(o) << (4 * k);
            cc.diag |= nib7(corner) << (4 * k);
            keep |= nib7(rin & fluid & ~ru) << (4 * k);
            ghm |= nib7(zero7(e4) ^ 0x80808080u) << (4 * k);
            const uint32_t y = e4 | (e4 >> 4);
            cc.g[k >> 1] |= ((y & 0xffu) | ((y >> 8) & 0xff00u)) << (16 * (k & 1));
        }

        wm = (ex0 & 1) ? Blk<K, R, HY, NS, 1>::template init<true>(P, FL, pv, fo, cc, lx0, m)
                       : Blk<K, R, HY, NS, 0>::template init<true>(P, FL, pv, fo, cc, lx0, m);

        // ---- the updated cells' neighbours that are straight ghosts of them,
        // from the flags every thread has put in LDS (N / S across the pair
        // boundary, E / W across the strip ends; the rest are the thread's own),
        // four cells at a time: each direction's neighbour flags laid out as the
        // cells' own (byte j of word k), matched against the edge kind that
        // makes this cell their fluid neighbour
        {
            auto fls = [&](int lx, int mm) -> uint32_t {   // a pair's flags, 0xffff outside the region
                const bool in = lx >= 0 && lx < EX && mm >= 0 && mm < HY;
                const int k = min(max(lx, 0), EX - 1) * HY + min(max(mm, 0), HY - 1);
                const uint32_t f = FL[k];
                return in ? f : 0xffffu;
            };
            uint32_t fN[R], fS[R];
    #pragma unroll
            for (int r = 0; r < R; r++) {
                fN[r] = fls(lx0 + r, m - 1) >> 8;      // the S cell of the pair above
                fS[r] = fls(lx0 + r, m + 1) & 0xffu;   // the N cell of the pair below
            }
            const uint32_t fE = fls(lx0 + R, m), fW = fls(lx0 - 1, m);
            uint32_t nsub = 0, ssub = 0, esub = 0, wsub = 0;
    #pragma unroll
            for (int k = 0; k < NW; k++) {
                const uint32_t w = fo[k];
                const uint32_t wn = ((w << 8) & 0xff00ff00u) | fN[2 * k] | (fN[2 * k + 1] << 16);
                const uint32_t ws = ((w >> 8) & 0x00ff00ffu) | (fS[2 * k] << 8) | (fS[2 * k + 1] << 24);
                const uint32_t we = (w >> 16) | ((k + 1 < NW ? fo[k + 1 < NW ? k + 1 : k] & 0xffffu : fE) << 16);
                const uint32_t ww = (w << 16) | (k > 0 ? fo[k > 0 ? k - 1 : k] >> 16 : fW);
                nsub |= nib7(straight7(wn, E_S) & upd7[k]) << (4 * k);
                ssub |= nib7(straight7(ws, E_N) & upd7[k]) << (4 * k);
                esub |= nib7(straight7(we, E_W) & upd7[k]) << (4 * k);
                wsub |= nib7(straight7(ww, E_E) & upd7[k]) << (4 * k);
            }
            using Sub = typename Codes<R>::Sub;
            cc.vsub = (Sub)nsub | (Sub)ssub << Codes<R>::SH;
            cc.hsub = (Sub)esub | (Sub)wsub << Codes<R>::SH;
            // the vertical substitutions from the own other cell (q = 1 reads N:
            // own_n; q = 0 reads S: own_s) or across the pair boundary (cross);
            // per lane when every updated cell of the slot has it (or none does)
            constexpr uint32_t Q1 = 0xaaaaaaaau & ((2 * R < 32) ? (1u << 2 * R) - 1 : ~0u);
            const uint32_t own_n = nsub & Q1, own_s = ssub & ~Q1;
            const uint32_t cross = (nsub & ~Q1) | (ssub & Q1);
            const uint32_t upd1 = cc.upd & Q1, upd0 = cc.upd & ~Q1;
            const bool lane_ok = cross == 0 && (own_n == 0 || own_n == upd1) &&
                                 (own_s == 0 || own_s == upd0);
            vown = (own_n != 0 ? 1u : 0u) | (own_s != 0 ? 2u : 0u);
            // the ghosts whose one fluid neighbour is a cell of this thread (and
            // updated): the other cell of the pair, the next or previous row
            constexpr uint32_t CELLS = (2 * R < 32) ? (1u << 2 * R) - 1 : ~0u;
            cpv = (own_n >> 1) | (own_s << 1);
            const uint32_t ce = (wsub >> 2), cw = (esub << 2) & CELLS;
            cph = (Sub)ce | (Sub)cw << Codes<R>::SH;
            wm |= __ballot(ghm != (cpv | ce | cw)) == 0 ? W_OWN : 0;
            const int vm = __ballot((nsub | ssub) != 0) == 0 ? 0 : (__ballot(!lane_ok) != 0 ? 2 : 1);
            wm |= vm << W_VSHIFT;
            wm |= __ballot(esub != 0) != 0 ? W_HX : 0;
            wm |= __ballot(wsub != 0) != 0 ? W_HW : 0;
            wm |= __ballot(keep != 0) != 0 ? W_SEL : 0;
        }
#pragma unroll
        for (int i = 0; i < 2 * R; i++)
            rest |= (cc.edge(i) == 0 ? 1u : 0u) << i;
        rest &= cc.out & ~cc.upd;
        if (CACHE && (mode & TB_CODES_PUT)) {
            static_assert(!CACHE || 2 * R <= 14, "rest and the workgroup bits share a word");
            bdiag = __syncthreads_or((wm & W_DIAG) != 0) != 0;
            blds = __syncthreads_or((wm & W_GHOST) != 0 && (wm & W_OWN) == 0) != 0;
            uint4* kp = reinterpret_cast<uint4*>(codes) + ((int64_t)t.idx * NT + tid) * 2;
            kp[0] = make_uint4(cc.upd | cc.out << 16, cc.diag | cpv << 16, (uint32_t)cc.vsub,
                               (uint32_t)cc.hsub);
            kp[1] = make_uint4((uint32_t)cph, cc.g[0],
                               Codes<R>::NG > 1 ? cc.g[Codes<R>::NG > 1 ? 1 : 0] : 0u,
                               vown | (uint32_t)wm << 8 | rest << 16 | (blds ? 1u << 30 : 0u) |
                                   (bdiag ? 1u << 31 : 0u));
        }
    }

    // the black corner ghosts for the first red half (before the same barrier)
    // (a workgroup without corner ghosts has nothing to refresh here, and
    // init's barrier already orders the first sweep's reads)
    if (bdiag) {
        if (ex0 & 1) Blk<K, R, HY, NS, 1>::start(P, pv, cc, lx0, m, wm);
        else Blk<K, R, HY, NS, 0>::start(P, pv, cc, lx0, m, wm);
        __syncthreads();
    }

    // ---- stop decision on the previous launch (identical in every tile)
    const int itermax = (budget > 0 && budget < d_itermax) ? budget : d_itermax;
    const int64_t sweeps = (int64_t)L * K;
    const int done = sweeps < itermax ? (int)sweeps : itermax;
    if (L > 0) {
        double res2 = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; w++) res2 = res2 + wsum[w];
        const bool conv = check && res2 <= tol2;
        // a residual that is not a number never meets the test: stop there
        // (the tick then reports it) instead of sweeping to itermax
        const bool nan = check && res2 != res2;
        if (conv || nan || done >= itermax) {
            if (t.idx == t.tb0 && tid == 0) {
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

    const double acc =
        (ex0 & 1) ? Blk<K, R, HY, NS, 1>::run(P, pv, rr, cc, vown, cpv, cph, lx0, m, nsw, check, wm,
                                              omw, mid, rdx2, rdy2, blds)
                  : Blk<K, R, HY, NS, 0>::run(P, pv, rr, cc, vown, cpv, cph, lx0, m, nsw, check, wm,
                                              omw, mid, rdx2, rdy2, blds);

    // ---- this tile's residual partial, before the stores: the barrier of
    // the sum then waits for the residual, not for every wave's write-back
    const double b = block_sum_to_first<NT>(acc, wsum);
    if (tid == 0) part_cur[t.idx] = b;

    // ---- write the output tile.  Output cells that are neither updated nor
    // ghosts keep their value in every sweep of the sequence (the sweeps may
    // have stored into their registers, see Blk::update): launch 0 copies
    // them from its input buffer to its output buffer, after which both
    // buffers hold them, and the later launches leave them out.
    uint32_t wout = cc.out;
    {
        if (L > 0) {
            wout &= ~rest;
        } else if (__ballot(rest != 0) != 0) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int64_t g = base + (int64_t)(ex0 + lx0 + r) * pitch + gy;
                if ((rest >> (2 * r)) & 1u) pv[r][0] = pin[g];
                if ((rest >> (2 * r + 1)) & 1u) pv[r][1] = pin[g + 1];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t g = base + (int64_t)(ex0 + lx0 + r) * pitch + gy;
        const bool o0 = (wout >> (2 * r)) & 1u, o1 = (wout >> (2 * r + 1)) & 1u;
        if (o0 && o1) *reinterpret_cast<double2*>(pout + g) = make_double2(pv[r][0], pv[r][1]);
        else if (o0) pout[g] = pv[r][0];
        else if (o1) pout[g + 1] = pv[r][1];
    }
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
    {K, NS * R - 2 * (2 * K + 4), 2 * HY - 2 * (2 * K + 4), NS * R, 2 * HY, NS * HY, 2 * R <= 16},
const TbVariant TB_VARIANTS[] = {STM_TB_VARIANTS(V_ENTRY)};
#undef V_ENTRY
const int N_TB_VARIANTS = sizeof(TB_VARIANTS) / sizeof(TB_VARIANTS[0]);

void launch_rb_tb(hipStream_t st, int variant, int ntb, const ChanDesc* cd, ChanState* cs,
                  const TbTile* tiles, const uint8_t* flags, const double* rhs, double* pa,
                  double* pb, const double* part_prev, double* part_cur, int32_t* active,
                  uint32_t* codes, int L, int mode) {
#define V_CASE(i, K, R, HY, NS)                                                              \
    case i:                                                                                  \
        k_rb_strip<K, R, HY, NS><<<ntb, NS * HY, 0, st>>>(cd, cs, tiles, flags, rhs, pa, pb, \
                                                          part_prev, part_cur, active,       \
                                                          codes, L, mode);                   \
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
