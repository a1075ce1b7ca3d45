// Host-side launchers of the gfx950 kernels (one translation unit per group).
#pragma once
#include <hip/hip_runtime.h>

#include "stm_internal.h"

namespace stm {

// kernels_step.hip
void launch_bc_velocity(hipStream_t, int nbtot, const BEntry*, double2* restore, double* u,
                        double* v);
void launch_fg(hipStream_t, int ntiles, const ChanDesc*, const ChanState*, const Tile*,
               const uint8_t* flags, const double* u, const double* v, double* f, double* g);
void launch_rhs(hipStream_t, int ntiles, const ChanDesc*, const ChanState*, const Tile*,
                const double* f, const double* g, double* rhs);
// launch_fg then launch_rhs in one pass (the same bits)
void launch_fg_rhs(hipStream_t, int ntiles, const ChanDesc*, const ChanState*, const Tile*,
                   const uint8_t* flags, const double* u, const double* v, double* f, double* g,
                   double* rhs);
void launch_restore(hipStream_t, int nbtot, const BEntry*, const double2* restore, double* u,
                    double* v);
void launch_ranges(hipStream_t, int ntiles, int nchan, const ChanDesc*, ChanState*, const Tile*,
                   const uint8_t* flags, const double* u, const double* v, const double* p,
                   double* part, int want_pressure);
// ADAP_UV on the fluid cells fused with the ranges; after launch_restore
void launch_adap_ranges(hipStream_t, int ntiles, int nchan, const ChanDesc*, ChanState*,
                        const Tile*, const uint8_t* flags, const double* f, const double* g,
                        const double* p, double* u, double* v, double* part,
                        int want_pressure);
void launch_delt(hipStream_t, int nchan, const ChanDesc*, ChanState*);
void launch_advance_time(hipStream_t, int nchan, ChanState*);
void launch_fill(hipStream_t, int64_t n, double* a, double val);
void launch_zero_cells(hipStream_t, int n, const int64_t* idx, double* u, double* v,
                       double* p);
void launch_copy(hipStream_t, int64_t bytes, const double* src, double* dst);

// kernels_pressure.hip
void launch_ghost(hipStream_t, int nbtot, const BEntry*, const ChanState*, double* p, int colour,
                  int only_active);
// rhs_mean and the sum of (rhs - mean)^2 per channel (partials in slots 0, 1)
void launch_rhs_stats(hipStream_t, int ntiles, int nchan, const ChanDesc*, ChanState*,
                      const Tile*, const uint8_t* flags, const double* rhs, double* part);
void launch_solve_begin(hipStream_t, int nchan, const ChanDesc*, ChanState*, const double* part,
                        int32_t* active);
void launch_initial_norm(hipStream_t, int nchan, const ChanDesc*, ChanState*, const double* rhs,
                         const double* p, double* scratch);
void launch_lex_solve(hipStream_t, int nchan, const ChanDesc*, ChanState*, const BEntry*,
                      const uint8_t* flags, const double* rhs, double* p, double* scratch);

// kernels_rbtb.hip: temporally blocked red-black SOR
struct TbVariant {
    int k;          // sweeps per launch
    int oxm, oym;   // largest output tile away from a channel's edges
    int ex, ey;     // loaded region; the halo is (ex - oxm) / 2 = 2k + 4
    int nt;         // threads per workgroup
    int codes;      // 1: keeps its per-thread cell codes in the codes table
};
extern const TbVariant TB_VARIANTS[];
extern const int N_TB_VARIANTS;
// mode bits: TB_CHECK, form the residual after the sweeps and stop on it;
// TB_CODES_PUT / TB_CODES_GET, compute the cell codes from the flags and
// store them in the codes table / load them from it (variants with codes).
enum : int { TB_CHECK = 1, TB_CODES_PUT = 2, TB_CODES_GET = 4 };
// codes table: 32 bytes per thread of every tile, at (tile idx * nt + thread)
constexpr int TB_CODE_BYTES = 32;
// Launch L of a blocked solve: reads the partials of launch L - 1 from
// part_prev, writes its own to part_cur (one double per tile).
void launch_rb_tb(hipStream_t, int variant, int ntb, const ChanDesc*, ChanState*, const TbTile*,
                  const uint8_t* flags, const double* rhs, double* pa, double* pb,
                  const double* part_prev, double* part_cur, int32_t* active,
                  uint32_t* codes, int L, int mode);
void launch_pbuf_fix(hipStream_t, int ntiles, int nchan, const ChanDesc*, ChanState*,
                     const Tile*, const double* pb, double* pa);

// kernels_view.hip: the GUI image of one channel (npix = nx * ny pixels)
void launch_render(hipStream_t, const ChanDesc*, const ChanState*, int ch, int64_t npix,
                   const uint8_t* flags, const double* u, const double* v, const double* p,
                   int mode, uint32_t* rgba);
// The static image buffer of stm_render_static: RENDER_STATIC_PIXELS pixels of
// 4 B (16 MiB) per device, devices 0 .. RENDER_STATIC_DEVICES - 1.  init looks
// its device address up (the device must be current; once per device);
// buffer returns it, or null before init.
constexpr int64_t RENDER_STATIC_PIXELS = (int64_t)4 << 20;
constexpr int RENDER_STATIC_DEVICES = 64;
hipError_t render_static_init(int device);
uint32_t* render_static_buffer(int device);

// kernels_learn.hip: the FP64 side of the learned correction and of capture
void launch_learn_residual(hipStream_t, int ntiles, const ChanDesc*, const ChanState*,
                           const LearnChan*, int first, const Tile*, const uint8_t* flags,
                           const double* rhs, const double* p, double* res, double* part);
void launch_learn_check(hipStream_t, int nch, const ChanDesc*, ChanState*, LearnChan*,
                        const double* part, int first);
// pixel blocks of an H x W frame (the size of cpart per channel)
int learn_coarse_blocks(int H, int W);
void launch_learn_restrict(hipStream_t, int nch, const ChanDesc*, LearnChan*,
                           const uint8_t* flags, const double* res, int H, int W, int f,
                           float* rc, float* mc, double2* cpart, int save);
// running channels -> batch entries (map), channels -> entries (inv), count
void launch_learn_compact(hipStream_t, int nch, const LearnChan*, int32_t* map, int32_t* inv,
                          int32_t* live);
void launch_learn_pack(hipStream_t, int nch, const LearnChan*, const float* rc, const float* mc,
                       int H, int W, __bf16* a0, __bf16* dx, int slot0, int dcap,
                       const int32_t* map = nullptr, const int32_t* live = nullptr);
void launch_learn_prolong(hipStream_t, int ntiles, const ChanDesc*, const Tile*,
                          const LearnChan*, const uint8_t* flags, const float* y,
                          const float* mc, int H, int W, int f, const int32_t* inv, double* cf);
void launch_learn_dots(hipStream_t, int ntiles, const ChanDesc*, const Tile*, const LearnChan*,
                       const uint8_t* flags, const double* res, const double* cf, const double* p,
                       const double* prev, const double* prev2, double* part);
void launch_learn_alpha(hipStream_t, int nch, const ChanDesc*, LearnChan*, const double* part,
                        int step, int f, int first);
void launch_learn_apply(hipStream_t, int ntiles, const ChanDesc*, const Tile*, const LearnChan*,
                        const uint8_t* flags, const double* cf, double* p, double* prev,
                        double* prev2);
void launch_learn_arm(hipStream_t, int nch, const ChanDesc*, ChanState*, const LearnChan*,
                      int32_t* active, int budget);
void launch_learn_collect(hipStream_t, int nch, ChanState*, LearnChan*, int final_phase);
void launch_learn_phase_end(hipStream_t, int nch, ChanState*, LearnChan*, int launches);
void launch_learn_target(hipStream_t, int ntiles, int nch, const ChanDesc*, const Tile*,
                         LearnChan*, const uint8_t* flags, const double* p, const double* p0,
                         double* part, int H, int W, int f, float* dt, int slot0, int dcap,
                         int which);

// kernels_sync.hip: the cost of a resident blocked solver's waits (measurement)
constexpr int SYNC_THREADS = 768;
void launch_sync_tile(hipStream_t, int nwg, double* out, int step);
void launch_sync_hold(hipStream_t, int nwg, double* out, int64_t ticks);
// hipSuccess when nwg workgroups of both waiting kernels fit on the device at once
hipError_t sync_coresident(int nwg, int device);
hipError_t launch_sync_grid(hipStream_t, int nwg, double* out, int steps, int write);
hipError_t launch_sync_neighbours(hipStream_t, int side, double* out, unsigned* flags,
                                  int steps, int write, unsigned spin_limit, unsigned* err);

}  // namespace stm
