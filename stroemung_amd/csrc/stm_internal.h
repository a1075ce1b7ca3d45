// Internal layout shared by the host runtime and the gfx950 kernels.
//
// A context holds one or more independent grids ("channels").  Every field
// (u, v, p, f, g, rhs and the u8 cell flags) lives in one packed device
// allocation per field; channel c occupies nx*pitch elements starting at
// element offset `off`, cell (x, y) at off + x*pitch + y.  y is the
// contiguous axis, as in the reference's ndarray C order
// (src/types.rs:12-19); pitch >= ny pads each x-row to a 128-byte multiple.
#pragma once
#include <stdint.h>

namespace stm {

// Cell flags: bit 4 = fluid, bits 5..7 = cell type, bits 0..3 = edge kind.
enum : uint8_t { F_FLUID = 0x10 };
enum CellType : uint8_t { T_FLUID = 0, T_NOSLIP = 1, T_INFLOW = 2, T_OUTFLOW = 3, T_MOVING = 4 };
enum Edge : uint8_t { E_NONE = 0, E_N = 1, E_NE = 2, E_E = 3, E_SE = 4, E_S = 5, E_SW = 6, E_W = 7, E_NW = 8 };

__host__ __device__ inline uint8_t make_flag(uint8_t type, uint8_t edge) {
    return type == T_FLUID ? (uint8_t)F_FLUID : (uint8_t)((type << 5) | edge);
}
__host__ __device__ inline bool flag_fluid(uint8_t f) { return (f & F_FLUID) != 0; }
__host__ __device__ inline uint8_t flag_edge(uint8_t f) { return f & 0x0f; }

// Solvers for the pressure equation.
enum Solver : int32_t { SOLVER_LEX = 0, SOLVER_RB = 1 };

// Immutable per-channel description (device copy in ChanDesc[]).
struct ChanDesc {
    int32_t nx, ny, pitch, solver;
    int64_t off;                 // element offset of cell (0, 0)
    int32_t boff, nb;            // slice of the boundary list
    int32_t tile0, ntiles;       // slice of the tile list
    int32_t itermax, check_every;
    int32_t tb0, ntb;            // slice of the blocked-solver tile list
    double delx, dely, gamma, reynolds, omega, eps, tau;
    double fluid_cells;
    double n_inner;              // interior fluid cells, x in [1, nx-2], y in [1, ny-2]
    // derived solver constants, computed on the host exactly as the oracle does
    double delx2, dely2;         // sq(delx), sq(dely)
    double omw, mid;             // 1 - omega, omega / ((2/dx2) + (2/dy2))
    double rdx2, rdy2;           // 1/dx2, 1/dy2 (red-black only)
    double eps2;                 // sq(eps)
};

// Mutable per-channel state (device copy in ChanState[]).
struct ChanState {
    double delt;                 // step size of the current / next tick
    double time;
    double init_norm2;           // Simulation::initial_norm_squared
    double norm2;                // last solve's returned norm
    double res2, tol2, rhs2;     // red-black: sums over interior fluid cells
    double rhs_mean;             // red-black: mean rhs over interior fluid cells
    double umax, vmax;           // max |u|, |v| over fluid faces (for delt)
    double pressure_range[2], speed_range[2];
    uint32_t iterations;         // ticks done
    uint32_t sor_iters;          // sweeps of the last solve
    int32_t has_init;
    int32_t active;              // red-black solve still running
    int32_t exhausted;           // last solve hit itermax
    int32_t pbuf;                // blocked solver: latest p is in the second buffer
    int32_t stop_launch;         // blocked solver: launch whose result ended the solve, -1 running
    int32_t budget;              // blocked solver: sweep limit below itermax, 0 = none
};

// Per-channel state of the learned correction (device copy in LearnChan[]).
struct LearnChan {
    double scale;                // network input scale: 1 / rms of the coarse residual
    double alpha;                // step length of the network's correction
    double beta;                 // two-direction step: length along the previous change
    double gamma;                // three-direction step: length along the one before
    double dmean;                // capture: mean of the pressure change
    double cscale[8];            // capture: `scale` of each stored residual (LEARN_POINTS)
    int32_t run;                 // the channel has not converged yet
    int32_t sweeps;              // sweeps summed over the phases of one solve
    int32_t hist;                // kept previous p's of this solve (0, 1, 2)
    int32_t pad;
};

// Training pairs one solve can store per channel.
constexpr int LEARN_POINTS = 8;

// One boundary cell (the reference's sorted_boundary_list entry,
// src/grid/mod.rs:64) with what the parallel SETBCOND needs.
struct BEntry {
    int64_t gidx;                // element index into the packed fields
    int32_t chan;
    int32_t pitch;
    uint8_t type, edge;
    uint8_t vw_rule;             // W/NW cells: v[W] is written first by (x-1, y+1)
    uint8_t colour;              // (x + y) & 1
    uint32_t pad;
    double bu, bv;               // wall / inflow velocity
    double vw_const;             // the value that writer stores into v[W]
};

struct Tile {
    int32_t chan, x0, y0, pad;
};

// Output tile of the blocked red-black solver: cells [x0, x0+xn) x [y0, y0+yn).
struct TbTile {
    int32_t chan, x0, y0, xn, yn;   // output cells [x0, x0 + xn) x [y0, y0 + yn)
    int32_t idx;                 // position in the full tile list (residual partial slot)
    int32_t rx0, ry0;            // origin of the loaded region (ry0 even)
    // the channel's geometry and tile slice (copies of its ChanDesc fields),
    // so a launch addresses its region from the tile alone
    int32_t nx, ny, pitch, tb0, ntb, pad;
    int64_t off;
};
static_assert(sizeof(TbTile) == 64, "one 64-byte scalar load per tile");

constexpr int TILE_X = 4;        // x-rows per tile
constexpr int TILE_Y = 64;       // y-cells per tile (one wavefront)
constexpr int TILE_THREADS = TILE_X * TILE_Y;

// Partial-sum slots per tile.
constexpr int NPART = 16;

}  // namespace stm
