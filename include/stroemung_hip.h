/*
 * stroemung_hip.h -- C ABI of libstroemung_hip.so, the MI355X (gfx950)
 * implementation of stroemung's NaSt2D time step.
 *
 * The library is a drop-in for the reference's time-step path: a host that
 * owns a `Simulation` (wickedchicken/stroemung src/simulation.rs:49-69) hands
 * its grid and fields to a context, calls stm_run_simulation_tick() where it
 * called Simulation::run_simulation_tick() (src/simulation.rs:324), and reads
 * the fields back when it needs them.
 *
 * Conventions
 *  - Every field is an nx*ny double array in the reference's layout: element
 *    (x, y) at x*ny + y, y contiguous, (0, 0) the upper-left cell, "north" =
 *    y-1 (src/types.rs:12-19, src/grid/mod.rs:168-170).  nx, ny are the
 *    reference's `size`, i.e. they include the outer ring of cells.
 *  - Functions return STM_OK (0) or a negative STM_ERR_* code and never abort
 *    or throw.  stm_last_error() describes the last failure of a context
 *    (or, for NULL, of the calling thread's last context-free call).
 *  - Caller buffers are borrowed for the duration of a call only.  The
 *    library owns all device memory.
 *  - A context is not thread-safe: one host thread per context.
 *  - A context holds one or more independent grids ("channels"); the
 *    reference's single simulation is a context with one channel.  All
 *    channels advance together in every tick.
 */
#ifndef STROEMUNG_HIP_H
#define STROEMUNG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  INVALID_GRID corresponds to the reference's
 * SimulationGridError::BoundaryTooThinError, STALE_BOUNDARY to
 * BoundaryListIncorrectError (src/grid/mod.rs:51-59). */
#define STM_OK 0
#define STM_ERR_NO_DEVICE (-1)
#define STM_ERR_HIP (-2)
#define STM_ERR_INVALID_GRID (-3)
#define STM_ERR_STALE_BOUNDARY (-4)
#define STM_ERR_NAN (-5)
#define STM_ERR_INVALID_ARG (-6)
#define STM_ERR_STATE (-7)

/* Cell types: the reference's Cell / BoundaryCell (src/cell.rs:7-23) plus a
 * moving wall (no reference counterpart) for lid-driven flows. */
#define STM_FLUID 0
#define STM_NOSLIP 1
#define STM_INFLOW 2
#define STM_OUTFLOW 3
#define STM_MOVING_WALL 4

/* Pressure solvers. */
#define STM_SOLVER_LEX 0 /* the reference's lexicographic SOR and stop rule, bit-exact */
#define STM_SOLVER_RB 1  /* red-black SOR to a fluid-masked relative residual */

/* The scalar fields of `Simulation` (src/simulation.rs:30-44), followed by
 * extensions with no reference counterpart. */
typedef struct stm_params {
    double delx, dely;            /* cell_size */
    double delt;                  /* fixed step; the first step when tau > 0 */
    double gamma;
    double reynolds;
    double sor_absolute_epsilon;  /* LEX: absolute eps; RB: relative eps */
    uint32_t max_iterations;
    double omega;
    int32_t has_initial_norm_squared;
    double initial_norm_squared;
    uint32_t iterations;
    double time;
    /* extensions */
    double tau;                   /* > 0: adaptive step (NaSt2D COMP_DELT), else fixed */
    int32_t solver;               /* STM_SOLVER_* */
    int32_t check_every;          /* RB: sweeps between residual checks (>= 1) */
} stm_params;

/* State of one channel after a call (read back with stm_get_state). */
typedef struct stm_state {
    double delt, time;
    double initial_norm_squared;
    double norm_squared;          /* value returned by the last tick */
    double residual_squared;      /* RB: sum of r^2 over interior fluid cells */
    double tolerance_squared;     /* RB: eps^2 * sum of (rhs - mean)^2 over interior fluid */
    double rhs_mean;              /* RB: mean rhs over interior fluid cells; the solve is for
                                   * rhs - mean (the all-Neumann problem needs zero mean) */
    double pressure_range[2];     /* SimulationGrid::pressure_range */
    double speed_range[2];        /* SimulationGrid::speed_range */
    double umax, vmax;            /* max |u|, |v| over fluid-cell faces */
    uint32_t iterations;          /* ticks done */
    uint32_t sor_iterations;      /* sweeps of the last solve */
    int32_t has_initial_norm_squared;
    int32_t exhausted;            /* last solve stopped at max_iterations */
} stm_state;

typedef struct stm_ctx stm_ctx;

/* ---- context lifetime (no reference counterpart: the reference is CPU-only) */
int stm_device_count(int *count);
int stm_create(int device, stm_ctx **out);
void stm_destroy(stm_ctx *ctx);
const char *stm_last_error(const stm_ctx *ctx);

/* ---- grid setup */

/* Replaces SimulationGrid::rebuild_boundary_list (src/grid/mod.rs:202-235)
 * with calculate_edges (270-332).  Host only; needs no device.  edge_out[i]
 * is -1 for fluid cells, else the edge kind 0 (none) .. 8 in the order
 * North, NorthEast, East, SouthEast, South, SouthWest, West, NorthWest;
 * list_out receives the flat indices of all boundary cells sorted by (x, y).
 * Returns STM_ERR_INVALID_GRID (too-thin boundary, cell named in err). */
int stm_rebuild_boundary_list(uint32_t nx, uint32_t ny, const uint8_t *cell_type,
                              int8_t *edge_out, uint32_t *list_out, uint32_t *n_out,
                              double *fluid_cells_out, char *err, size_t errlen);

/* Adds a grid with the scalars of `params` (SimulationGrid::try_from,
 * src/grid/mod.rs:127-153: classifies the boundary).  cell_velocity holds
 * 2*nx*ny doubles, the (u, v) of Inflow and moving-wall cells (NULL: zero).
 * Fields start at zero.  All channels must be added before the first
 * upload / construct. */
int stm_add_grid(stm_ctx *ctx, const stm_params *params, uint32_t nx, uint32_t ny,
                 const uint8_t *cell_type, const double *cell_velocity, uint32_t *channel_out);

/* Replaces the cell types of a channel after an edit and rebuilds its
 * boundary list (the GUI edit path, src/lib.rs:38-78). */
int stm_set_cells(stm_ctx *ctx, uint32_t channel, const uint8_t *cell_type,
                  const double *cell_velocity);

/* Updates the scalars of a channel (e.g. after loading a saved state). */
int stm_set_params(stm_ctx *ctx, uint32_t channel, const stm_params *params);

/* ---- state transfer: u, v, p (SimulationGrid::u/v/pressure) and f, g, rhs
 * (Simulation::f/g/rhs, which persist across ticks).  Any pointer may be
 * NULL to skip that field. */
int stm_upload(stm_ctx *ctx, uint32_t channel, const double *u, const double *v,
               const double *p, const double *f, const double *g, const double *rhs);
int stm_download(stm_ctx *ctx, uint32_t channel, double *u, double *v, double *p, double *f,
                 double *g, double *rhs);
int stm_get_state(stm_ctx *ctx, uint32_t channel, stm_state *out);

/* ---- the time step */

/* Simulation::try_from (src/simulation.rs:71-99) for every channel: pressure
 * and speed ranges, f = g = rhs = 0, F/G, RHS and initial_norm_squared
 * unless the params carried one. */
int stm_construct(stm_ctx *ctx);

/* Simulation::run_simulation_tick (src/simulation.rs:324-333) for every
 * channel.  sor_iterations / norm_squared receive one value per channel
 * (either may be NULL). */
int stm_run_simulation_tick(stm_ctx *ctx, uint32_t *sor_iterations, double *norm_squared);

/* nticks ticks without reading results back between them. */
int stm_run_ticks(stm_ctx *ctx, uint32_t nticks);

/* ---- the stages of a tick, for callers that drive them one by one */
int stm_set_boundary_u_and_v(stm_ctx *ctx);   /* src/grid/mod.rs:414-651 */
int stm_calculate_f_and_g(stm_ctx *ctx);      /* src/simulation.rs:122-202 */
int stm_calculate_rhs(stm_ctx *ctx);          /* src/simulation.rs:204-214 */
int stm_solve_pressure(stm_ctx *ctx);         /* src/simulation.rs:239-285 */
int stm_set_u_and_v(stm_ctx *ctx);            /* src/simulation.rs:287-322 */
int stm_copy_pressure_to_boundaries(stm_ctx *ctx); /* src/grid/mod.rs:343-412 */
/* Exactly n red-black sweeps (ghost refresh, red, black) on every channel,
 * with no residual test (no reference counterpart). */
int stm_rb_sweeps(stm_ctx *ctx, uint32_t n);

/* ---- red-black solver configuration (no reference counterpart)
 * The red-black solve runs as temporally blocked launches of K sweeps each,
 * with the residual checked after every launch; the variant fixes K and the
 * output tile.  params.check_every = K (1, 2, 4, 8) restricts the choice to
 * that K; check_every <= 0 lets the tuner pick any variant. */
/* variant >= 0 selects it; -1 only queries.  Reports the variant in use. */
int stm_solver_variant(stm_ctx *ctx, int32_t variant, int32_t *current, int32_t *sweeps,
                       int32_t *tile_x, int32_t *tile_y);
/* Times `launches` launches of each permitted variant on the current fields
 * (restored afterwards) and keeps the fastest per sweep. */
int stm_autotune(stm_ctx *ctx, uint32_t launches, int32_t *chosen, double *us_per_sweep);

/* ---- measurement (no reference counterpart) */
int stm_synchronize(stm_ctx *ctx);
/* Streaming device-to-device copy of `bytes` (read + write), `reps` times,
 * on the context's stream: the measured bandwidth ceiling in GB/s. */
int stm_stream_copy(stm_ctx *ctx, uint64_t bytes, uint32_t reps, double *gbs);
/* Time every `sample`-th launch of each kernel with HIP events on the
 * context's stream (0 disables). */
int stm_profile(stm_ctx *ctx, uint32_t sample);
/* Accumulated event time and timed launch count of kernel `name`, plus the
 * total number of its launches. */
int stm_kernel_stats(stm_ctx *ctx, const char *name, double *total_ms, uint64_t *timed,
                     uint64_t *launches);
int stm_reset_stats(stm_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
