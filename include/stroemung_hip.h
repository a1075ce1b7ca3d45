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
 * boundary list: the GUI edit path draw_cells (src/lib.rs:38-78).  As there,
 * u, v and p of every cell whose type changes are set to 0 (on the device),
 * and a too-thin boundary (STM_ERR_INVALID_GRID, BoundaryTooThinError) leaves
 * the grid and the fields as they were. */
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
 * (either may be NULL).  Returns STM_ERR_NAN, after the tick and with the
 * outputs filled, when a channel's residual norm is not a number; the
 * red-black solve stops at the first such residual, where the reference's
 * loop runs on to itermax and returns Ok with the NaN norm. */
int stm_run_simulation_tick(stm_ctx *ctx, uint32_t *sor_iterations, double *norm_squared);

/* nticks ticks without reading results back between them; STM_ERR_NAN as
 * above for the last tick. */
int stm_run_ticks(stm_ctx *ctx, uint32_t nticks);

/* ---- the stages of a tick, for callers that drive them one by one */
int stm_set_boundary_u_and_v(stm_ctx *ctx);   /* src/grid/mod.rs:414-651 */
int stm_calculate_f_and_g(stm_ctx *ctx);      /* src/simulation.rs:122-202 */
int stm_calculate_rhs(stm_ctx *ctx);          /* src/simulation.rs:204-214 */
int stm_solve_pressure(stm_ctx *ctx);         /* src/simulation.rs:239-285 */
int stm_set_u_and_v(stm_ctx *ctx);            /* src/simulation.rs:287-322 */
int stm_copy_pressure_to_boundaries(stm_ctx *ctx); /* src/grid/mod.rs:343-412 */
/* Exactly n red-black sweeps (ghost refresh, red, black) on every channel,
 * with no residual test, then a last ghost refresh so the ghosts hold the
 * final values as after a solve (no reference counterpart). */
int stm_rb_sweeps(stm_ctx *ctx, uint32_t n);

/* ---- red-black solver configuration (no reference counterpart)
 * The red-black solve runs as temporally blocked launches of K sweeps each,
 * with the residual checked after every launch; the variant fixes K and the
 * output tile.  params.check_every = K (1, 2, 4, 6, 8) restricts the choice to
 * that K; check_every <= 0 lets the tuner pick any variant. */
/* variant >= 0 selects it; -1 only queries.  Reports the variant in use. */
int stm_solver_variant(stm_ctx *ctx, int32_t variant, int32_t *current, int32_t *sweeps,
                       int32_t *tile_x, int32_t *tile_y);
/* Times `launches` launches of each permitted variant on the current fields
 * (restored afterwards) and keeps the fastest per sweep. */
int stm_autotune(stm_ctx *ctx, uint32_t launches, int32_t *chosen, double *us_per_sweep);
/* Runs one real tick per permitted variant (the simulation advances by that
 * many ticks) and keeps the variant with the most interior cell-sweeps per
 * second of the whole tick -- for batches of uneven grids, whose solves spend
 * most of their time with few channels still running. */
int stm_autotune_ticks(stm_ctx *ctx, int32_t *chosen, double *cell_sweeps_per_s);

/* ---- measurement (no reference counterpart) */
int stm_synchronize(stm_ctx *ctx);
/* Bytes one blocked launch of the current variant moves over every channel:
 * *unique = each cell once (p read, rhs read, flags read, p written: 25 B per
 * cell, ring included), the least any launch can move; *loaded = what its
 * tiles read and write, their regions overlapping by the halo (p, rhs and
 * flags for every region cell inside the channel, p for every output cell).
 * loaded / unique is the halo factor. */
int stm_solver_bytes(stm_ctx *ctx, double *unique_bytes, double *loaded_bytes);
/* Streaming device-to-device copy of `bytes` (read + write), `reps` times,
 * on the context's stream: the measured bandwidth ceiling in GB/s. */
int stm_stream_copy(stm_ctx *ctx, uint64_t bytes, uint32_t reps, double *gbs);
/* What keeping the blocked solver resident between its K-sweep steps would
 * cost, against relaunching it: `steps` (10 .. 100000) steps of 256
 * workgroups of 768 threads (the blocked kernel's shape on the 1026 x 1026
 * grid), each writing a 64 x 64 FP64 tile, on the context's stream.
 * Microseconds per step into us[0..7]: us[0] separate launches; us[1] the
 * same launches replayed from a hipGraph; us[2] one resident kernel with a
 * grid-wide barrier per step; us[3] one resident kernel whose workgroups
 * wait for their four neighbours' flags (release / acquire at agent scope);
 * us[4], us[5] the two waits without the tile write; us[6], us[7] separate
 * and graph-replayed launches of a kernel that holds every workgroup 10 us.
 * STM_ERR_STATE when the device cannot hold the 256 workgroups at once or a
 * neighbour wait reached its poll limit (every wait is bounded). */
int stm_sync_benchmark(stm_ctx *ctx, uint32_t steps, double *us);
/* Time every `sample`-th launch of each kernel with HIP events on the
 * context's stream (0 disables). */
int stm_profile(stm_ctx *ctx, uint32_t sample);
/* Accumulated event time and timed launch count of kernel `name`, plus the
 * total number of its launches. */
int stm_kernel_stats(stm_ctx *ctx, const char *name, double *total_ms, uint64_t *timed,
                     uint64_t *launches);
int stm_reset_stats(stm_ctx *ctx);

/* ---- the pressure network (no reference counterpart)
 * A 4-level U-Net (32 base channels, 3x3 convolutions, ReLU, 2x2 max pooling,
 * nearest upsampling with skip concatenation, 1x1 head) in BF16 with FP32
 * accumulation on the matrix cores and FP32 master weights trained with Adam.
 * It maps the normalised residual of the pressure equation (channel 0) and the
 * fluid fraction (channel 1) of an H x W grid to a correction of p (see
 * stm_set_learned for how a channel's grid maps onto the frame).  H and W
 * are multiples of 8.  The network lives on the context's device and
 * stream. */
typedef struct stm_net stm_net;

typedef struct stm_net_stats_t {
    double forward_ms;            /* event time of the forward passes inside solves */
    double train_ms;              /* event time of stm_net_train */
    double flops_forward_per_sample;
    uint64_t forward_calls, forward_samples;
    uint64_t train_steps, train_samples;
    uint64_t params;
    uint32_t dataset_size;
    uint32_t pad;
} stm_net_stats_t;

/* batch = capacity: the largest batch of any call (and the channel count of a
 * context that uses the net).  seed drives the He initialisation. */
int stm_net_create(stm_ctx *ctx, uint32_t height, uint32_t width, uint32_t batch, uint64_t seed,
                   stm_net **out);
void stm_net_destroy(stm_net *net);
const char *stm_net_last_error(const stm_net *net);
int stm_net_param_count(stm_net *net, uint64_t *count);
/* FP32 master parameters: per convolution W[cout][9][cin] (tap 3*(dh+1) +
 * (dw+1) weighs the input pixel at offset dh along x and dw along y) then
 * b[cout], padded to a multiple of 8; then the head's 32 weights and its
 * bias (padded to 8). */
int stm_net_get_params(stm_net *net, float *out);
int stm_net_set_params(stm_net *net, const float *in);
/* x: nb x H x W x 2 (fp32) -> y: nb x H x W (fp32) */
int stm_net_forward(stm_net *net, uint32_t nb, const float *x, float *y);
/* The masked relative squared error of the forward pass against t
 * (nb x H x W): loss = mean_s sum m (y - t)^2 / sum m t^2 with m = x[..., 1] > 0,
 * and its gradient with respect to every parameter (layout as above). */
int stm_net_gradients(stm_net *net, uint32_t nb, const float *x, const float *t, float *loss,
                      float *grads);
/* Device-resident training set of `capacity` samples, filled by capture. */
int stm_net_dataset_reserve(stm_net *net, uint32_t capacity);
int stm_net_dataset_size(stm_net *net, uint32_t *size);
/* Sample `index` of the training set: x (H x W x 2) and t (H x W), fp32. */
int stm_net_dataset_get(stm_net *net, uint32_t index, float *x, float *t);
/* `steps` Adam steps on random batches of the training set; *last_loss is the
 * mean loss of the last 8 steps. */
int stm_net_train(stm_net *net, uint32_t steps, uint32_t batch, float lr, uint64_t seed,
                  float *last_loss);
int stm_net_stats(stm_net *net, stm_net_stats_t *out);
/* Forward passes (and training steps when train_steps > 0) on the buffers
 * already resident, for timing only: ms per forward of nb samples and per
 * training step of nb samples. */
int stm_net_benchmark(stm_net *net, uint32_t nb, uint32_t reps, double *fwd_ms,
                      double *train_ms);

/* ---- learned pressure solve (no reference counterpart)
 * With a net set, every red-black solve of the context runs `cycles`
 * two-grid cycles first and then the plain red-black solve to the same eps.
 * A cycle restricts the residual r to a grid coarser by `coarsen` (1, 2, 4,
 * 8 or 16: pixel = mean of r over its coarsen^2 interior cells), normalises it,
 * runs the network there, prolongs its output y bilinearly to a correction z,
 * sets p += a z and runs `sweeps` red-black sweeps.  The network is trained
 * to output the restricted error times rms(R r) / h_c^2 (h_c^2 = coarsen^2
 * dx dy); step STM_STEP_NETWORK takes that magnitude (a = h_c^2 / rms(R r)),
 * STM_STEP_ENERGY the a that minimises the energy norm of the new error
 * (a = <r, z> / <z, -Lap z>), STM_STEP_MOMENTUM sets p += a z + b s with
 * (a, b) minimising that norm over span{z, s}, s being the change of p over
 * the previous cycle (a flexible-CG-like acceleration across cycles; the
 * first cycle of a solve takes the energy step); STM_STEP_MOMENTUM2 also keeps
 * the change of the cycle before (span{z, s1, s2}).  The stopping test and the
 * FP64 arithmetic of
 * the sweeps are those of the plain solver, so the result meets the same
 * residual.  net = NULL or cycles = 0 turns it off.  The net must come from
 * this context, hold one batch entry per channel and an H x W frame with
 * coarsen * H >= nx - 2 and coarsen * W >= ny - 2 (interior cell (x, y)
 * falls in pixel ((x - 1) / coarsen, (y - 1) / coarsen)). */
#define STM_STEP_ENERGY 0
#define STM_STEP_NETWORK 1
#define STM_STEP_MOMENTUM 2
#define STM_STEP_MOMENTUM2 3
int stm_set_learned(stm_ctx *ctx, stm_net *net, uint32_t cycles, uint32_t sweeps, int32_t step,
                    uint32_t coarsen);
/* ---- the GUI image (src/visualization.rs:29-105, render_simulation)
 * Colours every cell of a channel on the device from its pressure or speed
 * and the channel's pressure_range / speed_range, exactly as the reference's
 * render_simulation fills its macroquad Image (fluid: hue 240 at the low end
 * of the range to 0 at the high end, saturation 1, lightness 0.5; boundary
 * cells fixed grey or dark red), and copies back nx*ny RGBA pixels, pixel
 * (x, y) at (y * nx + x) * 4 -- 4 bytes per cell instead of the fields. */
#define STM_RENDER_PRESSURE 0
#define STM_RENDER_SPEED 1
int stm_render(stm_ctx *ctx, uint32_t channel, int32_t mode, uint8_t *rgba);
/* stm_render without allocating anything (no counterpart in the reference:
 * for callers that cannot allocate device memory -- a frame loop under a
 * harness that forbids it, code that must not call hipMalloc).  The image is
 * staged in a static device buffer of 4 * 2^22 pixels (16 MiB; a channel of
 * more cells is refused with STM_ERR_INVALID_ARG) that each device holds
 * until the process ends; the context must already hold its fields (after
 * stm_upload or stm_construct, else STM_ERR_STATE).  Same image as
 * stm_render, bit for bit.  Calls that overlap on one device share the
 * buffer and race. */
int stm_render_static(stm_ctx *ctx, uint32_t channel, int32_t mode, uint8_t *rgba);

/* on != 0: learned cycles after the first are captured once into a hipGraph
 * (per net, coarsening, sweeps, step rule and solver variant) and replayed;
 * results are bit-identical to launching them one by one.  Not used while
 * capturing training pairs or profiling (stm_profile > 0). */
int stm_set_graphs(stm_ctx *ctx, int32_t on);

/* With a net set, every red-black solve also stores `points` (1..8)
 * training pairs per channel in the net's training set (ring order): the
 * residual, restricted as for stm_set_learned, at `points` moments of the
 * solve, each with the restricted mean-free change of p from that moment to
 * the converged solution, in the units of the network's output.  A plain
 * solve is sampled at 0, s, 3 s, 7 s, ... sweeps (s = `sweeps`); a learned
 * solve (same net and coarsening) at the start of its first `points`
 * cycles, before the network's correction.  net = NULL turns it off. */
int stm_set_capture(stm_ctx *ctx, stm_net *net, uint32_t points, uint32_t sweeps,
                    uint32_t coarsen);

#ifdef __cplusplus
}
#endif
#endif
