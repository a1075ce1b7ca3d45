/*
 * CPU oracle for the stroemung NaSt2D time step — TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference Rust solver
 * (wickedchicken/stroemung v0.1.2, src/simulation.rs, src/math.rs,
 * src/grid/mod.rs) plus the pieces the reference does not have
 * (adaptive time step, moving-wall cells, red-black SOR with a
 * fluid-masked relative residual).  Only the test-suite, the smoke check and
 * the CPU-baseline leg of the benchmark may load it, and only as the
 * checker; the product library never links or calls it.
 *
 * Layout: every field is an nx*ny double array in the reference's ndarray
 * C order, element (x, y) at x*ny + y ("north" = y-1, "south" = y+1;
 * src/types.rs:12-19, src/grid/mod.rs:168-170).
 *
 * Parity status of each function (how it is pinned) is written next to it
 * in oracle.c.
 */
#ifndef STROEMUNG_ORACLE_H
#define STROEMUNG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cell types (src/cell.rs:7-23), plus a moving wall that the reference lacks. */
enum { ORC_FLUID = 0, ORC_NOSLIP = 1, ORC_INFLOW = 2, ORC_OUTFLOW = 3, ORC_MOVING = 4 };

/* Edge kinds (src/grid/mod.rs:19-49).  -1 marks a fluid cell. */
enum {
    ORC_EDGE_NONE = 0, ORC_EDGE_N = 1, ORC_EDGE_NE = 2, ORC_EDGE_E = 3, ORC_EDGE_SE = 4,
    ORC_EDGE_S = 5, ORC_EDGE_SW = 6, ORC_EDGE_W = 7, ORC_EDGE_NW = 8
};

/* Solvers for the pressure equation. */
enum { ORC_SOLVER_LEX = 0, ORC_SOLVER_RB = 1 };

/* One entry of BoundaryList::u_v_restore (src/grid/mod.rs:68). */
typedef struct {
    int32_t idx;
    int32_t has_u, has_v;
    double u, v;
} orc_restore;

/*
 * State of one simulation: the fields of `Simulation` (src/simulation.rs:49-69)
 * and `SimulationGrid` (src/grid/mod.rs:112-125).  Arrays are borrowed from
 * the caller.
 */
typedef struct {
    int32_t nx, ny;
    double delx, dely, delt, gamma, reynolds;
    double sor_absolute_epsilon;
    int32_t max_iterations;
    double omega;
    int32_t has_initial_norm_squared;
    double initial_norm_squared;
    uint32_t iterations;
    double time;
    /* extensions (no reference counterpart) */
    double tau;            /* > 0: adaptive time step, book formula */
    int32_t solver;        /* ORC_SOLVER_LEX (reference) or ORC_SOLVER_RB */
    int32_t check_every;   /* RB: sweeps between residual checks */
    int32_t masked_stop;   /* LEX: stop on the RB criterion instead of the reference's */
    int32_t fix_mean;      /* RB: use rhs_mean as given instead of computing it */
    double rhs_mean;       /* RB: mean of rhs over interior fluid cells (projection) */
    /* fields, nx*ny each */
    double *u, *v, *p, *f, *g, *rhs;
    const uint8_t *cell_type;   /* ORC_* */
    const double *cell_vel;     /* 2*nx*ny: (u, v) of inflow / moving cells */
    /* derived by orc_rebuild_boundary_list */
    int8_t *edge;               /* nx*ny */
    int32_t *blist;             /* sorted boundary list, flat indices */
    int32_t nb;
    double fluid_cells;
    orc_restore *restore;       /* capacity 2*nx*ny */
    int32_t nrestore;
    double pressure_range[2], speed_range[2];
    /* diagnostics of the last solve */
    double last_res2;           /* RB: masked mean-square residual */
    double last_tol2;           /* RB: eps^2 * mean-square rhs */
} orc_sim;

/* math.rs primitives on a 3x3 view given as v[a*3+b] = e[x-1+a][y-1+b]. */
double orc_du2dx(const double *u3, double delx, double gamma);
double orc_duvdx(const double *u3, const double *v3, double delx, double gamma);
double orc_duvdy(const double *u3, const double *v3, double dely, double gamma);
double orc_dv2dy(const double *v3, double dely, double gamma);
double orc_laplacian(const double *e3, double delx, double dely);
double orc_residual(const double *p3, double delx, double dely, double rhs);
double orc_calculate_f(const double *u3, const double *v3, double delx, double dely,
                       double delt, double gamma, double reynolds);
double orc_calculate_g(const double *u3, const double *v3, double delx, double dely,
                       double delt, double gamma, double reynolds);

/* Grid setup.  Returns 0, or -(1+idx) of the first too-thin cell. */
int orc_classify(int nx, int ny, const uint8_t *cell_type, int8_t *edge,
                 int32_t *blist, int32_t *nb, double *fluid_cells);
int orc_rebuild_boundary_list(orc_sim *s);

/* Stages of the tick. */
int orc_copy_pressure_to_boundaries(orc_sim *s);
int orc_set_boundary_u_and_v(orc_sim *s);
void orc_calculate_f_and_g(orc_sim *s);
void orc_calculate_rhs(orc_sim *s);
double orc_calculate_norm_squared(const orc_sim *s);
double orc_get_initial_norm_squared(orc_sim *s);
int orc_solve_sor(orc_sim *s, uint32_t *iters, double *norm2);
void orc_set_u_and_v(orc_sim *s);
void orc_calculate_pressure_range(orc_sim *s);
void orc_calculate_speed_range(orc_sim *s);
/* src/visualization.rs:7-105: the GUI's image, RGBA u8, pixel (x, y) at
 * (y * nx + x) * 4; mode 0 pressure, 1 speed.  The f32 colour to u8 step is
 * macroquad 0.4.13's Color -> [u8; 4] (component * 255, truncated,
 * saturating; NaN -> 0), restated: macroquad is not in the reference. */
void orc_render(const orc_sim *s, int mode, uint8_t *rgba);
int orc_construct(orc_sim *s);            /* Simulation::try_from */
int orc_run_simulation_tick(orc_sim *s, uint32_t *iters, double *norm2);

/* Extensions. */
void orc_max_uv(const orc_sim *s, double *umax, double *vmax);
double orc_comp_delt(const orc_sim *s, double umax, double vmax);
void orc_ghost_refresh(orc_sim *s, int colour);       /* colour -1: all */
void orc_rb_half_sweep(orc_sim *s, int colour);
void orc_rb_sweeps(orc_sim *s, int nsweeps);
void orc_rb_sweeps_mt(orc_sim *s, int nsweeps, int threads);
double orc_masked_residual_sum(orc_sim *s, double *rhs_sum2);
double orc_rhs_mean(const orc_sim *s);
void orc_set_threads(int threads);
int orc_solve_rb(orc_sim *s, uint32_t *iters, double *norm2);

#ifdef __cplusplus
}
#endif
#endif
