/*
 * CPU oracle for the stroemung NaSt2D time step — TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, as the checker.  The product (stroemung_amd/) never links or calls it.
 *
 * Built with -O3 -ffp-contract=off: Rust never fuses a*b+c, so neither may we.
 * Every expression below keeps the reference's association order; sq(x) is
 * x*x (what f64::powi(2) lowers to).
 *
 * Parity status (how each function is pinned; the tests live in tests/):
 *   primitives du2dx/duvdx/duvdy/dv2dy/laplacian, calculate_f/g
 *       pinned: the reference's own vectors, src/math.rs:195-398 and
 *       src/simulation.rs:452-554 (tests/golden/math_vectors.json).
 *   classify / rebuild_boundary_list
 *       pinned: src/grid/mod.rs:683-801 and deserialize_boundaries-2.snap.
 *   set_boundary_u_and_v, calculate_f_and_g, calculate_rhs, solve_sor,
 *   calculate_norm_squared, set_u_and_v, run_simulation_tick, construct
 *       pinned for N/S/E/W edges, Inflow-East, Outflow-West and NoSlip:
 *       the 4x3 tick trajectory snapshots after 1/101/201 ticks with their
 *       iteration counts and norms (src/simulation.rs:571-618) and
 *       initial_norm_squared = 899.9547140394143 (deserialize-2.snap:17).
 *       Diagonal edge kinds and the other Inflow/Outflow orientations appear
 *       in no reference golden; they are pinned by wall invariants (normal
 *       face velocity equals the wall velocity; ghost plus fluid tangential
 *       velocity is 0, or twice the wall velocity of a moving wall, for the
 *       straight and the diagonal kinds; an outflow cell copies its fluid
 *       faces; ghost pressure equals the neighbour mean) and by the net
 *       effect of the restore list after ADAP_UV, checked cell by cell
 *       (tests/test_oracle_extensions.py).
 *   calculate_pressure_range / calculate_speed_range
 *       pinned: a numpy brute force of the reference's folds, including an
 *       all-negative p (the max keeps its 0.0 start) and a grid with no
 *       fluid cell (tests/test_oracle_extensions.py).
 *   Extensions with no reference counterpart (moving wall, max_uv,
 *   comp_delt, red-black SOR, masked residual, solve_rb)
 *       pinned by closed forms, a manufactured Poisson problem with a known
 *       discrete solution, and convergence of red-black and lexicographic
 *       SOR to the same fixed point (tests/test_oracle_extensions.py).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline double sq(double x) { return x * x; }

/* ---------------------------------------------------------------- math.rs */

/* src/math.rs:19-33 */
double orc_du2dx(const double *u3, double delx, double gamma)
{
    double u_i_m1 = u3[0 * 3 + 1], u_i = u3[1 * 3 + 1], u_i_p1 = u3[2 * 3 + 1];
    double inner_left1 = sq(u_i + u_i_p1);
    double inner_right1 = sq(u_i_m1 + u_i);
    double left_side = inner_left1 - inner_right1;
    double inner_left2 = fabs(u_i + u_i_p1) * (u_i - u_i_p1);
    double inner_right2 = fabs(u_i_m1 + u_i) * (u_i_m1 - u_i);
    return (left_side + (gamma * (inner_left2 - inner_right2))) / (4.0 * delx);
}

/* src/math.rs:53-77 */
double orc_duvdx(const double *u3, const double *v3, double delx, double gamma)
{
    double u_i_j = u3[1 * 3 + 1], u_i_j_p1 = u3[1 * 3 + 2];
    double u_i_m1_j = u3[0 * 3 + 1], u_i_m1_j_p1 = u3[0 * 3 + 2];
    double v_i_j = v3[1 * 3 + 1], v_i_p1_j = v3[2 * 3 + 1], v_i_m1_j = v3[0 * 3 + 1];
    double inner_left1 = (u_i_j + u_i_j_p1) * (v_i_j + v_i_p1_j);
    double inner_right1 = (u_i_m1_j + u_i_m1_j_p1) * (v_i_m1_j + v_i_j);
    double left_side = inner_left1 - inner_right1;
    double inner_left2 = fabs(u_i_j + u_i_j_p1) * (v_i_j - v_i_p1_j);
    double inner_right2 = fabs(u_i_m1_j + u_i_m1_j_p1) * (v_i_m1_j - v_i_j);
    return (left_side + (gamma * (inner_left2 - inner_right2))) / (4.0 * delx);
}

/* src/math.rs:97-120 */
double orc_duvdy(const double *u3, const double *v3, double dely, double gamma)
{
    double u_i_j = u3[1 * 3 + 1], u_i_j_m1 = u3[1 * 3 + 0], u_i_j_p1 = u3[1 * 3 + 2];
    double v_i_j = v3[1 * 3 + 1], v_i_j_m1 = v3[1 * 3 + 0];
    double v_i_p1_j = v3[2 * 3 + 1], v_i_p1_j_m1 = v3[2 * 3 + 0];
    double inner_left1 = (v_i_j + v_i_p1_j) * (u_i_j + u_i_j_p1);
    double inner_right1 = (v_i_j_m1 + v_i_p1_j_m1) * (u_i_j_m1 + u_i_j);
    double left_side = inner_left1 - inner_right1;
    double inner_left2 = fabs(v_i_j + v_i_p1_j) * (u_i_j - u_i_j_p1);
    double inner_right2 = fabs(v_i_j_m1 + v_i_p1_j_m1) * (u_i_j_m1 - u_i_j);
    return (left_side + (gamma * (inner_left2 - inner_right2))) / (4.0 * dely);
}

/* src/math.rs:136-150 */
double orc_dv2dy(const double *v3, double dely, double gamma)
{
    double v_i_j = v3[1 * 3 + 1], v_i_j_p1 = v3[1 * 3 + 2], v_i_j_m1 = v3[1 * 3 + 0];
    double inner_left1 = sq(v_i_j + v_i_j_p1);
    double inner_right1 = sq(v_i_j_m1 + v_i_j);
    double left_side = inner_left1 - inner_right1;
    double inner_left2 = fabs(v_i_j + v_i_j_p1) * (v_i_j - v_i_j_p1);
    double inner_right2 = fabs(v_i_j_m1 + v_i_j) * (v_i_j_m1 - v_i_j);
    return (left_side + (gamma * (inner_left2 - inner_right2))) / (4.0 * dely);
}

/* src/math.rs:162-174 */
double orc_laplacian(const double *e3, double delx, double dely)
{
    double e_i_j = e3[1 * 3 + 1], e_i_j_m1 = e3[1 * 3 + 0], e_i_j_p1 = e3[1 * 3 + 2];
    double e_i_m1_j = e3[0 * 3 + 1], e_i_p1_j = e3[2 * 3 + 1];
    double d2edx2 = ((e_i_p1_j - (2. * e_i_j)) + e_i_m1_j) / sq(delx);
    double d2edy2 = ((e_i_j_p1 - (2. * e_i_j)) + e_i_j_m1) / sq(dely);
    return d2edx2 + d2edy2;
}

/* src/math.rs:176-186 */
double orc_residual(const double *p3, double delx, double dely, double rhs)
{
    double p_i_p1_j = p3[2 * 3 + 1], p_i_j = p3[1 * 3 + 1], p_i_m1_j = p3[0 * 3 + 1];
    double p_i_j_p1 = p3[1 * 3 + 2], p_i_j_m1 = p3[1 * 3 + 0];
    double part1 = ((p_i_p1_j - p_i_j) - (p_i_j - p_i_m1_j)) / sq(delx);
    double part2 = ((p_i_j_p1 - p_i_j) - (p_i_j - p_i_j_m1)) / sq(dely);
    return (part1 + part2) - rhs;
}

/* src/simulation.rs:349-363 */
double orc_calculate_f(const double *u3, const double *v3, double delx, double dely,
                       double delt, double gamma, double reynolds)
{
    return u3[4] + (delt * (((orc_laplacian(u3, delx, dely) / reynolds)
                             - orc_du2dx(u3, delx, gamma))
                            - orc_duvdy(u3, v3, dely, gamma)));
}

/* src/simulation.rs:378-392 */
double orc_calculate_g(const double *u3, const double *v3, double delx, double dely,
                       double delt, double gamma, double reynolds)
{
    return v3[4] + (delt * (((orc_laplacian(v3, delx, dely) / reynolds)
                             - orc_duvdx(u3, v3, delx, gamma))
                            - orc_dv2dy(v3, dely, gamma)));
}

/* ------------------------------------------------------------ grid/mod.rs */

#define IDX(s, x, y) ((x) * (s)->ny + (y))

/* src/grid/mod.rs:167-200 (neighbours) and 270-332 (calculate_edges). */
static int edge_of(int nx, int ny, const uint8_t *t, int x, int y)
{
    int up = (y > 0) && t[x * ny + (y - 1)] == ORC_FLUID;
    int down = (y < ny - 1) && t[x * ny + (y + 1)] == ORC_FLUID;
    int right = (x < nx - 1) && t[(x + 1) * ny + y] == ORC_FLUID;
    int left = (x > 0) && t[(x - 1) * ny + y] == ORC_FLUID;
    /* match (left, right, up, down), src/grid/mod.rs:297-331 */
    if (!left && !right && !up && !down) return ORC_EDGE_NONE;
    if (left && !right && !up && !down) return ORC_EDGE_W;
    if (left && !right && up && !down) return ORC_EDGE_NW;
    if (!left && !right && up && !down) return ORC_EDGE_N;
    if (!left && right && up && !down) return ORC_EDGE_NE;
    if (!left && right && !up && !down) return ORC_EDGE_E;
    if (!left && right && !up && down) return ORC_EDGE_SE;
    if (!left && !right && !up && down) return ORC_EDGE_S;
    if (left && !right && !up && down) return ORC_EDGE_SW;
    return -2; /* BoundaryTooThinError */
}

/* src/grid/mod.rs:202-235: every non-fluid cell, sorted by (x, y). */
int orc_classify(int nx, int ny, const uint8_t *cell_type, int8_t *edge,
                 int32_t *blist, int32_t *nb, double *fluid_cells)
{
    int n = 0;
    long fluid = 0;
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++) {
            int i = x * ny + y;
            if (cell_type[i] == ORC_FLUID) {
                fluid++;
                edge[i] = -1;
            } else {
                int e = edge_of(nx, ny, cell_type, x, y);
                if (e < 0) return -(1 + i);
                edge[i] = (int8_t)e;
                blist[n++] = i;
            }
        }
    *nb = n;
    *fluid_cells = (double)fluid;
    return 0;
}

int orc_rebuild_boundary_list(orc_sim *s)
{
    int rc = orc_classify(s->nx, s->ny, s->cell_type, s->edge, s->blist, &s->nb,
                          &s->fluid_cells);
    s->nrestore = 0;
    return rc;
}

/* Fluid neighbour indices of boundary cell i for its edge kind. */
static inline int nb_n(const orc_sim *s, int i) { (void)s; return i - 1; }
static inline int nb_s(const orc_sim *s, int i) { (void)s; return i + 1; }
static inline int nb_e(const orc_sim *s, int i) { return i + s->ny; }
static inline int nb_w(const orc_sim *s, int i) { return i - s->ny; }

/* src/grid/mod.rs:343-412 */
int orc_copy_pressure_to_boundaries(orc_sim *s)
{
    double *p = s->p;
    for (int k = 0; k < s->nb; k++) {
        int i = s->blist[k];
        int e = s->edge[i];
        if (e == ORC_EDGE_NONE) continue;
        if (s->cell_type[i] == ORC_FLUID) return -1; /* BoundaryListIncorrectError */
        switch (e) {
        case ORC_EDGE_N: p[i] = p[nb_n(s, i)]; break;
        case ORC_EDGE_NE: p[i] = (p[nb_n(s, i)] + p[nb_e(s, i)]) / 2.0; break;
        case ORC_EDGE_E: p[i] = p[nb_e(s, i)]; break;
        case ORC_EDGE_SE: p[i] = (p[nb_s(s, i)] + p[nb_e(s, i)]) / 2.0; break;
        case ORC_EDGE_S: p[i] = p[nb_s(s, i)]; break;
        case ORC_EDGE_SW: p[i] = (p[nb_s(s, i)] + p[nb_w(s, i)]) / 2.0; break;
        case ORC_EDGE_W: p[i] = p[nb_w(s, i)]; break;
        case ORC_EDGE_NW: p[i] = (p[nb_n(s, i)] + p[nb_w(s, i)]) / 2.0; break;
        }
    }
    return 0;
}

static void push_restore(orc_sim *s, int idx, int has_u, double u, int has_v, double v)
{
    orc_restore *r = &s->restore[s->nrestore++];
    r->idx = idx;
    r->has_u = has_u;
    r->u = u;
    r->has_v = has_v;
    r->v = v;
}

/*
 * src/grid/mod.rs:414-651, sequential in boundary-list order.
 * NoSlip (437-488) and Inflow (537-586) share one rule with (bu, bv) = 0 or
 * the inflow velocity.  The moving wall is an extension: like NoSlip, but
 * the ghost tangential velocity reflects about the wall velocity,
 * ghost = (2.0*b) - fluid.
 */
int orc_set_boundary_u_and_v(orc_sim *s)
{
    double *u = s->u, *v = s->v;
    s->nrestore = 0;
    for (int k = 0; k < s->nb; k++) {
        int i = s->blist[k];
        int e = s->edge[i];
        if (e == ORC_EDGE_NONE) {
            push_restore(s, i, 1, u[i], 1, v[i]);
            continue;
        }
        int t = s->cell_type[i];
        int N = nb_n(s, i), S = nb_s(s, i), E = nb_e(s, i), W = nb_w(s, i);
        if (t == ORC_NOSLIP || t == ORC_INFLOW) {
            double bu = 0.0, bv = 0.0;
            if (t == ORC_INFLOW) {
                bu = s->cell_vel[2 * i];
                bv = s->cell_vel[2 * i + 1];
            }
            switch (e) {
            case ORC_EDGE_N: u[i] = -u[N]; v[N] = bv; break;
            case ORC_EDGE_NE: u[i] = bu; v[N] = bv; v[i] = -v[E]; break;
            case ORC_EDGE_E: u[i] = bu; v[i] = -v[E]; break;
            case ORC_EDGE_SE: u[i] = bu; v[i] = bv; break;
            case ORC_EDGE_S: u[i] = -u[S]; v[i] = bv; break;
            case ORC_EDGE_SW: u[W] = bu; u[i] = -u[S]; v[i] = bv; break;
            case ORC_EDGE_W: u[W] = bu; v[i] = -v[W]; break;
            case ORC_EDGE_NW: u[W] = bu; u[i] = -u[N]; v[N] = bv; v[i] = -v[W]; break;
            }
        } else if (t == ORC_MOVING) {
            double bu = s->cell_vel[2 * i], bv = s->cell_vel[2 * i + 1];
            switch (e) {
            case ORC_EDGE_N: u[i] = (2.0 * bu) - u[N]; v[N] = bv; break;
            case ORC_EDGE_NE: u[i] = bu; v[N] = bv; v[i] = (2.0 * bv) - v[E]; break;
            case ORC_EDGE_E: u[i] = bu; v[i] = (2.0 * bv) - v[E]; break;
            case ORC_EDGE_SE: u[i] = bu; v[i] = bv; break;
            case ORC_EDGE_S: u[i] = (2.0 * bu) - u[S]; v[i] = bv; break;
            case ORC_EDGE_SW: u[W] = bu; u[i] = (2.0 * bu) - u[S]; v[i] = bv; break;
            case ORC_EDGE_W: u[W] = bu; v[i] = (2.0 * bv) - v[W]; break;
            case ORC_EDGE_NW:
                u[W] = bu; u[i] = (2.0 * bu) - u[N]; v[N] = bv; v[i] = (2.0 * bv) - v[W];
                break;
            }
        } else if (t == ORC_OUTFLOW) { /* src/grid/mod.rs:489-536 */
            switch (e) {
            case ORC_EDGE_N: u[i] = u[N]; v[i] = v[N]; break;
            case ORC_EDGE_NE: u[i] = u[N]; v[i] = v[E]; break;
            case ORC_EDGE_E: u[i] = u[E]; v[i] = v[E]; break;
            case ORC_EDGE_SE: u[i] = u[E]; v[i] = v[S]; break;
            case ORC_EDGE_S: u[i] = u[S]; v[i] = v[S]; break;
            case ORC_EDGE_SW: u[i] = u[W]; v[i] = v[S]; break;
            case ORC_EDGE_W: u[i] = u[W]; v[i] = v[W]; break;
            case ORC_EDGE_NW: u[i] = u[N]; v[i] = v[W]; break;
            }
        } else {
            return -1; /* BoundaryListIncorrectError */
        }
        /* src/grid/mod.rs:595-648: the extra entries target i itself. */
        push_restore(s, i, 1, u[i], 1, v[i]);
        switch (e) {
        case ORC_EDGE_N:
        case ORC_EDGE_NE: push_restore(s, i, 0, 0.0, 1, v[N]); break;
        case ORC_EDGE_SW:
        case ORC_EDGE_W: push_restore(s, i, 1, u[W], 0, 0.0); break;
        case ORC_EDGE_NW: push_restore(s, i, 1, u[W], 1, v[N]); break;
        default: break;
        }
    }
    return 0;
}

/* src/grid/mod.rs:237-251 */
void orc_calculate_pressure_range(orc_sim *s)
{
    double mn = DBL_MAX, mx = 0.0;
    for (int i = 0; i < s->nx * s->ny; i++)
        if (s->cell_type[i] == ORC_FLUID) {
            mn = fmin(mn, s->p[i]);
            mx = fmax(mx, s->p[i]);
        }
    s->pressure_range[0] = mn;
    s->pressure_range[1] = mx;
}

/* src/grid/mod.rs:253-268 */
void orc_calculate_speed_range(orc_sim *s)
{
    double mn = DBL_MAX, mx = 0.0;
    for (int i = 0; i < s->nx * s->ny; i++)
        if (s->cell_type[i] == ORC_FLUID) {
            double sp = sq(s->u[i]) + sq(s->v[i]);
            mn = fmin(mn, sp);
            mx = fmax(mx, sp);
        }
    s->speed_range[0] = sqrt(mn);
    s->speed_range[1] = sqrt(mx);
}

/* --------------------------------------------------------- visualization.rs */

/* src/visualization.rs:7-27, in f32 */
static void hsl_to_rgb(float hue, float sat, float light, float *r, float *g, float *b)
{
    const float c = (1.0f - fabsf(2.0f * light - 1.0f)) * sat;
    const float x = c * (1.0f - fabsf(fmodf(hue / 60.0f, 2.0f) - 1.0f));
    const float m = light - c / 2.0f;
    float rr, gg, bb;
    if (hue < 60.0f) { rr = c; gg = x; bb = 0.0f; }
    else if (hue < 120.0f) { rr = x; gg = c; bb = 0.0f; }
    else if (hue < 180.0f) { rr = 0.0f; gg = c; bb = x; }
    else if (hue < 240.0f) { rr = 0.0f; gg = x; bb = c; }
    else if (hue < 300.0f) { rr = x; gg = 0.0f; bb = c; }
    else { rr = c; gg = 0.0f; bb = x; }
    *r = rr + m;
    *g = gg + m;
    *b = bb + m;
}

/* Rust `(c * 255.) as u8` on f32: truncation toward zero, saturating, NaN -> 0 */
static uint8_t to_u8(float c)
{
    const float v = c * 255.0f;
    if (!(v > 0.0f)) return 0;
    if (v >= 255.0f) return 255;
    return (uint8_t)v;
}

/* src/visualization.rs:29-105 */
void orc_render(const orc_sim *s, int mode, uint8_t *rgba)
{
    for (int x = 0; x < s->nx; x++)
        for (int y = 0; y < s->ny; y++) {
            const int i = IDX(s, x, y);
            float r, g, b;
            if (s->cell_type[i] != ORC_FLUID) {
                r = 0.5f;
                g = mode ? 0.5f : 0.0f;
                b = mode ? 0.5f : 0.0f;
            } else {
                double t;
                if (mode) {
                    const double speed = sqrt(sq(s->u[i]) + sq(s->v[i]));
                    t = 240.0 - (speed - s->speed_range[0]) * 240.0 /
                                    (s->speed_range[1] - s->speed_range[0]);
                } else {
                    t = 240.0 - (s->p[i] - s->pressure_range[0]) * 240.0 /
                                    (s->pressure_range[1] - s->pressure_range[0]);
                }
                hsl_to_rgb((float)t, 1.0f, 0.5f, &r, &g, &b);
            }
            uint8_t *o = rgba + ((size_t)y * s->nx + x) * 4;
            o[0] = to_u8(r);
            o[1] = to_u8(g);
            o[2] = to_u8(b);
            o[3] = 255;
        }
}

/* --------------------------------------------------------- simulation.rs */

static void gather3(const orc_sim *s, const double *a, int x, int y, double *o)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) o[i * 3 + j] = a[IDX(s, x - 1 + i, y - 1 + j)];
}

/* src/simulation.rs:122-202 */
void orc_calculate_f_and_g(orc_sim *s)
{
    double u3[9], v3[9];
    for (int x = 1; x < s->nx - 1; x++)
        for (int y = 1; y < s->ny - 1; y++) {
            gather3(s, s->u, x, y, u3);
            gather3(s, s->v, x, y, v3);
            s->f[IDX(s, x, y)] = orc_calculate_f(u3, v3, s->delx, s->dely, s->delt,
                                                 s->gamma, s->reynolds);
            s->g[IDX(s, x, y)] = orc_calculate_g(u3, v3, s->delx, s->dely, s->delt,
                                                 s->gamma, s->reynolds);
        }
    for (int k = 0; k < s->nb; k++) {
        int i = s->blist[k];
        s->f[i] = s->u[i];
        s->g[i] = s->v[i];
        switch (s->edge[i]) {
        case ORC_EDGE_N: s->g[nb_n(s, i)] = s->v[nb_n(s, i)]; break;
        case ORC_EDGE_NW:
            s->f[nb_w(s, i)] = s->u[nb_w(s, i)];
            s->g[nb_n(s, i)] = s->v[nb_n(s, i)];
            break;
        case ORC_EDGE_W:
        case ORC_EDGE_SW: s->f[nb_w(s, i)] = s->u[nb_w(s, i)]; break;
        case ORC_EDGE_NE: s->g[nb_n(s, i)] = s->v[nb_n(s, i)]; break;
        default: break;
        }
    }
}

/* src/simulation.rs:204-214: x in [1, nx-1], y in [1, ny-1]. */
void orc_calculate_rhs(orc_sim *s)
{
    for (int x = 1; x < s->nx; x++)
        for (int y = 1; y < s->ny; y++) {
            double fc = s->f[IDX(s, x, y)], fw = s->f[IDX(s, x - 1, y)];
            double gc = s->g[IDX(s, x, y)], gn = s->g[IDX(s, x, y - 1)];
            s->rhs[IDX(s, x, y)] = (((fc - fw) / s->delx) + ((gc - gn) / s->dely)) / s->delt;
        }
}

/* src/simulation.rs:216-227: unmasked, x-major sequential fold from 0.0. */
double orc_calculate_norm_squared(const orc_sim *s)
{
    double acc = 0.0, p3[9];
    for (int x = 1; x < s->nx - 1; x++)
        for (int y = 1; y < s->ny - 1; y++) {
            gather3(s, s->p, x, y, p3);
            double r = orc_residual(p3, s->delx, s->dely, s->rhs[IDX(s, x, y)]);
            acc = acc + r * r;
        }
    return acc / s->fluid_cells;
}

/* src/simulation.rs:229-237 */
double orc_get_initial_norm_squared(orc_sim *s)
{
    if (s->has_initial_norm_squared) return s->initial_norm_squared;
    double n = orc_calculate_norm_squared(s);
    s->initial_norm_squared = n;
    s->has_initial_norm_squared = 1;
    return n;
}

/* src/simulation.rs:239-285 */
int orc_solve_sor(orc_sim *s, uint32_t *iters, double *norm2)
{
    double delx2 = sq(s->delx), dely2 = sq(s->dely);
    double one_minus_w = 1.0 - s->omega;
    double middle = s->omega / ((2.0 / delx2) + (2.0 / dely2));
    double epsilon_squared = sq(s->sor_absolute_epsilon);
    double norm_squared = 0.0;
    double *p = s->p;
    for (int it = 0; it < s->max_iterations; it++) {
        if (orc_copy_pressure_to_boundaries(s)) return -1;
        for (int x = 1; x < s->nx - 1; x++)
            for (int y = 1; y < s->ny - 1; y++) {
                int i = IDX(s, x, y);
                if (s->cell_type[i] != ORC_FLUID) continue;
                double p_i_j = p[i], p_i_m1_j = p[i - s->ny], p_i_p1_j = p[i + s->ny];
                double p_i_j_m1 = p[i - 1], p_i_j_p1 = p[i + 1];
                p[i] = (one_minus_w * p_i_j)
                     + middle * ((((p_i_p1_j + p_i_m1_j) / delx2)
                                  + ((p_i_j_p1 + p_i_j_m1) / dely2)) - s->rhs[i]);
            }
        if (s->masked_stop) {
            /* extension: the red-black solver's fluid-masked relative test,
             * on the unprojected rhs this solver iterates on */
            double rsum2 = 0.0;
            s->rhs_mean = 0.0;
            double res2 = orc_masked_residual_sum(s, &rsum2);
            norm_squared = res2 / s->fluid_cells;
            s->last_res2 = norm_squared;
            s->last_tol2 = sq(s->sor_absolute_epsilon) * rsum2 / s->fluid_cells;
            if (res2 <= sq(s->sor_absolute_epsilon) * rsum2) {
                *iters = (uint32_t)(it + 1);
                *norm2 = norm_squared;
                return 0;
            }
            continue;
        }
        double init = orc_get_initial_norm_squared(s);
        norm_squared = orc_calculate_norm_squared(s);
        if ((norm_squared < init) || (norm_squared < epsilon_squared)) {
            *iters = (uint32_t)(it + 1);
            *norm2 = norm_squared;
            return 0;
        }
    }
    orc_calculate_pressure_range(s);
    *iters = (uint32_t)s->max_iterations;
    *norm2 = norm_squared;
    return 0;
}

/* src/simulation.rs:287-322: x in [0, nx-2], y in [0, ny-2], then restore. */
void orc_set_u_and_v(orc_sim *s)
{
    for (int x = 0; x < s->nx - 1; x++)
        for (int y = 0; y < s->ny - 1; y++) {
            int i = IDX(s, x, y);
            double p_i_j = s->p[i], p_i_p1_j = s->p[i + s->ny], p_i_j_p1 = s->p[i + 1];
            s->u[i] = s->f[i] - ((s->delt / s->delx) * (p_i_p1_j - p_i_j));
            s->v[i] = s->g[i] - ((s->delt / s->dely) * (p_i_j_p1 - p_i_j));
        }
    for (int k = 0; k < s->nrestore; k++) {
        const orc_restore *r = &s->restore[k];
        if (r->has_u) s->u[r->idx] = r->u;
        if (r->has_v) s->v[r->idx] = r->v;
    }
    orc_calculate_speed_range(s);
}

/* src/simulation.rs:71-99 (TryFrom<UnfinalizedSimulation>) */
int orc_construct(orc_sim *s)
{
    int n = s->nx * s->ny;
    int rc = orc_rebuild_boundary_list(s);
    if (rc) return rc;
    orc_calculate_pressure_range(s);
    orc_calculate_speed_range(s);
    memset(s->f, 0, sizeof(double) * n);
    memset(s->g, 0, sizeof(double) * n);
    memset(s->rhs, 0, sizeof(double) * n);
    orc_calculate_f_and_g(s);
    orc_calculate_rhs(s);
    orc_get_initial_norm_squared(s);
    return 0;
}

/* ---------------------------------------------------------- extensions */

static inline int is_bnd(const orc_sim *s, int i) { return s->cell_type[i] != ORC_FLUID; }

/*
 * max|u|, max|v| over the faces of fluid cells: u and v of every fluid cell
 * (its east and south faces) and, where the west / north neighbour is a
 * boundary cell, that cell's u / v (the fluid cell's west / north face).
 */
void orc_max_uv(const orc_sim *s, double *umax, double *vmax)
{
    double um = 0.0, vm = 0.0;
    for (int x = 0; x < s->nx; x++)
        for (int y = 0; y < s->ny; y++) {
            int i = IDX(s, x, y);
            if (is_bnd(s, i)) continue;
            double a = fabs(s->u[i]), b = fabs(s->v[i]);
            if (a > um) um = a;
            if (b > vm) vm = b;
            if (x > 0 && is_bnd(s, i - s->ny)) {
                double c = fabs(s->u[i - s->ny]);
                if (c > um) um = c;
            }
            if (y > 0 && is_bnd(s, i - 1)) {
                double c = fabs(s->v[i - 1]);
                if (c > vm) vm = c;
            }
        }
    *umax = um;
    *vmax = vm;
}

/*
 * Adaptive time step, the stability bound of Griebel, Dornseifer and
 * Neunhoeffer (NaSt2D COMP_DELT):
 *   delt = tau * min(Re/2 / (1/dx^2 + 1/dy^2), dx/|u|max, dy/|v|max)
 * with the velocity terms dropped while the maximum is zero.
 */
double orc_comp_delt(const orc_sim *s, double umax, double vmax)
{
    if (!(s->tau > 0.0)) return s->delt;
    double dt = (s->reynolds / 2.0) / ((1.0 / sq(s->delx)) + (1.0 / sq(s->dely)));
    if (umax > 0.0) {
        double t = s->delx / umax;
        if (t < dt) dt = t;
    }
    if (vmax > 0.0) {
        double t = s->dely / vmax;
        if (t < dt) dt = t;
    }
    return s->tau * dt;
}

/* Ghost pressure of boundary cells whose colour (x+y)&1 matches (or all). */
void orc_ghost_refresh(orc_sim *s, int colour)
{
    double *p = s->p;
    for (int k = 0; k < s->nb; k++) {
        int i = s->blist[k];
        int x = i / s->ny, y = i % s->ny;
        if (colour >= 0 && ((x + y) & 1) != colour) continue;
        switch (s->edge[i]) {
        case ORC_EDGE_N: p[i] = p[nb_n(s, i)]; break;
        case ORC_EDGE_NE: p[i] = (p[nb_n(s, i)] + p[nb_e(s, i)]) * 0.5; break;
        case ORC_EDGE_E: p[i] = p[nb_e(s, i)]; break;
        case ORC_EDGE_SE: p[i] = (p[nb_s(s, i)] + p[nb_e(s, i)]) * 0.5; break;
        case ORC_EDGE_S: p[i] = p[nb_s(s, i)]; break;
        case ORC_EDGE_SW: p[i] = (p[nb_s(s, i)] + p[nb_w(s, i)]) * 0.5; break;
        case ORC_EDGE_W: p[i] = p[nb_w(s, i)]; break;
        case ORC_EDGE_NW: p[i] = (p[nb_n(s, i)] + p[nb_w(s, i)]) * 0.5; break;
        default: break;
        }
    }
}

/* Red-black constants; shared bit-for-bit with the GPU solver. */
typedef struct { double omw, mid, rdx2, rdy2; } rb_consts;

static rb_consts rb_make(const orc_sim *s)
{
    rb_consts c;
    double delx2 = sq(s->delx), dely2 = sq(s->dely);
    c.omw = 1.0 - s->omega;
    c.mid = s->omega / ((2.0 / delx2) + (2.0 / dely2));
    c.rdx2 = 1.0 / delx2;
    c.rdy2 = 1.0 / dely2;
    return c;
}

/*
 * One red-black half-sweep over fluid cells of colour (x+y)&1 == colour:
 *   p = fma(omw, p, mid * (fma(pE + pW, 1/dx^2, (pS + pN) * (1/dy^2)) - rhs))
 */
void orc_rb_half_sweep(orc_sim *s, int colour)
{
    rb_consts c = rb_make(s);
    double *p = s->p;
    int ny = s->ny;
    for (int x = 1; x < s->nx - 1; x++)
        for (int y = 1 + ((x + 1 + colour) & 1); y < ny - 1; y += 2) {
            int i = x * ny + y;
            if (s->cell_type[i] != ORC_FLUID) continue;
            double t = fma(p[i + ny] + p[i - ny], c.rdx2, (p[i + 1] + p[i - 1]) * c.rdy2);
            p[i] = fma(c.omw, p[i], c.mid * (t - (s->rhs[i] - s->rhs_mean)));
        }
}

/* Mean of rhs over interior fluid cells, summed x-major from 0.0.  The
 * pressure problem has Neumann conditions everywhere (the ghost copy treats
 * every boundary as a wall), so it is solvable only for rhs of zero mean;
 * with inflow and outflow the mean is not zero and no iteration can drive
 * the residual below it.  The red-black solver therefore solves for
 * rhs - mean. */
double orc_rhs_mean(const orc_sim *s)
{
    double acc = 0.0;
    long n = 0;
    for (int x = 1; x < s->nx - 1; x++)
        for (int y = 1; y < s->ny - 1; y++) {
            int i = x * s->ny + y;
            if (s->cell_type[i] != ORC_FLUID) continue;
            acc = acc + s->rhs[i];
            n++;
        }
    return n ? acc / (double)n : 0.0;
}

/*
 * orc_rb_sweeps with each half-sweep's rows of x split over `threads` OpenMP
 * threads.  A cell of one colour reads only cells of the other colour, so the
 * result is bit-identical to orc_rb_sweeps.  Not the reference algorithm (that
 * is the lexicographic orc_solve_sor): a multi-core CPU comparison for the
 * red-black solver.
 */
void orc_rb_sweeps_mt(orc_sim *s, int nsweeps, int threads)
{
    const rb_consts c = rb_make(s);
    double *p = s->p;
    const int nx = s->nx, ny = s->ny;
    const double mean = s->rhs_mean;
    for (int k = 0; k < nsweeps; k++) {
        orc_ghost_refresh(s, -1);
        for (int colour = 0; colour < 2; colour++) {
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int x = 1; x < nx - 1; x++)
                for (int y = 1 + ((x + 1 + colour) & 1); y < ny - 1; y += 2) {
                    const int i = x * ny + y;
                    if (s->cell_type[i] != ORC_FLUID) continue;
                    const double t = fma(p[i + ny] + p[i - ny], c.rdx2,
                                         (p[i + 1] + p[i - 1]) * c.rdy2);
                    p[i] = fma(c.omw, p[i], c.mid * (t - (s->rhs[i] - mean)));
                }
        }
    }
    if (nsweeps > 0) orc_ghost_refresh(s, -1);
}

/* One sweep = ghost refresh of every boundary cell, red half, black half;
 * after the last sweep the ghosts are refreshed once more, so they hold the
 * final values (as the red-black solve leaves them: its residual test
 * refreshes them). */
void orc_rb_sweeps(orc_sim *s, int nsweeps)
{
    for (int k = 0; k < nsweeps; k++) {
        orc_ghost_refresh(s, -1);
        orc_rb_half_sweep(s, 0);
        orc_rb_half_sweep(s, 1);
    }
    if (nsweeps > 0) orc_ghost_refresh(s, -1);
}

/*
 * Fluid-masked residual after a full ghost refresh:
 *   r = fma((pE-p)-(p-pW), 1/dx^2, ((pS-p)-(p-pN)) * (1/dy^2)) - rhs
 * Returns sum r^2 over interior fluid cells; *rhs_sum2 = sum rhs^2 there.
 */
double orc_masked_residual_sum(orc_sim *s, double *rhs_sum2)
{
    rb_consts c = rb_make(s);
    double *p = s->p;
    int ny = s->ny;
    double acc = 0.0, racc = 0.0;
    orc_ghost_refresh(s, -1);
    for (int x = 1; x < s->nx - 1; x++)
        for (int y = 1; y < ny - 1; y++) {
            int i = x * ny + y;
            if (s->cell_type[i] != ORC_FLUID) continue;
            double a = (p[i + ny] - p[i]) - (p[i] - p[i - ny]);
            double b = (p[i + 1] - p[i]) - (p[i] - p[i - 1]);
            double rh = s->rhs[i] - s->rhs_mean;
            double r = fma(a, c.rdx2, b * c.rdy2) - rh;
            acc += r * r;
            racc += rh * rh;
        }
    if (rhs_sum2) *rhs_sum2 = racc;
    return acc;
}

/*
 * Red-black SOR to the fluid-masked relative criterion
 *   sum r^2 <= eps^2 * sum rhs^2   (both over interior fluid cells),
 * checked every check_every sweeps and capped at max_iterations sweeps.
 * *norm2 = sum r^2 / fluid_cells.
 */
static int g_rb_threads = 1;

/* Threads for the red-black sweeps of orc_solve_rb (bit-identical to one:
 * orc_rb_sweeps_mt); the residual sums stay sequential. */
void orc_set_threads(int threads) { g_rb_threads = threads > 1 ? threads : 1; }

int orc_solve_rb(orc_sim *s, uint32_t *iters, double *norm2)
{
    int every = s->check_every > 0 ? s->check_every : 1;
    double rsum2 = 0.0;
    if (!s->fix_mean) s->rhs_mean = orc_rhs_mean(s);
    double res2 = orc_masked_residual_sum(s, &rsum2);
    double tol2 = sq(s->sor_absolute_epsilon) * rsum2;
    int it = 0;
    while (it < s->max_iterations) {
        int n = s->max_iterations - it < every ? s->max_iterations - it : every;
        if (g_rb_threads > 1) orc_rb_sweeps_mt(s, n, g_rb_threads);
        else orc_rb_sweeps(s, n);
        it += n;
        res2 = orc_masked_residual_sum(s, NULL);
        if (res2 <= tol2) break;
    }
    if (it >= s->max_iterations && !(res2 <= tol2)) orc_calculate_pressure_range(s);
    s->last_res2 = res2 / s->fluid_cells;
    s->last_tol2 = tol2 / s->fluid_cells;
    *iters = (uint32_t)it;
    *norm2 = res2 / s->fluid_cells;
    return 0;
}

/* src/simulation.rs:324-333, with the optional adaptive step in front. */
int orc_run_simulation_tick(orc_sim *s, uint32_t *iters, double *norm2)
{
    if (s->tau > 0.0) {
        double um, vm;
        orc_max_uv(s, &um, &vm);
        s->delt = orc_comp_delt(s, um, vm);
    }
    if (orc_set_boundary_u_and_v(s)) return -1;
    orc_calculate_f_and_g(s);
    orc_calculate_rhs(s);
    int rc = s->solver == ORC_SOLVER_RB ? orc_solve_rb(s, iters, norm2)
                                        : orc_solve_sor(s, iters, norm2);
    if (rc) return rc;
    orc_set_u_and_v(s);
    s->time += s->delt;
    s->iterations += 1;
    return 0;
}
