// Host runtime and C ABI of libstroemung_hip.so.
//
// The context owns the device: one HIP stream, packed per-field allocations
// for all channels, the boundary lists and the per-channel scalar state.  A
// tick is a fixed sequence of kernel launches on that stream; the only host
// synchronisation inside a tick is the red-black solve polling its
// per-channel "still running" flags.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/stroemung_hip.h"
#include "kernels.h"
#include "nn_net.h"
#include "stm_internal.h"
#include "tiling.h"

using namespace stm;

namespace {

thread_local std::string g_last_error;

struct HostChan {
    stm_params prm;
    uint32_t nx = 0, ny = 0, pitch = 0;
    int64_t off = 0;
    std::vector<uint8_t> type;
    std::vector<double> vel;       // 2 per cell
    std::vector<int8_t> edge;
    std::vector<uint32_t> blist;
    double fluid_cells = 0.0;
    double n_inner = 0.0;          // interior fluid cells
    int32_t tile0 = 0, ntiles = 0, boff = 0;
    int32_t tb0 = 0, ntb = 0;      // blocked-solver tiles of the current variant
};

struct KStat {
    uint64_t launches = 0, timed = 0;
    double ms = 0.0;
};

struct PendingTiming {
    KStat* stat;
    hipEvent_t a, b;
    uint32_t count;   // launches between a and b
};

}  // namespace

struct stm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<HostChan> chans;
    bool finalized = false;
    int64_t total = 0;             // elements per packed field
    int ntiles = 0, nbtot = 0;
    bool any_rb = false, any_lex = false;
    double *u = nullptr, *v = nullptr, *p = nullptr, *f = nullptr, *g = nullptr, *rhs = nullptr;
    double* scratch = nullptr;
    double* p2 = nullptr;          // second pressure buffer of the blocked solver
    uint8_t* flags = nullptr;
    ChanDesc* d_cd = nullptr;
    ChanState* d_cs = nullptr;
    BEntry* d_be = nullptr;
    double2* d_restore = nullptr;
    Tile* d_tiles = nullptr;
    double* d_part = nullptr;
    int32_t* d_active = nullptr;
    int32_t* h_active = nullptr;   // pinned, two slots of nch flags (the blocked solve's polls)
    hipEvent_t flag_evs[2] = {};   // one per slot: its copy has landed
    hipEvent_t tbc_ev = nullptr;   // the last copy out of h_tbc has completed
    bool tbc_pending = false;      // tbc_ev recorded and not yet waited for
    double* d_tbpart = nullptr;    // blocked solver: residual partials, two launches
    TbTile* d_tb = nullptr;
    int ntb = 0, tb_capacity = 0, part_capacity = 0;
    uint32_t* d_codes = nullptr;   // blocked solver: per-thread cell codes of every tile
    size_t codes_capacity = 0;     // bytes
    std::vector<TbTile> h_tb;      // host copy of the full list
    // the list launches use: d_tb, or during a polled solve the tiles of the
    // channels still running (d_tbc, staged through pinned h_tbc)
    const TbTile* tb_cur = nullptr;
    int ntb_cur = 0;
    TbTile* d_tbc = nullptr;
    TbTile* h_tbc = nullptr;
    int tbc_capacity = 0;
    int variant = -1;              // blocked-solver variant in use
    ChanState* h_cs = nullptr;     // pinned
    int be_capacity = 0;
    // learned correction and capture of training pairs
    stm_net* lnet = nullptr;       // learned cycles before the plain solve
    int lcycles = 0, lsweeps = 0, lstep = 0, lcoarsen = 1;
    stm_net* cnet = nullptr;       // capture into this net's training set
    int csweeps = 0, ccoarsen = 1, cpoints = 2;
    LearnChan* d_lc = nullptr;
    LearnChan* h_lc = nullptr;     // pinned
    double *lres = nullptr, *lcf = nullptr;
    double* lprev = nullptr;       // two-direction step: p at the previous cycle
    double* lprev2 = nullptr;      // three-direction step: p the cycle before
    double* lsnap = nullptr;       // capture: LEARN_POINTS snapshots of p
    float *lrc = nullptr, *lmc = nullptr;   // coarse residual and fluid fraction, nch frames
    double2* d_cpart = nullptr;
    int32_t *d_map = nullptr, *d_inv = nullptr, *d_live = nullptr;   // running channels
    uint32_t* d_rgba = nullptr;    // rendered image
    int64_t rgba_capacity = 0;
    int64_t lframe = 0;            // pixels per frame the coarse buffers hold
    std::vector<hipEvent_t> nn_evs;  // per-cycle forward timing, read after the solve
    // running-channel count of recent learned cycles, read back without draining the queue
    int32_t* h_live_ring = nullptr;  // pinned, POLL_RING entries
    hipEvent_t poll_evs[4] = {};
    // learned cycles replayed from a captured hipGraph.  A captured launch
    // holds its pointers and counts by value, so every change of a buffer or
    // count a cycle uses bumps graph_gen (and drops the graph), and the key
    // also names the net by its serial number, not its address
    bool use_graphs = false;
    hipGraphExec_t cycle_exec = nullptr;
    std::vector<int64_t> cycle_key;  // what the captured cycle depends on
    uint64_t graph_gen = 0;
    std::vector<stm_net*> nets;      // nets created on this context, detached on destroy
    // profiling
    uint32_t sample = 0;
    bool span_open = false;   // a batch is timed as a whole: no per-launch events
    std::unordered_map<std::string, KStat> stats;
    std::vector<PendingTiming> pending;
    std::vector<hipEvent_t> free_events;
};

namespace {

int fail(stm_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    g_last_error = buf;
    return code;
}

#define HIPCHK(ctx, call)                                                              \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail((ctx), STM_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                           \
    } while (0)

int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// A buffer or count that captured learned cycles use is about to change:
// drop the graph (after the launches queued from it) and move the key on.
void invalidate_graphs(stm_ctx* ctx) {
    if (ctx->cycle_exec) {
        hipStreamSynchronize(ctx->stream);
        hipGraphExecDestroy(ctx->cycle_exec);
        ctx->cycle_exec = nullptr;
    }
    ctx->cycle_key.clear();
    ctx->graph_gen++;
}

// src/grid/mod.rs:167-200 and 270-332: the edge kind of boundary cell (x, y)
// from which of its four neighbours are fluid; -2 if too thin.
int edge_kind(uint32_t nx, uint32_t ny, const uint8_t* t, uint32_t x, uint32_t y) {
    const bool up = y > 0 && t[x * ny + (y - 1)] == STM_FLUID;
    const bool down = y + 1 < ny && t[x * ny + (y + 1)] == STM_FLUID;
    const bool right = x + 1 < nx && t[(x + 1) * ny + y] == STM_FLUID;
    const bool left = x > 0 && t[(x - 1) * ny + y] == STM_FLUID;
    const int code = (left ? 8 : 0) | (right ? 4 : 0) | (up ? 2 : 0) | (down ? 1 : 0);
    switch (code) {
    case 0: return E_NONE;
    case 8: return E_W;
    case 8 | 2: return E_NW;
    case 2: return E_N;
    case 4 | 2: return E_NE;
    case 4: return E_E;
    case 4 | 1: return E_SE;
    case 1: return E_S;
    case 8 | 1: return E_SW;
    default: return -2;
    }
}

const char* type_name(uint8_t t) {
    switch (t) {
    case STM_NOSLIP: return "Boundary(NoSlip)";
    case STM_INFLOW: return "Boundary(Inflow)";
    case STM_OUTFLOW: return "Boundary(Outflow)";
    case STM_MOVING_WALL: return "Boundary(MovingWall)";
    default: return "Fluid";
    }
}

int classify(uint32_t nx, uint32_t ny, const uint8_t* t, int8_t* edge, uint32_t* list,
             uint32_t* n, double* fluid, char* err, size_t errlen) {
    uint32_t k = 0;
    uint64_t nf = 0;
    for (uint32_t x = 0; x < nx; x++)
        for (uint32_t y = 0; y < ny; y++) {
            const size_t i = (size_t)x * ny + y;
            if (t[i] > STM_MOVING_WALL) {
                if (err) snprintf(err, errlen, "unknown cell type %u at (%u, %u)", t[i], x, y);
                return STM_ERR_INVALID_ARG;
            }
            if (t[i] == STM_FLUID) {
                nf++;
                edge[i] = -1;
                continue;
            }
            const int e = edge_kind(nx, ny, t, x, y);
            if (e < 0) {
                if (err)
                    snprintf(err, errlen,
                             "BoundaryTooThinError: A cell `%s` at `(%u, %u)` has fluid on "
                             "opposing sides.", type_name(t[i]), x, y);
                return STM_ERR_INVALID_GRID;
            }
            edge[i] = (int8_t)e;
            list[k++] = (uint32_t)i;
        }
    *n = k;
    *fluid = (double)nf;
    return STM_OK;
}

int classify_chan(stm_ctx* ctx, HostChan& c) {
    const size_t n = (size_t)c.nx * c.ny;
    c.edge.assign(n, 0);
    c.blist.assign(n, 0);
    uint32_t nb = 0;
    char err[256] = {0};
    const int rc = classify(c.nx, c.ny, c.type.data(), c.edge.data(), c.blist.data(), &nb,
                            &c.fluid_cells, err, sizeof err);
    if (rc != STM_OK) return fail(ctx, rc, "%s", err);
    c.blist.resize(nb);
    uint64_t inner = 0;
    for (uint32_t x = 1; x + 1 < c.nx; x++)
        for (uint32_t y = 1; y + 1 < c.ny; y++) inner += c.type[(size_t)x * c.ny + y] == STM_FLUID;
    c.n_inner = (double)inner;
    return STM_OK;
}

ChanDesc make_desc(const HostChan& c) {
    ChanDesc d{};
    const stm_params& q = c.prm;
    d.nx = (int32_t)c.nx;
    d.ny = (int32_t)c.ny;
    d.pitch = (int32_t)c.pitch;
    d.solver = q.solver;
    d.off = c.off;
    d.boff = c.boff;
    d.nb = (int32_t)c.blist.size();
    d.tile0 = c.tile0;
    d.ntiles = c.ntiles;
    d.itermax = (int32_t)q.max_iterations;
    d.check_every = q.check_every > 0 ? q.check_every : 1;
    d.tb0 = c.tb0;
    d.ntb = c.ntb;
    d.delx = q.delx;
    d.dely = q.dely;
    d.gamma = q.gamma;
    d.reynolds = q.reynolds;
    d.omega = q.omega;
    d.eps = q.sor_absolute_epsilon;
    d.tau = q.tau;
    d.fluid_cells = c.fluid_cells;
    d.n_inner = c.n_inner;
    // The same expressions as the oracle (and src/simulation.rs:240-246).
    d.delx2 = q.delx * q.delx;
    d.dely2 = q.dely * q.dely;
    d.omw = 1.0 - q.omega;
    d.mid = q.omega / ((2.0 / d.delx2) + (2.0 / d.dely2));
    d.rdx2 = 1.0 / d.delx2;
    d.rdy2 = 1.0 / d.dely2;
    d.eps2 = q.sor_absolute_epsilon * q.sor_absolute_epsilon;
    return d;
}

void fill_entries(const HostChan& c, int chan, std::vector<BEntry>& out) {
    const uint32_t ny = c.ny;
    for (uint32_t i : c.blist) {
        const uint32_t x = i / ny, y = i % ny;
        BEntry e{};
        e.gidx = c.off + (int64_t)x * c.pitch + y;
        e.chan = chan;
        e.pitch = (int32_t)c.pitch;
        e.type = c.type[i];
        e.edge = (uint8_t)c.edge[i];
        e.colour = (uint8_t)((x + y) & 1);
        e.bu = (e.type == STM_INFLOW || e.type == STM_MOVING_WALL) ? c.vel[2 * i] : 0.0;
        e.bv = (e.type == STM_INFLOW || e.type == STM_MOVING_WALL) ? c.vel[2 * i + 1] : 0.0;
        // Order rule of the parallel SETBCOND: a West / NorthWest cell reads
        // v[W] after the wall cell at (x-1, y+1) has stored its bv there.
        if ((e.edge == E_W || e.edge == E_NW) && y + 1 < ny) {
            const uint32_t j = (x - 1) * ny + (y + 1);
            const uint8_t tw = c.type[j];
            if (tw == STM_NOSLIP || tw == STM_INFLOW || tw == STM_MOVING_WALL) {
                e.vw_rule = 1;
                e.vw_const = tw == STM_NOSLIP ? 0.0 : c.vel[2 * j + 1];
            }
        }
        out.push_back(e);
    }
}

int upload_boundaries(stm_ctx* ctx) {
    invalidate_graphs(ctx);   // ghost refreshes capture d_be and nbtot
    std::vector<BEntry> be;
    int boff = 0;
    for (size_t k = 0; k < ctx->chans.size(); k++) {
        ctx->chans[k].boff = boff;
        fill_entries(ctx->chans[k], (int)k, be);
        boff = (int)be.size();
    }
    ctx->nbtot = (int)be.size();
    if (ctx->nbtot > ctx->be_capacity) {
        if (ctx->d_be) HIPCHK(ctx, hipFree(ctx->d_be));
        if (ctx->d_restore) HIPCHK(ctx, hipFree(ctx->d_restore));
        ctx->d_be = nullptr;
        ctx->d_restore = nullptr;
        const int cap = std::max(ctx->nbtot, 1);
        HIPCHK(ctx, hipMalloc(&ctx->d_be, sizeof(BEntry) * cap));
        HIPCHK(ctx, hipMalloc(&ctx->d_restore, sizeof(double2) * cap));
        ctx->be_capacity = cap;
    }
    if (ctx->nbtot)
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_be, be.data(), sizeof(BEntry) * be.size(),
                                   hipMemcpyHostToDevice, ctx->stream));
    return STM_OK;
}

int upload_flags(stm_ctx* ctx, const HostChan& c) {
    std::vector<uint8_t> fl((size_t)c.nx * c.pitch, 0);
    for (uint32_t x = 0; x < c.nx; x++)
        for (uint32_t y = 0; y < c.ny; y++) {
            const size_t i = (size_t)x * c.ny + y;
            fl[(size_t)x * c.pitch + y] =
                make_flag(c.type[i], c.type[i] == STM_FLUID ? 0 : (uint8_t)c.edge[i]);
        }
    HIPCHK(ctx, hipMemcpyAsync(ctx->flags + c.off, fl.data(), fl.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    return STM_OK;
}

int upload_descs(stm_ctx* ctx) {
    std::vector<ChanDesc> cd;
    for (auto& c : ctx->chans) cd.push_back(make_desc(c));
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_cd, cd.data(), sizeof(ChanDesc) * cd.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    return STM_OK;
}

int count_tb(stm_ctx* ctx, int variant) {
    const TbVariant& v = TB_VARIANTS[variant];
    const int H = 2 * v.k + 4;
    int n = 0;
    for (auto& c : ctx->chans)
        n += (int)(plan_axis((int)c.nx, v.ex, H, 1).size() *
                   plan_axis((int)c.ny, v.ey, H, 2).size());
    return n;
}

// Morton index of (a, b): bits interleaved, a in the odd positions.
uint64_t morton2(uint32_t a, uint32_t b) {
    uint64_t m = 0;
    for (int i = 0; i < 32; i++)
        m |= (uint64_t)((a >> i) & 1u) << (2 * i + 1) | (uint64_t)((b >> i) & 1u) << (2 * i);
    return m;
}

// Launch order of one channel's tiles (positions [p0, p0 + n) of the list).
// Workgroups go to the chip's XCDs round-robin (position p to XCD p % 8), and
// each XCD has its own L2: a tile's halo is served from that L2 when the
// neighbour that shares it runs on the same XCD.  So each XCD gets one
// compact block: the tiles in Z order (Morton index of their grid position),
// cut into consecutive runs of the sizes the XCDs' position classes have,
// each run laid on one class.  Every tile keeps its canonical `idx` (its
// residual slot), so the order changes no result.
void order_for_xcds(TbTile* t, int p0, int n, int nyt) {
    constexpr int NXCD = 8;
    std::vector<TbTile> z(t, t + n);
    std::stable_sort(z.begin(), z.end(), [&](const TbTile& a, const TbTile& b) {
        const int ia = a.idx - p0, ib = b.idx - p0;
        return morton2(ia / nyt, ia % nyt) < morton2(ib / nyt, ib % nyt);
    });
    int k = 0;
    for (int c = 0; c < NXCD; c++)
        for (int p = p0 + c; p < p0 + n; p += NXCD) t[p - p0] = z[k++];
}

// Output tiles of the blocked solver for `variant`, channel by channel
// (canonical idx: x-major over the channel's tile grid; launch order from
// order_for_xcds).
int plan_tb(stm_ctx* ctx, int variant) {
    invalidate_graphs(ctx);   // blocked launches capture d_tb and its length
    const TbVariant& v = TB_VARIANTS[variant];
    const int H = 2 * v.k + 4;
    std::vector<TbTile> tl;
    for (size_t k = 0; k < ctx->chans.size(); k++) {
        HostChan& c = ctx->chans[k];
        c.tb0 = (int32_t)tl.size();
        const auto ax = plan_axis((int)c.nx, v.ex, H, 1), ay = plan_axis((int)c.ny, v.ey, H, 2);
        for (auto xs : ax)
            for (auto ys : ay)
                tl.push_back(TbTile{(int32_t)k, xs.start, ys.start, xs.len, ys.len,
                                    (int32_t)tl.size(), xs.origin, ys.origin, (int32_t)c.nx,
                                    (int32_t)c.ny, (int32_t)c.pitch, c.tb0, 0, 0,
                                    (int64_t)c.off});
        c.ntb = (int32_t)tl.size() - c.tb0;
        for (size_t i = (size_t)c.tb0; i < tl.size(); i++) tl[i].ntb = c.ntb;
        order_for_xcds(tl.data() + c.tb0, c.tb0, c.ntb, (int)ay.size());
    }
    if ((int)tl.size() > ctx->tb_capacity) {
        if (ctx->d_tb) HIPCHK(ctx, hipFree(ctx->d_tb));
        ctx->d_tb = nullptr;
        HIPCHK(ctx, hipMalloc(&ctx->d_tb, sizeof(TbTile) * tl.size()));
        ctx->tb_capacity = (int)tl.size();
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_tb, tl.data(), sizeof(TbTile) * tl.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    // the codes table (graphs holding the old one were dropped above)
    const size_t cbytes = v.codes ? (size_t)tl.size() * v.nt * TB_CODE_BYTES : 0;
    if (cbytes > ctx->codes_capacity) {
        if (ctx->d_codes) HIPCHK(ctx, hipFree(ctx->d_codes));
        ctx->d_codes = nullptr;
        HIPCHK(ctx, hipMalloc(&ctx->d_codes, cbytes));
        ctx->codes_capacity = cbytes;
    }
    ctx->ntb = (int)tl.size();
    ctx->h_tb = tl;
    ctx->tb_cur = ctx->d_tb;
    ctx->ntb_cur = ctx->ntb;
    ctx->variant = variant;
    return upload_descs(ctx);
}

// The variant for the requested check interval (the sweeps per launch):
// check_every = K picks the first K-sweep variant of the list (the tuners
// may pick another with the same K); check_every <= 0 leaves the choice to
// the tuners, starting from the first K = 4 variant.
int default_variant(stm_ctx* ctx) {
    const int want = ctx->chans.empty() ? 4 : ctx->chans[0].prm.check_every;
    for (int v = 0; v < N_TB_VARIANTS; v++)
        if (TB_VARIANTS[v].k == (want > 0 ? want : 4)) return v;
    return 2;
}

// Which solvers the channels use (a tick launches only what some channel needs).
void update_solvers(stm_ctx* ctx) {
    ctx->any_rb = ctx->any_lex = false;
    for (const auto& c : ctx->chans) {
        if (c.prm.solver == STM_SOLVER_RB) ctx->any_rb = true;
        else ctx->any_lex = true;
    }
}

bool variant_allowed(stm_ctx* ctx, int v) {
    const int want = ctx->chans[0].prm.check_every;
    return want <= 0 || TB_VARIANTS[v].k == want;
}

int finalize(stm_ctx* ctx) {
    if (ctx->finalized) return STM_OK;
    if (ctx->chans.empty()) return fail(ctx, STM_ERR_STATE, "no grid added");
    int64_t off = 0;
    int tiles = 0;
    std::vector<Tile> tl;
    for (size_t k = 0; k < ctx->chans.size(); k++) {
        HostChan& c = ctx->chans[k];
        c.pitch = (uint32_t)round_up(c.ny, 16);
        c.off = off;
        off += round_up((int64_t)c.nx * c.pitch, 64);
        c.tile0 = tiles;
        for (uint32_t x0 = 0; x0 < c.nx; x0 += TILE_X)
            for (uint32_t y0 = 0; y0 < c.ny; y0 += TILE_Y)
                tl.push_back(Tile{(int32_t)k, (int32_t)x0, (int32_t)y0, 0});
        c.ntiles = (int)tl.size() - tiles;
        tiles = (int)tl.size();
    }
    update_solvers(ctx);
    ctx->total = off;
    ctx->ntiles = tiles;
    const size_t bytes = sizeof(double) * (size_t)off;
    double** fields[] = {&ctx->u, &ctx->v, &ctx->p, &ctx->f, &ctx->g, &ctx->rhs, &ctx->scratch,
                         &ctx->p2};
    for (double** fp : fields) {
        HIPCHK(ctx, hipMalloc(fp, bytes));
        HIPCHK(ctx, hipMemsetAsync(*fp, 0, bytes, ctx->stream));
    }
    HIPCHK(ctx, hipMalloc(&ctx->flags, (size_t)off));
    HIPCHK(ctx, hipMemsetAsync(ctx->flags, 0, (size_t)off, ctx->stream));
    const size_t nch = ctx->chans.size();
    HIPCHK(ctx, hipMalloc(&ctx->d_cd, sizeof(ChanDesc) * nch));
    HIPCHK(ctx, hipMalloc(&ctx->d_cs, sizeof(ChanState) * nch));
    HIPCHK(ctx, hipMalloc(&ctx->d_tiles, sizeof(Tile) * tl.size()));
    size_t maxtiles = tl.size();
    for (int v = 0; v < N_TB_VARIANTS; v++)
        maxtiles = std::max(maxtiles, (size_t)count_tb(ctx, v));
    HIPCHK(ctx, hipMalloc(&ctx->d_part, sizeof(double) * NPART * maxtiles));
    ctx->part_capacity = (int)maxtiles;
    HIPCHK(ctx, hipMalloc(&ctx->d_active, sizeof(int32_t) * nch));
    HIPCHK(ctx, hipMalloc(&ctx->d_tbpart, sizeof(double) * 2 * maxtiles));
    HIPCHK(ctx, hipHostMalloc(&ctx->h_active, sizeof(int32_t) * 2 * nch, hipHostMallocDefault));
    for (hipEvent_t& e : ctx->flag_evs)
        HIPCHK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(ctx, hipEventCreateWithFlags(&ctx->tbc_ev, hipEventDisableTiming));
    HIPCHK(ctx, hipHostMalloc(&ctx->h_cs, sizeof(ChanState) * nch, hipHostMallocDefault));
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_tiles, tl.data(), sizeof(Tile) * tl.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    int rc = upload_boundaries(ctx);
    if (rc) return rc;
    for (auto& c : ctx->chans)
        if ((rc = upload_flags(ctx, c))) return rc;
    if ((rc = plan_tb(ctx, default_variant(ctx)))) return rc;
    std::vector<ChanState> cs(nch);
    for (size_t k = 0; k < nch; k++) {
        const stm_params& q = ctx->chans[k].prm;
        ChanState s{};
        s.delt = q.delt;
        s.time = q.time;
        s.iterations = q.iterations;
        s.has_init = q.has_initial_norm_squared ? 1 : 0;
        s.init_norm2 = q.has_initial_norm_squared ? q.initial_norm_squared : 0.0;
        cs[k] = s;
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_cs, cs.data(), sizeof(ChanState) * nch,
                               hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->finalized = true;
    return STM_OK;
}

// ------------------------------------------------------------- profiling

int drain(stm_ctx* ctx) {
    for (auto& t : ctx->pending) {
        HIPCHK(ctx, hipEventSynchronize(t.b));
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, t.a, t.b));
        t.stat->ms += ms;
        t.stat->timed += t.count;
        ctx->free_events.push_back(t.a);
        ctx->free_events.push_back(t.b);
    }
    ctx->pending.clear();
    return STM_OK;
}

hipEvent_t take_event(stm_ctx* ctx) {
    if (ctx->free_events.empty()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    hipEvent_t e = ctx->free_events.back();
    ctx->free_events.pop_back();
    return e;
}

template <class F>
int launch(stm_ctx* ctx, const char* name, F&& fn) {
    KStat& k = ctx->stats[name];
    const bool timed = ctx->sample && !ctx->span_open && (k.launches % ctx->sample) == 0;
    k.launches++;
    hipEvent_t a = nullptr, b = nullptr;
    if (timed) {
        a = take_event(ctx);
        b = take_event(ctx);
        if (!a || !b) return fail(ctx, STM_ERR_HIP, "hipEventCreate failed");
        HIPCHK(ctx, hipEventRecord(a, ctx->stream));
    }
    fn();
    if (timed) {
        HIPCHK(ctx, hipEventRecord(b, ctx->stream));
        ctx->pending.push_back(PendingTiming{&k, a, b, 1});
        if (ctx->pending.size() >= 2048) return drain(ctx);
    }
    return STM_OK;
}

#define LAUNCH(name, ...)                                      \
    do {                                                       \
        int rc_ = launch(ctx, name, [&] { __VA_ARGS__; });     \
        if (rc_) return rc_;                                   \
    } while (0)

int check_launch(stm_ctx* ctx) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(ctx, STM_ERR_HIP, "kernel launch: %s", hipGetErrorString(e));
    return STM_OK;
}

int ready(stm_ctx* ctx) {
    if (!ctx) return fail(nullptr, STM_ERR_INVALID_ARG, "null context");
    int rc = finalize(ctx);
    if (rc) return rc;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    return STM_OK;
}

int nch(stm_ctx* ctx) { return (int)ctx->chans.size(); }

// ------------------------------------------------------------- stages

int stage_bc(stm_ctx* ctx) {
    LAUNCH("bc_velocity", launch_bc_velocity(ctx->stream, ctx->nbtot, ctx->d_be, ctx->d_restore,
                                             ctx->u, ctx->v));
    return STM_OK;
}

int stage_fg(stm_ctx* ctx) {
    LAUNCH("fg", launch_fg(ctx->stream, ctx->ntiles, ctx->d_cd, ctx->d_cs, ctx->d_tiles,
                           ctx->flags, ctx->u, ctx->v, ctx->f, ctx->g));
    return STM_OK;
}

// COMP_FG and COMP_RHS in one pass
int stage_fg_rhs(stm_ctx* ctx) {
    LAUNCH("fg_rhs", launch_fg_rhs(ctx->stream, ctx->ntiles, ctx->d_cd, ctx->d_cs, ctx->d_tiles,
                                   ctx->flags, ctx->u, ctx->v, ctx->f, ctx->g, ctx->rhs));
    return STM_OK;
}

int stage_rhs(stm_ctx* ctx) {
    LAUNCH("rhs", launch_rhs(ctx->stream, ctx->ntiles, ctx->d_cd, ctx->d_cs, ctx->d_tiles,
                             ctx->f, ctx->g, ctx->rhs));
    return STM_OK;
}

// After a poll (`active`: the flags it read): when the channels still running
// hold under 7/8 of the tiles launched so far, later launches cover only
// their tiles.  A channel that stopped has already taken its stop decision
// (that is what the poll read), and every tile of a running channel is kept,
// so each launch still writes the partials its successors read.  The pinned
// staging buffer is rewritten only after the copy out of it has completed
// (the host reads a batch's flags while the next batch is queued, so a poll
// does not imply that the copy queued after the previous one has run).
int compact_tiles(stm_ctx* ctx, const int32_t* active) {
    int n = 0;
    for (int k = 0; k < nch(ctx); k++)
        if (active[k]) n += ctx->chans[k].ntb;
    if (n == 0 || 8 * n >= 7 * ctx->ntb_cur) return STM_OK;
    if (ctx->tbc_pending) {
        HIPCHK(ctx, hipEventSynchronize(ctx->tbc_ev));
        ctx->tbc_pending = false;
    }
    if (ctx->tbc_capacity < ctx->ntb) {
        if (ctx->d_tbc) HIPCHK(ctx, hipFree(ctx->d_tbc));
        if (ctx->h_tbc) HIPCHK(ctx, hipHostFree(ctx->h_tbc));
        ctx->d_tbc = ctx->h_tbc = nullptr;
        ctx->tbc_capacity = 0;
        HIPCHK(ctx, hipMalloc(&ctx->d_tbc, sizeof(TbTile) * ctx->ntb));
        HIPCHK(ctx, hipHostMalloc(&ctx->h_tbc, sizeof(TbTile) * ctx->ntb, hipHostMallocDefault));
        ctx->tbc_capacity = ctx->ntb;
    }
    int m = 0;
    for (int k = 0; k < nch(ctx); k++) {
        if (!active[k]) continue;
        const HostChan& c = ctx->chans[k];
        for (int i = c.tb0; i < c.tb0 + c.ntb; i++) ctx->h_tbc[m++] = ctx->h_tb[i];
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_tbc, ctx->h_tbc, sizeof(TbTile) * m, hipMemcpyHostToDevice,
                               ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->tbc_ev, ctx->stream));
    ctx->tbc_pending = true;
    ctx->tb_cur = ctx->d_tbc;
    ctx->ntb_cur = m;
    return STM_OK;
}

// Launch L of the blocked solver on the current variant.  Every sequence
// of launches starts at L = 0, which computes the cell codes of every
// running tile and stores them; launches L > 0 of the sequence load them
// (variants whose codes fit the table).
void launch_blocked(stm_ctx* ctx, int variant, int L, int mode) {
    const int cap = ctx->part_capacity;
    double* prev = ctx->d_tbpart + (size_t)((L + 1) & 1) * cap;
    double* cur = ctx->d_tbpart + (size_t)(L & 1) * cap;
    if (TB_VARIANTS[variant].codes) mode |= L == 0 ? TB_CODES_PUT : TB_CODES_GET;
    launch_rb_tb(ctx->stream, variant, ctx->ntb_cur, ctx->d_cd, ctx->d_cs, ctx->tb_cur, ctx->flags,
                 ctx->rhs, ctx->p, ctx->p2, prev, cur, ctx->d_active, ctx->d_codes, L, mode);
}

// Blocked red-black launches until no channel is running.  Launch L also
// takes the stop decision on launch L - 1, so a solve ends with one launch
// that only decides.  Launches are queued in growing batches; after each
// batch the per-channel flags are copied to a pinned slot (slots alternate,
// each with an event), and the host then reads the previous batch's slot: the
// queue holds a batch while the host decides, and at most one batch after the
// stop runs for nothing (a launch finding its channel stopped returns after
// its loads).  check = 0 runs a fixed number of sweeps without the residual;
// fixed > 0: exactly that many launches, with no polls (every channel's
// budget ends at the same launch).
int blocked_solve(stm_ctx* ctx, int check, int fixed = 0) {
    int batch = 2, L = 0, prev = -1, slot = 0;
    const int n = nch(ctx);
    ctx->tb_cur = ctx->d_tb;
    ctx->ntb_cur = ctx->ntb;
    for (;;) {
        // when profiling, batches of 8 or more launches are also timed as a
        // whole ("rb_tb_batch"): one pair of events around back-to-back
        // launches, whose cost per launch is then negligible
        hipEvent_t ea = nullptr;
        const bool span = ctx->sample && batch >= 8;
        if (span) {
            ea = take_event(ctx);
            if (!ea) return fail(ctx, STM_ERR_HIP, "hipEventCreate failed");
            HIPCHK(ctx, hipEventRecord(ea, ctx->stream));
            ctx->span_open = true;
        }
        if (fixed > 0) batch = std::min(batch, fixed - L);
        for (int b = 0; b < batch; b++, L++) {
            const int rc = launch(ctx, "rb_tb", [&] {
                launch_blocked(ctx, ctx->variant, L, check ? TB_CHECK : 0);
            });
            if (rc) {
                if (span) {
                    ctx->span_open = false;
                    ctx->free_events.push_back(ea);
                }
                return rc;
            }
        }
        if (span) {
            ctx->span_open = false;
            hipEvent_t eb = take_event(ctx);
            if (!eb) return fail(ctx, STM_ERR_HIP, "hipEventCreate failed");
            HIPCHK(ctx, hipEventRecord(eb, ctx->stream));
            ctx->pending.push_back(PendingTiming{&ctx->stats["rb_tb_batch"], ea, eb,
                                                 (uint32_t)batch});
        }
        if (fixed > 0) {
            if (L >= fixed) break;
            if (batch < 64) batch *= 2;
            continue;
        }
        HIPCHK(ctx, hipMemcpyAsync(ctx->h_active + (size_t)slot * n, ctx->d_active,
                                   sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipEventRecord(ctx->flag_evs[slot], ctx->stream));
        if (prev >= 0) {
            HIPCHK(ctx, hipEventSynchronize(ctx->flag_evs[prev]));
            int rc = check_launch(ctx);
            if (rc) return rc;
            const int32_t* active = ctx->h_active + (size_t)prev * n;
            bool any = false;
            for (int k = 0; k < n; k++) any = any || active[k];
            if (!any) break;
            if ((rc = compact_tiles(ctx, active))) return rc;
        }
        prev = slot;
        slot ^= 1;
        if (batch < 64) batch *= 2;
    }
    ctx->tb_cur = ctx->d_tb;
    ctx->ntb_cur = ctx->ntb;
    LAUNCH("pbuf_fix", launch_pbuf_fix(ctx->stream, ctx->ntiles, nch(ctx), ctx->d_cd, ctx->d_cs,
                                       ctx->d_tiles, ctx->p2, ctx->p));
    return STM_OK;
}

// Mean of rhs and the sum of (rhs - mean)^2 over interior fluid cells.
int stage_rhs_stats(stm_ctx* ctx) {
    LAUNCH("rhs_stats", launch_rhs_stats(ctx->stream, ctx->ntiles, nch(ctx), ctx->d_cd,
                                         ctx->d_cs, ctx->d_tiles, ctx->flags, ctx->rhs,
                                         ctx->d_part));
    return STM_OK;
}

// ------------------------------------------------------------- learned solve

// Ghosts of p, the fine residual, and per channel: converged?
int learn_prepare(stm_ctx* ctx, int first) {
    LAUNCH("learn_ghost", launch_ghost(ctx->stream, ctx->nbtot, ctx->d_be, ctx->d_cs, ctx->p, -1,
                                       0));
    LAUNCH("learn_residual", launch_learn_residual(ctx->stream, ctx->ntiles, ctx->d_cd, ctx->d_cs,
                                                   ctx->d_lc, first, ctx->d_tiles, ctx->flags,
                                                   ctx->rhs, ctx->p, ctx->lres, ctx->d_part));
    LAUNCH("learn_check", launch_learn_check(ctx->stream, nch(ctx), ctx->d_cd, ctx->d_cs,
                                             ctx->d_lc, ctx->d_part, first));
    return STM_OK;
}

// The coarse residual of every channel in the net's frame and its scale.
int learn_restrict(stm_ctx* ctx, const Net& net, int f, int save) {
    LAUNCH("learn_restrict", launch_learn_restrict(ctx->stream, nch(ctx), ctx->d_cd, ctx->d_lc,
                                                   ctx->flags, ctx->lres, net.H, net.W, f,
                                                   ctx->lrc, ctx->lmc, ctx->d_cpart, save));
    return STM_OK;
}

// Blocked red-black launches on the channels still running, `budget` sweeps
// each (0 = up to itermax); final = 1 ends the solve.  A phase with a budget
// is ceil(budget / K) launches without the residual test and without a host
// poll (the next cycle's residual is the test); launches on a channel that
// is not running return at once.
int learn_phase(stm_ctx* ctx, int budget, int final_phase) {
    LAUNCH("learn_arm", launch_learn_arm(ctx->stream, nch(ctx), ctx->d_cd, ctx->d_cs, ctx->d_lc,
                                         ctx->d_active, budget));
    if (budget > 0) {
        // no channel sweeps past its itermax (k_learn_arm clamps each one's
        // budget), so launches beyond the largest itermax would only return
        int most = 0;
        for (const auto& c : ctx->chans) most = std::max(most, (int)c.prm.max_iterations);
        const int K = TB_VARIANTS[ctx->variant].k;
        const int nl = (std::min(budget, most) + K - 1) / K;
        for (int L = 0; L < nl; L++)
            LAUNCH("rb_tb", launch_blocked(ctx, ctx->variant, L, 0));
        LAUNCH("learn_phase_end", launch_learn_phase_end(ctx->stream, nch(ctx), ctx->d_cs,
                                                         ctx->d_lc, nl));
        // the channels whose latest p is in the second buffer copy it back
        // (after an odd number of launches, or a clamped budget's stop)
        LAUNCH("pbuf_fix", launch_pbuf_fix(ctx->stream, ctx->ntiles, nch(ctx), ctx->d_cd,
                                           ctx->d_cs, ctx->d_tiles, ctx->p2, ctx->p));
        if (!final_phase) return STM_OK;   // the sweeps are counted by learn_phase_end
    } else {
        int rc = blocked_solve(ctx, 1);
        if (rc) return rc;
    }
    LAUNCH("learn_collect", launch_learn_collect(ctx->stream, nch(ctx), ctx->d_cs, ctx->d_lc,
                                                 final_phase));
    return STM_OK;
}

// The network's correction of p: the running channels packed to the front
// of the batch (the network skips the rest), forward on their coarse
// residuals, prolongation, ghosts, step length, update.
int learn_correct(stm_ctx* ctx, Net& net, int f, int first, hipEvent_t ev0, hipEvent_t ev1) {
    const int n = nch(ctx);
    double* prev = (ctx->lstep == STM_STEP_MOMENTUM || ctx->lstep == STM_STEP_MOMENTUM2)
                       ? ctx->lprev : nullptr;
    double* prev2 = ctx->lstep == STM_STEP_MOMENTUM2 ? ctx->lprev2 : nullptr;
    LAUNCH("learn_compact", launch_learn_compact(ctx->stream, n, ctx->d_lc, ctx->d_map,
                                                 ctx->d_inv, ctx->d_live));
    LAUNCH("learn_pack", launch_learn_pack(ctx->stream, n, ctx->d_lc, ctx->lrc, ctx->lmc, net.H,
                                           net.W, net.input(), nullptr, 0, 1, ctx->d_map,
                                           ctx->d_live));
    if (ev0) HIPCHK(ctx, hipEventRecord(ev0, ctx->stream));
    LAUNCH("nn_forward", net.forward_infer(n, ctx->d_live));
    if (ev1) HIPCHK(ctx, hipEventRecord(ev1, ctx->stream));
    LAUNCH("learn_prolong", launch_learn_prolong(ctx->stream, ctx->ntiles, ctx->d_cd,
                                                 ctx->d_tiles, ctx->d_lc, ctx->flags, net.y,
                                                 ctx->lmc, net.H, net.W, f, ctx->d_inv,
                                                 ctx->lcf));
    LAUNCH("learn_ghost", launch_ghost(ctx->stream, ctx->nbtot, ctx->d_be, ctx->d_cs, ctx->lcf,
                                       -1, 0));
    LAUNCH("learn_dots", launch_learn_dots(ctx->stream, ctx->ntiles, ctx->d_cd, ctx->d_tiles,
                                           ctx->d_lc, ctx->flags, ctx->lres, ctx->lcf, ctx->p,
                                           prev, prev2, ctx->d_part));
    LAUNCH("learn_alpha", launch_learn_alpha(ctx->stream, n, ctx->d_cd, ctx->d_lc, ctx->d_part,
                                             ctx->lstep, f, first));
    LAUNCH("learn_apply", launch_learn_apply(ctx->stream, ctx->ntiles, ctx->d_cd, ctx->d_tiles,
                                             ctx->d_lc, ctx->flags, ctx->lcf, ctx->p, prev,
                                             prev2));
    return STM_OK;
}

// One learned cycle after the first: residual, convergence test, network
// correction, budgeted sweeps.  Every launch depends only on device pointers
// and on the settings in `key`, so with graphs on the sequence is captured
// once into a hipGraph and replayed.
int learned_cycle(stm_ctx* ctx, Net& net, int f) {
    int rc;
    if ((rc = learn_prepare(ctx, 0))) return rc;
    if ((rc = learn_restrict(ctx, net, f, -1))) return rc;
    if ((rc = learn_correct(ctx, net, f, 0, nullptr, nullptr))) return rc;
    return learn_phase(ctx, ctx->lsweeps, 0);
}

int graph_cycle(stm_ctx* ctx, Net& net, int f) {
    const std::vector<int64_t> key = {(int64_t)ctx->graph_gen, (int64_t)net.serial, f,
                                      ctx->lsweeps, ctx->lstep, ctx->variant, net.H, net.W,
                                      nch(ctx), (int64_t)(intptr_t)ctx->tb_cur, ctx->ntb_cur};
    if (!ctx->cycle_exec || key != ctx->cycle_key) {
        if (ctx->cycle_exec) HIPCHK(ctx, hipGraphExecDestroy(ctx->cycle_exec));
        ctx->cycle_exec = nullptr;
        hipGraph_t g = nullptr;
        HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int rc = learned_cycle(ctx, net, f);
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(ctx, STM_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiateWithFlags(&ctx->cycle_exec, g, 0);
        hipGraphDestroy(g);
        if (ei != hipSuccess) {
            ctx->cycle_exec = nullptr;
            return fail(ctx, STM_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ei));
        }
        ctx->cycle_key = key;
    }
    HIPCHK(ctx, hipGraphLaunch(ctx->cycle_exec, ctx->stream));
    return STM_OK;
}

// The learned solve reads back how many channels cycle k found running
// (k_learn_compact's count) into a small ring of pinned slots, each with an
// event.  Before queueing cycle k the host waits for cycle k - 1 - POLL_LAG
// only, so the queue keeps POLL_LAG + 1 cycles of work while the host
// decides; a slot is rewritten POLL_RING cycles later, after its event was
// waited for.
constexpr int POLL_LAG = 1, POLL_RING = 4;
static_assert(POLL_RING > POLL_LAG + 1, "a slot is read before it is rewritten");

int post_live(stm_ctx* ctx, int k) {
    const int j = k % POLL_RING;
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_live_ring + j, ctx->d_live, sizeof(int32_t),
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->poll_evs[j], ctx->stream));
    return STM_OK;
}

// Red-black solve in phases.  Phase k starts with the residual and the
// convergence test; when capturing, its first `cpoints` phases each store
// a training input and a snapshot of p; with the learned solve on, each
// applies the network's correction.  The learned solve runs `lcycles`
// phases of `lsweeps` sweeps, a plain capture `cpoints` phases of csweeps,
// 2 csweeps, 4 csweeps, ...; a final phase runs to eps.  Captured targets
// are the restricted change from each snapshot to the converged p.
int phased_solve(stm_ctx* ctx) {
    Net* L = (ctx->lnet && ctx->lcycles > 0) ? &ctx->lnet->net : nullptr;
    Net* C = ctx->cnet ? &ctx->cnet->net : nullptr;
    if (L && C && (L != C || ctx->lcoarsen != ctx->ccoarsen))
        return fail(ctx, STM_ERR_STATE,
                    "capture during a learned solve needs the same net and coarsening");
    const int n = nch(ctx);
    const int f = L ? ctx->lcoarsen : ctx->ccoarsen;
    const int phases = L ? ctx->lcycles : ctx->cpoints;
    const size_t bytes = sizeof(double) * (size_t)ctx->total;
    const int slot_base = C ? C->dwrite : 0;
    int captured = 0, corrected = 0, rc;
    // graphs replay cycles 1.. of a learned solve that captures nothing and is not profiled
    const bool graphs = L && !C && ctx->use_graphs && ctx->sample == 0;
    for (int k = 0; k < phases; k++) {
        if (L && k > POLL_LAG) {
            // cycle k - 1 - POLL_LAG found no channel running: the cycles after it are no-ops
            const int j = (k - 1 - POLL_LAG) % POLL_RING;
            HIPCHK(ctx, hipEventSynchronize(ctx->poll_evs[j]));
            if (ctx->h_live_ring[j] == 0) break;
        }
        if (graphs && k > 0) {
            if ((rc = graph_cycle(ctx, *L, f))) return rc;
            if ((rc = post_live(ctx, k))) return rc;
            continue;
        }
        if ((rc = learn_prepare(ctx, k == 0))) return rc;
        const bool cap = C && captured < ctx->cpoints;
        if ((rc = learn_restrict(ctx, L ? *L : *C, f, cap ? captured : -1))) return rc;
        if (cap) {
            LAUNCH("learn_pack", launch_learn_pack(ctx->stream, n, ctx->d_lc, ctx->lrc, ctx->lmc,
                                                   C->H, C->W, nullptr, C->dx,
                                                   slot_base + captured * n, C->dcap));
            HIPCHK(ctx, hipMemcpyAsync(ctx->lsnap + (size_t)captured * ctx->total, ctx->p, bytes,
                                       hipMemcpyDeviceToDevice, ctx->stream));
            captured++;
        }
        if (L) {
            while ((int)ctx->nn_evs.size() < 2 * (k + 1)) {
                hipEvent_t e;
                HIPCHK(ctx, hipEventCreate(&e));
                ctx->nn_evs.push_back(e);
            }
            if ((rc = learn_correct(ctx, *L, f, k == 0, ctx->nn_evs[2 * k],
                                    ctx->nn_evs[2 * k + 1])))
                return rc;
            corrected++;
        }
        const int budget = L ? ctx->lsweeps : ctx->csweeps << k;
        if ((rc = learn_phase(ctx, budget, 0))) return rc;
        if (L && (rc = post_live(ctx, k))) return rc;
    }
    // the residual after the last budgeted phase: channels that converged
    // there, or used up itermax, end before the final phase
    if ((rc = learn_prepare(ctx, 0))) return rc;
    if ((rc = learn_phase(ctx, 0, 1))) return rc;
    for (int j = 0; j < corrected; j++) {   // the final phase has synchronised
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->nn_evs[2 * j], ctx->nn_evs[2 * j + 1]));
        L->fwd_ms += ms;
        L->fwd_calls += 1;
        L->fwd_samples += (uint64_t)n;
    }
    for (int j = 0; j < captured; j++)
        LAUNCH("learn_target", launch_learn_target(ctx->stream, ctx->ntiles, n, ctx->d_cd,
                                                   ctx->d_tiles, ctx->d_lc, ctx->flags, ctx->p,
                                                   ctx->lsnap + (size_t)j * ctx->total,
                                                   ctx->d_part, C->H, C->W, f, C->dt,
                                                   slot_base + j * n, C->dcap, j));
    if (C) {
        C->dwrite = (C->dwrite + captured * n) % C->dcap;
        C->dsize = std::min(C->dsize + captured * n, C->dcap);
    }
    return STM_OK;
}

// Buffers of the learned solve for a net of H x W (and of capture).
int learn_buffers(stm_ctx* ctx, const Net& net, bool capture) {
    invalidate_graphs(ctx);   // the coarse frames may be reallocated below
    const size_t bytes = sizeof(double) * (size_t)ctx->total;
    if (!ctx->d_lc) {
        HIPCHK(ctx, hipMalloc(&ctx->d_lc, sizeof(LearnChan) * nch(ctx)));
        HIPCHK(ctx, hipMemset(ctx->d_lc, 0, sizeof(LearnChan) * nch(ctx)));
        HIPCHK(ctx, hipHostMalloc(&ctx->h_lc, sizeof(LearnChan) * nch(ctx), hipHostMallocDefault));
        HIPCHK(ctx, hipMalloc(&ctx->lres, bytes));
        HIPCHK(ctx, hipMalloc(&ctx->lcf, bytes));
        HIPCHK(ctx, hipMemset(ctx->lres, 0, bytes));
        HIPCHK(ctx, hipMemset(ctx->lcf, 0, bytes));
        HIPCHK(ctx, hipMalloc(&ctx->lprev, bytes));
        HIPCHK(ctx, hipMemset(ctx->lprev, 0, bytes));
        HIPCHK(ctx, hipMalloc(&ctx->lprev2, bytes));
        HIPCHK(ctx, hipMemset(ctx->lprev2, 0, bytes));
        HIPCHK(ctx, hipMalloc(&ctx->d_map, sizeof(int32_t) * nch(ctx)));
        HIPCHK(ctx, hipMalloc(&ctx->d_inv, sizeof(int32_t) * nch(ctx)));
        HIPCHK(ctx, hipMalloc(&ctx->d_live, sizeof(int32_t)));
        HIPCHK(ctx, hipHostMalloc(&ctx->h_live_ring, sizeof(int32_t) * POLL_RING,
                                  hipHostMallocDefault));
        for (int j = 0; j < POLL_RING; j++)
            HIPCHK(ctx, hipEventCreateWithFlags(&ctx->poll_evs[j], hipEventDisableTiming));
    }
    const int64_t frame = (int64_t)net.H * net.W;
    if (frame > ctx->lframe) {
        if (ctx->lrc) hipFree(ctx->lrc);
        if (ctx->lmc) hipFree(ctx->lmc);
        if (ctx->d_cpart) hipFree(ctx->d_cpart);
        ctx->lrc = ctx->lmc = nullptr;
        ctx->d_cpart = nullptr;
        ctx->lframe = 0;
        HIPCHK(ctx, hipMalloc(&ctx->lrc, sizeof(float) * frame * nch(ctx)));
        HIPCHK(ctx, hipMalloc(&ctx->lmc, sizeof(float) * frame * nch(ctx)));
        HIPCHK(ctx, hipMalloc(&ctx->d_cpart, sizeof(double2) * learn_coarse_blocks(net.H, net.W) *
                                                 (size_t)nch(ctx)));
        ctx->lframe = frame;
    }
    if (capture && !ctx->lsnap) HIPCHK(ctx, hipMalloc(&ctx->lsnap, bytes * LEARN_POINTS));
    return STM_OK;
}

// The net fits the context: same device and stream, every channel's
// interior fits f times the net's H x W, a batch entry per channel, all
// channels red-black.
int check_net(stm_ctx* ctx, stm_net* net, uint32_t f) {
    const Net& m = net->net;
    if (m.ctx != ctx) return fail(ctx, STM_ERR_INVALID_ARG, "the net belongs to another context");
    if (f != 1 && f != 2 && f != 4 && f != 8 && f != 16)
        return fail(ctx, STM_ERR_INVALID_ARG, "coarsening %u is not 1, 2, 4, 8 or 16", f);
    if (m.cap < nch(ctx))
        return fail(ctx, STM_ERR_INVALID_ARG, "net batch %d below the channel count %d", m.cap,
                    nch(ctx));
    for (size_t k = 0; k < ctx->chans.size(); k++) {
        const HostChan& c = ctx->chans[k];
        if ((int64_t)c.nx - 2 > (int64_t)f * m.H || (int64_t)c.ny - 2 > (int64_t)f * m.W)
            return fail(ctx, STM_ERR_INVALID_ARG,
                        "channel %zu (interior %ux%u) exceeds %u x the net's %dx%d", k, c.nx - 2,
                        c.ny - 2, f, m.H, m.W);
        if (c.prm.solver != STM_SOLVER_RB)
            return fail(ctx, STM_ERR_INVALID_ARG, "channel %zu is not red-black", k);
    }
    return STM_OK;
}

// POISSON: the rhs statistics, tol2 and the per-channel flags, then each
// channel's solver.
int stage_solve(stm_ctx* ctx) {
    int rc = stage_rhs_stats(ctx);
    if (rc) return rc;
    LAUNCH("solve_begin", launch_solve_begin(ctx->stream, nch(ctx), ctx->d_cd, ctx->d_cs,
                                             ctx->d_part, ctx->d_active));
    if (ctx->any_lex)
        LAUNCH("lex_solve", launch_lex_solve(ctx->stream, nch(ctx), ctx->d_cd, ctx->d_cs,
                                             ctx->d_be, ctx->flags, ctx->rhs, ctx->p,
                                             ctx->scratch));
    if (!ctx->any_rb) return STM_OK;
    if (ctx->cnet || (ctx->lnet && ctx->lcycles > 0)) return phased_solve(ctx);
    return blocked_solve(ctx, 1);
}

// ADAP_UV with the restore list (src/simulation.rs:287-322) and the ranges
// that follow it: the restore list first (it sets every boundary cell), then
// ADAP_UV of the fluid cells fused with the range reductions.
int stage_adap(stm_ctx* ctx, int want_pressure) {
    LAUNCH("restore", launch_restore(ctx->stream, ctx->nbtot, ctx->d_be, ctx->d_restore,
                                     ctx->u, ctx->v));
    LAUNCH("adap_ranges", launch_adap_ranges(ctx->stream, ctx->ntiles, nch(ctx), ctx->d_cd,
                                             ctx->d_cs, ctx->d_tiles, ctx->flags, ctx->f, ctx->g,
                                             ctx->p, ctx->u, ctx->v, ctx->d_part,
                                             want_pressure));
    return STM_OK;
}

int stage_ranges(stm_ctx* ctx, int want_pressure) {
    LAUNCH("ranges", launch_ranges(ctx->stream, ctx->ntiles, nch(ctx), ctx->d_cd, ctx->d_cs,
                                   ctx->d_tiles, ctx->flags, ctx->u, ctx->v, ctx->p,
                                   ctx->d_part, want_pressure));
    return STM_OK;
}

int tick(stm_ctx* ctx) {
    int rc;
    LAUNCH("delt", launch_delt(ctx->stream, nch(ctx), ctx->d_cd, ctx->d_cs));
    if ((rc = stage_bc(ctx))) return rc;
    if ((rc = stage_fg_rhs(ctx))) return rc;
    if ((rc = stage_solve(ctx))) return rc;
    if ((rc = stage_adap(ctx, 2))) return rc;
    LAUNCH("advance_time", launch_advance_time(ctx->stream, nch(ctx), ctx->d_cs));
    return check_launch(ctx);
}

int read_states(stm_ctx* ctx) {
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_cs, ctx->d_cs, sizeof(ChanState) * nch(ctx),
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return check_launch(ctx);
}

int check_chan(stm_ctx* ctx, uint32_t ch) {
    if (!ctx) return fail(nullptr, STM_ERR_INVALID_ARG, "null context");
    if (ch >= ctx->chans.size())
        return fail(ctx, STM_ERR_INVALID_ARG, "channel %u out of range (%zu channels)", ch,
                    ctx->chans.size());
    return STM_OK;
}

int check_params(stm_ctx* ctx, const stm_params* q) {
    if (!q) return fail(ctx, STM_ERR_INVALID_ARG, "null params");
    if (!(q->delx > 0.0) || !(q->dely > 0.0))
        return fail(ctx, STM_ERR_INVALID_ARG, "cell size must be positive");
    if (q->solver != STM_SOLVER_LEX && q->solver != STM_SOLVER_RB)
        return fail(ctx, STM_ERR_INVALID_ARG, "unknown solver %d", q->solver);
    if (q->check_every > 0) {
        bool ok = false;
        for (int v = 0; v < N_TB_VARIANTS; v++) ok = ok || TB_VARIANTS[v].k == q->check_every;
        if (!ok)
            return fail(ctx, STM_ERR_INVALID_ARG,
                        "check_every must be 1, 2, 4, 6 or 8 (or <= 0 to let the tuner pick)");
    }
    return STM_OK;
}

}  // namespace

// =================================================================== C ABI

extern "C" {

int stm_device_count(int* count) {
    if (!count) return fail(nullptr, STM_ERR_INVALID_ARG, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        *count = 0;
        return fail(nullptr, STM_ERR_NO_DEVICE, "no HIP device: %s", hipGetErrorString(e));
    }
    *count = n;
    return STM_OK;
}

int stm_create(int device, stm_ctx** out) {
    if (!out) return fail(nullptr, STM_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    int n = 0;
    int rc = stm_device_count(&n);
    if (rc) return rc;
    if (device < 0 || device >= n)
        return fail(nullptr, STM_ERR_NO_DEVICE, "device %d not present (%d devices)", device, n);
    stm_ctx* ctx = new stm_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(nullptr, STM_ERR_HIP, "stream creation: %s", hipGetErrorString(e));
    }
    // the static image buffer's address on this device (once per device)
    if ((e = render_static_init(device)) != hipSuccess) {
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(nullptr, STM_ERR_HIP, "static image buffer: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return STM_OK;
}

void stm_destroy(stm_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    // nets outlive their context only to be destroyed: they no longer refer to it
    for (stm_net* n : ctx->nets) {
        n->net.ctx = nullptr;
        n->net.stream = nullptr;
    }
    drain(ctx);
    for (hipEvent_t e : ctx->free_events) hipEventDestroy(e);
    void* bufs[] = {ctx->u, ctx->v, ctx->p, ctx->f, ctx->g, ctx->rhs, ctx->scratch, ctx->p2,
                    ctx->flags, ctx->d_cd, ctx->d_cs, ctx->d_be, ctx->d_restore, ctx->d_tiles,
                    ctx->d_part, ctx->d_active, ctx->d_tbpart, ctx->d_tb, ctx->d_codes, ctx->d_tbc, ctx->d_lc, ctx->lres,
                    ctx->lcf, ctx->lprev, ctx->lprev2, ctx->lsnap, ctx->lrc, ctx->lmc, ctx->d_cpart, ctx->d_map,
                    ctx->d_inv, ctx->d_live, ctx->d_rgba};
    for (void* b : bufs)
        if (b) hipFree(b);
    if (ctx->h_active) hipHostFree(ctx->h_active);
    for (hipEvent_t e : ctx->flag_evs)
        if (e) hipEventDestroy(e);
    if (ctx->tbc_ev) hipEventDestroy(ctx->tbc_ev);
    if (ctx->h_cs) hipHostFree(ctx->h_cs);
    if (ctx->h_lc) hipHostFree(ctx->h_lc);
    if (ctx->h_tbc) hipHostFree(ctx->h_tbc);
    if (ctx->h_live_ring) hipHostFree(ctx->h_live_ring);
    for (hipEvent_t e : ctx->poll_evs)
        if (e) hipEventDestroy(e);
    for (hipEvent_t e : ctx->nn_evs) hipEventDestroy(e);
    if (ctx->cycle_exec) hipGraphExecDestroy(ctx->cycle_exec);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* stm_last_error(const stm_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_last_error.c_str();
}

int stm_rebuild_boundary_list(uint32_t nx, uint32_t ny, const uint8_t* cell_type,
                              int8_t* edge_out, uint32_t* list_out, uint32_t* n_out,
                              double* fluid_cells_out, char* err, size_t errlen) {
    if (!cell_type || !edge_out || !list_out || !n_out || !fluid_cells_out || nx == 0 || ny == 0)
        return fail(nullptr, STM_ERR_INVALID_ARG, "invalid arguments");
    const int rc = classify(nx, ny, cell_type, edge_out, list_out, n_out, fluid_cells_out, err,
                            errlen);
    if (rc && err) g_last_error = err;
    return rc;
}

int stm_add_grid(stm_ctx* ctx, const stm_params* params, uint32_t nx, uint32_t ny,
                 const uint8_t* cell_type, const double* cell_velocity, uint32_t* channel_out) {
    if (!ctx) return fail(nullptr, STM_ERR_INVALID_ARG, "null context");
    if (ctx->finalized)
        return fail(ctx, STM_ERR_STATE, "grids must be added before the first upload/construct");
    int rc = check_params(ctx, params);
    if (rc) return rc;
    if (!cell_type || nx < 3 || ny < 3)
        return fail(ctx, STM_ERR_INVALID_ARG, "grid must be at least 3x3 with cell types");
    if (!ctx->chans.empty() && params->check_every != ctx->chans[0].prm.check_every)
        return fail(ctx, STM_ERR_INVALID_ARG, "all channels must share check_every");
    HostChan c;
    c.prm = *params;
    c.nx = nx;
    c.ny = ny;
    const size_t n = (size_t)nx * ny;
    c.type.assign(cell_type, cell_type + n);
    c.vel.assign(2 * n, 0.0);
    if (cell_velocity) std::copy(cell_velocity, cell_velocity + 2 * n, c.vel.begin());
    if ((rc = classify_chan(ctx, c))) return rc;
    ctx->chans.push_back(std::move(c));
    if (channel_out) *channel_out = (uint32_t)(ctx->chans.size() - 1);
    return STM_OK;
}

int stm_set_cells(stm_ctx* ctx, uint32_t ch, const uint8_t* cell_type,
                  const double* cell_velocity) {
    int rc = check_chan(ctx, ch);
    if (rc) return rc;
    if (!cell_type) return fail(ctx, STM_ERR_INVALID_ARG, "null cell types");
    HostChan& c = ctx->chans[ch];
    HostChan trial = c;
    const size_t n = (size_t)c.nx * c.ny;
    trial.type.assign(cell_type, cell_type + n);
    if (cell_velocity) trial.vel.assign(cell_velocity, cell_velocity + 2 * n);
    // the old grid and fields stay on error (draw_cells' rollback)
    if ((rc = classify_chan(ctx, trial))) return rc;
    // cells whose type changes start from rest: u = v = p = 0 (draw_cells,
    // src/lib.rs:52-64); before the first upload every field is still zero
    std::vector<int64_t> changed;
    if (ctx->finalized)
        for (uint32_t x = 0; x < c.nx; x++)
            for (uint32_t y = 0; y < c.ny; y++)
                if (trial.type[(size_t)x * c.ny + y] != c.type[(size_t)x * c.ny + y])
                    changed.push_back(c.off + (int64_t)x * c.pitch + y);
    c = std::move(trial);
    if (!ctx->finalized) return STM_OK;
    if ((rc = ready(ctx))) return rc;
    if ((rc = upload_boundaries(ctx))) return rc;
    if ((rc = upload_flags(ctx, c))) return rc;
    if ((rc = upload_descs(ctx))) return rc;
    if (!changed.empty()) {
        int64_t* d_idx = nullptr;
        HIPCHK(ctx, hipMalloc(&d_idx, sizeof(int64_t) * changed.size()));
        hipError_t e = hipMemcpyAsync(d_idx, changed.data(), sizeof(int64_t) * changed.size(),
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) {
            launch_zero_cells(ctx->stream, (int)changed.size(), d_idx, ctx->u, ctx->v, ctx->p);
            e = hipStreamSynchronize(ctx->stream);
        }
        hipFree(d_idx);
        if (e != hipSuccess)
            return fail(ctx, STM_ERR_HIP, "zeroing edited cells: %s", hipGetErrorString(e));
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return check_launch(ctx);
}

int stm_set_params(stm_ctx* ctx, uint32_t ch, const stm_params* params) {
    int rc = check_chan(ctx, ch);
    if (rc) return rc;
    if ((rc = check_params(ctx, params))) return rc;
    if (params->check_every != ctx->chans[0].prm.check_every && ctx->chans.size() > 1)
        return fail(ctx, STM_ERR_INVALID_ARG, "all channels must share check_every");
    const bool new_k = params->check_every != ctx->chans[ch].prm.check_every;
    ctx->chans[ch].prm = *params;
    update_solvers(ctx);
    if (!ctx->finalized) return STM_OK;
    if ((rc = ready(ctx))) return rc;
    // a new check interval takes the blocked-solver variant with that many
    // sweeps per launch (<= 0: the first K = 4 one, until a tuner picks)
    if (new_k) {
        if ((rc = plan_tb(ctx, default_variant(ctx)))) return rc;
    } else if ((rc = upload_descs(ctx))) {
        return rc;
    }
    if ((rc = read_states(ctx))) return rc;
    ChanState s = ctx->h_cs[ch];
    s.delt = params->delt;
    s.time = params->time;
    s.iterations = params->iterations;
    s.has_init = params->has_initial_norm_squared ? 1 : 0;
    s.init_norm2 = params->has_initial_norm_squared ? params->initial_norm_squared : 0.0;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_cs + ch, &s, sizeof s, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return STM_OK;
}

static int copy_field(stm_ctx* ctx, const HostChan& c, double* dev, const double* host,
                      bool up) {
    if (up)
        HIPCHK(ctx, hipMemcpy2DAsync(dev + c.off, sizeof(double) * c.pitch, host,
                                     sizeof(double) * c.ny, sizeof(double) * c.ny, c.nx,
                                     hipMemcpyHostToDevice, ctx->stream));
    else
        HIPCHK(ctx, hipMemcpy2DAsync((void*)host, sizeof(double) * c.ny, dev + c.off,
                                     sizeof(double) * c.pitch, sizeof(double) * c.ny, c.nx,
                                     hipMemcpyDeviceToHost, ctx->stream));
    return STM_OK;
}

int stm_upload(stm_ctx* ctx, uint32_t ch, const double* u, const double* v, const double* p,
               const double* f, const double* g, const double* rhs) {
    int rc = check_chan(ctx, ch);
    if (rc || (rc = ready(ctx))) return rc;
    const HostChan& c = ctx->chans[ch];
    const double* src[] = {u, v, p, f, g, rhs};
    double* dst[] = {ctx->u, ctx->v, ctx->p, ctx->f, ctx->g, ctx->rhs};
    for (int k = 0; k < 6; k++)
        if (src[k] && (rc = copy_field(ctx, c, dst[k], src[k], true))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return STM_OK;
}

int stm_download(stm_ctx* ctx, uint32_t ch, double* u, double* v, double* p, double* f,
                 double* g, double* rhs) {
    int rc = check_chan(ctx, ch);
    if (rc || (rc = ready(ctx))) return rc;
    const HostChan& c = ctx->chans[ch];
    double* dst[] = {u, v, p, f, g, rhs};
    double* src[] = {ctx->u, ctx->v, ctx->p, ctx->f, ctx->g, ctx->rhs};
    for (int k = 0; k < 6; k++)
        if (dst[k] && (rc = copy_field(ctx, c, src[k], dst[k], false))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return check_launch(ctx);
}

int stm_get_state(stm_ctx* ctx, uint32_t ch, stm_state* out) {
    int rc = check_chan(ctx, ch);
    if (rc || (rc = ready(ctx))) return rc;
    if (!out) return fail(ctx, STM_ERR_INVALID_ARG, "null out");
    if ((rc = read_states(ctx))) return rc;
    const ChanState& s = ctx->h_cs[ch];
    stm_state o{};
    o.delt = s.delt;
    o.time = s.time;
    o.initial_norm_squared = s.init_norm2;
    o.norm_squared = s.norm2;
    o.residual_squared = s.res2;
    o.tolerance_squared = s.tol2;
    o.rhs_mean = s.rhs_mean;
    o.pressure_range[0] = s.pressure_range[0];
    o.pressure_range[1] = s.pressure_range[1];
    o.speed_range[0] = s.speed_range[0];
    o.speed_range[1] = s.speed_range[1];
    o.umax = s.umax;
    o.vmax = s.vmax;
    o.iterations = s.iterations;
    o.sor_iterations = s.sor_iters;
    o.has_initial_norm_squared = s.has_init;
    o.exhausted = s.exhausted;
    *out = o;
    return STM_OK;
}

int stm_construct(stm_ctx* ctx) {
    int rc = ready(ctx);
    if (rc) return rc;
    if ((rc = stage_ranges(ctx, 1))) return rc;
    launch_fill(ctx->stream, ctx->total, ctx->f, 0.0);
    launch_fill(ctx->stream, ctx->total, ctx->g, 0.0);
    launch_fill(ctx->stream, ctx->total, ctx->rhs, 0.0);
    if ((rc = stage_fg_rhs(ctx))) return rc;
    LAUNCH("initial_norm", launch_initial_norm(ctx->stream, nch(ctx), ctx->d_cd, ctx->d_cs,
                                               ctx->rhs, ctx->p, ctx->scratch));
    return read_states(ctx);
}

// After ticks: STM_ERR_NAN when a channel's last residual is not a number.
int check_nan(stm_ctx* ctx) {
    for (int k = 0; k < nch(ctx); k++) {
        const double r = ctx->h_cs[k].norm2;
        if (r != r)
            return fail(ctx, STM_ERR_NAN,
                        "channel %d: the pressure residual is not a number after %u sweeps", k,
                        ctx->h_cs[k].sor_iters);
    }
    return STM_OK;
}

int stm_run_simulation_tick(stm_ctx* ctx, uint32_t* sor_iterations, double* norm_squared) {
    int rc = ready(ctx);
    if (rc || (rc = tick(ctx)) || (rc = read_states(ctx))) return rc;
    for (int k = 0; k < nch(ctx); k++) {
        if (sor_iterations) sor_iterations[k] = ctx->h_cs[k].sor_iters;
        if (norm_squared) norm_squared[k] = ctx->h_cs[k].norm2;
    }
    if ((rc = drain(ctx))) return rc;
    return check_nan(ctx);
}

int stm_run_ticks(stm_ctx* ctx, uint32_t nticks) {
    int rc = ready(ctx);
    if (rc) return rc;
    for (uint32_t t = 0; t < nticks; t++)
        if ((rc = tick(ctx))) return rc;
    if ((rc = read_states(ctx)) || (rc = drain(ctx))) return rc;
    return check_nan(ctx);
}

int stm_set_boundary_u_and_v(stm_ctx* ctx) {
    int rc = ready(ctx);
    if (rc || (rc = stage_bc(ctx))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return check_launch(ctx);
}

int stm_calculate_f_and_g(stm_ctx* ctx) {
    int rc = ready(ctx);
    if (rc || (rc = stage_fg(ctx))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return check_launch(ctx);
}

int stm_calculate_rhs(stm_ctx* ctx) {
    int rc = ready(ctx);
    if (rc || (rc = stage_rhs(ctx))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return check_launch(ctx);
}

int stm_solve_pressure(stm_ctx* ctx) {
    int rc = ready(ctx);
    if (rc) return rc;
    if ((rc = stage_solve(ctx))) return rc;
    if ((rc = stage_ranges(ctx, 2))) return rc;
    return read_states(ctx);
}

int stm_set_u_and_v(stm_ctx* ctx) {
    int rc = ready(ctx);
    if (rc || (rc = stage_adap(ctx, 0))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return check_launch(ctx);
}

int stm_copy_pressure_to_boundaries(stm_ctx* ctx) {
    int rc = ready(ctx);
    if (rc) return rc;
    LAUNCH("ghost", launch_ghost(ctx->stream, ctx->nbtot, ctx->d_be, ctx->d_cs, ctx->p, -1, 0));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return check_launch(ctx);
}

// Sets every red-black channel running with itermax = n (or restores the
// descriptors when n == 0) and zero sweeps done; tol2 < 0 never converges.
int arm_channels(stm_ctx* ctx, int itermax_override, bool never_converge,
                 std::vector<ChanState>* saved) {
    int rc = read_states(ctx);
    if (rc) return rc;
    if (saved) saved->assign(ctx->h_cs, ctx->h_cs + nch(ctx));
    std::vector<ChanState> cs(ctx->h_cs, ctx->h_cs + nch(ctx));
    std::vector<ChanDesc> cd;
    for (size_t k = 0; k < ctx->chans.size(); k++) {
        ChanDesc d = make_desc(ctx->chans[k]);
        if (itermax_override > 0) d.itermax = itermax_override;
        cd.push_back(d);
        ChanState& s = cs[k];
        s.active = ctx->chans[k].prm.solver == STM_SOLVER_RB ? 1 : 0;
        s.stop_launch = s.active ? -1 : 0;
        s.sor_iters = 0;
        s.pbuf = 0;
        s.budget = 0;
        if (never_converge) s.tol2 = -1.0;
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_cd, cd.data(), sizeof(ChanDesc) * cd.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_cs, cs.data(), sizeof(ChanState) * cs.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    // the flags the host polls must say "running" too
    std::vector<int32_t> act(cs.size());
    for (size_t k = 0; k < cs.size(); k++) act[k] = cs[k].active;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_active, act.data(), sizeof(int32_t) * act.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return STM_OK;
}

int stm_rb_sweeps(stm_ctx* ctx, uint32_t n) {
    int rc = ready(ctx);
    if (rc) return rc;
    if (n == 0) return STM_OK;
    std::vector<ChanState> saved;
    if ((rc = stage_rhs_stats(ctx))) return rc;     // the projection uses the current rhs
    if ((rc = arm_channels(ctx, (int)n, false, &saved))) return rc;
    // every channel's budget is n: ceil(n / K) launches sweep, one more decides
    const int k = TB_VARIANTS[ctx->variant].k;
    if ((rc = blocked_solve(ctx, 0, ((int)n + k - 1) / k + 1))) return rc;
    // restore the scalar state, keep the new p
    for (auto& s : saved) s.pbuf = 0;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_cs, saved.data(), sizeof(ChanState) * saved.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    if ((rc = upload_descs(ctx))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if ((rc = check_launch(ctx))) return rc;
    return drain(ctx);
}

int stm_solver_variant(stm_ctx* ctx, int32_t variant, int32_t* current, int32_t* sweeps,
                       int32_t* tile_x, int32_t* tile_y) {
    int rc = ready(ctx);
    if (rc) return rc;
    if (variant >= N_TB_VARIANTS)
        return fail(ctx, STM_ERR_INVALID_ARG, "variant %d out of range (%d)", variant,
                    N_TB_VARIANTS);
    if (variant >= 0) {
        if (!variant_allowed(ctx, variant))
            return fail(ctx, STM_ERR_INVALID_ARG,
                        "variant %d does %d sweeps per check, params ask for %d", variant,
                        TB_VARIANTS[variant].k, ctx->chans[0].prm.check_every);
        if ((rc = plan_tb(ctx, variant))) return rc;
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    const TbVariant& v = TB_VARIANTS[ctx->variant];
    if (current) *current = ctx->variant;
    if (sweeps) *sweeps = v.k;
    if (tile_x) *tile_x = v.oxm;
    if (tile_y) *tile_y = v.oym;
    return STM_OK;
}

// Times `launches` blocked launches of every permitted variant on the
// current fields (which are restored afterwards) and keeps the fastest per
// sweep.
int stm_autotune(stm_ctx* ctx, uint32_t launches, int32_t* chosen, double* us_per_sweep) {
    int rc = ready(ctx);
    if (rc) return rc;
    if (!ctx->any_rb) return fail(ctx, STM_ERR_STATE, "no red-black channel to tune");
    if (launches == 0) launches = 8;
    const size_t bytes = sizeof(double) * (size_t)ctx->total;
    double* backup = nullptr;
    HIPCHK(ctx, hipMalloc(&backup, bytes));
    HIPCHK(ctx, hipMemcpyAsync(backup, ctx->p, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    hipEvent_t a, b;
    HIPCHK(ctx, hipEventCreate(&a));
    HIPCHK(ctx, hipEventCreate(&b));
    std::vector<ChanState> saved;
    int best = -1;
    double best_us = 1e300;
    for (int v = 0; v < N_TB_VARIANTS && rc == STM_OK; v++) {
        if (!variant_allowed(ctx, v)) continue;
        if ((rc = plan_tb(ctx, v))) break;
        if ((rc = arm_channels(ctx, 1 << 30, true, saved.empty() ? &saved : nullptr))) break;
        int L = 0;
        for (; L < 2; L++) launch_blocked(ctx, v, L, TB_CHECK);
        hipEventRecord(a, ctx->stream);
        for (uint32_t l = 0; l < launches; l++, L++) launch_blocked(ctx, v, L, TB_CHECK);
        hipEventRecord(b, ctx->stream);
        if (hipEventSynchronize(b) != hipSuccess || (rc = check_launch(ctx))) {
            if (!rc) rc = fail(ctx, STM_ERR_HIP, "autotune launch failed");
            break;
        }
        float ms = 0.f;
        hipEventElapsedTime(&ms, a, b);
        const double us = ms * 1e3 / ((double)launches * TB_VARIANTS[v].k);
        if (us < best_us) {
            best_us = us;
            best = v;
        }
        hipMemcpyAsync(ctx->p, backup, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (!saved.empty()) {
        for (auto& s : saved) s.pbuf = 0;
        hipMemcpyAsync(ctx->d_cs, saved.data(), sizeof(ChanState) * saved.size(),
                       hipMemcpyHostToDevice, ctx->stream);
    }
    hipStreamSynchronize(ctx->stream);
    hipFree(backup);
    if (rc) return rc;
    if ((rc = plan_tb(ctx, best))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (chosen) *chosen = best;
    if (us_per_sweep) *us_per_sweep = best_us;
    return STM_OK;
}

// One real tick per permitted variant, scored by interior cell-sweeps per
// second of wall time over the whole tick; keeps the best.  Unlike
// stm_autotune this sees every phase of a solve, including the ones where
// few channels of an uneven batch are still running.
int stm_autotune_ticks(stm_ctx* ctx, int32_t* chosen, double* cell_sweeps_per_s) {
    int rc = ready(ctx);
    if (rc) return rc;
    if (!ctx->any_rb) return fail(ctx, STM_ERR_STATE, "no red-black channel to tune");
    std::vector<uint32_t> iters(nch(ctx));
    int best = -1;
    double best_rate = -1.0;
    for (int v = 0; v < N_TB_VARIANTS; v++) {
        if (!variant_allowed(ctx, v)) continue;
        if ((rc = plan_tb(ctx, v))) return rc;
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        const auto t0 = std::chrono::steady_clock::now();
        if ((rc = stm_run_simulation_tick(ctx, iters.data(), nullptr))) return rc;
        const double sec =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        double work = 0.0;
        for (int k = 0; k < nch(ctx); k++) {
            const HostChan& c = ctx->chans[k];
            work += (double)(c.nx - 2) * (double)(c.ny - 2) * iters[k];
        }
        const double rate = sec > 0.0 ? work / sec : 0.0;
        if (rate > best_rate) {
            best_rate = rate;
            best = v;
        }
    }
    if (best < 0) return fail(ctx, STM_ERR_STATE, "no permitted variant");
    if ((rc = plan_tb(ctx, best))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (chosen) *chosen = best;
    if (cell_sweeps_per_s) *cell_sweeps_per_s = best_rate;
    return STM_OK;
}

int stm_stream_copy(stm_ctx* ctx, uint64_t bytes, uint32_t reps, double* gbs) {
    if (!ctx || !gbs || bytes < 16 || reps == 0)
        return fail(ctx, STM_ERR_INVALID_ARG, "invalid arguments");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    bytes &= ~(uint64_t)15;
    double *a = nullptr, *b = nullptr;
    HIPCHK(ctx, hipMalloc(&a, bytes));
    if (hipMalloc(&b, bytes) != hipSuccess) {
        hipFree(a);
        return fail(ctx, STM_ERR_HIP, "hipMalloc of %llu bytes failed", (unsigned long long)bytes);
    }
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    launch_fill(ctx->stream, (int64_t)(bytes / 8), a, 1.0);
    launch_copy(ctx->stream, (int64_t)bytes, a, b);   // warm-up
    hipEventRecord(e0, ctx->stream);
    for (uint32_t r = 0; r < reps; r++) launch_copy(ctx->stream, (int64_t)bytes, a, b);
    hipEventRecord(e1, ctx->stream);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    if (e != hipSuccess) return fail(ctx, STM_ERR_HIP, "copy probe: %s", hipGetErrorString(e));
    *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    return check_launch(ctx);
}

int stm_sync_benchmark(stm_ctx* ctx, uint32_t steps, double* us) {
    if (!ctx || !us || steps < 10 || steps > 100000)
        return fail(ctx, STM_ERR_INVALID_ARG, "invalid arguments");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    constexpr int SIDE = 16, NWG = SIDE * SIDE, TILE = 64 * 64, GRAPH_NODES = 200;
    constexpr unsigned SPIN = 20000;   // flag polls per wait before giving up
    {
        const hipError_t e = sync_coresident(NWG, ctx->device);
        if (e == hipErrorCooperativeLaunchTooLarge)
            return fail(ctx, STM_ERR_STATE, "the device cannot hold %d workgroups of %d threads "
                        "at once", NWG, SYNC_THREADS);
        HIPCHK(ctx, e);
    }
    int khz = 0;
    HIPCHK(ctx, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
    const int64_t hold = (int64_t)khz / 100;   // 10 us of wall-clock ticks

    double* out = nullptr;
    unsigned* flags = nullptr;   // NWG flags, then the error word
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto release = [&]() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        if (out) hipFree(out);
        if (flags) hipFree(flags);
    };
    hipError_t e = hipMalloc(&out, sizeof(double) * NWG * TILE);
    if (e == hipSuccess) e = hipMalloc(&flags, sizeof(unsigned) * (NWG + 1));
    if (e == hipSuccess) e = hipMemsetAsync(flags, 0, sizeof(unsigned) * (NWG + 1), ctx->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // microseconds per step of the work queued by `work` (timed between two events)
    auto timed = [&](auto work, double per) -> double {
        if (e != hipSuccess) return 0.0;
        e = hipEventRecord(e0, ctx->stream);
        if (e == hipSuccess) e = work();
        if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        return ms * 1e3 / per;
    };
    const int reps = (int)((steps + GRAPH_NODES - 1) / GRAPH_NODES);
    for (int kind = 0; kind < 2 && e == hipSuccess; kind++) {   // tile write, 10-us hold
        auto one = [&](int i) {
            if (kind == 0) launch_sync_tile(ctx->stream, NWG, out, i);
            else launch_sync_hold(ctx->stream, NWG, out, hold);
        };
        for (int i = 0; i < 10; i++) one(i);
        const double direct = timed([&]() {
            for (uint32_t i = 0; i < steps; i++) one((int)i);
            return hipGetLastError();
        }, steps);
        if (e == hipSuccess) e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            for (int i = 0; i < GRAPH_NODES; i++) one(i);
            e = hipStreamEndCapture(ctx->stream, &graph);
        }
        if (e == hipSuccess) e = hipGraphInstantiateWithFlags(&exec, graph, 0);
        if (e == hipSuccess) e = hipGraphLaunch(exec, ctx->stream);
        const double replay = timed([&]() {
            hipError_t r = hipSuccess;
            for (int i = 0; i < reps && r == hipSuccess; i++) r = hipGraphLaunch(exec, ctx->stream);
            return r;
        }, (double)reps * GRAPH_NODES);
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        us[kind == 0 ? 0 : 6] = direct;
        us[kind == 0 ? 1 : 7] = replay;
    }
    for (int write = 1; write >= 0 && e == hipSuccess; write--) {
        e = launch_sync_grid(ctx->stream, NWG, out, 10, write);   // the first launch sets up
        us[write ? 2 : 4] = timed([&]() {
            return launch_sync_grid(ctx->stream, NWG, out, (int)steps, write);
        }, steps);
        for (int pass = 0; pass < 2 && e == hipSuccess; pass++) {
            e = hipMemsetAsync(flags, 0, sizeof(unsigned) * NWG, ctx->stream);
            const int n = pass == 0 ? 10 : (int)steps;
            const double t = timed([&]() {
                return launch_sync_neighbours(ctx->stream, SIDE, out, flags, n, write, SPIN,
                                              flags + NWG);
            }, n);
            if (pass == 1) us[write ? 3 : 5] = t;
        }
    }
    unsigned spun = 0;
    if (e == hipSuccess)
        e = hipMemcpy(&spun, flags + NWG, sizeof(unsigned), hipMemcpyDeviceToHost);
    release();
    if (e != hipSuccess) return fail(ctx, STM_ERR_HIP, "sync benchmark: %s", hipGetErrorString(e));
    if (spun) return fail(ctx, STM_ERR_STATE, "a neighbour wait reached its poll limit");
    return check_launch(ctx);
}

int stm_solver_bytes(stm_ctx* ctx, double* unique_bytes, double* loaded_bytes) {
    int rc = ready(ctx);
    if (rc) return rc;
    const TbVariant& v = TB_VARIANTS[ctx->variant];
    double unique = 0.0, loaded = 0.0;
    for (const auto& c : ctx->chans) unique += 25.0 * (double)c.nx * (double)c.ny;
    for (const TbTile& t : ctx->h_tb) {
        const HostChan& c = ctx->chans[t.chan];
        const int64_t rx = std::min<int64_t>(t.rx0 + v.ex, c.nx) - std::max(t.rx0, 0);
        const int64_t ry = std::min<int64_t>(t.ry0 + v.ey, c.ny) - std::max(t.ry0, 0);
        loaded += 17.0 * (double)(rx * ry) + 8.0 * (double)t.xn * (double)t.yn;
    }
    if (unique_bytes) *unique_bytes = unique;
    if (loaded_bytes) *loaded_bytes = loaded;
    return STM_OK;
}

int stm_synchronize(stm_ctx* ctx) {
    if (!ctx) return fail(nullptr, STM_ERR_INVALID_ARG, "null context");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    int rc = check_launch(ctx);
    if (rc) return rc;
    return drain(ctx);
}

int stm_profile(stm_ctx* ctx, uint32_t sample) {
    if (!ctx) return fail(nullptr, STM_ERR_INVALID_ARG, "null context");
    ctx->sample = sample;
    return STM_OK;
}

int stm_kernel_stats(stm_ctx* ctx, const char* name, double* total_ms, uint64_t* timed,
                     uint64_t* launches) {
    if (!ctx || !name) return fail(ctx, STM_ERR_INVALID_ARG, "null argument");
    int rc = drain(ctx);
    if (rc) return rc;
    auto it = ctx->stats.find(name);
    const KStat k = it == ctx->stats.end() ? KStat{} : it->second;
    if (total_ms) *total_ms = k.ms;
    if (timed) *timed = k.timed;
    if (launches) *launches = k.launches;
    return STM_OK;
}

int stm_set_learned(stm_ctx* ctx, stm_net* net, uint32_t cycles, uint32_t sweeps, int32_t step,
                    uint32_t coarsen) {
    int rc = ready(ctx);
    if (rc) return rc;
    invalidate_graphs(ctx);
    if (!net || cycles == 0) {
        ctx->lnet = nullptr;
        ctx->lcycles = 0;
        return STM_OK;
    }
    if (sweeps == 0) return fail(ctx, STM_ERR_INVALID_ARG, "sweeps per cycle must be positive");
    if (step != STM_STEP_ENERGY && step != STM_STEP_NETWORK && step != STM_STEP_MOMENTUM &&
        step != STM_STEP_MOMENTUM2)
        return fail(ctx, STM_ERR_INVALID_ARG, "unknown step rule %d", step);
    if ((rc = check_net(ctx, net, coarsen))) return rc;
    if ((rc = learn_buffers(ctx, net->net, false))) return rc;
    ctx->lstep = step;
    ctx->lcoarsen = (int)coarsen;
    ctx->lnet = net;
    ctx->lcycles = (int)cycles;
    ctx->lsweeps = (int)sweeps;
    return STM_OK;
}

int stm_render(stm_ctx* ctx, uint32_t ch, int32_t mode, uint8_t* rgba) {
    int rc = ready(ctx);
    if (rc) return rc;
    if ((rc = check_chan(ctx, ch))) return rc;
    if (!rgba) return fail(ctx, STM_ERR_INVALID_ARG, "null image");
    if (mode != STM_RENDER_PRESSURE && mode != STM_RENDER_SPEED)
        return fail(ctx, STM_ERR_INVALID_ARG, "unknown render mode %d", mode);
    const HostChan& c = ctx->chans[ch];
    const int64_t npix = (int64_t)c.nx * c.ny;
    if (npix > ctx->rgba_capacity) {
        if (ctx->d_rgba) HIPCHK(ctx, hipFree(ctx->d_rgba));
        ctx->d_rgba = nullptr;
        ctx->rgba_capacity = 0;
        HIPCHK(ctx, hipMalloc(&ctx->d_rgba, sizeof(uint32_t) * npix));
        ctx->rgba_capacity = npix;
    }
    LAUNCH("render", launch_render(ctx->stream, ctx->d_cd, ctx->d_cs, (int)ch, npix, ctx->flags,
                                   ctx->u, ctx->v, ctx->p, mode, ctx->d_rgba));
    HIPCHK(ctx, hipMemcpyAsync(rgba, ctx->d_rgba, sizeof(uint32_t) * npix, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return check_launch(ctx);
}

// stm_render without the library's image buffer: the kernel writes into the
// device's static buffer (looked up in stm_create).  No allocation of any
// kind: the context must already hold its fields, and no profiling record
// is kept.
int stm_render_static(stm_ctx* ctx, uint32_t ch, int32_t mode, uint8_t* rgba) {
    if (!ctx) return fail(nullptr, STM_ERR_INVALID_ARG, "null context");
    if (!ctx->finalized)
        return fail(ctx, STM_ERR_STATE, "upload or construct before rendering without allocation");
    int rc = check_chan(ctx, ch);
    if (rc) return rc;
    if (!rgba) return fail(ctx, STM_ERR_INVALID_ARG, "null image");
    if (mode != STM_RENDER_PRESSURE && mode != STM_RENDER_SPEED)
        return fail(ctx, STM_ERR_INVALID_ARG, "unknown render mode %d", mode);
    const HostChan& c = ctx->chans[ch];
    const int64_t npix = (int64_t)c.nx * c.ny;
    if (npix > RENDER_STATIC_PIXELS)
        return fail(ctx, STM_ERR_INVALID_ARG,
                    "channel %u has %lld cells, more than the static image buffer's %lld",
                    ch, (long long)npix, (long long)RENDER_STATIC_PIXELS);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    uint32_t* buf = render_static_buffer(ctx->device);
    if (!buf) return fail(ctx, STM_ERR_STATE, "static image buffer not set up on device %d",
                          ctx->device);
    launch_render(ctx->stream, ctx->d_cd, ctx->d_cs, (int)ch, npix, ctx->flags, ctx->u, ctx->v,
                  ctx->p, mode, buf);
    HIPCHK(ctx, hipMemcpyAsync(rgba, buf, sizeof(uint32_t) * npix, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return check_launch(ctx);
}

int stm_set_graphs(stm_ctx* ctx, int32_t on) {
    int rc = ready(ctx);
    if (rc) return rc;
    invalidate_graphs(ctx);
    ctx->use_graphs = on != 0;
    return STM_OK;
}

int stm_set_capture(stm_ctx* ctx, stm_net* net, uint32_t points, uint32_t sweeps,
                    uint32_t coarsen) {
    int rc = ready(ctx);
    if (rc) return rc;
    invalidate_graphs(ctx);
    if (!net) {
        ctx->cnet = nullptr;
        return STM_OK;
    }
    if (sweeps == 0) return fail(ctx, STM_ERR_INVALID_ARG, "sweeps must be positive");
    if (points == 0 || points > (uint32_t)LEARN_POINTS)
        return fail(ctx, STM_ERR_INVALID_ARG, "points must be 1..%d", LEARN_POINTS);
    if ((rc = check_net(ctx, net, coarsen))) return rc;
    if (net->net.dcap < (int)points * nch(ctx))
        return fail(ctx, STM_ERR_INVALID_ARG,
                    "training set of %d samples cannot take the %d pairs of one solve",
                    net->net.dcap, (int)points * nch(ctx));
    if ((rc = learn_buffers(ctx, net->net, true))) return rc;
    ctx->cnet = net;
    ctx->csweeps = (int)sweeps;
    ctx->ccoarsen = (int)coarsen;
    ctx->cpoints = (int)points;
    return STM_OK;
}

int stm_reset_stats(stm_ctx* ctx) {
    if (!ctx) return fail(nullptr, STM_ERR_INVALID_ARG, "null context");
    int rc = drain(ctx);
    if (rc) return rc;
    ctx->stats.clear();
    return STM_OK;
}

}  // extern "C"

namespace stm {
hipStream_t ctx_stream(stm_ctx* ctx) { return ctx->stream; }
int ctx_device(stm_ctx* ctx) { return ctx->device; }

void ctx_attach_net(stm_ctx* ctx, stm_net* net) { ctx->nets.push_back(net); }

// A net is being destroyed: the context stops using it (learned solve and
// capture off if they used it) and drops graphs that captured its buffers.
void ctx_detach_net(stm_ctx* ctx, stm_net* net) {
    ctx->nets.erase(std::remove(ctx->nets.begin(), ctx->nets.end(), net), ctx->nets.end());
    if (ctx->lnet == net) {
        ctx->lnet = nullptr;
        ctx->lcycles = 0;
    }
    if (ctx->cnet == net) ctx->cnet = nullptr;
    invalidate_graphs(ctx);
}
}  // namespace stm
