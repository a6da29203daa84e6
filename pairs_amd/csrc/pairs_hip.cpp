// C ABI of the LJ molecular-dynamics hot path (include/pairs_hip.h).
//
// The context replaces the reference's PairsSimulation object
// (runtime/pairs.hpp:27-297): it owns every device array, keeps counts on the
// device so that rebuild kernels chain without host round trips, and replays
// hipGraphs of the fused per-step kernels between neighbour rebuilds.
// Device storage is in cell-ordered physical slots (md_kernels.hpp); every
// upload and download goes through the reference-index relabeling, so the
// ABI speaks the reference's particle indices only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pairs_hip.h"
#include "../../include/pairs_hip_batch.h"
#include "../../include/pairs_hip_mlp.h"
#include "lj_ctx.hpp"
#include "mlp_model.hpp"

// Scratch of pairs_hip_energy_noalloc: the per-wave energy and virial
// partials (rows 1 and 2 of the thermo layout, NOALLOC_WAVES each) and the
// reduced row.  One copy per GPU, held until the process ends; host code
// uses only the device address resolved per device (g_noalloc_ptr), never
// the symbol itself.
constexpr int NOALLOC_WAVES = 16384;  // locals <= 64 * NOALLOC_WAVES = 1048576
constexpr int NOALLOC_DOUBLES = 3 * NOALLOC_WAVES + 8;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static double *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {nullptr};

// resolve this device's copy of the scratch array (once per device, with the
// device current); 0 or PAIRS_HIP_ERR_HIP
static int noalloc_resolve(int dev, std::string &err) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        err = "device ordinal beyond the scratch table";
        return PAIRS_HIP_ERR_HIP;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[dev]) return 0;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return PAIRS_HIP_ERR_HIP;
    }
    g_noalloc_ptr[dev] = (double *) ptr;
    return 0;
}

// An executable graph is destroyed only once the stream has drained: runs
// enqueue work ahead of the host (rebuild counters are checked a rebuild
// later), so a replay of it may still be pending.
static void destroy_exec(pairs_hip_ctx *ctx, hipGraphExec_t &g) {
    if (!g) return;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    hipGraphExecDestroy(g);
    g = nullptr;
}

static void drop_reb_graphs(pairs_hip_ctx *ctx) {
    for (int p = 0; p < 2; p++) destroy_exec(ctx, ctx->reb_graph[p]);
}

static void drop_graphs(pairs_hip_ctx *ctx) {
    for (int p = 0; p < 2; p++) {
        for (int k = 0; k < pairs_hip_ctx::CHAIN_MAX; k++) destroy_exec(ctx, ctx->step_graph[p][k]);
        destroy_exec(ctx, ctx->start_graph[p]);
    }
    drop_reb_graphs(ctx);
}

template <typename T>
static int dalloc(pairs_hip_ctx *ctx, T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    HIP_OK(hipMalloc((void **) p, n * sizeof(T)));
    return 0;
}

static void dfree(void *p) {
    if (p) hipFree(p);
}

// rows of the per-wave partials (force waves of 64 locals; blocks of 256 of
// the stand-alone integrator)
// one partial per wave of the reduction kernels: the fewest locals per wave
// is 16 (the MLP list passes of a small system, four lanes per local)
static size_t pstride_of(const DevState &d, const Geom &g) {
    (void) g;
    return (size_t) d.cap / 16 + 8;
}

static void free_particle_arrays(DevState &d) {
    for (int b = 0; b < 2; b++) { dfree(d.X[b]); dfree(d.Y[b]); dfree(d.Z[b]); d.X[b] = d.Y[b] = d.Z[b] = nullptr; }
    for (int b = 0; b < 2; b++) { dfree(d.P[b]); d.P[b] = nullptr; }
    double **dd[] = {&d.vx, &d.vy, &d.vz, &d.fx, &d.fy, &d.fz, &d.mass, &d.alt_vx, &d.alt_vy, &d.alt_vz,
                     &d.alt_mass, &d.partials};
    for (auto p : dd) { dfree(*p); *p = nullptr; }
    int **ii[] = {&d.type, &d.flags, &d.uid, &d.shape, &d.ref, &d.phys, &d.alt_type, &d.alt_flags, &d.alt_uid,
                  &d.alt_shape, &d.alt_ref, &d.alt_cell, &d.numneighs, &d.particle_cell, &d.slot, &d.perm,
                  &d.exchg_flag, &d.tile_cnt, &d.sync_src, &d.sync_mult, &d.img_cnt, &d.img_slot, &d.nrev};
    for (auto p : ii) { dfree(*p); *p = nullptr; }
}

static int alloc_particle_arrays(pairs_hip_ctx *ctx, DevState &d, int cap) {
    int r = 0;
    for (int b = 0; b < 2; b++) { r |= dalloc(ctx, &d.X[b], cap); r |= dalloc(ctx, &d.Y[b], cap); r |= dalloc(ctx, &d.Z[b], cap); }
    r |= dalloc(ctx, &d.P[0], cap);
    r |= dalloc(ctx, &d.P[1], cap);
    double **dd[] = {&d.vx, &d.vy, &d.vz, &d.fx, &d.fy, &d.fz, &d.mass, &d.alt_vx, &d.alt_vy, &d.alt_vz, &d.alt_mass};
    for (auto p : dd) r |= dalloc(ctx, p, cap);
    int **ii[] = {&d.type, &d.flags, &d.uid, &d.shape, &d.ref, &d.phys, &d.alt_type, &d.alt_flags, &d.alt_uid,
                  &d.alt_shape, &d.alt_ref, &d.alt_cell, &d.numneighs, &d.particle_cell, &d.slot, &d.perm,
                  &d.exchg_flag, &d.sync_src, &d.nrev};
    for (auto p : ii) r |= dalloc(ctx, p, cap);
    r |= dalloc(ctx, &d.sync_mult, 3 * (size_t) cap);
    r |= dalloc(ctx, &d.img_cnt, cap);
    r |= dalloc(ctx, &d.img_slot, (size_t) IMG_MAX * cap);
    r |= dalloc(ctx, &d.tile_cnt, 2 * ((size_t) cap / 256 + 2));
    d.cap = cap;
    d.pstride = (int) pstride_of(d, ctx->geom);
    r |= dalloc(ctx, &d.partials, 3 * (size_t) d.pstride);
    if (r) return PAIRS_HIP_ERR_HIP;
    return 0;
}

static int alloc_neighbor_arrays(pairs_hip_ctx *ctx, DevState &d, int ncap) {
    dfree(d.nl);
    d.nl = nullptr;
    d.nl_stride = d.cap;
    // half lists: as many rows again for the mirror parts
    const int rows = (ncap + 3) / 4 * 4;
    d.rev0 = ctx->cfg.half ? rows : 0;
    if (dalloc(ctx, &d.nl, (size_t) rows * (ctx->cfg.half ? 2 : 1) * d.nl_stride)) return PAIRS_HIP_ERR_HIP;
    d.ncap = ncap;
    return 0;
}

static int alloc_send_arrays(pairs_hip_ctx *ctx, DevState &d, int send_cap) {
    dfree(d.send_map); dfree(d.send_mult); dfree(d.scratch_i); dfree(d.stage);
    d.send_map = d.send_mult = d.scratch_i = nullptr;
    d.stage = nullptr;
    int r = 0;
    r |= dalloc(ctx, &d.send_map, send_cap);
    r |= dalloc(ctx, &d.send_mult, 3 * (size_t) send_cap);
    r |= dalloc(ctx, &d.scratch_i, 4 * (size_t) send_cap);
    r |= dalloc(ctx, &d.stage, 10 * (size_t) send_cap);
    if (r) return PAIRS_HIP_ERR_HIP;
    d.send_cap = send_cap;
    return 0;
}

static int zero_particle_arrays(pairs_hip_ctx *ctx) {
    DevState &d = ctx->d;
    hipStream_t s = ctx->stream;
    const int n = d.cap;
    for (int *p : {d.flags, d.uid, d.numneighs, d.type, d.alt_flags, d.alt_uid, d.alt_type}) launch_fill_i32(s, p, 0, n);
    launch_fill_i32(s, d.shape, SHAPE_POINT_MASS, n);
    launch_fill_i32(s, d.alt_shape, SHAPE_POINT_MASS, n);
    launch_fill_f64(s, d.mass, 1.0, n);
    launch_fill_f64(s, d.alt_mass, 1.0, n);
    for (double *p : {d.vx, d.vy, d.vz, d.fx, d.fy, d.fz, d.alt_vx, d.alt_vy, d.alt_vz}) launch_fill_f64(s, p, 0.0, n);
    for (int b = 0; b < 2; b++) for (double *p : {d.X[b], d.Y[b], d.Z[b]}) launch_fill_f64(s, p, 0.0, n);
    launch_identity(s, d.ref, n);
    launch_identity(s, d.phys, n);
    CHECK_LAUNCH();
    return 0;
}

// segments: zc cells of one z-column, sized for about home_target (40 by
// default) home particles per force workgroup (overrides: PAIRS_HIP_ZC, PAIRS_HIP_HOME)
static void setup_segments(pairs_hip_ctx *ctx, int zc) {
    Geom &g = ctx->geom;
    if (zc <= 0) {
        const double vol = g.len[0] * g.len[1] * g.len[2];
        const double s3 = g.spacing * g.spacing * g.spacing;
        const double per_cell = ctx->d.nlocal > 0 && vol > 0 ? ctx->d.nlocal * s3 / vol : 16.0;
        zc = (int) std::lround((double) ctx->home_target / std::max(per_cell, 1e-3));
    }
    zc = std::max(1, std::min(zc, std::max(1, g.dim_cells[2] - 1)));
    g.zc = zc;
    g.nchunks = std::max(1, (g.dim_cells[2] - 1 + zc - 1) / zc);
    g.nsegs = g.dim_cells[0] * g.dim_cells[1] * g.nchunks;
}

// build_cell_lists_stencil (src/pairs/sim/cell_lists.py:46-87)
static int setup_geometry(pairs_hip_ctx *ctx) {
    Geom &g = ctx->geom;
    const pairs_hip_md_config &c = ctx->cfg;
    const double s = c.cell_spacing;
    int total = 1;
    for (int d = 0; d < 3; d++) {
        g.lo[d] = c.box_lo[d];
        g.hi[d] = c.box_hi[d];
        g.len[d] = c.box_hi[d] - c.box_lo[d];
        const double dmax = g.hi[d] + s, dmin = g.lo[d] - s;
        g.dim_cells[d] = (int) std::ceil((dmax - dmin) / s) + 1;
        total *= g.dim_cells[d];
    }
    g.spacing = s;
    g.ncells = total + 1;
    g.nstencil = 0;
    for (int a = -1; a < 2; a++)
        for (int b = -1; b < 2; b++)
            for (int e = -1; e < 2; e++) g.stencil[g.nstencil++] = (a * g.dim_cells[1] + b) * g.dim_cells[2] + e;
    setup_segments(ctx, ctx->zc_override);
    return 0;
}

// host-side bookkeeping of a counters snapshot already in ctx->h_counters
static void take_counters(pairs_hip_ctx *ctx) {
    ctx->nghost = ctx->h_counters[C_NGHOST];
    ctx->nsend_all = ctx->h_counters[C_NSEND_ALL];
    ctx->ghost_grid = std::max(1, std::min((ctx->nsend_all + 255) / 256 + 1, 4096));
}

static int read_counters(pairs_hip_ctx *ctx) {
    HIP_OK(hipMemcpyAsync(ctx->h_counters, ctx->d.counters, sizeof(int) * C_COUNT, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    take_counters(ctx);
    return 0;
}

// overflow left by the last device work, as a positive return code
static int take_overflow(pairs_hip_ctx *ctx) {
    const int *h = ctx->h_counters;
    int code = 0;
    if (h[C_OVF] == OVF_SEND) { ctx->req[PAIRS_HIP_OVERFLOW_SEND] = h[C_REQ_SEND]; code = PAIRS_HIP_OVERFLOW_SEND; }
    else if (h[C_OVF] == OVF_PART) { ctx->req[PAIRS_HIP_OVERFLOW_PARTICLES] = h[C_REQ_PART]; code = PAIRS_HIP_OVERFLOW_PARTICLES; }
    else if (h[C_OVF] == OVF_TILE) { code = PAIRS_HIP_OVERFLOW_CELLS; ctx->req[PAIRS_HIP_OVERFLOW_CELLS] = h[C_TILE_MAX]; }
    else if (h[C_REQ_NEIGH] > 0) { ctx->req[PAIRS_HIP_OVERFLOW_NEIGHBORS] = h[C_REQ_NEIGH]; code = PAIRS_HIP_OVERFLOW_NEIGHBORS; }
    if (code) {
        launch_set_counters(ctx->stream, ctx->d.counters, C_OVF, 1, 0);
        launch_set_counters(ctx->stream, ctx->d.counters, C_REQ_NEIGH, 3, 0);
    }
    return code;
}

static int download_raw(pairs_hip_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return 0;
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// physical slot of every reference index in [0, n)
static int host_phys(pairs_hip_ctx *ctx, int n, std::vector<int> &phys) {
    phys.resize(n);
    return download_raw(ctx, phys.data(), ctx->d.phys, n * sizeof(int));
}

template <typename T>
static int ref_order(pairs_hip_ctx *ctx, const T *dev, int n, const std::vector<int> &phys, T *out, int stride,
                     int comp) {
    if (n <= 0) return 0;
    int maxp = 0;
    for (int r = 0; r < n; r++) maxp = std::max(maxp, phys[r]);
    std::vector<T> tmp(maxp + 1);
    if (download_raw(ctx, tmp.data(), dev, (maxp + 1) * sizeof(T))) return PAIRS_HIP_ERR_HIP;
    for (int r = 0; r < n; r++) out[(size_t) r * stride + comp] = tmp[phys[r]];
    return 0;
}

extern "C" {

const char *pairs_hip_last_error(const pairs_hip_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pairs_hip_last_required(const pairs_hip_ctx *ctx, int which) {
    if (!ctx || which < 0 || which > 4) return -1;
    return ctx->req[which];
}

int pairs_hip_create(pairs_hip_ctx **out, const pairs_hip_md_config *cfg) {
    if (!out || !cfg) return PAIRS_HIP_ERR_ARG;
    *out = nullptr;
    if (cfg->particle_capacity <= 0 || cfg->neighbor_capacity <= 2 || cfg->send_capacity <= 0 ||
        cfg->ncells_capacity <= 0 || cfg->ntypes < 1 || cfg->ntypes > 4 || cfg->cell_spacing <= 0.0)
        return PAIRS_HIP_ERR_ARG;
    pairs_hip_ctx *ctx = new pairs_hip_ctx();
    ctx->cfg = *cfg;
    ctx->cfg.epsilon = nullptr;
    ctx->cfg.sigma6 = nullptr;
    LJParams &p = ctx->prm;
    p.rc2 = cfg->cutoff * cfg->cutoff;
    p.dt = cfg->dt;
    p.hdt = cfg->dt * 0.5;
    p.ntypes = cfg->ntypes;
    for (int k = 0; k < 16; k++) {
        p.eps[k] = (k < cfg->ntypes * cfg->ntypes && cfg->epsilon) ? cfg->epsilon[k] : 1.0;
        p.sig6[k] = (k < cfg->ntypes * cfg->ntypes && cfg->sigma6) ? cfg->sigma6[k] : 1.0;
    }
    if (const char *z = getenv("PAIRS_HIP_ZC")) ctx->zc_override = atoi(z);
    memset(&ctx->d, 0, sizeof(DevState));
    auto fail = [&](int code) { pairs_hip_destroy(ctx); return code; };
    if (cfg->stream) {
        ctx->stream = (hipStream_t) cfg->stream;
    } else {
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return fail(PAIRS_HIP_ERR_HIP);
        ctx->own_stream = true;
    }
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(PAIRS_HIP_ERR_HIP);
    ctx->device = dev;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
        ctx->max_lds = lds;
    if (noalloc_resolve(dev, ctx->err)) return fail(PAIRS_HIP_ERR_HIP);
    DevState &d = ctx->d;
    d.max_lds = ctx->max_lds;
    if (const char *e = getenv("PAIRS_HIP_HOME")) ctx->home_target = std::max(8, atoi(e));
    setup_geometry(ctx);
    // the cell grid is fixed by the box; its arrays are sized to fit it
    // (the generated code grows ncells_capacity to twice the need on overflow)
    d.ncells_cap = std::max(cfg->ncells_capacity, ctx->geom.ncells + 2);
    if (alloc_particle_arrays(ctx, d, cfg->particle_capacity)) return fail(PAIRS_HIP_ERR_HIP);
    if (alloc_neighbor_arrays(ctx, d, cfg->neighbor_capacity)) return fail(PAIRS_HIP_ERR_HIP);
    if (alloc_send_arrays(ctx, d, cfg->send_capacity)) return fail(PAIRS_HIP_ERR_HIP);
    if (dalloc(ctx, &d.lcount, (size_t) d.ncells_cap + 1) || dalloc(ctx, &d.gcount, (size_t) d.ncells_cap + 1) ||
        dalloc(ctx, &d.lstart, (size_t) d.ncells_cap + 1) || dalloc(ctx, &d.gstart, (size_t) d.ncells_cap + 1) ||
        dalloc(ctx, &d.counters, C_COUNT) || dalloc(ctx, &d.thermo, 5 * 4096))
        return fail(PAIRS_HIP_ERR_HIP);
    launch_fill_i32(ctx->stream, d.counters, 0, C_COUNT);
    for (int *q : {d.lcount, d.gcount, d.lstart, d.gstart}) launch_fill_i32(ctx->stream, q, 0, d.ncells_cap + 1);
    if (zero_particle_arrays(ctx)) return fail(PAIRS_HIP_ERR_HIP);
    d.cur = 0;
    d.tile_cap = 64;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(PAIRS_HIP_ERR_HIP);
    *out = ctx;
    return 0;
}

int pairs_hip_destroy(pairs_hip_ctx *ctx) {
    if (!ctx) return 0;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    drop_graphs(ctx);
    for (auto e : ctx->ev_begin) hipEventDestroy(e);
    for (auto e : ctx->ev_end) hipEventDestroy(e);
    DevState &d = ctx->d;
    free_particle_arrays(d);
    dfree(d.nl); dfree(d.send_map); dfree(d.send_mult); dfree(d.scratch_i); dfree(d.stage);
    dfree(d.lcount); dfree(d.gcount); dfree(d.lstart); dfree(d.gstart); dfree(d.counters); dfree(d.thermo);
    dfree(ctx->d_eatom); dfree(ctx->d_expand);
    dfree(ctx->mlp_G); dfree(ctx->mlp_D); dfree(ctx->mlp_E); dfree(ctx->mlp_f); dfree(ctx->mlp_red);
    dfree(ctx->mlp_rho); dfree(ctx->mlp_U);
    dfree(ctx->ckpt);
    if (ctx->h_spec) hipHostFree(ctx->h_spec);
    if (ctx->spec_ev) hipEventDestroy(ctx->spec_ev);
    if (ctx->reb_ev) hipEventDestroy(ctx->reb_ev);
    if (ctx->side) { hipStreamSynchronize(ctx->side); hipStreamDestroy(ctx->side); }
    if (ctx->own_stream && ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

// Download the locals (reference order) needed to rebuild state elsewhere.
struct HostLocals {
    int n = 0;
    std::vector<double> x, v, m;
    std::vector<int> t, f, u;
};

static int fetch_locals(pairs_hip_ctx *ctx, HostLocals &h) {
    DevState &d = ctx->d;
    if (read_counters(ctx)) return PAIRS_HIP_ERR_HIP;
    const int n = ctx->h_counters[C_NLOCAL];
    h.n = n;
    std::vector<int> phys;
    if (host_phys(ctx, n, phys)) return PAIRS_HIP_ERR_HIP;
    h.x.resize(3 * (size_t) n); h.v.resize(3 * (size_t) n); h.m.resize(n); h.t.resize(n); h.f.resize(n); h.u.resize(n);
    const int c = d.cur;
    const double *pos[3] = {d.X[c], d.Y[c], d.Z[c]}, *vel[3] = {d.vx, d.vy, d.vz};
    for (int k = 0; k < 3; k++) {
        if (ref_order(ctx, pos[k], n, phys, h.x.data(), 3, k)) return PAIRS_HIP_ERR_HIP;
        if (ref_order(ctx, vel[k], n, phys, h.v.data(), 3, k)) return PAIRS_HIP_ERR_HIP;
    }
    if (ref_order(ctx, (const double *) d.mass, n, phys, h.m.data(), 1, 0) ||
        ref_order(ctx, (const int *) d.type, n, phys, h.t.data(), 1, 0) ||
        ref_order(ctx, (const int *) d.flags, n, phys, h.f.data(), 1, 0) ||
        ref_order(ctx, (const int *) d.uid, n, phys, h.u.data(), 1, 0))
        return PAIRS_HIP_ERR_HIP;
    return 0;
}

int pairs_hip_reserve(pairs_hip_ctx *ctx, int particle_capacity, int neighbor_capacity, int send_capacity) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    DevState &d = ctx->d;
    if (particle_capacity > d.cap) {
        // grow by re-uploading the locals in reference order (ghosts, cells and
        // lists are rebuilt by the next rebuild anyway)
        HostLocals h;
        if (fetch_locals(ctx, h)) return PAIRS_HIP_ERR_HIP;
        free_particle_arrays(d);
        if (alloc_particle_arrays(ctx, d, particle_capacity)) return PAIRS_HIP_ERR_HIP;
        if (zero_particle_arrays(ctx)) return PAIRS_HIP_ERR_HIP;
        if (alloc_neighbor_arrays(ctx, d, std::max(neighbor_capacity, d.ncap))) return PAIRS_HIP_ERR_HIP;
        int r = pairs_hip_upload(ctx, h.n, h.x.data(), h.v.data(), h.m.data(), h.t.data(), h.f.data(), h.u.data());
        if (r) return r;
    } else if (neighbor_capacity > d.ncap) {
        if (alloc_neighbor_arrays(ctx, d, neighbor_capacity)) return PAIRS_HIP_ERR_HIP;
        launch_fill_i32(ctx->stream, d.numneighs, 0, d.cap);
    }
    if (send_capacity > d.send_cap) {
        if (alloc_send_arrays(ctx, d, send_capacity)) return PAIRS_HIP_ERR_HIP;
        // the ghost maps are gone with the old arrays: a rebuild must follow
        launch_set_counters(ctx->stream, d.counters, C_NSEND_ALL, 1, 0);
        launch_set_counters(ctx->stream, d.counters, C_NGHOST, 1, 0);
    }
    CHECK_LAUNCH();
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int pairs_hip_upload(pairs_hip_ctx *ctx, int n, const double *x, const double *v, const double *m, const int *type,
                     const int *flags, const int *uid) {
    if (ctx) ctx->xy_valid[0] = ctx->xy_valid[1] = false;
    if (!ctx || n < 0 || !x || !v) return PAIRS_HIP_ERR_ARG;
    DevState &d = ctx->d;
    if (n > d.cap) {
        ctx->req[PAIRS_HIP_OVERFLOW_PARTICLES] = n;
        return PAIRS_HIP_OVERFLOW_PARTICLES;
    }
    drop_graphs(ctx);
    std::vector<double> hx[3], hv[3];
    for (int k = 0; k < 3; k++) { hx[k].resize(n); hv[k].resize(n); }
    std::vector<double> hm(n);
    std::vector<int> ht(n), hf(n), hu(n), hs(n, SHAPE_POINT_MASS);
    for (int i = 0; i < n; i++) {
        const int t = type ? type[i] : 0;
        if (t < 0 || t >= ctx->cfg.ntypes) { ctx->err = "type out of range"; return PAIRS_HIP_ERR_ARG; }
        for (int k = 0; k < 3; k++) { hx[k][i] = x[3 * i + k]; hv[k][i] = v[3 * i + k]; }
        hm[i] = m ? m[i] : 1.0;
        ht[i] = t;
        hf[i] = flags ? flags[i] : 0;
        hu[i] = uid ? uid[i] : 0;
    }
    d.cur = 0;
    if (n > 0) {
        for (int b = 0; b < 2; b++) {
            HIP_OK(hipMemcpyAsync(d.X[b], hx[0].data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            HIP_OK(hipMemcpyAsync(d.Y[b], hx[1].data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            HIP_OK(hipMemcpyAsync(d.Z[b], hx[2].data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        }
        HIP_OK(hipMemcpyAsync(d.vx, hv[0].data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipMemcpyAsync(d.vy, hv[1].data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipMemcpyAsync(d.vz, hv[2].data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipMemcpyAsync(d.mass, hm.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipMemcpyAsync(d.type, ht.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipMemcpyAsync(d.flags, hf.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipMemcpyAsync(d.uid, hu.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipMemcpyAsync(d.shape, hs.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    }
    launch_identity(ctx->stream, d.ref, d.cap);
    launch_identity(ctx->stream, d.phys, d.cap);
    launch_fill_f64(ctx->stream, d.fx, 0.0, d.cap);
    launch_fill_f64(ctx->stream, d.fy, 0.0, d.cap);
    launch_fill_f64(ctx->stream, d.fz, 0.0, d.cap);
    launch_fill_i32(ctx->stream, d.counters, 0, C_COUNT);
    launch_set_counters(ctx->stream, d.counters, C_NLOCAL, 1, n);
    CHECK_LAUNCH();
    HIP_OK(hipStreamSynchronize(ctx->stream));
    d.nlocal = n;
    ctx->nghost = 0;
    ctx->nsend_all = 0;
    setup_segments(ctx, ctx->zc_override);
    return 0;
}

// Park-Miller minimal standard generator without masking (runtime/copper_fcc_lattice.hpp:11-26)
static double park_miller(int *seed) {
    const int IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836;
    const int k = (*seed) / IQ;
    *seed = IA * (*seed - k * IQ) - IR * k;
    if (*seed < 0) *seed += IM;
    return (1.0 / IM) * (*seed);
}

// pairs::copper_fcc_lattice (runtime/copper_fcc_lattice.hpp:64-145), on the host
int pairs_hip_copper_fcc_lattice(pairs_hip_ctx *ctx, int nx, int ny, int nz, double rho, int ntypes) {
    if (ctx) ctx->xy_valid[0] = ctx->xy_valid[1] = false;
    if (!ctx || nx <= 0 || ny <= 0 || nz <= 0 || rho <= 0.0 || ntypes < 1) return PAIRS_HIP_ERR_ARG;
    const double *lo = ctx->cfg.box_lo, *hi = ctx->cfg.box_hi;
    const double xprd = hi[0] - lo[0], yprd = hi[1] - lo[1], zprd = hi[2] - lo[2];
    const double alat = std::pow((4.0 / rho), (1.0 / 3.0));
    int ilo = (int) (0.0 / (0.5 * alat) - 1), ihi = (int) (xprd / (0.5 * alat) + 1);
    int jlo = (int) (0.0 / (0.5 * alat) - 1), jhi = (int) (yprd / (0.5 * alat) + 1);
    int klo = (int) (0.0 / (0.5 * alat) - 1), khi = (int) (zprd / (0.5 * alat) + 1);
    ilo = std::max(ilo, 0); ihi = std::min(ihi, 2 * nx - 1);
    jlo = std::max(jlo, 0); jhi = std::min(jhi, 2 * ny - 1);
    klo = std::max(klo, 0); khi = std::min(khi, 2 * nz - 1);
    const double SMALL = 1e-5;
    std::vector<double> x, v, m;
    std::vector<int> t;
    int sx = 0, sy = 0, sz = 0, ox = 0, oy = 0, oz = 0;
    const int sub = 8;
    while (oz * sub <= khi) {
        const int k = oz * sub + sz, j = oy * sub + sy, i = ox * sub + sx;
        if (((i + j + k) % 2 == 0) && i >= ilo && i <= ihi && j >= jlo && j <= jhi && k >= klo && k <= khi) {
            const double xt = 0.5 * alat * i, yt = 0.5 * alat * j, zt = 0.5 * alat * k;
            if (xt >= lo[0] && xt < hi[0] - SMALL && yt >= lo[1] && yt < hi[1] - SMALL && zt >= lo[2] &&
                zt < hi[2] - SMALL) {
                int n = k * (2 * ny) * (2 * nx) + j * (2 * nx) + i + 1;
                for (int q = 0; q < 5; q++) park_miller(&n);
                const double vx = park_miller(&n);
                for (int q = 0; q < 5; q++) park_miller(&n);
                const double vy = park_miller(&n);
                for (int q = 0; q < 5; q++) park_miller(&n);
                const double vz = park_miller(&n);
                x.insert(x.end(), {xt, yt, zt});
                v.insert(v.end(), {vx, vy, vz});
                m.push_back(1.0);
                t.push_back(rand() % ntypes);
            }
        }
        sx++;
        if (sx == sub) { sx = 0; sy++; }
        if (sy == sub) { sy = 0; sz++; }
        if (sz == sub) { sz = 0; ox++; }
        if (ox * sub > ihi) { ox = 0; oy++; }
        if (oy * sub > jhi) { oy = 0; oz++; }
    }
    const int n = (int) m.size();
    const int r = pairs_hip_upload(ctx, n, x.data(), v.data(), m.data(), t.data(), nullptr, nullptr);
    return r ? r : n;
}

// pairs::adjust_thermo (runtime/thermo.hpp:53-97), on the host in the same order
int pairs_hip_adjust_thermo(pairs_hip_ctx *ctx, double temp) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    DevState &d = ctx->d;
    HostLocals h;
    if (fetch_locals(ctx, h)) return PAIRS_HIP_ERR_HIP;
    const int n = h.n;
    if (n <= 1) return PAIRS_HIP_ERR_STATE;
    std::vector<double> &v = h.v;
    double vxt = 0.0, vyt = 0.0, vzt = 0.0;
    for (int i = 0; i < n; i++) { vxt += v[3 * i]; vyt += v[3 * i + 1]; vzt += v[3 * i + 2]; }
    vxt /= n; vyt /= n; vzt /= n;
    for (int i = 0; i < n; i++) { v[3 * i] -= vxt; v[3 * i + 1] -= vyt; v[3 * i + 2] -= vzt; }
    const double dof = (n * 3 - 3);
    const double t_scale = 1.0 / dof;
    double t = 0.0;
    for (int i = 0; i < n; i++)
        t += h.m[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    t = t * t_scale;
    const double factor = std::sqrt(temp / t);
    for (int i = 0; i < 3 * n; i++) v[i] *= factor;
    std::vector<int> phys;
    if (host_phys(ctx, n, phys)) return PAIRS_HIP_ERR_HIP;
    std::vector<double> comp[3];
    for (int k = 0; k < 3; k++) {
        comp[k].resize(d.cap);
        double *dev = k == 0 ? d.vx : (k == 1 ? d.vy : d.vz);
        if (download_raw(ctx, comp[k].data(), dev, d.cap * sizeof(double))) return PAIRS_HIP_ERR_HIP;
        for (int r = 0; r < n; r++) comp[k][phys[r]] = v[3 * r + k];
        HIP_OK(hipMemcpyAsync(dev, comp[k].data(), d.cap * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int pairs_hip_count(pairs_hip_ctx *ctx, int what) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    if (read_counters(ctx)) return PAIRS_HIP_ERR_HIP;
    switch (what) {
        case PAIRS_HIP_NLOCAL: return ctx->h_counters[C_NLOCAL];
        case PAIRS_HIP_NGHOST: return ctx->h_counters[C_NGHOST];
        case PAIRS_HIP_NCELLS: return ctx->geom.ncells;
        case PAIRS_HIP_NSEND_ALL: return ctx->h_counters[C_NSEND_ALL];
        case PAIRS_HIP_NEIGHBOR_CAPACITY: return ctx->d.ncap;
        case PAIRS_HIP_PARTICLE_CAPACITY: return ctx->d.cap;
        case PAIRS_HIP_SEND_CAPACITY: return ctx->d.send_cap;
        case PAIRS_HIP_TILE_CAPACITY: return ctx->d.tile_cap;
        case PAIRS_HIP_REBUILDS_DEFERRED: return (int) std::min(ctx->spec_count, (long long) INT_MAX);
        case PAIRS_HIP_REBUILD_ROLLBACKS: return (int) std::min(ctx->spec_rollbacks, (long long) INT_MAX);
        case PAIRS_HIP_NEIGHBOR_TOTAL: {
            const int n = ctx->d.nlocal;
            std::vector<int> nn(n);
            if (download_raw(ctx, nn.data(), ctx->d.numneighs, n * sizeof(int))) return PAIRS_HIP_ERR_HIP;
            long long s = 0;
            for (int x : nn) s += x;
            return (int) std::min<long long>(s, 0x7fffffff);
        }
    }
    return PAIRS_HIP_ERR_ARG;
}

int pairs_hip_download(pairs_hip_ctx *ctx, int what, void *host, int count) {
    if (!ctx || !host || count < 0) return PAIRS_HIP_ERR_ARG;
    if (read_counters(ctx)) return PAIRS_HIP_ERR_HIP;
    DevState &d = ctx->d;
    const int nlocal = ctx->h_counters[C_NLOCAL], nall = nlocal + ctx->h_counters[C_NGHOST];
    const int nsend = ctx->h_counters[C_NSEND_ALL];
    auto lim = [&](int avail) { return std::min(count, avail); };
    std::vector<int> phys;
    const int c = d.cur;
    switch (what) {
        case PAIRS_HIP_POSITION:
        case PAIRS_HIP_VELOCITY:
        case PAIRS_HIP_FORCE: {
            const int n = lim(what == PAIRS_HIP_FORCE ? nlocal : nall);
            if (host_phys(ctx, n, phys)) return PAIRS_HIP_ERR_HIP;
            const double *src[3];
            if (what == PAIRS_HIP_POSITION) { src[0] = d.X[c]; src[1] = d.Y[c]; src[2] = d.Z[c]; }
            else if (what == PAIRS_HIP_VELOCITY) { src[0] = d.vx; src[1] = d.vy; src[2] = d.vz; }
            else { src[0] = d.fx; src[1] = d.fy; src[2] = d.fz; }
            for (int k = 0; k < 3; k++)
                if (ref_order(ctx, src[k], n, phys, (double *) host, 3, k)) return PAIRS_HIP_ERR_HIP;
            return n;
        }
        case PAIRS_HIP_MASS: {
            const int n = lim(nall);
            if (host_phys(ctx, n, phys)) return PAIRS_HIP_ERR_HIP;
            return ref_order(ctx, (const double *) d.mass, n, phys, (double *) host, 1, 0) ? PAIRS_HIP_ERR_HIP : n;
        }
        case PAIRS_HIP_TYPE:
        case PAIRS_HIP_FLAGS:
        case PAIRS_HIP_UID:
        case PAIRS_HIP_PARTICLE_CELL: {
            const int n = lim(nall);
            if (host_phys(ctx, n, phys)) return PAIRS_HIP_ERR_HIP;
            const int *src = what == PAIRS_HIP_TYPE ? d.type
                             : what == PAIRS_HIP_FLAGS ? d.flags
                             : what == PAIRS_HIP_UID ? d.uid : d.particle_cell;
            return ref_order(ctx, src, n, phys, (int *) host, 1, 0) ? PAIRS_HIP_ERR_HIP : n;
        }
        case PAIRS_HIP_NUMNEIGHS: {
            const int n = lim(nlocal);
            if (host_phys(ctx, n, phys)) return PAIRS_HIP_ERR_HIP;
            return ref_order(ctx, (const int *) d.numneighs, n, phys, (int *) host, 1, 0) ? PAIRS_HIP_ERR_HIP : n;
        }
        case PAIRS_HIP_NEIGHBORLISTS: {
            const int n = lim(nlocal);
            if (n == 0) return 0;
            const size_t need = (size_t) nlocal * d.ncap;
            if (ctx->expand_cap < need) {
                dfree(ctx->d_expand);
                ctx->d_expand = nullptr;
                if (dalloc(ctx, &ctx->d_expand, need)) return PAIRS_HIP_ERR_HIP;
                ctx->expand_cap = need;
            }
            launch_fill_i32(ctx->stream, ctx->d_expand, -1, (int) need);
            launch_expand_neighbors(ctx->stream, d, ctx->geom, ctx->d_expand, d.ncap);
            CHECK_LAUNCH();
            std::vector<int> rows(need);
            if (download_raw(ctx, rows.data(), ctx->d_expand, need * sizeof(int))) return PAIRS_HIP_ERR_HIP;
            if (host_phys(ctx, n, phys)) return PAIRS_HIP_ERR_HIP;
            int *o = (int *) host;
            for (int r = 0; r < n; r++)
                memcpy(o + (size_t) r * d.ncap, rows.data() + (size_t) phys[r] * d.ncap, d.ncap * sizeof(int));
            return n;
        }
        case PAIRS_HIP_CELL_SIZES:
        case PAIRS_HIP_CELL_PARTICLES: {
            const int nc = ctx->geom.ncells;
            std::vector<int> ls(nc + 1), gs(nc + 1), ref(nall);
            if (download_raw(ctx, ls.data(), d.lstart, (nc + 1) * sizeof(int)) ||
                download_raw(ctx, gs.data(), d.gstart, (nc + 1) * sizeof(int)) ||
                download_raw(ctx, ref.data(), d.ref, nall * sizeof(int)))
                return PAIRS_HIP_ERR_HIP;
            if (what == PAIRS_HIP_CELL_SIZES) {
                const int m = lim(nc);
                for (int cc = 0; cc < m; cc++) ((int *) host)[cc] = (ls[cc + 1] - ls[cc]) + (gs[cc + 1] - gs[cc]);
                return m;
            }
            const int m = lim(nall);
            int k = 0;
            for (int cc = 0; cc < nc && k < m; cc++) {
                for (int p = ls[cc]; p < ls[cc + 1] && k < m; p++) ((int *) host)[k++] = ref[p];
                for (int p = nlocal + gs[cc]; p < nlocal + gs[cc + 1] && k < m; p++) ((int *) host)[k++] = ref[p];
            }
            return k;
        }
        case PAIRS_HIP_SEND_MAP: {
            const int n = lim(nsend);
            return download_raw(ctx, host, d.send_map, n * sizeof(int)) ? PAIRS_HIP_ERR_HIP : n;
        }
        case PAIRS_HIP_SEND_MULT: {
            const int n = lim(nsend);
            return download_raw(ctx, host, d.send_mult, 3 * (size_t) n * sizeof(int)) ? PAIRS_HIP_ERR_HIP : n;
        }
        case PAIRS_HIP_NSEND: {
            const int n = lim(6);
            memcpy(host, ctx->h_counters + C_NSEND, n * sizeof(int));
            return n;
        }
        case PAIRS_HIP_ATOM_ENERGY: {
            const int n = lim(std::min(nlocal, ctx->eatom_cap));
            if (!ctx->d_eatom || n == 0) return 0;
            if (host_phys(ctx, n, phys)) return PAIRS_HIP_ERR_HIP;
            return ref_order(ctx, (const double *) ctx->d_eatom, n, phys, (double *) host, 1, 0) ? PAIRS_HIP_ERR_HIP : n;
        }
    }
    return PAIRS_HIP_ERR_ARG;
}

int pairs_hip_setup_cells(pairs_hip_ctx *ctx) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    return setup_geometry(ctx);
}

static int finish_module(pairs_hip_ctx *ctx) {
    CHECK_LAUNCH();
    if (read_counters(ctx)) return PAIRS_HIP_ERR_HIP;
    return take_overflow(ctx);
}

int pairs_hip_exchange(pairs_hip_ctx *ctx) {
    if (ctx) ctx->xy_valid[0] = ctx->xy_valid[1] = false;
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    drop_graphs(ctx);
    launch_exchange(ctx->stream, ctx->d, ctx->geom, ctx->cfg.periodic);
    launch_set_counters(ctx->stream, ctx->d.counters, C_NGHOST, 1, 0);  // the ghosts are gone (comm.py:105-118)
    return finish_module(ctx);
}

int pairs_hip_borders(pairs_hip_ctx *ctx) {
    if (ctx) ctx->xy_valid[0] = ctx->xy_valid[1] = false;
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    drop_graphs(ctx);
    launch_borders(ctx->stream, ctx->d, ctx->geom, ctx->cfg.periodic);
    return finish_module(ctx);
}

int pairs_hip_sync_ghosts(pairs_hip_ctx *ctx) {
    if (ctx) ctx->xy_valid[0] = ctx->xy_valid[1] = false;
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    launch_sync_inplace(ctx->stream, ctx->d, ctx->geom, ctx->ghost_grid, 0);
    CHECK_LAUNCH();
    return 0;
}

int pairs_hip_bin_cells(pairs_hip_ctx *ctx) {
    if (ctx) ctx->xy_valid[0] = ctx->xy_valid[1] = false;
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    drop_graphs(ctx);
    launch_bin_sort(ctx->stream, ctx->d, ctx->geom);
    return finish_module(ctx);
}

// LDS entries the neighbour build stages per segment tile: the largest tile
// of the last build plus 1/8 and 64 entries of headroom (tiles change only at a
// rebuild), never above what a workgroup's LDS holds; grows only, so captured
// graphs stay valid while tiles fluctuate
static void set_tile_cap(pairs_hip_ctx *ctx) {
    const int tmax = ctx->h_counters[C_TILE_MAX];
    int cap = tmax + tmax / 8 + 64;
    cap = (cap + 63) / 64 * 64;
    const int lim = build_tile_limit(ctx->d, ctx->geom, ctx->cfg.half);
    const int grown = std::min(std::max(cap, ctx->d.tile_cap), lim);
    if (grown != ctx->d.tile_cap) drop_reb_graphs(ctx);  // they stage tile_cap entries
    ctx->d.tile_cap = grown;
}

static void enqueue_build(pairs_hip_ctx *ctx, hipStream_t s, bool reset_tile = true) {
    const double rl = ctx->cfg.cell_spacing;
    if (reset_tile) launch_set_counters(s, ctx->d.counters, C_TILE_MAX, 1, 0);
    launch_build_neighbors(s, ctx->d, ctx->geom, ctx->cfg.half, rl * rl);
}

static void enqueue_build(pairs_hip_ctx *ctx) { enqueue_build(ctx, ctx->stream); }

// a segment's tile outgrew the LDS entries the build staged (counters just
// read): stage more, or cut the segments when a workgroup cannot hold it
static int grow_tile(pairs_hip_ctx *ctx) {
    launch_set_counters(ctx->stream, ctx->d.counters, C_OVF, 1, 0);
    drop_graphs(ctx);
    const int tmax = ctx->h_counters[C_TILE_MAX];
    if (tmax <= build_tile_limit(ctx->d, ctx->geom, ctx->cfg.half) && tmax > ctx->d.tile_cap) {
        set_tile_cap(ctx);
        return 0;
    }
    if (ctx->geom.zc <= 1) { ctx->err = "cell tile exceeds the workgroup's LDS"; return PAIRS_HIP_ERR_STATE; }
    setup_segments(ctx, ctx->geom.zc - 1);
    return 0;
}

// the build enqueued, counters read: repeat the build while its tile grows,
// then report the other overflows
static int finish_build(pairs_hip_ctx *ctx) {
    for (int attempt = 0; attempt < 64; attempt++) {
        if (ctx->h_counters[C_OVF] != OVF_TILE) {
            set_tile_cap(ctx);
            return take_overflow(ctx);
        }
        if (int r = grow_tile(ctx)) return r;
        enqueue_build(ctx);
        CHECK_LAUNCH();
        if (read_counters(ctx)) return PAIRS_HIP_ERR_HIP;
    }
    ctx->err = "neighbour build did not converge";
    return PAIRS_HIP_ERR_STATE;
}

static int build_neighbors(pairs_hip_ctx *ctx) {
    enqueue_build(ctx);
    CHECK_LAUNCH();
    if (read_counters(ctx)) return PAIRS_HIP_ERR_HIP;
    return finish_build(ctx);
}

int pairs_hip_build_neighbors(pairs_hip_ctx *ctx) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    drop_graphs(ctx);
    return build_neighbors(ctx);
}

int pairs_hip_lj(pairs_hip_ctx *ctx) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    DevState &d = ctx->d;
    launch_lj_full(ctx->stream, d, ctx->geom, ctx->prm, d.cur, d.cur, 0, 0, 0, 0, !ctx->xy_valid[d.cur]);
    ctx->xy_valid[d.cur] = true;
    CHECK_LAUNCH();
    return 0;
}

int pairs_hip_vv_initial(pairs_hip_ctx *ctx) {
    if (ctx) ctx->xy_valid[0] = ctx->xy_valid[1] = false;
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    launch_vv_initial(ctx->stream, ctx->d, ctx->prm);
    CHECK_LAUNCH();
    return 0;
}

int pairs_hip_vv_final(pairs_hip_ctx *ctx) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    DevState &d = ctx->d;
    launch_vv_post(ctx->stream, d, ctx->geom, ctx->prm, d.cur, d.cur, 1, 0, 0);
    CHECK_LAUNCH();
    return 0;
}

int pairs_hip_lj_older(pairs_hip_ctx *ctx) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    if (ctx->cfg.half) {
        ctx->err = "the older API's lj module runs on full neighbour lists (cfg.half = 0)";
        return PAIRS_HIP_ERR_STATE;
    }
    launch_lj_older(ctx->stream, ctx->d, ctx->geom, ctx->prm);
    ctx->xy_valid[ctx->d.cur] = true;  // packed by the launch
    CHECK_LAUNCH();
    return 0;
}

int pairs_hip_euler(pairs_hip_ctx *ctx) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    ctx->xy_valid[0] = ctx->xy_valid[1] = false;
    launch_euler_md(ctx->stream, ctx->d, ctx->prm.dt);
    CHECK_LAUNCH();
    return 0;
}

static int ensure_eatom(pairs_hip_ctx *ctx) {
    if (ctx->eatom_cap < ctx->d.cap) {
        dfree(ctx->d_eatom);
        ctx->d_eatom = nullptr;
        if (dalloc(ctx, &ctx->d_eatom, ctx->d.cap)) return PAIRS_HIP_ERR_HIP;
        ctx->eatom_cap = ctx->d.cap;
    }
    return 0;
}

int pairs_hip_energy(pairs_hip_ctx *ctx, double *pe, double *virial) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    if (ensure_eatom(ctx)) return PAIRS_HIP_ERR_HIP;
    double out[5];
    launch_energy(ctx->stream, ctx->d, ctx->geom, ctx->prm, ctx->cfg.half, ctx->d_eatom);
    launch_thermo_reduce(ctx->stream, ctx->d, 0, force_waves(ctx->d), 2, ctx->geom, ctx->d.thermo);
    CHECK_LAUNCH();
    if (download_raw(ctx, out, ctx->d.thermo, 5 * sizeof(double))) return PAIRS_HIP_ERR_HIP;
    if (pe) *pe = out[2];
    if (virial) *virial = out[3];
    return 0;
}

int pairs_hip_energy_noalloc(pairs_hip_ctx *ctx, double *pe, double *virial) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    const int nw = force_waves(ctx->d);
    if (nw > NOALLOC_WAVES) {
        ctx->err = "pairs_hip_energy_noalloc: " + std::to_string(ctx->d.nlocal) +
                   " locals exceed its scratch (at most " + std::to_string(64 * NOALLOC_WAVES) + ")";
        return PAIRS_HIP_ERR_ARG;
    }
    double *ptr = ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES ? g_noalloc_ptr[ctx->device] : nullptr;
    if (!ptr) {
        ctx->err = "pairs_hip_energy_noalloc: the scratch address of this device is not resolved";
        return PAIRS_HIP_ERR_STATE;
    }
    // the same kernels as pairs_hip_energy, without the per-atom energies and
    // with the partials in the scratch array: every partial the reduction
    // reads is written first by the energy pass (one per launched wave), so
    // nothing needs clearing
    DevState d = ctx->d;
    d.partials = ptr;
    d.pstride = NOALLOC_WAVES;
    double *row = ptr + 3 * (size_t) NOALLOC_WAVES;
    launch_energy(ctx->stream, d, ctx->geom, ctx->prm, ctx->cfg.half, nullptr);
    launch_thermo_reduce(ctx->stream, d, 0, nw, 2, ctx->geom, row);
    CHECK_LAUNCH();
    double out[5];
    if (download_raw(ctx, out, row, 5 * sizeof(double))) return PAIRS_HIP_ERR_HIP;
    if (pe) *pe = out[2];
    if (virial) *virial = out[3];
    return 0;
}

int pairs_hip_thermo(pairs_hip_ctx *ctx, double out[5]) {
    if (!ctx || !out) return PAIRS_HIP_ERR_ARG;
    if (ensure_eatom(ctx)) return PAIRS_HIP_ERR_HIP;
    launch_energy(ctx->stream, ctx->d, ctx->geom, ctx->prm, ctx->cfg.half, ctx->d_eatom);
    launch_ke_partial(ctx->stream, ctx->d);
    launch_thermo_reduce(ctx->stream, ctx->d, (ctx->d.nlocal + 255) / 256, force_waves(ctx->d), 3, ctx->geom,
                         ctx->d.thermo);
    CHECK_LAUNCH();
    return download_raw(ctx, out, ctx->d.thermo, 5 * sizeof(double));
}

int pairs_hip_sync(pairs_hip_ctx *ctx) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// per-slot MLP work arrays sized to the particle capacity (re-bound and the
// step graphs dropped when they grow)
static int ensure_mlp_work(pairs_hip_ctx *ctx) {
    const int cap = ctx->d.cap;
    if (ctx->mlp_cap < cap) {
        HIP_OK(hipStreamSynchronize(ctx->stream));
        drop_graphs(ctx);
        dfree(ctx->mlp_G); dfree(ctx->mlp_D); dfree(ctx->mlp_E); dfree(ctx->mlp_f);
        ctx->mlp_G = ctx->mlp_D = ctx->mlp_E = nullptr;
        ctx->mlp_f = nullptr;
        if (dalloc(ctx, &ctx->mlp_G, (size_t) cap * MLP_ND) || dalloc(ctx, &ctx->mlp_D, (size_t) cap * MLP_ND) ||
            dalloc(ctx, &ctx->mlp_E, cap) || dalloc(ctx, &ctx->mlp_f, (size_t) 3 * cap))
            return PAIRS_HIP_ERR_HIP;
        ctx->mlp_cap = cap;
    }
    if (!ctx->mlp_red && dalloc(ctx, &ctx->mlp_red, 2048)) return PAIRS_HIP_ERR_HIP;
    ctx->mlp.G = ctx->mlp_G;
    ctx->mlp.D = ctx->mlp_D;
    ctx->mlp.E = ctx->mlp_E;
    ctx->mlp.cap = ctx->mlp_cap;
    return 0;
}

static inline bool every(int ts, int n) { return n > 0 && (((ts + 1) % n) == 0 || ts == 0); }

// exchange + borders + cells + neighbour lists, retried with grown capacities
// on overflow like the generated resize loops (transformations/modules.py:159-205)
// exchange, borders, cell binning and the neighbour build of a rebuild
static void enqueue_rebuild(pairs_hip_ctx *ctx, hipStream_t s) {
    launch_exchange(s, ctx->d, ctx->geom, ctx->cfg.periodic);
    launch_borders(s, ctx->d, ctx->geom, ctx->cfg.periodic);
    launch_bin_sort(s, ctx->d, ctx->geom);  // flips d.cur; clears the image and tile flags
    // the neighbour build runs right behind the cell phase; every rebuild kernel
    // returns at once when an earlier one has set an overflow
    enqueue_build(ctx, s, false);
    // the packed (x, y, z) copy of the new buffer, for the force step after it
    launch_pack_p(s, ctx->d, ctx->d.cur);
}

// The rebuild's ~37 small kernels have grids that depend only on
// capacities, and their counts live on the device: they are replayed from a
// graph (one per position buffer, captured on the context's stream), so they
// run back to back and cost one launch on the host.
// The first launch of an executable graph prepares its dispatch packets on
// the host (several times the cost of a later launch, with the device idle
// when the graph opens a run); uploading it when it is instantiated moves
// that work to set-up.  An upload the runtime refuses leaves it to the first
// launch.
static void upload_graph(pairs_hip_ctx *ctx, hipGraphExec_t exec) {
    if (hipGraphUpload(exec, ctx->stream) != hipSuccess) (void) hipGetLastError();
}

// capture (without running) the rebuild that starts on buffer cur when missing
static int capture_rebuild_graph(pairs_hip_ctx *ctx, int cur) {
    hipGraphExec_t *slot = &ctx->reb_graph[cur];
    if (!*slot) {
        const int cur0 = ctx->d.cur;
        ctx->d.cur = cur;
        hipGraph_t g;
        HIP_OK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        enqueue_rebuild(ctx, ctx->stream);
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        ctx->d.cur = cur0;
        if (ec != hipSuccess) {
            ctx->err = std::string("rebuild capture: ") + hipGetErrorString(ec);
            return PAIRS_HIP_ERR_HIP;
        }
        const hipError_t e = hipGraphInstantiate(slot, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) {
            ctx->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
            *slot = nullptr;
            return PAIRS_HIP_ERR_HIP;
        }
        upload_graph(ctx, *slot);
    }
    return 0;
}

static int launch_rebuild(pairs_hip_ctx *ctx, hipStream_t s) {
    const int cur = ctx->d.cur;
    if (!ctx->use_graphs) {
        enqueue_rebuild(ctx, s);
        CHECK_LAUNCH();
        return 0;
    }
    if (int r = capture_rebuild_graph(ctx, cur)) return r;
    HIP_OK(hipGraphLaunch(ctx->reb_graph[cur], s));
    ctx->d.cur = 1 - cur;
    return 0;
}

// An overflow of the send map or of the particle arrays stops the cell phase
// before the permutation that writes the other position buffer: the state to
// retry from is still in the buffer the rebuild started on (the exchange
// dimensions already done moved particles there in place, and repeating them
// moves nothing).  Later overflows (tile, neighbour rows) come after the
// permutation, whose buffer then holds the state.
static bool overflow_before_permute(const int *h) { return h[C_OVF] == OVF_SEND || h[C_OVF] == OVF_PART; }

static int rebuild(pairs_hip_ctx *ctx) {
    if (ctx) ctx->xy_valid[0] = ctx->xy_valid[1] = false;
    int last = 0;
    for (int attempt = 0; attempt < 8; attempt++) {
        // one counter read for the whole rebuild: after an overflow it is
        // repeated with grown capacities
        const int cur0 = ctx->d.cur;
        if (int r = launch_rebuild(ctx, ctx->stream)) return r;
        if (read_counters(ctx)) return PAIRS_HIP_ERR_HIP;
        if (overflow_before_permute(ctx->h_counters)) ctx->d.cur = cur0;
        // images refresh the ghosts from the position kernels unless some
        // local has too many of them (then a separate refresh runs)
        ctx->d.img_on = ctx->h_counters[C_IMG_OVF] == 0 ? 1 : 0;
        int code = ctx->h_counters[C_OVF] == OVF_TILE ? finish_build(ctx) : take_overflow(ctx);
        if (code == 0) {
            set_tile_cap(ctx);
            ctx->xy_valid[ctx->d.cur] = true;  // packed at the end of the rebuild
            if (ctx->mlp_on) code = ensure_mlp_work(ctx);
            if (code <= 0) return code;
        }
        last = code;
        int pc = 0, nc = 0, sc = 0;
        if (code == PAIRS_HIP_OVERFLOW_PARTICLES) pc = 2 * ctx->req[code];
        if (code == PAIRS_HIP_OVERFLOW_NEIGHBORS) nc = 2 * ctx->req[code];
        if (code == PAIRS_HIP_OVERFLOW_SEND) sc = 2 * ctx->req[code];
        int r = pairs_hip_reserve(ctx, pc, nc, sc);
        if (r) return r;
    }
    ctx->err = "capacity growth did not converge: overflow " + std::to_string(last) + " (required " +
               std::to_string(last > 0 && last < 5 ? ctx->req[last] : 0) + "; capacities: particles " +
               std::to_string(ctx->d.cap) + ", neighbours " + std::to_string(ctx->d.ncap) + ", send " +
               std::to_string(ctx->d.send_cap) + ", tile " + std::to_string(ctx->d.tile_cap) + ")";
    return PAIRS_HIP_ERR_STATE;
}

// ---- rebuilds checked later (pairs_hip_run)
// The reference copies its resize counters to the host after every module
// and re-runs the module on an overflow (src/pairs/transformations/
// devices.py:66-79, modules.py:159-205).  Overflows are rare after the first
// rebuild, so pairs_hip_run does not wait for them: it checkpoints the locals'
// state, enqueues the rebuild, a copy of its counters into pinned memory and
// the steps after it, and checks the copy at the next rebuild or at the end of
// the run.  Every kernel stays in bounds after an overflow (counts are clamped
// to capacities and later rebuild kernels return at once), so the steps run
// on a faulty rebuild only compute values that are thrown away: the
// checkpoint is restored and the rebuild repeated with the host loop.

static CkptArrays ckpt_live(DevState &d, int cur) {
    return CkptArrays{d.X[cur], d.Y[cur], d.Z[cur], d.vx, d.vy, d.vz, d.mass, d.type, d.flags,
                      d.uid, d.shape, d.ref, d.phys, d.counters};
}

static CkptArrays ckpt_saved(pairs_hip_ctx *ctx) {
    const size_t n = (size_t) ctx->ckpt_n;
    double *c = ctx->ckpt;
    int *k = (int *) (c + 7 * n);
    return CkptArrays{c, c + n, c + 2 * n, c + 3 * n, c + 4 * n, c + 5 * n, c + 6 * n,
                      k, k + n, k + 2 * n, k + 3 * n, k + 4 * n, k + 5 * n, k + 6 * n};
}

static int ensure_spec(pairs_hip_ctx *ctx) {
    if (!ctx->h_spec) HIP_OK(hipHostMalloc((void **) &ctx->h_spec, sizeof(int) * C_COUNT));
    if (!ctx->spec_ev) HIP_OK(hipEventCreateWithFlags(&ctx->spec_ev, hipEventDisableTiming));
    if (!ctx->reb_ev) HIP_OK(hipEventCreateWithFlags(&ctx->reb_ev, hipEventDisableTiming));
    if (!ctx->side) HIP_OK(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    const int n = ctx->d.nlocal;
    if (ctx->ckpt_n < n) {
        HIP_OK(hipStreamSynchronize(ctx->stream));
        dfree(ctx->ckpt);
        ctx->ckpt_n = 0;
        // 7 doubles and 6 ints per local, then the counters
        const size_t words = 7 * (size_t) n + (6 * (size_t) n + C_COUNT + 1) / 2;
        if (dalloc(ctx, &ctx->ckpt, words)) return PAIRS_HIP_ERR_HIP;
        ctx->ckpt_n = n;
        // a checkpoint of no particles (it copies the counters only, which
        // every checkpoint rewrites): the kernel's first launch, several
        // times the host cost of a later one, happens here and not at the
        // first rebuild of a run
        launch_checkpoint(ctx->stream, ckpt_live(ctx->d, ctx->d.cur), ckpt_saved(ctx), 0);
        CHECK_LAUNCH();
    }
    return 0;
}

static int spec_rebuild(pairs_hip_ctx *ctx, int ts, int nthermo) {
    if (int r = ensure_spec(ctx)) return r;
    DevState &d = ctx->d;
    ctx->xy_valid[0] = ctx->xy_valid[1] = false;
    launch_checkpoint(ctx->stream, ckpt_live(d, d.cur), ckpt_saved(ctx), d.nlocal);
    ctx->spec_cur = d.cur;
    ctx->spec_ts = ts;
    ctx->spec_nthermo = nthermo;
    ctx->spec_ev_used = ctx->ev_used;
    if (int r = launch_rebuild(ctx, ctx->stream)) return r;  // flips d.cur
    ctx->xy_valid[d.cur] = true;                               // packed at its end
    // the steps after the rebuild only read the counters; the next writer
    // (the next rebuild, or any host path after the run) comes after the
    // check, which waits for the copy
    HIP_OK(hipEventRecord(ctx->reb_ev, ctx->stream));
    HIP_OK(hipStreamWaitEvent(ctx->side, ctx->reb_ev, 0));
    HIP_OK(hipMemcpyAsync(ctx->h_spec, d.counters, sizeof(int) * C_COUNT, hipMemcpyDeviceToHost, ctx->side));
    HIP_OK(hipEventRecord(ctx->spec_ev, ctx->side));
    CHECK_LAUNCH();
    ctx->spec_pending = true;
    ctx->spec_count++;
    return 0;
}

// 0: the pending rebuild was clean; 1: it overflowed (or left a local with
// more images than the table holds) and must be repeated
static int spec_check(pairs_hip_ctx *ctx) {
    ctx->spec_pending = false;
    HIP_OK(hipEventSynchronize(ctx->spec_ev));
    std::memcpy(ctx->h_counters, ctx->h_spec, sizeof(int) * C_COUNT);
    const int *h = ctx->h_counters;
    if (h[C_OVF] != 0 || h[C_REQ_NEIGH] > 0 || h[C_IMG_OVF] != 0) return 1;
    take_counters(ctx);
    set_tile_cap(ctx);
    return 0;
}

// back to the state before the pending rebuild: nothing enqueued after it
// runs any more, the locals' state and the counters are restored
static int spec_rollback(pairs_hip_ctx *ctx) {
    HIP_OK(hipStreamSynchronize(ctx->stream));
    DevState &d = ctx->d;
    d.cur = ctx->spec_cur;
    ctx->xy_valid[0] = ctx->xy_valid[1] = false;
    launch_checkpoint(ctx->stream, ckpt_saved(ctx), ckpt_live(d, d.cur), d.nlocal);
    CHECK_LAUNCH();
    HIP_OK(hipStreamSynchronize(ctx->stream));
    ctx->spec_rollbacks++;
    return 0;
}

int pairs_hip_set_speculative_rebuilds(pairs_hip_ctx *ctx, int enable) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    ctx->spec_enabled = enable != 0;
    return 0;
}

int pairs_hip_step(pairs_hip_ctx *ctx, int ts, int nsteps) {
    if (!ctx || ts < 0 || ts > nsteps) return PAIRS_HIP_ERR_ARG;
    DevState &d = ctx->d;
    hipStream_t s = ctx->stream;
    if (ts > 0) launch_vv_initial(s, d, ctx->prm);
    ctx->xy_valid[0] = ctx->xy_valid[1] = false;
    if (every(ts, ctx->cfg.reneighbor_every)) {
        drop_graphs(ctx);
        int r = rebuild(ctx);
        if (r) return r;
    } else {
        launch_sync_inplace(s, d, ctx->geom, ctx->ghost_grid, 0);
    }
    if (ctx->mlp_on) launch_mlp_forces(s, d, ctx->geom, ctx->mlp, d.cur, 0, ctx->ghost_grid);
    else {
        launch_lj_full(s, d, ctx->geom, ctx->prm, d.cur, d.cur, 0, 0, 0, 0, 1);
        ctx->xy_valid[d.cur] = true;
    }
    if (ts > 0) launch_vv_post(s, d, ctx->geom, ctx->prm, d.cur, d.cur, 1, 0, 0);
    CHECK_LAUNCH();
    return 0;
}

// one non-rebuild step of the fused double-buffered form:
//   sync ghosts of ts into buffer cur (stale multi-hop values from the other)
//   force(ts) + final_integrate(ts) + initial_integrate(ts+1) -> other buffer
// force(ts) on buffer cur, then final_integrate(ts) + initial_integrate(ts+1)
// into buffer next (fused into the LJ kernel; half lists walk their mirror
// part in the same kernel)
static void launch_force_post(pairs_hip_ctx *ctx, int cur, int next, int vv2, int vv1, int reduce, bool timed) {
    DevState &d = ctx->d;
    hipStream_t s = ctx->stream;
    if (timed && ctx->mlp_on) hipEventRecord(ctx->ev_begin[ctx->ev_used], s);
    if (ctx->mlp_on) {
        // the force pass integrates too, except on thermo steps (whose
        // kinetic partials come from the stand-alone integrator); either
        // writes the next buffer's packed copy, so the descriptor pass packs
        // only when the previous step did not leave one
        double4 *Pn = (vv1 && next != cur) ? d.P[next] : nullptr;
        const MLPVV vv{&ctx->prm, next, vv2, vv1, Pn};
        launch_mlp_forces(s, d, ctx->geom, ctx->mlp, cur, reduce, ctx->ghost_grid, reduce ? nullptr : &vv,
                          !ctx->xy_valid[cur]);
        if (timed) hipEventRecord(ctx->ev_end[ctx->ev_used], s);
        if (reduce) launch_vv_post(s, d, ctx->geom, ctx->prm, cur, next, vv2, vv1, reduce, Pn);
        ctx->xy_valid[cur] = true;
        if (Pn) ctx->xy_valid[next] = true;
        return;
    } else {
        // the timed region is the force kernel itself (after a pack of
        // (x, y) when the previous kernel did not leave one)
        launch_lj_full(s, d, ctx->geom, ctx->prm, cur, next, 1, vv2, vv1, reduce, !ctx->xy_valid[cur],
                       timed ? ctx->ev_begin[ctx->ev_used] : nullptr);
        if (timed) hipEventRecord(ctx->ev_end[ctx->ev_used], s);
        ctx->xy_valid[cur] = true;
        if (vv1 && next != cur) ctx->xy_valid[next] = true;  // written by the fused kernel
    }
}

static void launch_fused_step(pairs_hip_ctx *ctx, int cur, int vv2, int vv1, int reduce, bool timed) {
    DevState &d = ctx->d;
    const int other = 1 - cur;
    if (!d.img_on) launch_sync_pp(ctx->stream, d, ctx->geom, cur, other, ctx->ghost_grid);
    launch_force_post(ctx, cur, other, vv2, vv1, reduce, timed);
}

static int ensure_events(pairs_hip_ctx *ctx, int n) {
    while ((int) ctx->ev_begin.size() < n) {
        hipEvent_t a, b;
        HIP_OK(hipEventCreate(&a));
        HIP_OK(hipEventCreate(&b));
        ctx->ev_begin.push_back(a);
        ctx->ev_end.push_back(b);
    }
    return 0;
}

int pairs_hip_set_graphs(pairs_hip_ctx *ctx, int enable) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    ctx->use_graphs = enable != 0;
    if (!ctx->use_graphs) drop_graphs(ctx);
    return 0;
}

int pairs_hip_set_force_timing(pairs_hip_ctx *ctx, int every_n) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    ctx->time_force = every_n > 0;
    ctx->time_every = every_n > 0 ? every_n : 10;
    return 0;
}

int pairs_hip_force_kernel_time(pairs_hip_ctx *ctx, double *avg_ms, int *launches) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    if (avg_ms) *avg_ms = ctx->force_ms_avg;
    if (launches) *launches = ctx->force_launches;
    return 0;
}

static bool step_graph_stale(const pairs_hip_ctx *ctx, int cur, int n) {
    return !ctx->step_graph[cur][n - 1] || ctx->graph_ghost_grid[cur][n - 1] < ctx->ghost_grid ||
           ctx->graph_img[cur][n - 1] != ctx->d.img_on;
}

// Capture the kernels that enqueue(ctx) launches on the context's stream
// into slot, without running them, with the ghost grid given headroom so
// that later rebuilds with a few more ghosts reuse the graph.  The host's
// record of which packed copies are current is left as it was.
static int capture_graph(pairs_hip_ctx *ctx, hipGraphExec_t &slot, int &grid_out, int &img_out,
                         const std::function<void()> &enqueue) {
    destroy_exec(ctx, slot);
    const int saved = ctx->ghost_grid;
    const bool xy0 = ctx->xy_valid[0], xy1 = ctx->xy_valid[1];
    ctx->ghost_grid = saved + saved / 4 + 4;
    hipGraph_t g;
    HIP_OK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    enqueue();
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
    hipError_t e = ec;
    if (ec == hipSuccess) {
        e = hipGraphInstantiate(&slot, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e == hipSuccess) upload_graph(ctx, slot);
    }
    grid_out = ctx->ghost_grid;
    img_out = ctx->d.img_on;
    ctx->ghost_grid = saved;
    ctx->xy_valid[0] = xy0;
    ctx->xy_valid[1] = xy1;
    if (e != hipSuccess) {
        ctx->err = std::string("step graph capture: ") + hipGetErrorString(e);
        slot = nullptr;
        return PAIRS_HIP_ERR_HIP;
    }
    return 0;
}

// capture (without running) n steady steps from buffer cur when missing or
// stale: the buffers alternate from step to step, so the chain is the single
// step's kernels repeated with the parities swapped.  A chain runs only after
// a fused step, which left the packed copy of its first buffer current.
static int capture_step_graph(pairs_hip_ctx *ctx, int cur, int n) {
    if (!step_graph_stale(ctx, cur, n)) return 0;
    return capture_graph(ctx, ctx->step_graph[cur][n - 1], ctx->graph_ghost_grid[cur][n - 1],
                         ctx->graph_img[cur][n - 1], [&] {
                             ctx->xy_valid[0] = ctx->xy_valid[1] = true;
                             for (int q = 0; q < n; q++) launch_fused_step(ctx, cur ^ (q & 1), 1, 1, 0, false);
                         });
}

// n >= 1 steady steps from buffer cur: one chain, or chains of CHAIN_MAX
// steps and the rest
static int graph_steps(pairs_hip_ctx *ctx, int cur, int n) {
    while (n > 0) {
        const int m = std::min(n, pairs_hip_ctx::CHAIN_MAX);
        if (int r = capture_step_graph(ctx, cur, m)) return r;
        HIP_OK(hipGraphLaunch(ctx->step_graph[cur][m - 1], ctx->stream));
        cur ^= m & 1;
        n -= m;
    }
    return 0;
}

// The first step of a run that is neither a rebuild nor a thermo or timed
// step, from buffer cur, as one launch.  When every ghost is an image of a
// local (img_on): initial_integrate(ts) of the locals into the other buffer
// with their ghost images and the packed copy (an image of a ghost takes
// that ghost's position before the refresh: the stale multi-hop
// synchronize), then the fused force step from the other buffer back into
// cur -- two kernels, and the run continues on buffer cur.  Otherwise:
// initial_integrate in place, the ghosts refreshed in place, the packed copy
// and the fused force step into the other buffer, which the run continues on.
// start_img[cur] records which form the graph holds.
static int capture_start_graph(pairs_hip_ctx *ctx, int cur) {
    if (ctx->start_graph[cur] && ctx->start_ghost_grid[cur] >= ctx->ghost_grid && ctx->start_img[cur] == ctx->d.img_on)
        return 0;
    const int cur0 = ctx->d.cur;
    ctx->d.cur = cur;
    const int r = capture_graph(ctx, ctx->start_graph[cur], ctx->start_ghost_grid[cur], ctx->start_img[cur], [&] {
        DevState &d = ctx->d;
        const int other = 1 - cur;
        if (d.img_on) {
            launch_vv_post(ctx->stream, d, ctx->geom, ctx->prm, cur, other, 0, 1, 0, d.P[other]);
            ctx->xy_valid[other] = true;
            launch_force_post(ctx, other, cur, 1, 1, 0, false);
        } else {
            launch_vv_initial(ctx->stream, d, ctx->prm);
            launch_sync_inplace(ctx->stream, d, ctx->geom, ctx->ghost_grid, 0);
            ctx->xy_valid[cur] = false;
            launch_force_post(ctx, cur, other, 1, 1, 0, false);
        }
    });
    ctx->d.cur = cur0;
    return r;
}

int pairs_hip_prepare_graphs(pairs_hip_ctx *ctx) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    if (!ctx->use_graphs) return 0;
    if (ctx->time_force && ensure_events(ctx, 4096)) return PAIRS_HIP_ERR_HIP;
    // the deferred checks' pinned copy, events, second stream and checkpoint
    if (ctx->spec_enabled && !ctx->mlp_on)
        if (int r = ensure_spec(ctx)) return r;
    // the chains a run between two rebuilds uses (longer or other lengths,
    // cut by thermo or timed steps, are captured when first needed)
    const int nmax = std::max(1, std::min(pairs_hip_ctx::CHAIN_MAX, ctx->cfg.reneighbor_every - 1));
    for (int b = 0; b < 2; b++) {
        if (int r = capture_rebuild_graph(ctx, b)) return r;
        if (ctx->h_counters[C_NLOCAL] > 0 && ctx->d.nl && !ctx->mlp_on) {
            if (int r = capture_start_graph(ctx, b)) return r;
            for (int n = 1; n <= nmax; n++)
                if (int r = capture_step_graph(ctx, b, n)) return r;
        }
    }
    return 0;
}

int pairs_hip_run(pairs_hip_ctx *ctx, int ts_begin, int ts_end, int nsteps, double *thermo_out, int thermo_capacity) {
    if (!ctx || ts_begin < 0 || ts_end > nsteps + 1 || ts_begin > ts_end) return PAIRS_HIP_ERR_ARG;
    DevState &d = ctx->d;
    hipStream_t s = ctx->stream;
    int nthermo = 0;
    bool prev_in_other = false;  // previous ghosts live in the other buffer
    ctx->ev_used = 0;
    if (ctx->time_force && ensure_events(ctx, 4096)) return PAIRS_HIP_ERR_HIP;
    bool redo = false;  // repeating a rebuild step whose deferred check failed
    for (int ts = ts_begin;; ts++) {
        // a rebuild still unchecked is checked before the next one and at the
        // end; when it overflowed, the run continues from its step again
        if (ctx->spec_pending && (ts >= ts_end || every(ts, ctx->cfg.reneighbor_every))) {
            const int c = spec_check(ctx);
            if (c < 0) return c;
            if (c > 0) {
                if (int r = spec_rollback(ctx)) return r;
                ts = ctx->spec_ts;
                nthermo = ctx->spec_nthermo;
                ctx->ev_used = ctx->spec_ev_used;
                redo = true;
            }
        }
        if (ts >= ts_end) break;
        const bool R = every(ts, ctx->cfg.reneighbor_every);
        const bool T = every(ts, ctx->cfg.thermo_every);
        const int vv2 = ts > 0, vv1 = (ts + 1 < ts_end) && (ts + 1 <= nsteps);
        auto timed_at = [&](int t) {
            return ctx->time_force && !every(t, ctx->cfg.reneighbor_every) && ((t % ctx->time_every) == 0) &&
                   ctx->ev_used < 4096;
        };
        const bool timed = timed_at(ts);
        if (ts == ts_begin && ts > 0 && !redo) {  // (a repeat starts from the checkpoint, taken after it)
            if (ctx->use_graphs && !ctx->mlp_on && !R && !T && !timed && vv1) {
                if (int r = capture_start_graph(ctx, d.cur)) return r;
                HIP_OK(hipGraphLaunch(ctx->start_graph[d.cur], s));
                ctx->xy_valid[0] = ctx->xy_valid[1] = true;  // packed, and written by the fused kernel
                if (!ctx->start_img[d.cur]) d.cur ^= 1;       // the image form ends on buffer cur
                prev_in_other = true;
                continue;
            }
            launch_vv_initial(s, d, ctx->prm);
            ctx->xy_valid[d.cur] = false;
        }
        auto steady = [&](int t) {
            return !every(t, ctx->cfg.reneighbor_every) && !every(t, ctx->cfg.thermo_every) && t > 0 &&
                   t + 1 < ts_end && t + 1 <= nsteps && !timed_at(t);
        };
        if (prev_in_other && ctx->use_graphs && steady(ts)) {
            // this step and the steady ones after it replay as chains
            int n = 1;
            while (steady(ts + n)) n++;
            if (int r = graph_steps(ctx, d.cur, n)) return r;
            d.cur ^= n & 1;
            ctx->xy_valid[d.cur] = true;  // the replayed force kernel wrote it
            ts += n - 1;
            continue;
        }
        if (R) {
            const bool spec = !redo && ctx->spec_enabled && ctx->use_graphs && !ctx->mlp_on && d.img_on;
            int r = spec ? spec_rebuild(ctx, ts, nthermo) : rebuild(ctx);
            if (r) return r;
            redo = false;
            prev_in_other = false;
        }
        const int cur = d.cur, other = 1 - cur;
        if (!R && !prev_in_other) {
            launch_sync_inplace(s, d, ctx->geom, ctx->ghost_grid, 0);
            ctx->xy_valid[d.cur] = false;
        }
        if (!R && prev_in_other) {
            launch_fused_step(ctx, cur, vv2, vv1, T ? 1 : 0, timed);
        } else {
            // ghosts already current in buffer cur: force + integrators only
            launch_force_post(ctx, cur, other, vv2, vv1, T ? 1 : 0, timed);
        }
        if (timed) ctx->ev_used++;
        if (T && nthermo < 4096) {
            // partials of the fused LJ kernel (one per wave), or of the MLP
            // force kernel (one per wave of 64) and the integrator (per 256)
            const int nb_ke = ctx->mlp_on ? (d.nlocal + 255) / 256 : lj_waves(d);
            const int nb_f = ctx->mlp_on ? mlp_waves(d) : lj_waves(d);
            launch_thermo_reduce(s, d, nb_ke, nb_f, 3, ctx->geom, d.thermo + 5 * nthermo);
            nthermo++;
        }
        CHECK_LAUNCH();
        if (vv1) { d.cur = other; prev_in_other = true; }
    }
    HIP_OK(hipStreamSynchronize(s));
    if (thermo_out && nthermo > 0) {
        const int n = std::min(nthermo, thermo_capacity);
        HIP_OK(hipMemcpy(thermo_out, d.thermo, 5 * n * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (ctx->time_force && ctx->ev_used > 0) {
        double tot = 0.0;
        for (int k = 0; k < ctx->ev_used; k++) {
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, ctx->ev_begin[k], ctx->ev_end[k]));
            tot += ms;
        }
        ctx->force_ms_avg = tot / ctx->ev_used;
        ctx->force_launches = ctx->ev_used;
    }
    return nthermo;
}

// ---------------------------------------------------------------- MLP force field (include/pairs_hip_mlp.h)

static int mlp_prepare(pairs_hip_ctx *ctx, pairs_hip_mlp *m) {
    if (ctx->cfg.half) {
        ctx->err = "the MLP potential runs on full neighbour lists (cfg.half = 0)";
        return PAIRS_HIP_ERR_STATE;
    }
    if (int r = ensure_mlp_work(ctx)) return r;
    if (m) HIP_OK(hipStreamSynchronize(m->stream));  // the model's own stream (training) is done
    return 0;
}

static MLPBind mlp_bind(pairs_hip_ctx *ctx, pairs_hip_mlp *m) {
    MLPBind b = ctx->mlp;
    b.p = m->p;
    b.sc = m->sc;
    return b;
}

int pairs_hip_mlp_descriptors(pairs_hip_ctx *ctx, float *G, int count) {
    if (!ctx || !G || count < ctx->d.nlocal) return PAIRS_HIP_ERR_ARG;
    if (int r = mlp_prepare(ctx, nullptr)) return r;
    DevState &d = ctx->d;
    const int n = d.nlocal;
    launch_mlp_descriptors(ctx->stream, d, ctx->geom, d.cur, ctx->mlp_G);
    CHECK_LAUNCH();
    std::vector<int> phys;
    if (host_phys(ctx, n, phys)) return PAIRS_HIP_ERR_HIP;
    std::vector<float> tmp((size_t) n * MLP_ND);
    if (download_raw(ctx, tmp.data(), ctx->mlp_G, tmp.size() * sizeof(float))) return PAIRS_HIP_ERR_HIP;
    for (int r = 0; r < n; r++)
        std::memcpy(G + (size_t) r * MLP_ND, tmp.data() + (size_t) phys[r] * MLP_ND, MLP_ND * sizeof(float));
    return 0;
}

int pairs_hip_mlp_add_snapshot(pairs_hip_mlp *m, pairs_hip_ctx *ctx) {
    if (!ctx || !m) return PAIRS_HIP_ERR_ARG;
    if (int r = mlp_prepare(ctx, m)) return r;
    if (ensure_eatom(ctx)) return PAIRS_HIP_ERR_HIP;
    DevState &d = ctx->d;
    const int n = d.nlocal;
    if (m->ds_rows + n > m->ds_cap) {
        ctx->err = "the model's training set is full";
        return PAIRS_HIP_ERR_STATE;
    }
    hipStream_t s = ctx->stream;
    launch_mlp_descriptors(s, d, ctx->geom, d.cur, ctx->mlp_G);
    launch_energy(s, d, ctx->geom, ctx->prm, 0, ctx->d_eatom);
    launch_mlp_snapshot(s, ctx->mlp_G, ctx->d_eatom, n, m->ds_G + (size_t) m->ds_rows * MLP_ND,
                        m->ds_label + m->ds_rows);
    CHECK_LAUNCH();
    HIP_OK(hipStreamSynchronize(s));
    m->ds_rows += n;
    return n;
}

int pairs_hip_mlp_forces(pairs_hip_ctx *ctx, pairs_hip_mlp *m, double *energy, double *virial) {
    if (!ctx || !m) return PAIRS_HIP_ERR_ARG;
    if (int r = mlp_prepare(ctx, m)) return r;
    DevState &d = ctx->d;
    launch_mlp_forces(ctx->stream, d, ctx->geom, mlp_bind(ctx, m), d.cur, 1, ctx->ghost_grid);
    launch_thermo_reduce(ctx->stream, d, 0, mlp_waves(d), 2, ctx->geom, d.thermo);
    CHECK_LAUNCH();
    double out[5];
    if (download_raw(ctx, out, d.thermo, 5 * sizeof(double))) return PAIRS_HIP_ERR_HIP;
    if (energy) *energy = out[2];
    if (virial) *virial = out[3];
    return 0;
}

int pairs_hip_mlp_force_error(pairs_hip_ctx *ctx, pairs_hip_mlp *m, double *rel_rms) {
    if (!ctx || !m || !rel_rms) return PAIRS_HIP_ERR_ARG;
    if (int r = mlp_prepare(ctx, m)) return r;
    DevState &d = ctx->d;
    hipStream_t s = ctx->stream;
    const int n = d.nlocal;
    if (n <= 0) { *rel_rms = 0.0; return 0; }
    const size_t cap = (size_t) ctx->mlp_cap;
    launch_mlp_forces(s, d, ctx->geom, mlp_bind(ctx, m), d.cur, 0, ctx->ghost_grid);
    HIP_OK(hipMemcpyAsync(ctx->mlp_f, d.fx, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(ctx->mlp_f + cap, d.fy, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(ctx->mlp_f + 2 * cap, d.fz, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    launch_lj_full(s, d, ctx->geom, ctx->prm, d.cur, d.cur, 0, 0, 0, 0, !ctx->xy_valid[d.cur]);
    ctx->xy_valid[d.cur] = true;
    const int nb = std::min((n + 255) / 256, 1024);
    launch_force_diff(s, ctx->mlp_f, ctx->mlp_f + cap, ctx->mlp_f + 2 * cap, d.fx, d.fy, d.fz, n, ctx->mlp_red, nb);
    CHECK_LAUNCH();
    std::vector<double> part(2 * nb);
    if (download_raw(ctx, part.data(), ctx->mlp_red, part.size() * sizeof(double))) return PAIRS_HIP_ERR_HIP;
    double num = 0.0, den = 0.0;
    for (int b = 0; b < nb; b++) { num += part[2 * b]; den += part[2 * b + 1]; }
    *rel_rms = den > 0.0 ? std::sqrt(num / den) : std::sqrt(num);
    return 0;
}

int pairs_hip_mlp_config_grad(pairs_hip_ctx *ctx, pairs_hip_mlp *m, double w_energy, double w_force, float *grad,
                              double *loss) {
    if (!ctx || !m || !(w_energy >= 0.0) || !(w_force >= 0.0)) return PAIRS_HIP_ERR_ARG;
    if (ctx->mlp_on) {
        ctx->err = "pairs_hip_mlp_config_grad labels the configuration with LJ: detach the model first";
        return PAIRS_HIP_ERR_STATE;
    }
    if (int r = mlp_prepare(ctx, m)) return r;
    if (ensure_eatom(ctx)) return PAIRS_HIP_ERR_HIP;
    DevState &d = ctx->d;
    const int n = d.nlocal;
    if (n <= 0) {
        ctx->err = "no particles";
        return PAIRS_HIP_ERR_STATE;
    }
    if (ctx->mlp_train_cap < d.cap) {
        HIP_OK(hipStreamSynchronize(ctx->stream));
        dfree(ctx->mlp_rho); dfree(ctx->mlp_U);
        ctx->mlp_rho = nullptr;
        ctx->mlp_U = nullptr;
        if (dalloc(ctx, &ctx->mlp_rho, d.cap) || dalloc(ctx, &ctx->mlp_U, (size_t) d.cap * MLP_ND))
            return PAIRS_HIP_ERR_HIP;
        ctx->mlp_train_cap = d.cap;
    }
    hipStream_t s = ctx->stream;
    const size_t cap = (size_t) ctx->mlp_cap;
    double *lx = ctx->mlp_f, *ly = ctx->mlp_f + cap, *lz = ctx->mlp_f + 2 * cap;
    // labels: LJ forces and per-atom LJ energies of this configuration
    launch_lj_full(s, d, ctx->geom, ctx->prm, d.cur, d.cur, 0, 0, 0, 0, !ctx->xy_valid[d.cur]);
    ctx->xy_valid[d.cur] = true;
    HIP_OK(hipMemcpyAsync(lx, d.fx, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(ly, d.fy, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(lz, d.fz, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    launch_energy(s, d, ctx->geom, ctx->prm, 0, ctx->d_eatom);
    const int nb = std::min((n + 255) / 256, 1024);
    MLPCfgGrad t{};
    t.nwaves = m->nwaves;
    t.partial = m->partial;
    t.loss = m->loss;
    launch_mlp_config_grad(s, d, ctx->geom, mlp_bind(ctx, m), d.cur, lx, ly, lz, ctx->d_eatom, w_energy, w_force,
                           ctx->mlp_rho, ctx->mlp_U, ctx->mlp_red, nb, ctx->ghost_grid, t);
    launch_mlp_accum(s, m->partial, m->nwaves, m->gacc, grad ? m->gout : nullptr);
    // leave the LJ forces in the force array, as after pairs_hip_lj
    HIP_OK(hipMemcpyAsync(d.fx, lx, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(d.fy, ly, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(d.fz, lz, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    CHECK_LAUNCH();
    m->gcount++;
    if (grad) HIP_OK(hipMemcpyAsync(grad, m->gout, MLP_NP * sizeof(float), hipMemcpyDeviceToHost, s));
    if (loss) {
        std::vector<double> part(2 * nb);
        std::vector<float> wl(m->nwaves);
        HIP_OK(hipMemcpyAsync(part.data(), ctx->mlp_red, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(wl.data(), m->loss, wl.size() * sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        double num = 0.0, den = 0.0, le = 0.0;
        for (int b = 0; b < nb; b++) { num += part[2 * b]; den += part[2 * b + 1]; }
        for (float x : wl) le += x;
        loss[0] = le / n;
        loss[1] = den > 0.0 ? num / den : num;
    }
    HIP_OK(hipStreamSynchronize(s));
    return 0;
}

int pairs_hip_mlp_attach(pairs_hip_ctx *ctx, pairs_hip_mlp *m, double max_rel_rms, double *rel_rms) {
    if (!ctx || !m) return PAIRS_HIP_ERR_ARG;
    if (int r = mlp_prepare(ctx, m)) return r;
    if (rel_rms) *rel_rms = -1.0;
    drop_graphs(ctx);
    if (max_rel_rms >= 0.0) {
        double rel = 0.0;
        if (int r = pairs_hip_mlp_force_error(ctx, m, &rel)) return r;
        if (rel_rms) *rel_rms = rel;
        if (!(rel <= max_rel_rms)) {  // fallback: keep LJ (its forces are in the force array)
            ctx->mlp_on = false;
            return 1;
        }
    }
    ctx->mlp = mlp_bind(ctx, m);
    ctx->mlp_on = true;
    DevState &d = ctx->d;
    launch_mlp_forces(ctx->stream, d, ctx->geom, ctx->mlp, d.cur, 0, ctx->ghost_grid);
    CHECK_LAUNCH();
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int pairs_hip_mlp_detach(pairs_hip_ctx *ctx) {
    if (!ctx) return PAIRS_HIP_ERR_ARG;
    drop_graphs(ctx);
    ctx->mlp_on = false;
    return 0;
}

int pairs_hip_mlp_active(const pairs_hip_ctx *ctx) { return ctx && ctx->mlp_on ? 1 : 0; }

}  // extern "C"

// ---------------------------------------------------------------- batched ensemble (include/pairs_hip_batch.h)

struct pairs_hip_batch {
    std::vector<pairs_hip_ctx *> rep;
    hipStream_t stream = nullptr;
    std::string err;
    // descriptors of the fused force kernel, one per replica, for the two
    // buffer phases after a rebuild (phase p: replica r reads buffer base_cur[r] ^ p)
    std::vector<int> base_cur;
    LJArgs *h_desc = nullptr;          // pinned staging, 3 * R
    LJArgs *d_desc[3] = {nullptr, nullptr, nullptr};  // phase 0, phase 1, one-off
    // block tables of the force kernel (one block per wave of 64 locals):
    // block k belongs to replica d_rep[k] and is its wave k - d_base[replica]
    int *d_rep = nullptr, *d_base = nullptr;
    int nblocks = 0, blocks_cap = 0;
    // the same for the neighbour build (one block per segment)
    int *d_srep = nullptr, *d_sbase = nullptr;
    int nsblocks = 0, sblocks_cap = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_blocks[2] = {0, 0}, graph_wg[2] = {0, 0};
    int max_wg = 0;                    // workgroups of the persistent force launch (0: all resident)
    std::vector<double> thermo;        // last thermo row per replica
    bool ntypes1 = true;
    bool unit = true;                  // every replica: one type, epsilon = sigma6 = 1
    int pending_home = 0;              // segment size to apply at the next rebuild
    int *h_cnt = nullptr;              // pinned, [R][C_COUNT] counters of every replica
    RepDesc *h_rdesc = nullptr, *d_rdesc = nullptr;    // every replica's state for the batched kernels
    bool rdesc_valid = false;          // d_rdesc holds every replica's current arrays
    double *d_thermo = nullptr, *h_thermo = nullptr;   // [R][5] thermo rows
    BuildArgs *h_bdesc = nullptr, *d_bdesc = nullptr;  // every replica's neighbour-build arguments
    const int **h_cptr = nullptr, **d_cptr = nullptr;  // every replica's device counters
    int *d_call = nullptr;             // [R][C_COUNT] gathered counters
    bool tables_valid = false;         // the block tables match the replicas' waves and segments
};

static void batch_drop_graphs(pairs_hip_batch *b) {
    if (b->stream && (b->graph[0] || b->graph[1])) hipStreamSynchronize(b->stream);  // no replay pending
    for (int p = 0; p < 2; p++) {
        if (b->graph[p]) hipGraphExecDestroy(b->graph[p]);
        b->graph[p] = nullptr;
    }
}

#define BATCH_OK(expr)                                                                                 \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            b->err = std::string(#expr) + ": " + hipGetErrorString(e_);                                \
            return PAIRS_HIP_ERR_HIP;                                                                  \
        }                                                                                              \
    } while (0)

// block -> replica tables after (re)segmentation
// one block table: replica r owns blocks [base[r], base[r + 1]), count(r) of them
template <typename Count>
static int batch_table(pairs_hip_batch *b, Count count, int *&d_rep, int *d_base, int &cap, int &nblocks) {
    const int R = (int) b->rep.size();
    std::vector<int> base(R + 1);
    int tot = 0;
    for (int r = 0; r < R; r++) {
        base[r] = tot;
        tot += count(b->rep[r]);
    }
    base[R] = tot;
    if (tot > cap) {
        if (d_rep) hipFree(d_rep);
        d_rep = nullptr;
        cap = tot + tot / 4 + 64;
        BATCH_OK(hipMalloc((void **) &d_rep, cap * sizeof(int)));
    }
    std::vector<int> rep(std::max(tot, 1));
    for (int r = 0; r < R; r++)
        for (int k = base[r]; k < base[r + 1]; k++) rep[k] = r;
    BATCH_OK(hipMemcpyAsync(d_rep, rep.data(), tot * sizeof(int), hipMemcpyHostToDevice, b->stream));
    BATCH_OK(hipMemcpyAsync(d_base, base.data(), (R + 1) * sizeof(int), hipMemcpyHostToDevice, b->stream));
    BATCH_OK(hipStreamSynchronize(b->stream));
    nblocks = tot;
    return 0;
}

// the force kernel's wave table and the neighbour build's segment table
static int batch_tables(pairs_hip_batch *b) {
    if (int e = batch_table(b, [](pairs_hip_ctx *c) { return force_waves(c->d); }, b->d_rep, b->d_base,
                            b->blocks_cap, b->nblocks))
        return e;
    if (int e = batch_table(b, [](pairs_hip_ctx *c) { return c->geom.nsegs; }, b->d_srep, b->d_sbase,
                            b->sblocks_cap, b->nsblocks))
        return e;
    b->tables_valid = true;
    return 0;
}

// descriptors of both phases (steady steps: final and initial integrate on)
static int batch_descs(pairs_hip_batch *b) {
    const int R = (int) b->rep.size();
    b->base_cur.resize(R);
    for (int p = 0; p < 2; p++)
        for (int r = 0; r < R; r++) {
            pairs_hip_ctx *c = b->rep[r];
            const int cur = c->d.cur ^ p;
            b->h_desc[p * R + r] = lj_args(c->d, c->geom, c->prm, cur, 1 - cur, 1, 1, 1, nullptr, true);
        }
    for (int r = 0; r < R; r++) b->base_cur[r] = b->rep[r]->d.cur;
    for (int p = 0; p < 2; p++)
        BATCH_OK(hipMemcpyAsync(b->d_desc[p], b->h_desc + p * R, R * sizeof(LJArgs), hipMemcpyHostToDevice, b->stream));
    BATCH_OK(hipStreamSynchronize(b->stream));
    return 0;
}

// every replica's rebuild kernels back to back, one synchronisation per
// phase; replicas that overflow fall back to their own retry loop
static int batch_rebuild(pairs_hip_batch *b) {
    const int R = (int) b->rep.size();
    hipStream_t s = b->stream;
    if (b->pending_home > 0) {
        // a new segment size: segments and lists change together, here
        for (pairs_hip_ctx *c : b->rep) {
            c->home_target = b->pending_home;
            if (c->zc_override <= 0) setup_segments(c, 0);
            drop_graphs(c);
        }
        batch_drop_graphs(b);
        b->pending_home = 0;
        b->tables_valid = false;
    }
    if (!b->tables_valid)
        if (int e = batch_tables(b)) return e;
    // every replica's cell phase (exchange, borders, binning), each kernel
    // one launch over all replicas; then one neighbour build over every
    // replica's segments
    int max_cap = 0, max_ncells = 0;
    for (int r = 0; r < R; r++) {
        pairs_hip_ctx *c = b->rep[r];
        RepDesc &rd = b->h_rdesc[r];
        rd.d = c->d;
        rd.g = c->geom;
        for (int k = 0; k < 3; k++) rd.periodic[k] = c->cfg.periodic[k] ? 1 : 0;
        max_cap = std::max(max_cap, c->d.cap);
        max_ncells = std::max(max_ncells, c->geom.ncells);
        c->xy_valid[0] = c->xy_valid[1] = false;
    }
    BATCH_OK(hipMemcpyAsync(b->d_rdesc, b->h_rdesc, R * sizeof(RepDesc), hipMemcpyHostToDevice, s));
    b->rdesc_valid = true;
    launch_cells_batch(s, b->d_rdesc, R, max_cap, max_ncells);
    std::vector<int> cur0(R);
    for (int r = 0; r < R; r++) {
        cur0[r] = b->rep[r]->d.cur;
        b->rep[r]->d.cur = 1 - cur0[r];  // the permutation moved every replica
    }
    size_t lds = 0;
    for (int r = 0; r < R; r++) {
        pairs_hip_ctx *c = b->rep[r];
        b->h_bdesc[r] = build_args(c->d, c->geom, 0, c->cfg.cell_spacing * c->cfg.cell_spacing);
        lds = std::max(lds, build_lds_bytes(c->d, c->geom, 0));
    }
    BATCH_OK(hipMemcpyAsync(b->d_bdesc, b->h_bdesc, R * sizeof(BuildArgs), hipMemcpyHostToDevice, s));
    launch_build_batch(s, b->d_bdesc, b->d_srep, b->d_sbase, b->nsblocks, lds, 0);
    BATCH_OK(hipGetLastError());
    // every replica's counters gathered on the device, one copy to the host
    for (int r = 0; r < R; r++) b->h_cptr[r] = b->rep[r]->d.counters;
    BATCH_OK(hipMemcpyAsync(b->d_cptr, b->h_cptr, R * sizeof(int *), hipMemcpyHostToDevice, s));
    launch_gather_counters(s, b->d_cptr, R, b->d_call);
    BATCH_OK(hipMemcpyAsync(b->h_cnt, b->d_call, (size_t) R * C_COUNT * sizeof(int), hipMemcpyDeviceToHost, s));
    BATCH_OK(hipStreamSynchronize(s));
    for (int r = 0; r < R; r++) {
        std::memcpy(b->rep[r]->h_counters, b->h_cnt + (size_t) r * C_COUNT, sizeof(int) * C_COUNT);
        take_counters(b->rep[r]);
    }
    std::vector<char> fallback(R, 0);
    for (int r = 0; r < R; r++) {
        pairs_hip_ctx *c = b->rep[r];
        c->d.img_on = c->h_counters[C_IMG_OVF] == 0 ? 1 : 0;
        if (c->h_counters[C_OVF] != 0 || c->h_counters[C_REQ_NEIGH] > 0) {
            // an overflow anywhere: the replica repeats its rebuild on its
            // own, growing what it needs, from the buffer that holds its
            // state (the one it started on when the cell phase stopped
            // before the permutation)
            if (overflow_before_permute(c->h_counters)) c->d.cur = cur0[r];
            // its arrays, segments or tile capacity may change: the batched
            // thermo descriptors must be rebuilt before their next use
            b->rdesc_valid = false;
            launch_set_counters(s, c->d.counters, C_OVF, 1, 0);
            launch_set_counters(s, c->d.counters, C_REQ_NEIGH, 3, 0);
            fallback[r] = 1;
        } else {
            set_tile_cap(c);
        }
        if (fallback[r]) {
            drop_graphs(c);
            b->tables_valid = false;  // the replica's segments may change
            // launched directly: a replica's own rebuild graphs are never
            // replayed in batched runs
            const bool ug = c->use_graphs;
            c->use_graphs = false;
            const int code = rebuild(c);
            c->use_graphs = ug;
            if (code) { b->err = "replica " + std::to_string(r) + ": " + c->err; return code; }
        }
    }
    int e = b->tables_valid ? 0 : batch_tables(b);
    if (!e) e = batch_descs(b);
    return e;
}

static int batch_graph_launch(pairs_hip_batch *b, int phase) {
    if (!b->graph[phase] || b->graph_blocks[phase] != b->nblocks || b->graph_wg[phase] != b->max_wg) {
        if (b->graph[phase]) {
            BATCH_OK(hipStreamSynchronize(b->stream));  // no replay pending
            hipGraphExecDestroy(b->graph[phase]);
        }
        b->graph[phase] = nullptr;
        hipGraph_t g;
        BATCH_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        for (size_t r = 0; r < b->rep.size(); r++) {
            pairs_hip_ctx *c = b->rep[r];
            if (!c->d.img_on) {
                const int cur = b->base_cur[r] ^ phase;
                launch_sync_pp(b->stream, c->d, c->geom, cur, 1 - cur, c->ghost_grid);
            }
        }
        launch_lj_full_batch(b->stream, b->d_desc[phase], b->d_rep, b->d_base, b->nblocks, b->ntypes1, b->unit, 0,
                             b->max_wg);
        BATCH_OK(hipStreamEndCapture(b->stream, &g));
        const hipError_t e = hipGraphInstantiate(&b->graph[phase], g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) { b->graph[phase] = nullptr; b->err = "batch graph instantiation failed"; return PAIRS_HIP_ERR_HIP; }
        b->graph_blocks[phase] = b->nblocks;
        b->graph_wg[phase] = b->max_wg;
    }
    BATCH_OK(hipGraphLaunch(b->graph[phase], b->stream));
    return 0;
}

extern "C" {

int pairs_hip_batch_create(pairs_hip_batch **out, int nreplicas, const pairs_hip_md_config *cfgs) {
    if (!out || nreplicas <= 0 || !cfgs) return PAIRS_HIP_ERR_ARG;
    *out = nullptr;
    for (int r = 1; r < nreplicas; r++)
        if (cfgs[r].reneighbor_every != cfgs[0].reneighbor_every || cfgs[r].thermo_every != cfgs[0].thermo_every ||
            cfgs[r].half || cfgs[0].half || cfgs[r].ntypes != cfgs[0].ntypes)
            return PAIRS_HIP_ERR_ARG;
    pairs_hip_batch *b = new pairs_hip_batch();
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) { delete b; return PAIRS_HIP_ERR_HIP; }
    b->ntypes1 = cfgs[0].ntypes == 1;
    for (int r = 0; r < nreplicas; r++) {
        pairs_hip_md_config c = cfgs[r];
        c.stream = (void *) b->stream;
        pairs_hip_ctx *ctx = nullptr;
        const int e = pairs_hip_create(&ctx, &c);
        if (e) { pairs_hip_batch_destroy(b); return e; }
        b->rep.push_back(ctx);
        b->unit = b->unit && b->ntypes1 && ctx->prm.eps[0] == 1.0 && ctx->prm.sig6[0] == 1.0;
    }
    const size_t R = nreplicas;
    if (hipHostMalloc((void **) &b->h_desc, 2 * R * sizeof(LJArgs)) != hipSuccess ||
        hipMalloc((void **) &b->d_desc[0], R * sizeof(LJArgs)) != hipSuccess ||
        hipMalloc((void **) &b->d_desc[1], R * sizeof(LJArgs)) != hipSuccess ||
        hipMalloc((void **) &b->d_desc[2], R * sizeof(LJArgs)) != hipSuccess ||
        hipMalloc((void **) &b->d_base, (R + 1) * sizeof(int)) != hipSuccess ||
        hipMalloc((void **) &b->d_sbase, (R + 1) * sizeof(int)) != hipSuccess) {
        pairs_hip_batch_destroy(b);
        return PAIRS_HIP_ERR_HIP;
    }
    if (hipHostMalloc((void **) &b->h_cnt, R * C_COUNT * sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **) &b->h_rdesc, R * sizeof(RepDesc)) != hipSuccess ||
        hipMalloc((void **) &b->d_rdesc, R * sizeof(RepDesc)) != hipSuccess ||
        hipHostMalloc((void **) &b->h_thermo, 5 * R * sizeof(double)) != hipSuccess ||
        hipMalloc((void **) &b->d_thermo, 5 * R * sizeof(double)) != hipSuccess ||
        hipHostMalloc((void **) &b->h_bdesc, R * sizeof(BuildArgs)) != hipSuccess ||
        hipMalloc((void **) &b->d_bdesc, R * sizeof(BuildArgs)) != hipSuccess ||
        hipHostMalloc((void **) &b->h_cptr, R * sizeof(int *)) != hipSuccess ||
        hipMalloc((void **) &b->d_cptr, R * sizeof(int *)) != hipSuccess ||
        hipMalloc((void **) &b->d_call, R * C_COUNT * sizeof(int)) != hipSuccess) {
        pairs_hip_batch_destroy(b);
        return PAIRS_HIP_ERR_HIP;
    }
    b->thermo.assign(5 * R, 0.0);
    *out = b;
    return 0;
}

int pairs_hip_batch_destroy(pairs_hip_batch *b) {
    if (!b) return 0;
    if (b->stream) hipStreamSynchronize(b->stream);
    batch_drop_graphs(b);
    for (pairs_hip_ctx *c : b->rep) pairs_hip_destroy(c);
    if (b->h_desc) hipHostFree(b->h_desc);
    for (int p = 0; p < 3; p++) if (b->d_desc[p]) hipFree(b->d_desc[p]);
    if (b->d_rep) hipFree(b->d_rep);
    if (b->d_base) hipFree(b->d_base);
    if (b->d_srep) hipFree(b->d_srep);
    if (b->d_sbase) hipFree(b->d_sbase);
    if (b->h_cnt) hipHostFree(b->h_cnt);
    if (b->h_rdesc) hipHostFree(b->h_rdesc);
    if (b->d_rdesc) hipFree(b->d_rdesc);
    if (b->h_thermo) hipHostFree(b->h_thermo);
    if (b->d_thermo) hipFree(b->d_thermo);
    if (b->h_bdesc) hipHostFree(b->h_bdesc);
    if (b->d_bdesc) hipFree(b->d_bdesc);
    if (b->h_cptr) hipHostFree(b->h_cptr);
    if (b->d_cptr) hipFree(b->d_cptr);
    if (b->d_call) hipFree(b->d_call);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

const char *pairs_hip_batch_last_error(const pairs_hip_batch *b) { return b ? b->err.c_str() : "null batch"; }

pairs_hip_ctx *pairs_hip_batch_replica(pairs_hip_batch *b, int r) {
    if (!b || r < 0 || r >= (int) b->rep.size()) return nullptr;
    return b->rep[r];
}

int pairs_hip_batch_run(pairs_hip_batch *b, int ts_begin, int ts_end, int nsteps) {
    if (!b || ts_begin < 0 || ts_end > nsteps + 1 || ts_begin > ts_end) return PAIRS_HIP_ERR_ARG;
    const int R = (int) b->rep.size();
    hipStream_t s = b->stream;
    pairs_hip_ctx *c0 = b->rep[0];
    bool prev_in_other = false;
    int phase = 0;
    if (b->nblocks == 0) {  // first use: tables and descriptors from the current segmentation
        int e = batch_tables(b);
        if (!e) e = batch_descs(b);
        if (e) return e;
    }
    for (int ts = ts_begin; ts < ts_end; ts++) {
        const bool Rb = every(ts, c0->cfg.reneighbor_every);
        const bool T = every(ts, c0->cfg.thermo_every);
        if (ts == ts_begin && ts > 0)
            for (pairs_hip_ctx *c : b->rep) launch_vv_initial(s, c->d, c->prm);
        const int vv2 = ts > 0, vv1 = (ts + 1 < ts_end) && (ts + 1 <= nsteps);
        if (Rb) {
            const int e = batch_rebuild(b);
            if (e) return e;
            prev_in_other = false;
            phase = 0;
        }
        if (!Rb && !prev_in_other) {
            // the first step of a run: ghosts refreshed in place, and the
            // descriptors follow each replica's current buffer
            for (pairs_hip_ctx *c : b->rep) launch_sync_inplace(s, c->d, c->geom, c->ghost_grid, 0);
            const int e = batch_descs(b);
            if (e) return e;
            phase = 0;
        }
        if (prev_in_other && !Rb) {
            for (int r = 0; r < R; r++) {
                pairs_hip_ctx *c = b->rep[r];
                if (!c->d.img_on && (!vv1 || T || !vv2)) {
                    const int cur = b->base_cur[r] ^ phase;
                    launch_sync_pp(s, c->d, c->geom, cur, 1 - cur, c->ghost_grid);
                }
            }
        }
        if (vv1 && vv2 && !T && prev_in_other && !Rb) {
            const int e = batch_graph_launch(b, phase);
            if (e) return e;
        } else {
            int max_cap = 0;
            for (int r = 0; r < R; r++) {
                pairs_hip_ctx *c = b->rep[r];
                const int cur = b->base_cur[r] ^ phase;
                b->h_desc[r] = lj_args(c->d, c->geom, c->prm, cur, 1 - cur, vv2, vv1, 1, nullptr, true);
                max_cap = std::max(max_cap, c->d.cap);
            }
            BATCH_OK(hipMemcpyAsync(b->d_desc[2], b->h_desc, R * sizeof(LJArgs), hipMemcpyHostToDevice, s));
            BATCH_OK(hipStreamSynchronize(s));  // h_desc is reused by the next descriptor build
            // positions may have moved without their packed copy (a rebuild,
            // an in-place ghost refresh or initial integrate): pack them;
            // this launch writes the next buffer's copy for the steps after
            launch_pack_p_batch(s, b->d_desc[2], R, max_cap);
            launch_lj_full_batch(s, b->d_desc[2], b->d_rep, b->d_base, b->nblocks, b->ntypes1, b->unit, T ? 1 : 0,
                                 b->max_wg);
            if (T) {
                if (!b->rdesc_valid) {  // the replicas' arrays and sizes for the batched reduction
                    BATCH_OK(hipStreamSynchronize(s));
                    for (int r = 0; r < R; r++) {
                        pairs_hip_ctx *c = b->rep[r];
                        b->h_rdesc[r].d = c->d;
                        b->h_rdesc[r].g = c->geom;
                        for (int k = 0; k < 3; k++) b->h_rdesc[r].periodic[k] = c->cfg.periodic[k] ? 1 : 0;
                    }
                    BATCH_OK(hipMemcpyAsync(b->d_rdesc, b->h_rdesc, R * sizeof(RepDesc), hipMemcpyHostToDevice, s));
                    b->rdesc_valid = true;
                }
                launch_thermo_batch(s, b->d_rdesc, R, b->d_thermo);
            }
        }
        {
            hipError_t e_ = hipGetLastError();
            if (e_ != hipSuccess) { b->err = std::string("kernel launch: ") + hipGetErrorString(e_); return PAIRS_HIP_ERR_HIP; }
        }
        if (T) {
            BATCH_OK(hipMemcpyAsync(b->h_thermo, b->d_thermo, 5 * (size_t) R * sizeof(double), hipMemcpyDeviceToHost, s));
            BATCH_OK(hipStreamSynchronize(s));
            std::memcpy(b->thermo.data(), b->h_thermo, 5 * (size_t) R * sizeof(double));
        }
        if (vv1) {
            for (pairs_hip_ctx *c : b->rep) c->d.cur = 1 - c->d.cur;
            phase ^= 1;
            prev_in_other = true;
        }
    }
    BATCH_OK(hipStreamSynchronize(s));
    return 0;
}

int pairs_hip_batch_run_each(pairs_hip_batch *b, int ts_begin, int ts_end, int nsteps) {
    if (!b) return PAIRS_HIP_ERR_ARG;
    for (size_t r = 0; r < b->rep.size(); r++) {
        double th[5 * 16];
        const int n = pairs_hip_run(b->rep[r], ts_begin, ts_end, nsteps, th, 16);
        if (n < 0) { b->err = "replica " + std::to_string(r) + ": " + b->rep[r]->err; return n; }
        if (n > 0) for (int k = 0; k < 5; k++) b->thermo[5 * r + k] = th[5 * (std::min(n, 16) - 1) + k];
    }
    b->nblocks = 0;  // per-replica runs may have re-segmented: rebuild tables on the next batched run
    b->rdesc_valid = false;
    return 0;
}

int pairs_hip_batch_set_home(pairs_hip_batch *b, int home) {
    if (!b || home < 8) return PAIRS_HIP_ERR_ARG;
    b->pending_home = home;
    b->rdesc_valid = false;
    return 0;
}

int pairs_hip_batch_set_workgroups(pairs_hip_batch *b, int workgroups) {
    if (!b || workgroups < 0) return PAIRS_HIP_ERR_ARG;
    b->max_wg = workgroups;
    return 0;
}

int pairs_hip_batch_autotune(pairs_hip_batch *b, int ts_begin, int steps, int nsteps, int ncand, const int *cand,
                             double *ms_per_step, int *chosen, int *ts_next) {
    if (!b || ncand <= 0 || !cand || steps <= 0) return PAIRS_HIP_ERR_ARG;
    const int every_r = b->rep[0]->cfg.reneighbor_every;
    int best = -1;
    double best_ms = 1e300;
    int ts = ts_begin;
    hipEvent_t e0, e1;
    BATCH_OK(hipEventCreate(&e0));
    BATCH_OK(hipEventCreate(&e1));
    for (int k = 0; k < ncand; k++) {
        // advance to the next rebuild step (where a new segment size takes
        // effect), rebuild with the candidate, then time `steps` steps
        int tr = ts;
        while (tr <= nsteps && !every(tr, every_r)) tr++;
        if (tr + 1 + steps > nsteps + 1) break;
        int e = tr > ts ? pairs_hip_batch_run(b, ts, tr, nsteps) : 0;
        if (!e) e = pairs_hip_batch_set_home(b, cand[k]);
        if (!e) e = pairs_hip_batch_run(b, tr, tr + 1, nsteps);
        if (e) return e;
        BATCH_OK(hipEventRecord(e0, b->stream));
        e = pairs_hip_batch_run(b, tr + 1, tr + 1 + steps, nsteps);
        if (e) return e;
        BATCH_OK(hipEventRecord(e1, b->stream));
        BATCH_OK(hipEventSynchronize(e1));
        float ms = 0.f;
        BATCH_OK(hipEventElapsedTime(&ms, e0, e1));
        ts = tr + 1 + steps;
        const double per = ms / steps;
        if (ms_per_step) ms_per_step[k] = per;
        if (per < best_ms) { best_ms = per; best = k; }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (ts_next) *ts_next = ts;
    if (best < 0) { b->err = "autotune: not enough timesteps for the candidates"; return PAIRS_HIP_ERR_ARG; }
    if (chosen) *chosen = cand[best];
    return pairs_hip_batch_set_home(b, cand[best]);
}

int pairs_hip_batch_thermo(pairs_hip_batch *b, double *out) {
    if (!b || !out) return PAIRS_HIP_ERR_ARG;
    std::memcpy(out, b->thermo.data(), b->thermo.size() * sizeof(double));
    return 0;
}

}  // extern "C"
