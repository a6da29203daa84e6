// C ABI of the DEM hot path (include/pairs_hip_dem.h).
//
// The context replaces the reference's PairsSimulation for examples/dem.py
// (runtime/pairs.hpp:27-297): it owns every device array in the reference's
// particle order, keeps all counts in device memory, and replays one captured
// hipGraph per step (the step's kernel sequence does not depend on counts).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/pairs_hip_dem.h"
#include "dem_kernels.hpp"

using namespace dk;

struct pairs_hip_dem_ctx {
    pairs_hip_dem_config cfg;
    DemGeom geom;
    DemParams prm;
    DemPtrs p;
    int *counters = nullptr;
    int *h_counters = nullptr;  // pinned
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    int req[5] = {0, 0, 0, 0, 0};
    int nlocal = 0;             // host mirror of the locals (constant on one rank)
    int ncells_cap = 0;
    bool use_graphs = true;
    hipGraphExec_t graph = nullptr;
    int time_every = 0;
    // timed steps: before prepare, before the contact search, after it, after
    // the slot and the pair kernels, after the sums, after euler
    hipEvent_t ev[7] = {};
    double t_phase[6] = {};
    double t_sum = 0.0;
    int t_count = 0;
    std::mt19937 generator;     // runtime/dem_sc_grid.hpp:13-20: default seeded, shared by all grids
    // resize-and-retry: the locals' state at the start of a chunk of steps
    int check_every = 100;      // steps per chunk (one counter read per chunk)
    char *snap = nullptr;       // device checkpoint
    size_t snap_bytes = 0;
    int snap_n = 0, snap_k = 0;
    int snap_counters[D_COUNT] = {};
    CopyDesc *h_copy = nullptr; // pinned copy tables (checkpoint, restore)
    CopyDesc *d_copy = nullptr;
    int resizes = 0;            // capacity growths so far
};

#define HIP_OK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                              \
            return -1;                                                                                 \
        }                                                                                              \
    } while (0)

#define CHECK_LAUNCH()                                                                                 \
    do {                                                                                               \
        hipError_t e_ = hipGetLastError();                                                             \
        if (e_ != hipSuccess) {                                                                        \
            ctx->err = std::string("kernel launch: ") + hipGetErrorString(e_);                         \
            return -1;                                                                                 \
        }                                                                                              \
    } while (0)

template <typename T>
static int dalloc(pairs_hip_dem_ctx *ctx, T **p, size_t n) {
    *p = nullptr;
    HIP_OK(hipMalloc((void **) p, std::max<size_t>(n, 1) * sizeof(T)));
    HIP_OK(hipMemsetAsync(*p, 0, std::max<size_t>(n, 1) * sizeof(T), ctx->stream));
    return 0;
}

static void drop_graph(pairs_hip_dem_ctx *ctx) {
    if (ctx->graph) {
        if (ctx->stream) hipStreamSynchronize(ctx->stream);  // no replay pending
        hipGraphExecDestroy(ctx->graph);
    }
    ctx->graph = nullptr;
}

static void free_arrays(DemPtrs &p) {
    void *ptrs[] = {p.x, p.y, p.z, p.vx, p.vy, p.vz, p.wx, p.wy, p.wz, p.fx, p.fy, p.fz, p.tx, p.ty, p.tz, p.r, p.m,
                    p.nx, p.ny, p.nz, p.inv_i, p.rot_m, p.rot_q, p.uid, p.flags, p.type, p.shape, p.num_contacts,
                    p.c_uid, p.c_used, p.c_stick, p.c_tsd, p.c_ivm, p.cell_count, p.cell_start, p.cell_slot,
                    p.cell_particles, p.particle_cell, p.nshapes, p.send_map, p.send_mult, p.exchg_flag, p.copy_to,
                    p.blk, p.scan_tmp, p.buf, p.hbuf, p.cand_n, p.cand, p.cslot, p.cf};
    for (void *q : ptrs)
        if (q) hipFree(q);
    const int cap = p.cap, ccap = p.ccap, send_cap = p.send_cap;
    p = DemPtrs{};
    p.cap = cap; p.ccap = ccap; p.send_cap = send_cap;
}

static void free_all(pairs_hip_dem_ctx *ctx) {
    free_arrays(ctx->p);
    if (ctx->counters) hipFree(ctx->counters);
    if (ctx->h_counters) hipHostFree(ctx->h_counters);
    if (ctx->snap) hipFree(ctx->snap);
    if (ctx->d_copy) hipFree(ctx->d_copy);
    if (ctx->h_copy) hipHostFree(ctx->h_copy);
}

// every array of the particle, contact and send capacities (zeroed)
static int alloc_arrays(pairs_hip_dem_ctx *ctx, DemPtrs &p, size_t C, size_t K, size_t S) {
    p.cap = (int) C; p.ccap = (int) K; p.send_cap = (int) S;
    int r = 0;
    for (double **q : {&p.x, &p.y, &p.z, &p.vx, &p.vy, &p.vz, &p.wx, &p.wy, &p.wz, &p.fx, &p.fy, &p.fz, &p.tx, &p.ty,
                       &p.tz, &p.r, &p.m, &p.nx, &p.ny, &p.nz})
        r |= dalloc(ctx, q, C);
    r |= dalloc(ctx, &p.inv_i, 9 * C);
    r |= dalloc(ctx, &p.rot_m, 9 * C);
    r |= dalloc(ctx, &p.rot_q, 4 * C);
    for (int **q : {&p.uid, &p.flags, &p.type, &p.shape, &p.num_contacts, &p.particle_cell, &p.exchg_flag,
                    &p.cell_slot})
        r |= dalloc(ctx, q, C);
    r |= dalloc(ctx, &p.c_uid, K * C);
    r |= dalloc(ctx, &p.c_used, K * C);
    r |= dalloc(ctx, &p.c_stick, K * C);
    r |= dalloc(ctx, &p.c_tsd, 3 * K * C);
    r |= dalloc(ctx, &p.c_ivm, K * C);
    const size_t NC = (size_t) ctx->ncells_cap;
    r |= dalloc(ctx, &p.cell_count, NC + 1);
    r |= dalloc(ctx, &p.cell_start, NC + 1);
    r |= dalloc(ctx, &p.nshapes, 2 * NC);
    r |= dalloc(ctx, &p.cell_particles, C);
    r |= dalloc(ctx, &p.scan_tmp, NC / 1024 + 8);
    r |= dalloc(ctx, &p.send_map, S);
    r |= dalloc(ctx, &p.send_mult, 3 * S);
    r |= dalloc(ctx, &p.copy_to, S);
    r |= dalloc(ctx, &p.blk, 2 * (C / 256 + 2));
    r |= dalloc(ctx, &p.buf, 40 * S);
    r |= dalloc(ctx, &p.hbuf, (2 + 5 * K) * S);
    r |= dalloc(ctx, &p.cand_n, C);
    r |= dalloc(ctx, &p.cand, DCAND * C);
    r |= dalloc(ctx, &p.cslot, DCAND * C);
    r |= dalloc(ctx, &p.cf, 6 * DCAND * C);
    return r ? -1 : 0;
}

// The state a chunk of steps starts from: the locals' rows of every array a
// step reads before writing (positions, velocities, angular velocities,
// radius, mass, normals, inertia, rotation, uid, flags, type, shape, contact
// counts and the contact history), packed at pitch n, and the counters.
// Ghosts, cells and forces are rebuilt by every step.
template <typename F>
static void state_fields(DemPtrs &p, F &&f) {
    for (double *a : {p.x, p.y, p.z, p.vx, p.vy, p.vz, p.wx, p.wy, p.wz, p.r, p.m, p.nx, p.ny, p.nz})
        f((void *) a, 8, 1);
    f((void *) p.inv_i, 8, 9);
    f((void *) p.rot_m, 8, 9);
    f((void *) p.rot_q, 8, 4);
    for (int *a : {p.uid, p.flags, p.type, p.shape, p.num_contacts}) f((void *) a, 4, 1);
    f((void *) p.c_uid, 4, -1);
    f((void *) p.c_used, 4, -1);
    f((void *) p.c_stick, 4, -1);
    f((void *) p.c_tsd, 8, -3);
    f((void *) p.c_ivm, 8, -1);
}
constexpr int MAX_COPY = 32;

// checkpoint (to_snap) or restore the locals' state; rows of the contact
// arrays: min(k, capacity) (restore: rows past the checkpoint stay zero)
static int snap_copy(pairs_hip_dem_ctx *ctx, bool to_snap) {
    DemPtrs &p = ctx->p;
    const int n = to_snap ? ctx->h_counters[D_NLOCAL] : ctx->snap_n;
    const int k = to_snap ? p.ccap : ctx->snap_k;
    if (to_snap) {
        size_t bytes = 0;
        state_fields(p, [&](void *, int el, int rows) { bytes += (size_t) el * n * (rows > 0 ? rows : -rows * k); });
        bytes = std::max<size_t>(bytes, 256);
        if (bytes > ctx->snap_bytes) {
            if (ctx->snap) hipFree(ctx->snap);
            ctx->snap = nullptr;
            ctx->snap_bytes = 0;
            HIP_OK(hipMalloc((void **) &ctx->snap, bytes + bytes / 4));
            ctx->snap_bytes = bytes + bytes / 4;
        }
        ctx->snap_n = n;
        ctx->snap_k = k;
        std::memcpy(ctx->snap_counters, ctx->h_counters, sizeof(ctx->snap_counters));
    }
    if (!ctx->h_copy) {
        HIP_OK(hipHostMalloc((void **) &ctx->h_copy, 2 * MAX_COPY * sizeof(CopyDesc)));
        HIP_OK(hipMalloc((void **) &ctx->d_copy, 2 * MAX_COPY * sizeof(CopyDesc)));
    }
    CopyDesc *tab = ctx->h_copy + (to_snap ? 0 : MAX_COPY);
    int nt = 0;
    size_t off = 0;
    state_fields(p, [&](void *a, int el, int rows) {
        const int nrows = rows > 0 ? rows : -rows * std::min(k, p.ccap);
        const int srows = rows > 0 ? rows : -rows * k;  // rows held in the checkpoint
        char *dev = (char *) a, *sn = ctx->snap + off;
        CopyDesc &c = tab[nt++];
        c.src = to_snap ? dev : sn;
        c.dst = to_snap ? sn : dev;
        c.words = (int) ((size_t) el * n / 4);
        c.rows = nrows;
        c.src_pitch = to_snap ? (size_t) el * p.cap / 4 : (size_t) el * n / 4;
        c.dst_pitch = to_snap ? (size_t) el * n / 4 : (size_t) el * p.cap / 4;
        off += (size_t) el * n * srows;
    });
    HIP_OK(hipMemcpyAsync(ctx->d_copy + (to_snap ? 0 : MAX_COPY), tab, nt * sizeof(CopyDesc), hipMemcpyHostToDevice,
                          ctx->stream));
    launch_copy_rows(ctx->stream, ctx->d_copy + (to_snap ? 0 : MAX_COPY), nt, n);
    CHECK_LAUNCH();
    if (!to_snap) {
        // the counters of the checkpoint, overflow flags cleared
        std::memcpy(ctx->h_counters, ctx->snap_counters, sizeof(ctx->snap_counters));
        ctx->h_counters[D_OVF] = 0;
        ctx->h_counters[D_REQ_SEND] = ctx->h_counters[D_REQ_CONTACT] = ctx->h_counters[D_REQ_PART] = 0;
        HIP_OK(hipMemcpyAsync(ctx->counters, ctx->h_counters, D_COUNT * sizeof(int), hipMemcpyHostToDevice,
                              ctx->stream));
    }
    // the pinned table and counters must not change before the copies ran
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// after an overflow: every array re-allocated with the overflowing capacity
// at twice the required size (the reference's grow rule,
// src/pairs/transformations/modules.py:159-205), the checkpoint restored into
// them and the step graph dropped
static int grow_and_restore(pairs_hip_dem_ctx *ctx, int code) {
    DemPtrs &p = ctx->p;
    size_t C = p.cap, K = p.ccap, S = p.send_cap;
    const int need = ctx->req[code];
    if (code == PAIRS_HIP_DEM_OVERFLOW_PARTICLES) C = std::max(C + 1, 2 * (size_t) need);
    else if (code == PAIRS_HIP_DEM_OVERFLOW_CONTACTS) K = std::max(K + 1, 2 * (size_t) need);
    else if (code == PAIRS_HIP_DEM_OVERFLOW_SEND) S = std::max(S + 1, 2 * (size_t) need);
    else return code;  // the cell grid is fixed by the domain
    HIP_OK(hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    free_arrays(p);
    if (alloc_arrays(ctx, p, C, K, S)) {
        ctx->err = "re-allocation after a capacity overflow failed";
        return -1;
    }
    if (snap_copy(ctx, false)) return -1;
    ctx->resizes++;
    return 0;
}

// build_cell_lists_stencil (src/pairs/sim/cell_lists.py:46-87)
static int setup_geometry(pairs_hip_dem_ctx *ctx) {
    DemGeom &g = ctx->geom;
    const pairs_hip_dem_config &c = ctx->cfg;
    const double s = c.cell_width;
    long total = 1;
    for (int d = 0; d < 3; d++) {
        g.lo[d] = c.box_lo[d];
        g.hi[d] = c.box_hi[d];
        g.len[d] = c.box_hi[d] - c.box_lo[d];
        g.dim_cells[d] = (int) std::ceil(((g.hi[d] + s) - (g.lo[d] - s)) / s) + 1;
        g.periodic[d] = c.periodic[d] ? 1 : 0;
        total *= g.dim_cells[d];
    }
    g.width = s;
    g.ncells = (int) total + 1;
    if (g.ncells + 1 > ctx->ncells_cap) {
        ctx->req[PAIRS_HIP_DEM_OVERFLOW_CELLS] = g.ncells + 1;
        ctx->err = "ncells_capacity below the cell grid";
        return PAIRS_HIP_DEM_OVERFLOW_CELLS;
    }
    int k = 0;
    for (int a = -1; a < 2; a++)
        for (int b = -1; b < 2; b++)
            for (int e = -1; e < 2; e++) g.stencil[k++] = (a * g.dim_cells[1] + b) * g.dim_cells[2] + e;
    return 0;
}

extern "C" {

int pairs_hip_dem_create(pairs_hip_dem_ctx **out, const pairs_hip_dem_config *cfg) {
    if (!out || !cfg) return -2;
    *out = nullptr;
    if (cfg->particle_capacity <= 0 || cfg->contact_capacity <= 0 || cfg->send_capacity <= 0 ||
        cfg->ncells_capacity <= 1 || cfg->cell_width <= 0.0 || cfg->ntypes <= 0 || cfg->ntypes > 4)
        return -2;
    pairs_hip_dem_ctx *ctx = new pairs_hip_dem_ctx();
    ctx->cfg = *cfg;
    auto fail = [&](int code) { free_all(ctx); if (ctx->own_stream) hipStreamDestroy(ctx->stream); *out = nullptr; delete ctx; return code; };
    if (cfg->stream) {
        ctx->stream = (hipStream_t) cfg->stream;
    } else {
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return fail(-1);
        ctx->own_stream = true;
    }
    ctx->ncells_cap = cfg->ncells_capacity;
    DemParams &prm = ctx->prm;
    prm.dt = cfg->dt;
    prm.pi = 3.141592653589793;
    prm.kappa = cfg->kappa;
    prm.sqrt_kappa = std::sqrt(cfg->kappa);
    prm.ln_e = cfg->ln_restitution;
    prm.t_c = cfg->collision_time;
    prm.kn_fixed = cfg->stiffness_norm > 0.0 ? cfg->stiffness_norm : 0.0;
    prm.sqrt_pl = std::sqrt(prm.pi * prm.pi + prm.ln_e * prm.ln_e);
    prm.drho = cfg->density_particle - cfg->density_fluid;
    prm.grav = cfg->gravity;
    prm.ntypes = cfg->ntypes;
    for (int k = 0; k < 16; k++) { prm.mu_s[k] = cfg->friction_static[k]; prm.mu_d[k] = cfg->friction_dynamic[k]; }
    int r = alloc_arrays(ctx, ctx->p, (size_t) cfg->particle_capacity, (size_t) cfg->contact_capacity,
                         (size_t) cfg->send_capacity);
    r |= dalloc(ctx, &ctx->counters, D_COUNT);
    if (r) return fail(-1);
    if (hipHostMalloc((void **) &ctx->h_counters, D_COUNT * sizeof(int)) != hipSuccess) return fail(-1);
    if (int e = setup_geometry(ctx)) return fail(e);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(-1);
    *out = ctx;
    return 0;
}

int pairs_hip_dem_destroy(pairs_hip_dem_ctx *ctx) {
    if (!ctx) return 0;
    hipStreamSynchronize(ctx->stream);
    drop_graph(ctx);
    for (hipEvent_t e : ctx->ev)
        if (e) hipEventDestroy(e);
    free_all(ctx);
    if (ctx->own_stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

const char *pairs_hip_dem_last_error(const pairs_hip_dem_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pairs_hip_dem_last_required(const pairs_hip_dem_ctx *ctx, int which) {
    if (!ctx || which < 1 || which > 4) return -2;
    return ctx->req[which];
}

// host staging of appended locals: write rows [nlocal, nlocal + n)
static int upload_rows(pairs_hip_dem_ctx *ctx, int n, const std::vector<int> &uid, const std::vector<int> &type,
                       const std::vector<int> &flags, const std::vector<int> &shape, const std::vector<double> &r,
                       const std::vector<double> &m, const std::vector<double> &pos, const std::vector<double> &vel,
                       const std::vector<double> &nrm) {
    if (n == 0) return 0;
    const int o = ctx->nlocal;
    if (o + n > ctx->p.cap) {
        ctx->req[PAIRS_HIP_DEM_OVERFLOW_PARTICLES] = o + n;
        ctx->err = "particle capacity exceeded";
        return PAIRS_HIP_DEM_OVERFLOW_PARTICLES;
    }
    DemPtrs &p = ctx->p;
    hipStream_t s = ctx->stream;
    auto up_i = [&](int *d, const std::vector<int> &h) { return hipMemcpyAsync(d + o, h.data(), n * sizeof(int), hipMemcpyHostToDevice, s); };
    auto up_d = [&](double *d, const double *h) { return hipMemcpyAsync(d + o, h, n * sizeof(double), hipMemcpyHostToDevice, s); };
    HIP_OK(up_i(p.uid, uid)); HIP_OK(up_i(p.type, type)); HIP_OK(up_i(p.flags, flags)); HIP_OK(up_i(p.shape, shape));
    HIP_OK(up_d(p.r, r.data())); HIP_OK(up_d(p.m, m.data()));
    std::vector<double> col(n);
    double *dst[3][3] = {{p.x, p.y, p.z}, {p.vx, p.vy, p.vz}, {p.nx, p.ny, p.nz}};
    const std::vector<double> *src[3] = {&pos, &vel, &nrm};
    for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++) {
            for (int i = 0; i < n; i++) col[i] = (*src[a])[3 * i + d];
            HIP_OK(up_d(dst[a][d], col.data()));
            HIP_OK(hipStreamSynchronize(s));  // col is reused
        }
    ctx->nlocal = o + n;
    ctx->h_counters[D_NLOCAL] = ctx->nlocal;
    HIP_OK(hipMemcpyAsync(ctx->counters + D_NLOCAL, &ctx->h_counters[D_NLOCAL], sizeof(int), hipMemcpyHostToDevice, s));
    HIP_OK(hipStreamSynchronize(s));
    drop_graph(ctx);
    return 0;
}

static bool within_subdomain(const pairs_hip_dem_ctx *ctx, const double *q) {
    // Regular6DStencil::isWithinSubdomain on one rank, SMALL = 1e-5
    // (runtime/domain/regular_6d_stencil.cpp:94-98)
    const DemGeom &g = ctx->geom;
    return q[0] >= g.lo[0] && q[0] < g.hi[0] - 0.00001 && q[1] >= g.lo[1] && q[1] < g.hi[1] - 0.00001 &&
           q[2] >= g.lo[2] && q[2] < g.hi[2] - 0.00001;
}

int pairs_hip_dem_sc_grid(pairs_hip_dem_ctx *ctx, double xmax, double ymax, double zmax, double spacing,
                          double diameter, double min_diameter, double max_diameter, double initial_velocity,
                          double particle_density, int ntypes) {
    if (!ctx || spacing <= 0.0 || ntypes <= 0) return -2;
    auto real_random = [&](double a, double b) {
        std::uniform_real_distribution<double> dist(a, b);
        return dist(ctx->generator);
    };
    auto in_box = [](const double *q, const double *box) {
        return q[0] >= box[0] && q[0] < box[3] && q[1] >= box[1] && q[1] < box[4] && q[2] >= box[2] && q[2] < box[5];
    };
    std::vector<int> uid, type, flags, shape;
    std::vector<double> r, m, pos, vel, nrm;
    const double gen[6] = {0.0, 0.0, diameter, xmax, ymax, zmax};
    const double ref[3] = {spacing * 0.5, spacing * 0.5, spacing * 0.5};
    const int iret = (int) std::ceil((0.0 - ref[0]) / spacing);
    const int jret = (int) std::ceil((0.0 - ref[1]) / spacing);
    const int kret = (int) std::ceil((diameter - ref[2]) / spacing);
    int i = iret, j = jret, k = kret, last_uid = 1;
    double q[3] = {ref[0] + i * spacing, ref[1] + j * spacing, ref[2] + k * spacing};
    while (in_box(q, gen)) {
        const double dia = real_random(min_diameter, max_diameter);
        if (within_subdomain(ctx, q)) {
            const double rad = dia * 0.5;
            uid.push_back(last_uid);
            r.push_back(rad);
            m.push_back(((4.0 / 3.0) * M_PI) * rad * rad * rad * particle_density);
            for (int d = 0; d < 3; d++) { pos.push_back(q[d]); nrm.push_back(0.0); }
            vel.push_back(0.1 * real_random(-initial_velocity, initial_velocity));
            vel.push_back(0.1 * real_random(-initial_velocity, initial_velocity));
            vel.push_back(-initial_velocity);
            type.push_back(rand() % ntypes);
            flags.push_back(0);
            shape.push_back(PAIRS_HIP_DEM_SPHERE);
        }
        ++i;
        q[0] = ref[0] + i * spacing; q[1] = ref[1] + j * spacing; q[2] = ref[2] + k * spacing;
        if (!in_box(q, gen)) {
            i = iret;
            j++;
            q[0] = ref[0] + i * spacing; q[1] = ref[1] + j * spacing; q[2] = ref[2] + k * spacing;
            if (!in_box(q, gen)) {
                j = jret;
                k++;
                q[0] = ref[0] + i * spacing; q[1] = ref[1] + j * spacing; q[2] = ref[2] + k * spacing;
                if (!in_box(q, gen)) break;
            }
        }
        last_uid++;
    }
    const int e = upload_rows(ctx, (int) uid.size(), uid, type, flags, shape, r, m, pos, vel, nrm);
    return e ? e : ctx->nlocal;
}

int pairs_hip_dem_jittered_grid(pairs_hip_dem_ctx *ctx, int nx, int ny, int nz, const double origin[3],
                                const double spacing[3], double jitter, double r_min, double r_max,
                                double density, unsigned long long seed) {
    if (!ctx || !origin || !spacing || nx < 0 || ny < 0 || nz < 0 || r_min <= 0.0 || r_max < r_min || jitter < 0.0)
        return -2;
    std::mt19937_64 gen(seed);
    auto unit = [&]() { return (double) (gen() >> 11) * 0x1.0p-53; };
    std::vector<int> uid, type, flags, shape;
    std::vector<double> r, m, pos, vel, nrm;
    int next_uid = 1;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int site[3] = {i, j, k};
                double q[3];
                for (int d = 0; d < 3; d++) q[d] = (origin[d] + site[d] * spacing[d]) + (2.0 * unit() - 1.0) * jitter;
                const double rad = r_min + (r_max - r_min) * unit();
                const int u = next_uid++;
                if (!within_subdomain(ctx, q)) continue;
                uid.push_back(u);
                r.push_back(rad);
                m.push_back(((4.0 / 3.0) * M_PI) * rad * rad * rad * density);
                for (int d = 0; d < 3; d++) { pos.push_back(q[d]); nrm.push_back(0.0); vel.push_back(0.0); }
                type.push_back(0);
                flags.push_back(0);
                shape.push_back(PAIRS_HIP_DEM_SPHERE);
            }
    const int e = upload_rows(ctx, (int) uid.size(), uid, type, flags, shape, r, m, pos, vel, nrm);
    return e ? e : ctx->nlocal;
}

int pairs_hip_dem_add_halfspace(pairs_hip_dem_ctx *ctx, int uid, int type, double mass, const double position[3],
                                const double normal[3], int flags) {
    if (!ctx || !position || !normal) return -2;
    // read_particle_data keeps a row inside the subdomain or flagged
    // Infinite | Fixed | Global (runtime/read_from_file.hpp:108-110)
    if (!within_subdomain(ctx, position) && !(flags & (DFLAG_INFINITE | DFLAG_FIXED | DFLAG_GLOBAL))) return ctx->nlocal;
    std::vector<double> pos(position, position + 3), nrm(normal, normal + 3), vel(3, 0.0);
    const int e = upload_rows(ctx, 1, {uid}, {type}, {flags}, {PAIRS_HIP_DEM_HALFSPACE}, {0.0}, {mass}, pos, vel, nrm);
    return e ? e : ctx->nlocal;
}

int pairs_hip_dem_add_spheres(pairs_hip_dem_ctx *ctx, int n, const int *uid, const int *type, const double *radius,
                              const double *mass, const double *position, const double *linear_velocity,
                              const int *flags) {
    if (!ctx || n < 0 || (n > 0 && (!uid || !type || !radius || !mass || !position || !linear_velocity || !flags)))
        return -2;
    std::vector<int> u(uid, uid + n), t(type, type + n), f(flags, flags + n), sh(n, PAIRS_HIP_DEM_SPHERE);
    std::vector<double> r(radius, radius + n), m(mass, mass + n), pos(position, position + 3 * n),
        vel(linear_velocity, linear_velocity + 3 * n), nrm(3 * (size_t) n, 0.0);
    const int e = upload_rows(ctx, n, u, t, f, sh, r, m, pos, vel, nrm);
    return e ? e : ctx->nlocal;
}

int pairs_hip_dem_setup(pairs_hip_dem_ctx *ctx) {
    if (!ctx) return -2;
    launch_mass_inertia(ctx->stream, ctx->p, ctx->nlocal);
    CHECK_LAUNCH();
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return 0;
}

static int read_counters(pairs_hip_dem_ctx *ctx) {
    HIP_OK(hipMemcpyAsync(ctx->h_counters, ctx->counters, D_COUNT * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return 0;
}

static int overflow_code(pairs_hip_dem_ctx *ctx) {
    const int *h = ctx->h_counters;
    switch (h[D_OVF]) {
    case DOVF_SEND: ctx->req[PAIRS_HIP_DEM_OVERFLOW_SEND] = h[D_REQ_SEND]; ctx->err = "send capacity exceeded"; return PAIRS_HIP_DEM_OVERFLOW_SEND;
    case DOVF_PART: ctx->req[PAIRS_HIP_DEM_OVERFLOW_PARTICLES] = h[D_REQ_PART]; ctx->err = "particle capacity exceeded"; return PAIRS_HIP_DEM_OVERFLOW_PARTICLES;
    case DOVF_CONTACT: ctx->req[PAIRS_HIP_DEM_OVERFLOW_CONTACTS] = h[D_REQ_CONTACT]; ctx->err = "contact capacity exceeded"; return PAIRS_HIP_DEM_OVERFLOW_CONTACTS;
    default: return 0;
    }
}

static void enqueue_step(pairs_hip_dem_ctx *ctx, bool timed) {
    if (timed) hipEventRecord(ctx->ev[0], ctx->stream);
    launch_prepare(ctx->stream, ctx->p, ctx->counters, ctx->geom);
    if (timed) hipEventRecord(ctx->ev[1], ctx->stream);
    launch_contacts(ctx->stream, ctx->p, ctx->counters, ctx->geom, ctx->prm, timed ? ctx->ev + 2 : nullptr);
    if (timed) hipEventRecord(ctx->ev[5], ctx->stream);
    launch_euler(ctx->stream, ctx->p, ctx->counters, ctx->prm);
    if (timed) hipEventRecord(ctx->ev[6], ctx->stream);
}

static int ensure_graph(pairs_hip_dem_ctx *ctx) {
    if (!ctx->use_graphs || ctx->graph) return 0;
    hipGraph_t g;
    HIP_OK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(ctx, false);
    HIP_OK(hipStreamEndCapture(ctx->stream, &g));
    const hipError_t e = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) { ctx->graph = nullptr; ctx->err = "graph instantiation failed"; return -1; }
    return 0;
}

// steps [a, b) without a counter read; t0: the first timestep of the run
// (the timing cadence)
static int enqueue_steps(pairs_hip_dem_ctx *ctx, int a, int b, int t0) {
    if (int e = ensure_graph(ctx)) return e;
    for (int ts = a; ts < b; ts++) {
        const bool timed = ctx->time_every > 0 && (ts - t0) % ctx->time_every == ctx->time_every - 1;
        if (timed) {
            enqueue_step(ctx, true);
            CHECK_LAUNCH();
            HIP_OK(hipEventSynchronize(ctx->ev[6]));
            for (int k = 0; k < 6; k++) {
                float ms = 0.0f;
                HIP_OK(hipEventElapsedTime(&ms, ctx->ev[k], ctx->ev[k + 1]));
                ctx->t_phase[k] += ms;
            }
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[5]));
            ctx->t_sum += ms;
            ctx->t_count++;
        } else if (ctx->use_graphs) {
            HIP_OK(hipGraphLaunch(ctx->graph, ctx->stream));
        } else {
            enqueue_step(ctx, false);
            CHECK_LAUNCH();
        }
    }
    return 0;
}

// Steps [a, b) in chunks of check_every: the locals' state is checkpointed
// at the start of each chunk and the counters are read at its end; a chunk
// that overflowed a capacity (every later kernel of it then did nothing) is
// run again from its checkpoint with that capacity grown to twice the
// required size, as the reference re-runs a module after a resize
// (src/pairs/transformations/modules.py:159-205).  The results do not depend
// on the capacities, so the trajectory is that of a run that never overflowed.
static int run_checked(pairs_hip_dem_ctx *ctx, int a, int b, int t0) {
    if (read_counters(ctx)) return -1;
    if (int code = overflow_code(ctx)) return code;  // a previous call left an overflow
    for (int c0 = a; c0 < b;) {
        const int c1 = std::min(b, c0 + std::max(ctx->check_every, 1));
        if (snap_copy(ctx, true)) return -1;
        const double ts_sum = ctx->t_sum;
        const int ts_count = ctx->t_count;
        double tp[6];
        std::memcpy(tp, ctx->t_phase, sizeof(tp));
        int code = 0;
        for (int attempt = 0; attempt < 16; attempt++) {
            if (int e = enqueue_steps(ctx, c0, c1, t0)) return e;
            if (read_counters(ctx)) return -1;
            code = overflow_code(ctx);
            if (code == 0) break;
            ctx->t_sum = ts_sum;
            ctx->t_count = ts_count;
            std::memcpy(ctx->t_phase, tp, sizeof(tp));
            if (int e = grow_and_restore(ctx, code)) return e;
        }
        if (code) return code;
        c0 = c1;
    }
    return 0;
}

int pairs_hip_dem_step(pairs_hip_dem_ctx *ctx, int ts) {
    if (!ctx) return -2;
    const bool g = ctx->use_graphs;
    ctx->use_graphs = false;
    const int e = run_checked(ctx, ts, ts + 1, ts);
    ctx->use_graphs = g;
    return e;
}

int pairs_hip_dem_run(pairs_hip_dem_ctx *ctx, int ts_begin, int ts_end) {
    if (!ctx || ts_end < ts_begin) return -2;
    ctx->t_sum = 0.0;
    ctx->t_count = 0;
    for (double &t : ctx->t_phase) t = 0.0;
    return run_checked(ctx, ts_begin, ts_end, ts_begin);
}

int pairs_hip_dem_set_check_every(pairs_hip_dem_ctx *ctx, int steps) {
    if (!ctx || steps < 1) return -2;
    ctx->check_every = steps;
    return 0;
}

int pairs_hip_dem_set_graphs(pairs_hip_dem_ctx *ctx, int enable) {
    if (!ctx) return -2;
    ctx->use_graphs = enable != 0;
    if (!ctx->use_graphs) drop_graph(ctx);
    return 0;
}

int pairs_hip_dem_sync(pairs_hip_dem_ctx *ctx) {
    if (!ctx) return -2;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int pairs_hip_dem_count(pairs_hip_dem_ctx *ctx, int what) {
    if (!ctx) return -2;
    if (read_counters(ctx)) return -1;
    switch (what) {
    case PAIRS_HIP_DEM_NLOCAL: return ctx->h_counters[D_NLOCAL];
    case PAIRS_HIP_DEM_NGHOST: return ctx->h_counters[D_NGHOST];
    case PAIRS_HIP_DEM_NCELLS: return ctx->geom.ncells;
    case PAIRS_HIP_DEM_PARTICLE_CAPACITY: return ctx->p.cap;
    case PAIRS_HIP_DEM_CONTACT_CAPACITY: return ctx->p.ccap;
    case PAIRS_HIP_DEM_SEND_CAPACITY: return ctx->p.send_cap;
    case PAIRS_HIP_DEM_RESIZES: return ctx->resizes;
    case PAIRS_HIP_DEM_CONTACTS: {
        const int n = ctx->h_counters[D_NLOCAL];
        std::vector<int> nc(std::max(n, 1));
        HIP_OK(hipMemcpy(nc.data(), ctx->p.num_contacts, n * sizeof(int), hipMemcpyDeviceToHost));
        long s = 0;
        for (int i = 0; i < n; i++) s += nc[i];
        return (int) s;
    }
    default: return -2;
    }
}

int pairs_hip_dem_download(pairs_hip_dem_ctx *ctx, int what, void *host, int count) {
    if (!ctx || !host || count < 0) return -2;
    if (read_counters(ctx)) return -1;
    const int nl = ctx->h_counters[D_NLOCAL], nall = nl + ctx->h_counters[D_NGHOST];
    const DemPtrs &p = ctx->p;
    const size_t C = p.cap;
    // at most what exists: locals and ghosts, or the locals of local-only arrays
    const int n_all = std::min(count, nall), n_loc = std::min(count, nl);
    auto get_d = [&](const double *d, double *h, int n) { return hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost); };
    auto get_i = [&](const int *d, int *h, int n) -> int {
        HIP_OK(hipMemcpy(h, d, n * sizeof(int), hipMemcpyDeviceToHost));
        return n;
    };
    auto interleave = [&](std::initializer_list<const double *> cols, double *out, int n) -> int {
        const int nc = (int) cols.size();
        std::vector<double> tmp(std::max(n, 1));
        int c = 0;
        for (const double *col : cols) {
            HIP_OK(get_d(col, tmp.data(), n));
            for (int i = 0; i < n; i++) out[(size_t) i * nc + c] = tmp[i];
            c++;
        }
        return n;
    };
    double *hd = (double *) host;
    int *hi = (int *) host;
    switch (what) {
    case PAIRS_HIP_DEM_POSITION: return interleave({p.x, p.y, p.z}, hd, n_all);
    case PAIRS_HIP_DEM_LINEAR_VELOCITY: return interleave({p.vx, p.vy, p.vz}, hd, n_all);
    case PAIRS_HIP_DEM_ANGULAR_VELOCITY: return interleave({p.wx, p.wy, p.wz}, hd, n_all);
    case PAIRS_HIP_DEM_FORCE: return interleave({p.fx, p.fy, p.fz}, hd, n_loc);
    case PAIRS_HIP_DEM_TORQUE: return interleave({p.tx, p.ty, p.tz}, hd, n_loc);
    case PAIRS_HIP_DEM_RADIUS: HIP_OK(get_d(p.r, hd, n_all)); return n_all;
    case PAIRS_HIP_DEM_MASS: HIP_OK(get_d(p.m, hd, n_all)); return n_all;
    case PAIRS_HIP_DEM_ROTATION_QUAT:
        return interleave({p.rot_q, p.rot_q + C, p.rot_q + 2 * C, p.rot_q + 3 * C}, hd, n_loc);
    case PAIRS_HIP_DEM_ROTATION_MATRIX:
        return interleave({p.rot_m, p.rot_m + C, p.rot_m + 2 * C, p.rot_m + 3 * C, p.rot_m + 4 * C, p.rot_m + 5 * C,
                           p.rot_m + 6 * C, p.rot_m + 7 * C, p.rot_m + 8 * C}, hd, n_loc);
    case PAIRS_HIP_DEM_UID: return get_i(p.uid, hi, n_all);
    case PAIRS_HIP_DEM_FLAGS: return get_i(p.flags, hi, n_all);
    case PAIRS_HIP_DEM_TYPE: return get_i(p.type, hi, n_all);
    case PAIRS_HIP_DEM_SHAPE: return get_i(p.shape, hi, n_all);
    case PAIRS_HIP_DEM_NUM_CONTACTS: return get_i(p.num_contacts, hi, n_loc);
    default: return -2;
    }
}

int pairs_hip_dem_contacts(pairs_hip_dem_ctx *ctx, int *num_contacts, int *partner_uid, int *sticking,
                           double *tangential_displacement, double *impact_velocity) {
    if (!ctx || !num_contacts || !partner_uid || !sticking || !tangential_displacement || !impact_velocity) return -2;
    if (read_counters(ctx)) return -1;
    const int nl = ctx->h_counters[D_NLOCAL];
    const DemPtrs &p = ctx->p;
    const size_t C = p.cap, K = p.ccap;
    HIP_OK(hipMemcpy(num_contacts, p.num_contacts, nl * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> ti(std::max<size_t>(nl, 1));
    std::vector<double> td(std::max<size_t>(nl, 1));
    for (size_t k = 0; k < K; k++) {
        HIP_OK(hipMemcpy(ti.data(), p.c_uid + k * C, nl * sizeof(int), hipMemcpyDeviceToHost));
        for (int i = 0; i < nl; i++) partner_uid[i * K + k] = ti[i];
        HIP_OK(hipMemcpy(ti.data(), p.c_stick + k * C, nl * sizeof(int), hipMemcpyDeviceToHost));
        for (int i = 0; i < nl; i++) sticking[i * K + k] = ti[i];
        HIP_OK(hipMemcpy(td.data(), p.c_ivm + k * C, nl * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < nl; i++) impact_velocity[i * K + k] = td[i];
        for (int d = 0; d < 3; d++) {
            HIP_OK(hipMemcpy(td.data(), p.c_tsd + (3 * k + d) * C, nl * sizeof(double), hipMemcpyDeviceToHost));
            for (int i = 0; i < nl; i++) tangential_displacement[(i * K + k) * 3 + d] = td[i];
        }
    }
    return 0;
}

int pairs_hip_dem_set_timing(pairs_hip_dem_ctx *ctx, int every_n) {
    if (!ctx || every_n < 0) return -2;
    ctx->time_every = every_n;
    if (every_n > 0 && !ctx->ev[0])
        for (hipEvent_t &e : ctx->ev) HIP_OK(hipEventCreate(&e));
    return 0;
}

int pairs_hip_dem_contact_kernel_time(pairs_hip_dem_ctx *ctx, double *avg_ms, int *launches) {
    if (!ctx || !avg_ms || !launches) return -2;
    *avg_ms = ctx->t_count ? ctx->t_sum / ctx->t_count : 0.0;
    *launches = ctx->t_count;
    return 0;
}

int pairs_hip_dem_phase_times(pairs_hip_dem_ctx *ctx, double *avg_ms) {
    if (!ctx || !avg_ms) return -2;
    for (int k = 0; k < 6; k++) avg_ms[k] = ctx->t_count ? ctx->t_phase[k] / ctx->t_count : 0.0;
    return 0;
}

}  // extern "C"
