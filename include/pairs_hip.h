/*
 * pairs_hip.h -- C ABI of the MI355X (gfx950) particle hot path.
 *
 * This is the drop-in boundary for the modules that the reference's code
 * generator emits as CUDA kernels for the LJ molecular-dynamics path
 * (examples/md.py).  A host program (hand written, or emitted by the HIP code
 * generation target in pairs_amd/dsl.py, emit_cpp) calls these functions in the order of
 * the reference's timestep (src/pairs/sim/simulation.py:382-452):
 *
 *   pre-step  initial_integrate                  -> pairs_hip_vv_initial
 *   every R   exchange (3 dims)                  -> pairs_hip_exchange
 *             borders  (3 dims) | synchronize    -> pairs_hip_borders | pairs_hip_sync_ghosts
 *             build_cell_lists + partition       -> pairs_hip_bin_cells
 *             build_neighbor_lists               -> pairs_hip_build_neighbors
 *   always    reset_volatile_properties + LJ     -> pairs_hip_lj
 *             final_integrate                    -> pairs_hip_vv_final
 *   every T   compute_thermo                     -> pairs_hip_thermo
 *
 * or hands the whole loop to pairs_hip_run (fused, graph-captured kernels).
 *
 * Conventions (all functions):
 *   - extern "C", no exceptions cross the ABI, plain pointers and sizes;
 *   - return 0 on success, a positive PAIRS_HIP_OVERFLOW_* code when a
 *     capacity would be exceeded (the required size is then available from
 *     pairs_hip_last_required; the caller grows the capacity with
 *     pairs_hip_reserve and repeats the call -- the resize protocol of
 *     src/pairs/transformations/modules.py:159-205), and a negative code on a
 *     HIP or argument error (text from pairs_hip_last_error);
 *   - work is ordered on the context's stream; host buffers are borrowed only
 *     for the duration of the call; downloads synchronise.
 *   - one context per host thread; calls on one context are not thread-safe.
 */
#ifndef PAIRS_HIP_H
#define PAIRS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pairs_hip_ctx pairs_hip_ctx;

enum {
    PAIRS_HIP_OK = 0,
    PAIRS_HIP_OVERFLOW_PARTICLES = 1, /* particle_capacity (locals + ghosts) */
    PAIRS_HIP_OVERFLOW_NEIGHBORS = 2, /* neighbor_capacity per particle      */
    PAIRS_HIP_OVERFLOW_SEND = 3,      /* send_capacity (ghost/exchange maps) */
    PAIRS_HIP_OVERFLOW_CELLS = 4,     /* ncells_capacity                     */
    PAIRS_HIP_ERR_HIP = -1,
    PAIRS_HIP_ERR_ARG = -2,
    PAIRS_HIP_ERR_STATE = -3
};

/* Flags of a particle (src/pairs/sim/flags.py:1-5). */
enum { PAIRS_HIP_FLAG_INFINITE = 1, PAIRS_HIP_FLAG_GHOST = 2, PAIRS_HIP_FLAG_FIXED = 4, PAIRS_HIP_FLAG_GLOBAL = 8 };

/* What pairs_hip_download / pairs_hip_count read. */
enum {
    PAIRS_HIP_POSITION = 0,        /* f64 [n][3], locals then ghosts                 */
    PAIRS_HIP_VELOCITY = 1,        /* f64 [n][3]                                     */
    PAIRS_HIP_FORCE = 2,           /* f64 [nlocal][3]                                */
    PAIRS_HIP_MASS = 3,            /* f64 [n]                                        */
    PAIRS_HIP_TYPE = 4,            /* i32 [n]                                        */
    PAIRS_HIP_FLAGS = 5,           /* i32 [n]                                        */
    PAIRS_HIP_UID = 6,             /* i32 [n]                                        */
    PAIRS_HIP_NUMNEIGHS = 7,       /* i32 [nlocal]                                   */
    PAIRS_HIP_NEIGHBORLISTS = 8,   /* i32 [nlocal][neighbor_capacity] (AoS, CPU form) */
    PAIRS_HIP_PARTICLE_CELL = 9,   /* i32 [nlocal + nghost]                          */
    PAIRS_HIP_CELL_SIZES = 10,     /* i32 [ncells]                                   */
    PAIRS_HIP_CELL_PARTICLES = 11, /* i32 [nlocal + nghost], CSR by cell, index order */
    PAIRS_HIP_SEND_MAP = 12,       /* i32 [nsend_all]                                */
    PAIRS_HIP_SEND_MULT = 13,      /* i32 [nsend_all][3]                             */
    PAIRS_HIP_NSEND = 14,          /* i32 [6]                                        */
    PAIRS_HIP_ATOM_ENERGY = 15     /* f64 [nlocal]  (after pairs_hip_energy)         */
};
enum {
    PAIRS_HIP_NLOCAL = 0, PAIRS_HIP_NGHOST = 1, PAIRS_HIP_NCELLS = 2, PAIRS_HIP_NSEND_ALL = 3,
    PAIRS_HIP_NEIGHBOR_CAPACITY = 4, PAIRS_HIP_PARTICLE_CAPACITY = 5, PAIRS_HIP_SEND_CAPACITY = 6,
    PAIRS_HIP_NEIGHBOR_TOTAL = 7,
    PAIRS_HIP_TILE_CAPACITY = 8,  /* cell-tile entries a force workgroup stages (grows at rebuilds) */
    PAIRS_HIP_REBUILDS_DEFERRED = 9,  /* rebuilds of pairs_hip_run whose counters were checked later */
    PAIRS_HIP_REBUILD_ROLLBACKS = 10  /* of those, repeated after an overflow (state restored)      */
};

typedef struct pairs_hip_md_config {
    /* capacities (src/pairs/__init__.py:16-17, sim/cell_lists.py:30-31, sim/comm.py:24-27) */
    int particle_capacity;  /* locals + ghosts                         (reference 800000) */
    int neighbor_capacity;  /* per particle                            (reference 100)    */
    int send_capacity;      /* ghost / exchange send map               (reference 200000) */
    int ncells_capacity;    /*                                          (reference 100000) */
    /* domain (src/pairs/sim/domain.py:7-16, runtime/domain/regular_6d_stencil.cpp:56-82) */
    double box_lo[3], box_hi[3];
    int periodic[3];
    /* interaction (examples/md.py:25-31,59,63) */
    double cell_spacing;    /* cell edge == neighbour-list radius rc + skin */
    double cutoff;          /* force cutoff rc                               */
    double dt;
    int half;               /* 1: half list (compute_half, Q6), 0: full list */
    int ntypes;             /* feature "type": tables are [ntypes][ntypes]    */
    const double *epsilon;  /* ntypes*ntypes, index t_i*ntypes + t_j          */
    const double *sigma6;
    int reneighbor_every;   /* rebuild when ts==0 || (ts+1)%R==0             */
    int thermo_every;       /* thermo when ts==0 || (ts+1)%T==0 (0: never)   */
    void *stream;           /* hipStream_t, or NULL for a context-owned one  */
} pairs_hip_md_config;

/* Context: replaces pairs::PairsSimulation's registries and device copies
 * (runtime/pairs.hpp:27-297, runtime/pairs.cpp:31-327, runtime/devices/cuda.cu). */
int pairs_hip_create(pairs_hip_ctx **ctx, const pairs_hip_md_config *cfg);
int pairs_hip_destroy(pairs_hip_ctx *ctx);
const char *pairs_hip_last_error(const pairs_hip_ctx *ctx);
/* required size after a positive return (resizes[] of modules.py:70-126) */
int pairs_hip_last_required(const pairs_hip_ctx *ctx, int which_overflow);
/* grow capacities (0 keeps the current value); contents are preserved
 * (pairs->realloc* of runtime/pairs.hpp:328-503) */
int pairs_hip_reserve(pairs_hip_ctx *ctx, int particle_capacity, int neighbor_capacity, int send_capacity);

/* Initial state.  pairs::copper_fcc_lattice (runtime/copper_fcc_lattice.hpp:64-145)
 * generated on the host and uploaded; returns the number of locals (>= 0). */
int pairs_hip_copper_fcc_lattice(pairs_hip_ctx *ctx, int nx, int ny, int nz, double rho, int ntypes);
/* pairs::adjust_thermo (runtime/thermo.hpp:53-97) */
int pairs_hip_adjust_thermo(pairs_hip_ctx *ctx, double temperature);
/* Upload n locals (AoS [n][3] doubles).  m, type, flags, uid may be NULL
 * (defaults 1.0, 0, 0, 0).  Ghosts are cleared. */
int pairs_hip_upload(pairs_hip_ctx *ctx, int n, const double *position, const double *velocity,
                     const double *mass, const int *type, const int *flags, const int *uid);
/* Copy one array (PAIRS_HIP_POSITION ...) to the host; count elements of its
 * leading dimension are written (at most what exists).  Returns the count. */
int pairs_hip_download(pairs_hip_ctx *ctx, int what, void *host, int count);
int pairs_hip_count(pairs_hip_ctx *ctx, int what);

/* build_cell_lists_stencil (src/pairs/sim/cell_lists.py:46-87) */
int pairs_hip_setup_cells(pairs_hip_ctx *ctx);
/* Comm.exchange, all three dimensions (src/pairs/sim/comm.py:100-151,445-520) */
int pairs_hip_exchange(pairs_hip_ctx *ctx);
/* Comm.borders, all three dimensions (src/pairs/sim/comm.py:56-98,223-366) */
int pairs_hip_borders(pairs_hip_ctx *ctx);
/* Comm.synchronize (src/pairs/sim/comm.py:45-54,369-442), stale multi-hop ghosts */
int pairs_hip_sync_ghosts(pairs_hip_ctx *ctx);
/* BuildCellLists + PartitionCellLists (src/pairs/sim/cell_lists.py:90-171) */
int pairs_hip_bin_cells(pairs_hip_ctx *ctx);
/* BuildNeighborLists (src/pairs/sim/neighbor_lists.py:21-48) */
int pairs_hip_build_neighbors(pairs_hip_ctx *ctx);
/* reset_volatile_properties + lennard_jones (src/pairs/sim/properties.py:61-70,
 * examples/md.py:5-8 via src/pairs/sim/interaction.py:201-292, ir/apply.py:111-125) */
int pairs_hip_lj(pairs_hip_ctx *ctx);
/* initial_integrate / final_integrate (examples/md.py:11-17) */
int pairs_hip_vv_initial(pairs_hip_ctx *ctx);
int pairs_hip_vv_final(pairs_hip_ctx *ctx);
/* The modules of the older API (examples/lj_onetype.py:5-13):
 * reset_volatile_properties + `lj`, whose pair force is
 * delta * 48.0 * sr6 * (sr6 - 0.5) * sr2 * epsilon evaluated left to right
 * (full lists only), and `euler`: velocity += dt * force / mass, then
 * position += dt * velocity, for the non-fixed locals. */
int pairs_hip_lj_older(pairs_hip_ctx *ctx);
int pairs_hip_euler(pairs_hip_ctx *ctx);
/* compute_thermo (runtime/thermo.hpp:11-51): out[0] = T, out[1] = kinetic P,
 * out[2] = potential energy, out[3] = virial W, out[4] = P with virial
 * (out[2..4] have no reference counterpart: U = sum 4eps[(s/r)^12-(s/r)^6]
 * over pairs r < rc, W = sum r_ij . F_ij, P = (N T + W/3) / V). */
int pairs_hip_thermo(pairs_hip_ctx *ctx, double out[5]);
/* per-atom energies and totals of the current list (no reference counterpart) */
int pairs_hip_energy(pairs_hip_ctx *ctx, double *pe, double *virial);
/* pairs_hip_energy's totals without allocating device memory (no counterpart
 * in the reference): for callers that cannot call hipMalloc (tight serving
 * loops, harnesses that forbid it).  No per-atom energies are kept: the
 * per-wave partials go to a static device array of 393 KB per GPU, so the
 * context may hold at most 1048576 locals (more: PAIRS_HIP_ERR_ARG, with the
 * reason in pairs_hip_last_error).  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other threads, other streams) race.  Same
 * kernels and summation order as pairs_hip_energy, so the same bits. */
int pairs_hip_energy_noalloc(pairs_hip_ctx *ctx, double *pe, double *virial);

/* One iteration ts of the reference loop (src/pairs/sim/timestep.py:9-72). */
int pairs_hip_step(pairs_hip_ctx *ctx, int ts, int nsteps);
/* Iterations [ts_begin, ts_end) of a run of nsteps (the loop runs ts=0..nsteps).
 * Steps between rebuilds run as replayed hipGraphs of fused kernels; the state
 * on return is that after iteration ts_end-1.  thermo_out, if not NULL,
 * receives 5 doubles (as pairs_hip_thermo) per thermo step, in order.
 * Returns the number of thermo steps (>= 0) or an error code. */
int pairs_hip_run(pairs_hip_ctx *ctx, int ts_begin, int ts_end, int nsteps, double *thermo_out,
                  int thermo_capacity);
int pairs_hip_sync(pairs_hip_ctx *ctx);
/* 1: replay hipGraphs in pairs_hip_run (default), 0: launch eagerly */
int pairs_hip_set_graphs(pairs_hip_ctx *ctx, int enable);
/* Counter checks of pairs_hip_run's rebuilds (1, default: deferred; 0: after
 * every rebuild).  Replaces the reference's per-module copy of its resize
 * counters to the host and the re-run on an overflow (src/pairs/
 * transformations/devices.py:66-79, modules.py:159-205): with 1 a rebuild's
 * counters are copied behind it into pinned memory and checked at the next
 * rebuild or at the end of the run, so the host does not wait for the device
 * at every rebuild; an overflow restores the state checkpointed before that
 * rebuild and repeats it with grown capacities, so results are those of 0. */
int pairs_hip_set_speculative_rebuilds(pairs_hip_ctx *ctx, int enable);
/* Set-up step with no reference counterpart (the reference launches every
 * module eagerly): capture and instantiate, without running them, the graphs
 * pairs_hip_run replays -- the first step of a run, the chains of steady
 * steps between two rebuilds and the rebuild, for both position buffers --
 * and allocate what its deferred rebuild checks use, so that no capture or
 * allocation falls inside a timed loop.  pairs_hip_run otherwise captures
 * each at its first use (and again when capacities or ghost counts outgrow
 * it).  Call after the first rebuild. */
int pairs_hip_prepare_graphs(pairs_hip_ctx *ctx);
/* Time the force kernel with HIP events on the context stream on every
 * every_n-th non-rebuild step of pairs_hip_run (those steps launch eagerly);
 * 0 disables. */
int pairs_hip_set_force_timing(pairs_hip_ctx *ctx, int every_n);
/* Average duration in ms of the force kernel over the timed launches of the
 * last pairs_hip_run and the number of launches timed. */
int pairs_hip_force_kernel_time(pairs_hip_ctx *ctx, double *avg_ms, int *launches);

#ifdef __cplusplus
}
#endif
#endif /* PAIRS_HIP_H */
