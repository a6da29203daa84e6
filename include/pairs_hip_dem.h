/*
 * pairs_hip_dem.h -- C ABI of the MI355X (gfx950) DEM hot path.
 *
 * The drop-in boundary for the modules the reference's code generator emits
 * for examples/dem.py: spheres and half-spaces with linear spring-dashpot
 * contacts and per-contact tangential history.  A host program calls them in
 * the order of the reference's timestep (rebuild every step, the default
 * reneighbor_every = 1 of src/pairs/sim/simulation.py:89):
 *
 *   exchange (periodic dims, with contact-history migration) \
 *   borders (periodic dims)                                   |
 *   build_cell_lists + partition_cell_lists                   |-> pairs_hip_dem_step
 *   reset_contact_history_usage_status, reset_volatile        |   (or pairs_hip_dem_run,
 *   gravity, linear_spring_dashpot, euler                     |    graph-replayed)
 *   clear_unused_contact_history                             /
 *
 * Conventions are those of pairs_hip.h: extern "C", return 0 on success, a
 * negative code on a HIP or argument error (text from
 * pairs_hip_dem_last_error); work is ordered on the context's stream;
 * downloads synchronise; one context per host thread.
 *
 * Resize protocol (src/pairs/transformations/modules.py:159-205): the step
 * functions checkpoint the locals' state at the start of every chunk of
 * steps (pairs_hip_dem_set_check_every, default 100) and read the device
 * counters at its end; a chunk in which a particle, contact or send capacity
 * would have been exceeded is run again from its checkpoint with that
 * capacity grown to twice the required size, so a run continues as if the
 * capacities had been large enough from the start.  A positive
 * PAIRS_HIP_DEM_OVERFLOW_* code (required size from
 * pairs_hip_dem_last_required) is returned only by set-up calls that exceed
 * a capacity, or when growing cannot resolve an overflow (the cell grid is
 * fixed by the domain).
 */
#ifndef PAIRS_HIP_DEM_H
#define PAIRS_HIP_DEM_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pairs_hip_dem_ctx pairs_hip_dem_ctx;

enum {
    PAIRS_HIP_DEM_OVERFLOW_PARTICLES = 1, /* particle_capacity (locals + ghosts)   */
    PAIRS_HIP_DEM_OVERFLOW_CONTACTS = 2,  /* contact_capacity per particle         */
    PAIRS_HIP_DEM_OVERFLOW_SEND = 3,      /* send_capacity                         */
    PAIRS_HIP_DEM_OVERFLOW_CELLS = 4      /* ncells_capacity                       */
};

/* Shapes (src/pairs/sim/shapes.py) and what pairs_hip_dem_download reads. */
enum { PAIRS_HIP_DEM_SPHERE = 0, PAIRS_HIP_DEM_HALFSPACE = 1 };
enum {
    PAIRS_HIP_DEM_POSITION = 0,         /* f64 [n][3], locals then ghosts      */
    PAIRS_HIP_DEM_LINEAR_VELOCITY = 1,  /* f64 [n][3]                          */
    PAIRS_HIP_DEM_ANGULAR_VELOCITY = 2, /* f64 [n][3]                          */
    PAIRS_HIP_DEM_FORCE = 3,            /* f64 [nlocal][3]                     */
    PAIRS_HIP_DEM_TORQUE = 4,           /* f64 [nlocal][3]                     */
    PAIRS_HIP_DEM_RADIUS = 5,           /* f64 [n]                             */
    PAIRS_HIP_DEM_MASS = 6,             /* f64 [n]                             */
    PAIRS_HIP_DEM_ROTATION_QUAT = 7,    /* f64 [nlocal][4]                     */
    PAIRS_HIP_DEM_ROTATION_MATRIX = 8,  /* f64 [nlocal][9]                     */
    PAIRS_HIP_DEM_UID = 9,              /* i32 [n]                             */
    PAIRS_HIP_DEM_FLAGS = 10,           /* i32 [n]                             */
    PAIRS_HIP_DEM_TYPE = 11,            /* i32 [n]                             */
    PAIRS_HIP_DEM_SHAPE = 12,           /* i32 [n]                             */
    PAIRS_HIP_DEM_NUM_CONTACTS = 13     /* i32 [nlocal]                        */
};
enum {
    PAIRS_HIP_DEM_NLOCAL = 0,
    PAIRS_HIP_DEM_NGHOST = 1,
    PAIRS_HIP_DEM_NCELLS = 2,
    PAIRS_HIP_DEM_CONTACTS = 3,
    PAIRS_HIP_DEM_PARTICLE_CAPACITY = 4, /* current capacities (they grow on overflow) */
    PAIRS_HIP_DEM_CONTACT_CAPACITY = 5,
    PAIRS_HIP_DEM_SEND_CAPACITY = 6,
    PAIRS_HIP_DEM_RESIZES = 7            /* capacity growths so far                   */
};

typedef struct pairs_hip_dem_config {
    int particle_capacity;  /* locals + ghosts (particle_capacity, examples/dem.py:154) */
    int contact_capacity;   /* contacts per particle (neighbor_capacity = 20)         */
    int send_capacity;      /* ghost / exchange map entries                           */
    int ncells_capacity;
    double box_lo[3], box_hi[3];
    int periodic[3];        /* psim.pbc([True, True, False])                          */
    double cell_width;      /* linkedCellWidth: cells and ghost layer                 */
    /* examples/dem.py:113-140 */
    double dt, density_particle, density_fluid, gravity;
    double kappa;           /* tangential / normal stiffness ratio                    */
    double ln_restitution;  /* lnDryResCoeff                                          */
    double collision_time;
    /* > 0: a fixed normal stiffness k_n (N/m) instead of the one the collision
       time gives, with the damping that keeps the restitution coefficient,
       gamma_n = -2 ln(e) sqrt(m_eff k_n) / sqrt(pi^2 + ln(e)^2) -- the same
       spring-dashpot kernel re-parameterised (no reference counterpart:
       examples/dem.py:18-74 derives both from collision_time); 0: from
       collision_time as the reference does */
    double stiffness_norm;
    int ntypes;
    double friction_static[16], friction_dynamic[16]; /* [t_i * ntypes + t_j] */
    void *stream;           /* hipStream_t or NULL (the context creates one)         */
} pairs_hip_dem_config;

/* Simulation object + device arrays (replaces PairsSimulation, runtime/pairs.hpp:27-297). */
int pairs_hip_dem_create(pairs_hip_dem_ctx **ctx, const pairs_hip_dem_config *cfg);
int pairs_hip_dem_destroy(pairs_hip_dem_ctx *ctx);
const char *pairs_hip_dem_last_error(const pairs_hip_dem_ctx *ctx);
int pairs_hip_dem_last_required(const pairs_hip_dem_ctx *ctx, int which_overflow);

/* pairs::dem_sc_grid (runtime/dem_sc_grid.hpp:62-172): spheres on a simple
   cubic grid, diameters and transverse velocities from the default-seeded
   std::mt19937 of the host's C++ library (the generator the reference uses);
   appended to the locals.  Returns the number of locals, or < 0. */
int pairs_hip_dem_sc_grid(pairs_hip_dem_ctx *ctx, double xmax, double ymax, double zmax, double spacing,
                          double diameter, double min_diameter, double max_diameter, double initial_velocity,
                          double particle_density, int ntypes);
/* Spheres at seeded random non-overlapping positions (no reference
   counterpart; a build-defined initial state): a grid of nx * ny * nz sites
   origin + (i, j, k) * spacing (x fastest), each moved by a uniform jitter in
   [-jitter, jitter] per coordinate, radii uniform in [r_min, r_max], mass
   4/3 pi r^3 density, at rest, uid from 1 in site order, type 0.  Uniform
   numbers are (g() >> 11) * 2^-53 of std::mt19937_64(seed), four per site in
   the order x, y, z jitter then radius.  Non-overlapping when every spacing
   exceeds 2 r_max + 2 jitter.  Sites outside the domain are dropped (as
   read_particle_data drops rows).  Returns the number of locals, or < 0. */
int pairs_hip_dem_jittered_grid(pairs_hip_dem_ctx *ctx, int nx, int ny, int nz, const double origin[3],
                                const double spacing[3], double jitter, double r_min, double r_max,
                                double density, unsigned long long seed);
/* One row of pairs::read_particle_data (runtime/read_from_file.hpp:33-115) for a
   half-space with attributes uid, type, mass, position, normal, flags. */
int pairs_hip_dem_add_halfspace(pairs_hip_dem_ctx *ctx, int uid, int type, double mass, const double position[3],
                                const double normal[3], int flags);
/* n spheres (read_particle_data with uid, type, radius, mass, position,
   linear velocity, flags; shape = sphere), appended to the locals. */
int pairs_hip_dem_add_spheres(pairs_hip_dem_ctx *ctx, int n, const int *uid, const int *type,
                              const double *radius, const double *mass, const double *position,
                              const double *linear_velocity, const int *flags);
/* update_mass_and_inertia (examples/dem.py:6-15) + build_cell_lists_stencil. */
int pairs_hip_dem_setup(pairs_hip_dem_ctx *ctx);

/* One timestep (the generated loop body); timesteps [ts_begin, ts_end).
   Both follow the resize protocol above (a step is its own chunk). */
int pairs_hip_dem_step(pairs_hip_dem_ctx *ctx, int ts);
int pairs_hip_dem_run(pairs_hip_dem_ctx *ctx, int ts_begin, int ts_end);
/* Steps per checkpoint and counter read of pairs_hip_dem_run (default 100). */
int pairs_hip_dem_set_check_every(pairs_hip_dem_ctx *ctx, int steps);
int pairs_hip_dem_set_graphs(pairs_hip_dem_ctx *ctx, int enable);
int pairs_hip_dem_sync(pairs_hip_dem_ctx *ctx);

int pairs_hip_dem_count(pairs_hip_dem_ctx *ctx, int what);
/* Copy one array (PAIRS_HIP_DEM_POSITION ...) to the host: count elements
   of its leading dimension are written, at most what exists (locals and
   ghosts; locals only for force, torque, rotation and contact counts).
   Returns the count written. */
int pairs_hip_dem_download(pairs_hip_dem_ctx *ctx, int what, void *host, int count);
/* Contact history of the locals, row-major [nlocal][contact capacity]
   (the current one: pairs_hip_dem_count(PAIRS_HIP_DEM_CONTACT_CAPACITY)):
   uid of the partner, is_sticking, tangential spring displacement [3],
   impact velocity magnitude (src/pairs/sim/contact_history.py). */
int pairs_hip_dem_contacts(pairs_hip_dem_ctx *ctx, int *num_contacts, int *partner_uid, int *sticking,
                           double *tangential_displacement, double *impact_velocity);
/* Average duration of the contact kernel over the timed launches of the
   last pairs_hip_dem_run (set every_n > 0 first; 0 disables timing). */
int pairs_hip_dem_set_timing(pairs_hip_dem_ctx *ctx, int every_n);
int pairs_hip_dem_contact_kernel_time(pairs_hip_dem_ctx *ctx, double *avg_ms, int *launches);
/* The same timed steps split into phases, avg_ms[6]: prepare (exchange,
   borders, cell lists), contact search, history slots, spring-dashpot per
   contact, per-particle sums + history clean-up, euler.
   No reference counterpart (the reference's timers are host-side). */
int pairs_hip_dem_phase_times(pairs_hip_dem_ctx *ctx, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
