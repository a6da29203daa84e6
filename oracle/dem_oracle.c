/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the reference's generated CPU program for the DEM
 * path (examples/dem.py lowered by src/pairs): spheres on a simple-cubic grid
 * falling onto half-spaces, linear spring-dashpot contacts with per-contact
 * tangential history, explicit Euler with quaternion rotation.  It exists to
 * check the HIP product path and to serve as the timed CPU baseline; nothing in
 * pairs_amd/ links, imports or calls it.
 *
 * Every floating-point expression keeps the operation tree the reference's code
 * generator emits (one C statement per IR node; examples/dem.py:6-90), so with
 * -ffp-contract=off the results are bit-identical to the generated program.
 *
 * Pinning (tests/test_oracle_dem.py):
 *   - odem_sc_grid, odem_add_halfspace, odem_setup: the initial state of
 *     tests/golden/dem_sc168.npz (bit-exact), the closed-form sphere mass and
 *     inertia; the grid generator's std::mt19937 is checked against the
 *     published 10000th output of the default-seeded generator.
 *   - odem_jittered_grid (build-defined initial state, no reference
 *     counterpart): its MT19937-64 against the C++ standard's check value
 *     (10000th output of the default seed 5489 is 9981545732273789042); site
 *     count, bounds and non-overlap of the spheres it places.
 *   - the fixed-stiffness contact form (odem_set_stiffness; the reference
 *     derives k_n from a collision time): equal, to rounding, to the
 *     collision-time form at the collision time that gives the same k_n for
 *     the pair's effective mass; a sphere resting on a plane sinks W / k_n.
 *   - odem_step (exchange with contact-history migration, borders, cells with
 *     shape partition, gravity, contacts, Euler, history compaction): the
 *     golden trajectory of the reference's generated program (positions,
 *     velocities, angular velocities, forces, torques, quaternions and the
 *     uid-keyed contact history), plus closed forms: a head-on two-sphere
 *     collision returns the restitution coefficient, a sphere at rest on a
 *     plane carries exactly its weight, free fall follows x = x0 + v t + a t^2/2.
 *
 * Reference quirks kept on purpose (they change results):
 *   - exchange packs the contact history of a leaver AFTER the hole filling
 *     (src/pairs/sim/comm.py:100-151 order), so a leaver whose slot was refilled
 *     carries the filler's history;
 *   - the contact-history pack writes every contact's uid to the same buffer
 *     slot and the unpack reads that slot for every contact, so a migrated
 *     particle's contacts all carry the uid of its last contact;
 *   - the shape partition of a cell is an in-place two-pointer swap
 *     (src/pairs/sim/cell_lists.py:137-171), which reverses two half-spaces.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FLAG_INFINITE 1
#define FLAG_FIXED 4
#define FLAG_GLOBAL 8
#define SHAPE_SPHERE 0
#define SHAPE_HALFSPACE 1

typedef struct odem {
    int cap, ccap, cellcap, send_cap, ncells_cap;
    int nlocal, nghost;
    /* properties, AoS as the reference (sim/properties.py) */
    int *uid, *shape, *flags, *type;
    double *x, *m, *v, *w, *f, *t, *r, *nrm, *inv_inertia, *rot_m, *rot_q;
    /* contact history (sim/contact_history.py) */
    int *num_contacts, *contact_lists, *contact_used, *is_sticking;
    double *tsd, *ivm;
    /* domain, cells */
    double lo[3], hi[3], len[3];
    int periodic[3], pbc[6];
    double spacing; /* linked-cell width == ghost layer width */
    int dim_cells[3], ncells, nstencil, stencil[27];
    int *cell_sizes, *cell_particles, *particle_cell, *nshapes;
    /* communication */
    int nsend_all, nsend[6];
    int *send_map, *send_mult, *exchg_flag, *exchg_copy_to;
    double *send_buf;
    /* parameters (examples/dem.py:113-140) */
    double dt, rho_p, rho_f, grav, pi, kappa, ln_e, t_c;
    double kn_fixed, sqrt_pl; /* fixed normal stiffness (> 0) and sqrt(pi^2 + ln^2 e) */
    int ntypes;
    double mu_s[16], mu_d[16];
    int overflow;
    /* std::mt19937 state of runtime/dem_sc_grid.hpp:13-20 */
    uint32_t mt[624];
    int mti;
} odem_t;

static void *zalloc(size_t n) { return calloc(1, n ? n : 1); }

odem_t *odem_create(int cap, int ccap, int cellcap, int send_cap, int ncells_cap, const double *lo,
                    const double *hi, const int *periodic, double spacing, double dt, double rho_p, double rho_f,
                    double grav, double kappa, double ln_e, double t_c, int ntypes, const double *mu_s,
                    const double *mu_d) {
    odem_t *s = (odem_t *) zalloc(sizeof(odem_t));
    s->cap = cap; s->ccap = ccap; s->cellcap = cellcap; s->send_cap = send_cap; s->ncells_cap = ncells_cap;
    for (int d = 0; d < 3; d++) {
        s->lo[d] = lo[d]; s->hi[d] = hi[d]; s->len[d] = hi[d] - lo[d];
        s->periodic[d] = periodic[d];
        s->pbc[2 * d] = 1; s->pbc[2 * d + 1] = -1; /* one rank (regular_6d_stencil.cpp:72-73) */
    }
    s->spacing = spacing; s->dt = dt; s->rho_p = rho_p; s->rho_f = rho_f; s->grav = grav;
    s->pi = 3.141592653589793; s->kappa = kappa; s->ln_e = ln_e; s->t_c = t_c; s->ntypes = ntypes;
    for (int k = 0; k < ntypes * ntypes; k++) { s->mu_s[k] = mu_s[k]; s->mu_d[k] = mu_d[k]; }
    size_t n = (size_t) cap;
    s->uid = zalloc(n * sizeof(int)); s->shape = zalloc(n * sizeof(int));
    s->flags = zalloc(n * sizeof(int)); s->type = zalloc(n * sizeof(int));
    s->x = zalloc(3 * n * sizeof(double)); s->m = zalloc(n * sizeof(double));
    s->v = zalloc(3 * n * sizeof(double)); s->w = zalloc(3 * n * sizeof(double));
    s->f = zalloc(3 * n * sizeof(double)); s->t = zalloc(3 * n * sizeof(double));
    s->r = zalloc(n * sizeof(double)); s->nrm = zalloc(3 * n * sizeof(double));
    s->inv_inertia = zalloc(9 * n * sizeof(double)); s->rot_m = zalloc(9 * n * sizeof(double));
    s->rot_q = zalloc(4 * n * sizeof(double));
    s->num_contacts = zalloc(n * sizeof(int));
    s->contact_lists = zalloc(n * ccap * sizeof(int)); s->contact_used = zalloc(n * ccap * sizeof(int));
    s->is_sticking = zalloc(n * ccap * sizeof(int));
    s->tsd = zalloc(3 * n * ccap * sizeof(double)); s->ivm = zalloc(n * ccap * sizeof(double));
    s->cell_sizes = zalloc((size_t) ncells_cap * sizeof(int));
    s->cell_particles = zalloc((size_t) ncells_cap * cellcap * sizeof(int));
    s->nshapes = zalloc((size_t) ncells_cap * 2 * sizeof(int));
    s->particle_cell = zalloc(n * sizeof(int));
    s->send_map = zalloc((size_t) send_cap * sizeof(int)); s->send_mult = zalloc(3 * (size_t) send_cap * sizeof(int));
    s->exchg_flag = zalloc(n * sizeof(int)); s->exchg_copy_to = zalloc((size_t) send_cap * sizeof(int));
    s->send_buf = zalloc((size_t) send_cap * (40 + 6 * (size_t) ccap + 2) * sizeof(double));
    /* std::mt19937 default seed 5489 */
    s->mt[0] = 5489u;
    for (int i = 1; i < 624; i++) s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t) i;
    s->mti = 624;
    return s;
}

void odem_destroy(odem_t *s) {
    if (!s) return;
    void *ps[] = {s->uid, s->shape, s->flags, s->type, s->x, s->m, s->v, s->w, s->f, s->t, s->r, s->nrm,
                  s->inv_inertia, s->rot_m, s->rot_q, s->num_contacts, s->contact_lists, s->contact_used,
                  s->is_sticking, s->tsd, s->ivm, s->cell_sizes, s->cell_particles, s->nshapes, s->particle_cell,
                  s->send_map, s->send_mult, s->exchg_flag, s->exchg_copy_to, s->send_buf};
    for (size_t k = 0; k < sizeof(ps) / sizeof(ps[0]); k++) free(ps[k]);
    free(s);
}

/* ---- initial state ----------------------------------------------------------- */

/* MT19937 (Matsumoto & Nishimura 1998), the generator std::mt19937 defines */
static uint32_t mt_next(odem_t *s) {
    if (s->mti >= 624) {
        for (int k = 0; k < 624; k++) {
            uint32_t y = (s->mt[k] & 0x80000000u) | (s->mt[(k + 1) % 624] & 0x7fffffffu);
            s->mt[k] = s->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        s->mti = 0;
    }
    uint32_t y = s->mt[s->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

uint32_t odem_mt_next(odem_t *s) { return mt_next(s); }

/* MT19937-64 (Nishimura 2000; the generator std::mt19937_64 defines): state
   of 312 words, seeded by x_i = 6364136223846793005 (x_{i-1} ^ (x_{i-1} >> 62)) + i */
typedef struct { uint64_t mt[312]; int mti; } mt64_t;

static void mt64_seed(mt64_t *g, uint64_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 312; i++) g->mt[i] = 6364136223846793005ull * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t) i;
    g->mti = 312;
}

static uint64_t mt64_next(mt64_t *g) {
    if (g->mti >= 312) {
        for (int k = 0; k < 312; k++) {
            const uint64_t y = (g->mt[k] & 0xFFFFFFFF80000000ull) | (g->mt[(k + 1) % 312] & 0x7FFFFFFFull);
            g->mt[k] = g->mt[(k + 156) % 312] ^ (y >> 1) ^ ((y & 1ull) ? 0xB5026F5AA96619E9ull : 0ull);
        }
        g->mti = 0;
    }
    uint64_t x = g->mt[g->mti++];
    x ^= (x >> 29) & 0x5555555555555555ull;
    x ^= (x << 17) & 0x71D67FFFEDA60000ull;
    x ^= (x << 37) & 0xFFF7EEE000000000ull;
    x ^= x >> 43;
    return x;
}

/* the n-th output (n >= 1) of MT19937-64 seeded with seed (for checks) */
uint64_t odem_mt64_nth(uint64_t seed, int n) {
    mt64_t g;
    mt64_seed(&g, seed);
    uint64_t x = 0;
    for (int k = 0; k < n; k++) x = mt64_next(&g);
    return x;
}

/* fixed normal stiffness for the contact model (0 = from the collision time) */
void odem_set_stiffness(odem_t *s, double kn) {
    s->kn_fixed = kn > 0.0 ? kn : 0.0;
    s->sqrt_pl = sqrt(s->pi * s->pi + s->ln_e * s->ln_e);
}

/* std::uniform_real_distribution<double>(a, b) over std::mt19937 as libstdc++
   (GCC 11) defines it: generate_canonical<double, 53> takes two 32-bit draws,
   sum = u0 + u1 * 2^32, r = sum / 2^64 (r >= 1 -> nextafter(1, 0)), then
   r * (b - a) + a.  runtime/dem_sc_grid.hpp:22-44 */
static double real_random(odem_t *s, double a, double b) {
    double sum = (double) mt_next(s);
    sum = sum + (double) mt_next(s) * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r * (b - a) + a;
}

/* Regular6DStencil::isWithinSubdomain (one rank), SMALL = 1e-5
   (runtime/domain/regular_6d_stencil.cpp:94-98) */
static int within_subdomain(const odem_t *s, const double *p) {
    return p[0] >= s->lo[0] && p[0] < s->hi[0] - 0.00001 && p[1] >= s->lo[1] && p[1] < s->hi[1] - 0.00001 &&
           p[2] >= s->lo[2] && p[2] < s->hi[2] - 0.00001;
}

static int within_aabb(const double *p, const double *box) {
    return p[0] >= box[0] && p[0] < box[3] && p[1] >= box[1] && p[1] < box[4] && p[2] >= box[2] && p[2] < box[5];
}

/* pairs::dem_sc_grid (runtime/dem_sc_grid.hpp:62-172), one rank */
int odem_sc_grid(odem_t *s, double xmax, double ymax, double zmax, double spacing, double diameter,
                 double min_diameter, double max_diameter, double initial_velocity, double particle_density,
                 int ntypes) {
    int last_uid = 1, np = s->nlocal;
    const double gen[6] = {0.0, 0.0, diameter, xmax, ymax, zmax};
    const double refp[3] = {spacing * 0.5, spacing * 0.5, spacing * 0.5};
    const int iret = (int) ceil((0.0 - refp[0]) / spacing);
    const int jret = (int) ceil((0.0 - refp[1]) / spacing);
    const int kret = (int) ceil((diameter - refp[2]) / spacing);
    int i = iret, j = jret, k = kret;
    double p[3] = {refp[0] + i * spacing, refp[1] + j * spacing, refp[2] + k * spacing};
    while (within_aabb(p, gen)) {
        const double dia = real_random(s, min_diameter, max_diameter);
        if (within_subdomain(s, p)) {
            if (np >= s->cap) { s->overflow |= 1; return np; }
            const double rad = dia * 0.5;
            s->uid[np] = last_uid;
            s->r[np] = rad;
            s->m[np] = (((4.0 / 3.0) * 3.141592653589793) * rad * rad * rad) * particle_density;
            for (int d = 0; d < 3; d++) s->x[3 * np + d] = p[d];
            s->v[3 * np + 0] = 0.1 * real_random(s, -initial_velocity, initial_velocity);
            s->v[3 * np + 1] = 0.1 * real_random(s, -initial_velocity, initial_velocity);
            s->v[3 * np + 2] = -initial_velocity;
            s->type[np] = rand() % ntypes;
            s->flags[np] = 0;
            s->shape[np] = SHAPE_SPHERE;
            np++;
        }
        ++i;
        p[0] = refp[0] + i * spacing; p[1] = refp[1] + j * spacing; p[2] = refp[2] + k * spacing;
        if (!within_aabb(p, gen)) {
            i = iret;
            j++;
            p[0] = refp[0] + i * spacing; p[1] = refp[1] + j * spacing; p[2] = refp[2] + k * spacing;
            if (!within_aabb(p, gen)) {
                j = jret;
                k++;
                p[0] = refp[0] + i * spacing; p[1] = refp[1] + j * spacing; p[2] = refp[2] + k * spacing;
                if (!within_aabb(p, gen)) break;
            }
        }
        last_uid++;
    }
    s->nlocal = np;
    return np;
}

/* Spheres at seeded random non-overlapping positions, the build-defined
   initial state of pairs_hip_dem_jittered_grid (no reference counterpart):
   sites origin + (i, j, k) * spacing, x fastest, each coordinate moved by
   (2u - 1) * jitter, radius r_min + (r_max - r_min) u, u = (g() >> 11) 2^-53
   of MT19937-64(seed), four per site (x, y, z, radius); at rest; uid from 1
   in site order; sites outside the domain dropped. */
int odem_jittered_grid(odem_t *s, int nx, int ny, int nz, const double *origin, const double *spacing,
                       double jitter, double r_min, double r_max, double density, uint64_t seed) {
    mt64_t g;
    mt64_seed(&g, seed);
    int uid = 1, np = s->nlocal;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int site[3] = {i, j, k};
                double q[3];
                for (int d = 0; d < 3; d++) {
                    const double u = (double) (mt64_next(&g) >> 11) * 0x1.0p-53;
                    q[d] = (origin[d] + site[d] * spacing[d]) + (2.0 * u - 1.0) * jitter;
                }
                const double u = (double) (mt64_next(&g) >> 11) * 0x1.0p-53;
                const double rad = r_min + (r_max - r_min) * u;
                const int id = uid++;
                if (!within_subdomain(s, q)) continue;
                if (np >= s->cap) { s->overflow |= 1; s->nlocal = np; return np; }
                s->uid[np] = id;
                s->r[np] = rad;
                s->m[np] = (((4.0 / 3.0) * 3.141592653589793) * rad * rad * rad) * density;
                for (int d = 0; d < 3; d++) { s->x[3 * np + d] = q[d]; s->v[3 * np + d] = 0.0; s->w[3 * np + d] = 0.0; }
                s->type[np] = 0;
                s->flags[np] = 0;
                s->shape[np] = SHAPE_SPHERE;
                np++;
            }
    s->nlocal = np;
    return np;
}

/* one row of read_particle_data (runtime/read_from_file.hpp:33-115) with the
   attribute list uid, type, mass, position, normal, flags (examples/dem.py:186-189) */
int odem_add_halfspace(odem_t *s, int uid, int type, double mass, const double *pos, const double *normal,
                       int flags) {
    int n = s->nlocal;
    if (n >= s->cap) { s->overflow |= 1; return n; }
    s->uid[n] = uid; s->type[n] = type; s->m[n] = mass;
    for (int d = 0; d < 3; d++) { s->x[3 * n + d] = pos[d]; s->nrm[3 * n + d] = normal[d]; }
    s->flags[n] = flags;
    if (within_subdomain(s, pos) || (flags & (FLAG_INFINITE | FLAG_FIXED | FLAG_GLOBAL))) {
        s->shape[n] = SHAPE_HALFSPACE;
        s->nlocal = n + 1;
    }
    return s->nlocal;
}

/* one sphere as read_particle_data would place it (attributes uid, type,
   radius, position, linear velocity; mass from the density as dem_sc_grid) */
int odem_add_sphere(odem_t *s, int uid, int type, double radius, double density, const double *pos,
                    const double *vel) {
    int n = s->nlocal;
    if (n >= s->cap) { s->overflow |= 1; return n; }
    s->uid[n] = uid; s->type[n] = type; s->r[n] = radius;
    s->m[n] = (((4.0 / 3.0) * 3.141592653589793) * radius * radius * radius) * density;
    for (int d = 0; d < 3; d++) { s->x[3 * n + d] = pos[d]; s->v[3 * n + d] = vel[d]; s->w[3 * n + d] = 0.0; }
    s->flags[n] = 0;
    s->shape[n] = SHAPE_SPHERE;
    s->nlocal = n + 1;
    return s->nlocal;
}

/* update_mass_and_inertia (examples/dem.py:6-15): identity rotation; spheres
   get the cofactor inverse of diag(0.4 m r^2), other shapes infinite mass */
void odem_setup(odem_t *s) {
    for (int i = 0; i < s->nlocal; i++) {
        double *R = s->rot_m + 9 * i, *q = s->rot_q + 4 * i, *Ii = s->inv_inertia + 9 * i;
        for (int k = 0; k < 9; k++) R[k] = (k % 4 == 0) ? 1.0 : 0.0;
        q[0] = 1.0; q[1] = 0.0; q[2] = 0.0; q[3] = 0.0;
        if (s->shape[i] == SHAPE_SPHERE) {
            double a = 0.4 * s->m[i];
            a = a * s->r[i];
            a = a * s->r[i];
            const double M[9] = {a, 0.0, 0.0, 0.0, a, 0.0, 0.0, 0.0, a};
            /* determinant and adjugate as the generated code orders them */
            double det = M[0] * (M[4] * M[8] - M[7] * M[5]);
            det = det + M[1] * (M[5] * M[6] - M[8] * M[3]);
            det = det + M[2] * (M[3] * M[7] - M[4] * M[6]);
            const double id = 1.0 / det;
            Ii[0] = id * (M[4] * M[8] - M[5] * M[7]);
            Ii[1] = id * (M[7] * M[2] - M[8] * M[1]);
            Ii[2] = id * (M[1] * M[5] - M[2] * M[4]);
            Ii[3] = id * (M[5] * M[6] - M[3] * M[8]);
            Ii[4] = id * (M[8] * M[0] - M[6] * M[2]);
            Ii[5] = id * (M[2] * M[3] - M[0] * M[5]);
            Ii[6] = id * (M[3] * M[7] - M[4] * M[6]);
            Ii[7] = id * (M[6] * M[1] - M[7] * M[0]);
            Ii[8] = id * (M[0] * M[4] - M[1] * M[3]);
        } else {
            s->m[i] = INFINITY;
            for (int k = 0; k < 9; k++) Ii[k] = 0.0;
        }
    }
}

/* build_cell_lists_stencil (src/pairs/sim/cell_lists.py:46-87) */
void odem_setup_cells(odem_t *s) {
    const double sp = s->spacing;
    int total = 1;
    for (int d = 0; d < 3; d++) {
        s->dim_cells[d] = (int) ceil(((s->hi[d] + sp) - (s->lo[d] - sp)) / sp) + 1;
        total *= s->dim_cells[d];
    }
    s->ncells = total + 1;
    if (s->ncells + 1 >= s->ncells_cap) s->overflow |= 2;
    s->nstencil = 0;
    for (int a = -1; a < 2; a++)
        for (int b = -1; b < 2; b++)
            for (int c = -1; c < 2; c++) s->stencil[s->nstencil++] = (a * s->dim_cells[1] + b) * s->dim_cells[2] + c;
}

/* ---- communication ----------------------------------------------------------- */

/* Determine{Exchange,Ghost}Particles (src/pairs/sim/comm.py:154-259) for one
   dimension: two passes over [0, nlocal+nghost), "< min+offset" then
   "> max-offset", skipping Infinite|Global; non-periodic dimensions of one
   rank send nothing (the generated guard pbc == 0 is false there) */
static void determine(odem_t *s, int dim, double offset, int exchange) {
    const int nall = s->nlocal + s->nghost;
    if (exchange) for (int i = 0; i < s->nlocal; i++) s->exchg_flag[i] = 0;
    if (!s->periodic[dim]) return;
    for (int side = 0; side < 2; side++) {
        const int jj = 2 * dim + side;
        for (int i = 0; i < nall; i++) {
            if (s->flags[i] & (FLAG_INFINITE | FLAG_GLOBAL)) continue;
            const double xd = s->x[3 * i + dim];
            const int sel = side == 0 ? (xd < s->lo[dim] + offset) : (xd > s->hi[dim] - offset);
            if (!sel) continue;
            const int idx = s->nsend_all++;
            if (idx >= s->send_cap) { s->overflow |= 4; continue; }
            s->send_map[idx] = i;
            if (exchange) s->exchg_flag[i] = 1;
            for (int d = 0; d < 3; d++) s->send_mult[3 * idx + d] = d == dim ? s->pbc[jj] : 0;
            s->nsend[jj]++;
        }
    }
}

static inline double shifted(const odem_t *s, int idx, int d, double x) {
    return x + s->send_mult[3 * idx + d] * s->len[d];
}

static void copy_contacts(odem_t *s, int dst, int src) {
    const int nc = s->num_contacts[src];
    for (int k = 0; k < nc; k++) {
        const size_t a = (size_t) dst * s->ccap + k, b = (size_t) src * s->ccap + k;
        s->contact_lists[a] = s->contact_lists[b];
        s->contact_used[a] = s->contact_used[b];
        s->is_sticking[a] = s->is_sticking[b];
        for (int d = 0; d < 3; d++) s->tsd[3 * a + d] = s->tsd[3 * b + d];
        s->ivm[a] = s->ivm[b];
    }
    s->num_contacts[dst] = nc;
}

/* Comm.exchange for one dimension (src/pairs/sim/comm.py:100-151, 445-520)
   with contact-history migration (pack/unpack_contact_history) */
static void exchange_dim(odem_t *s, int dim) {
    s->nsend_all = 0;
    s->nghost = 0;
    for (int k = 0; k < 6; k++) s->nsend[k] = 0;
    determine(s, dim, 0.0, 1);
    const int n = s->nsend_all;
    /* pack all non-volatile properties (40 doubles) */
    for (int i = 0; i < n; i++) {
        const int m = s->send_map[i];
        double *b = s->send_buf + 40 * (size_t) i;
        b[0] = s->uid[m]; b[1] = s->shape[m]; b[2] = s->flags[m];
        for (int d = 0; d < 3; d++) b[3 + d] = shifted(s, i, d, s->x[3 * m + d]);
        b[6] = s->m[m];
        for (int d = 0; d < 3; d++) { b[7 + d] = s->v[3 * m + d]; b[10 + d] = s->w[3 * m + d]; }
        b[13] = s->r[m];
        for (int d = 0; d < 3; d++) b[14 + d] = s->nrm[3 * m + d];
        for (int k = 0; k < 9; k++) { b[17 + k] = s->inv_inertia[9 * m + k]; b[26 + k] = s->rot_m[9 * m + k]; }
        for (int k = 0; k < 4; k++) b[35 + k] = s->rot_q[4 * m + k];
        b[39] = s->type[m];
    }
    /* remove_exchanged_particles_pt1 */
    int pos = s->nlocal - 1;
    for (int i = 0; i < n; i++) {
        if (s->send_map[i] < s->nlocal - n) {
            while (s->exchg_flag[pos] == 1) pos--;
            s->exchg_copy_to[i] = pos;
            pos--;
        } else {
            s->exchg_copy_to[i] = -1;
        }
    }
    /* remove_exchanged_particles_pt2: copy guard "src > 0" (comm.py:481) */
    for (int i = 0; i < n; i++) {
        const int src = s->exchg_copy_to[i];
        if (src > 0) {
            const int dst = s->send_map[i];
            s->uid[dst] = s->uid[src]; s->shape[dst] = s->shape[src]; s->flags[dst] = s->flags[src];
            for (int d = 0; d < 3; d++) s->x[3 * dst + d] = s->x[3 * src + d];
            s->m[dst] = s->m[src];
            for (int d = 0; d < 3; d++) { s->v[3 * dst + d] = s->v[3 * src + d]; s->w[3 * dst + d] = s->w[3 * src + d]; }
            s->r[dst] = s->r[src];
            for (int d = 0; d < 3; d++) s->nrm[3 * dst + d] = s->nrm[3 * src + d];
            for (int k = 0; k < 9; k++) {
                s->inv_inertia[9 * dst + k] = s->inv_inertia[9 * src + k];
                s->rot_m[9 * dst + k] = s->rot_m[9 * src + k];
            }
            for (int k = 0; k < 4; k++) s->rot_q[4 * dst + k] = s->rot_q[4 * src + k];
            s->type[dst] = s->type[src];
            copy_contacts(s, dst, src);
        }
    }
    s->nlocal -= n;
    /* unpack the properties at nlocal + i */
    for (int i = 0; i < n; i++) {
        const double *b = s->send_buf + 40 * (size_t) i;
        const int g = s->nlocal + i;
        s->uid[g] = (int) b[0]; s->shape[g] = (int) b[1]; s->flags[g] = (int) b[2];
        for (int d = 0; d < 3; d++) s->x[3 * g + d] = b[3 + d];
        s->m[g] = b[6];
        for (int d = 0; d < 3; d++) { s->v[3 * g + d] = b[7 + d]; s->w[3 * g + d] = b[10 + d]; }
        s->r[g] = b[13];
        for (int d = 0; d < 3; d++) s->nrm[3 * g + d] = b[14 + d];
        for (int k = 0; k < 9; k++) { s->inv_inertia[9 * g + k] = b[17 + k]; s->rot_m[9 * g + k] = b[26 + k]; }
        for (int k = 0; k < 4; k++) s->rot_q[4 * g + k] = b[35 + k];
        s->type[g] = (int) b[39];
    }
    /* pack_contact_history: read from the leaver's slot after the hole
       filling; every contact's uid lands in the particle's first slot, so the
       last contact's uid is the one sent */
    double *cb = s->send_buf + 40 * (size_t) n;
    size_t off = 0;
    for (int i = 0; i < n; i++) {
        const int m = s->send_map[i];
        const int nc = s->num_contacts[m];
        double *b = cb + off;
        b[0] = nc;
        double last_uid = 0.0;
        for (int k = 0; k < nc; k++) {
            const size_t a = (size_t) m * s->ccap + k;
            last_uid = s->contact_lists[a];
            b[1 + 6 * k + 1] = s->is_sticking[a];
            for (int d = 0; d < 3; d++) b[1 + 6 * k + 2 + d] = s->tsd[3 * a + d];
            b[1 + 6 * k + 5] = s->ivm[a];
        }
        if (nc > 0) b[1] = last_uid;
        off += 1 + 6 * (size_t) (nc > 0 ? nc : 0);
    }
    /* unpack_contact_history at nlocal + i; every contact gets the uid of
       the first slot; contact_used = 1 */
    off = 0;
    for (int i = 0; i < n; i++) {
        const double *b = cb + off;
        const int g = s->nlocal + i;
        const int nc = (int) b[0];
        s->num_contacts[g] = nc;
        for (int k = 0; k < nc; k++) {
            const size_t a = (size_t) g * s->ccap + k;
            s->contact_lists[a] = (int) b[1];
            s->contact_used[a] = 1;
            s->is_sticking[a] = (int) b[1 + 6 * k + 1];
            for (int d = 0; d < 3; d++) s->tsd[3 * a + d] = b[1 + 6 * k + 2 + d];
            s->ivm[a] = b[1 + 6 * k + 5];
        }
        off += 1 + 6 * (size_t) (nc > 0 ? nc : 0);
    }
    s->nlocal += n;
}

/* Comm.borders (src/pairs/sim/comm.py:56-98): per dimension, pack uid, type,
   mass, radius, position(+shift), linear and angular velocity, shape; unpack at
   nlocal + send index.  Ghost flags are not sent. */
static void borders(odem_t *s) {
    s->nsend_all = 0;
    s->nghost = 0;
    for (int dim = 0; dim < 3; dim++) {
        const int start = s->nsend_all;
        s->nsend[2 * dim] = s->nsend[2 * dim + 1] = 0;
        determine(s, dim, s->spacing, 0);
        const int end = s->nsend_all;
        for (int i = start; i < end; i++) {
            const int m = s->send_map[i], g = s->nlocal + i;
            double xs[3];
            for (int d = 0; d < 3; d++) xs[d] = shifted(s, i, d, s->x[3 * m + d]);
            s->uid[g] = s->uid[m]; s->type[g] = s->type[m]; s->m[g] = s->m[m]; s->r[g] = s->r[m];
            for (int d = 0; d < 3; d++) {
                s->x[3 * g + d] = xs[d];
                s->v[3 * g + d] = s->v[3 * m + d];
                s->w[3 * g + d] = s->w[3 * m + d];
            }
            s->shape[g] = s->shape[m];
        }
        s->nghost += end - start;
    }
}

/* ---- cells -------------------------------------------------------------------- */

static int cell_of(const odem_t *s, const double *x) {
    int idx = 0;
    for (int d = 0; d < 3; d++) {
        const double q = (x[d] - (s->lo[d] - s->spacing)) / s->spacing;
        int c = (int) q >= 0 ? (int) q : 0;
        c = c < s->dim_cells[d] ? c : s->dim_cells[d] - 1;
        idx = d == 0 ? c : idx * s->dim_cells[d] + c;
    }
    return idx + 1;
}

/* BuildCellLists + PartitionCellLists (src/pairs/sim/cell_lists.py:90-171) */
static void build_cells(odem_t *s) {
    for (int c = 0; c < s->ncells; c++) s->cell_sizes[c] = 0;
    const int nall = s->nlocal + s->nghost;
    for (int i = 0; i < nall; i++) {
        int flat = 0;
        if (!(s->flags[i] & FLAG_INFINITE)) flat = cell_of(s, s->x + 3 * i);
        if (flat >= 0 && flat < s->ncells) {
            s->particle_cell[i] = flat;
            const int slot = s->cell_sizes[flat];
            if (slot + 1 >= s->cellcap) { s->overflow |= 8; continue; }
            s->cell_sizes[flat] = slot + 1;
            s->cell_particles[(size_t) flat * s->cellcap + slot] = i;
        }
    }
    for (int c = 0; c < s->ncells; c++) {
        int *cp = s->cell_particles + (size_t) c * s->cellcap;
        int lo = 0, hi;
        for (int sh = 0; sh < 2; sh++) {
            const int first = lo;
            hi = s->cell_sizes[c] - 1;
            s->nshapes[2 * c + sh] = 0;
            while (lo <= hi) {
                const int p = cp[lo];
                if (s->shape[p] != sh) {
                    if (lo != hi) { cp[lo] = cp[hi]; cp[hi] = p; }
                    hi--;
                } else {
                    lo++;
                    s->nshapes[2 * c + sh] = lo - first;
                }
            }
        }
    }
}

/* ---- forces --------------------------------------------------------------------- */

/* contact slot of j's uid in i's history (last match), appended with zeroed
   history if absent (src/pairs/mapping/funcs.py:230-270) */
static int contact_slot(odem_t *s, int i, int uid_j) {
    int slot = -1;
    const int nc = s->num_contacts[i];
    for (int k = 0; k < nc; k++)
        if (s->contact_lists[(size_t) i * s->ccap + k] == uid_j) slot = k;
    if (slot == -1) {
        slot = nc;
        if (nc >= s->ccap) { s->overflow |= 16; return -1; }
        s->num_contacts[i] = nc + 1;
        const size_t a = (size_t) i * s->ccap + slot;
        s->contact_lists[a] = uid_j;
        s->is_sticking[a] = 0;
        s->tsd[3 * a] = 0.0; s->tsd[3 * a + 1] = 0.0; s->tsd[3 * a + 2] = 0.0;
        s->ivm[a] = 0.0;
    }
    s->contact_used[(size_t) i * s->ccap + slot] = 1;
    return slot;
}

/* linear_spring_dashpot (examples/dem.py:18-74) for one contact with normal n,
   penetration delta (> 0) and contact point cp; returns the partial force and
   accumulates force and torque */
static void spring_dashpot(odem_t *s, int i, int j, int slot, const double *n, double delta, const double *cp,
                           double *facc, double *tacc) {
    const size_t a = (size_t) i * s->ccap + slot;
    const int sticking = s->is_sticking[a] == 1;
    const double *xi = s->x + 3 * i, *xj = s->x + 3 * j;
    const double *vi = s->v + 3 * i, *vj = s->v + 3 * j, *wi = s->w + 3 * i, *wj = s->w + 3 * j;
    double ai[3], aj[3], vwi[3], vwj[3], rel[3], reln[3], relt[3];
    for (int d = 0; d < 3; d++) { ai[d] = cp[d] - xi[d]; aj[d] = cp[d] - xj[d]; }
    vwi[0] = vi[0] + (wi[1] * ai[2] - wi[2] * ai[1]);
    vwi[1] = vi[1] + (wi[2] * ai[0] - wi[0] * ai[2]);
    vwi[2] = vi[2] + (wi[0] * ai[1] - wi[1] * ai[0]);
    vwj[0] = vj[0] + (wj[1] * aj[2] - wj[2] * aj[1]);
    vwj[1] = vj[1] + (wj[2] * aj[0] - wj[0] * aj[2]);
    vwj[2] = vj[2] + (wj[0] * aj[1] - wj[1] * aj[0]);
    for (int d = 0; d < 3; d++) rel[d] = -(vwi[d] - vwj[d]);
    const double rn = (rel[0] * n[0] + rel[1] * n[1]) + rel[2] * n[2];
    for (int d = 0; d < 3; d++) { reln[d] = rn * n[d]; relt[d] = rel[d] - reln[d]; }
    const double relt_len = sqrt((relt[0] * relt[0] + relt[1] * relt[1]) + relt[2] * relt[2]);
    const int cond_a = sticking && relt_len < 1e-08;
    const double meff = 1.0 / (1.0 / s->m[i] + 1.0 / s->m[j]);
    /* stiffness from the collision time (examples/dem.py:34-35), or fixed */
    const double kn = s->kn_fixed > 0.0 ? s->kn_fixed
                                        : (meff * (s->pi * s->pi + s->ln_e * s->ln_e)) / (s->t_c * s->t_c);
    const double kt = s->kappa * kn;
    /* rotate the spring into the tangent plane, keep its length */
    const double *h = s->tsd + 3 * a;
    const double hn = (h[0] * n[0] + h[1] * n[1]) + h[2] * n[2];
    double rot[3], ntsd[3];
    for (int d = 0; d < 3; d++) rot[d] = h[d] - n[d] * hn;
    const double rot2 = (rot[0] * rot[0] + rot[1] * rot[1]) + rot[2] * rot[2];
    const double h2 = (h[0] * h[0] + h[1] * h[1]) + h[2] * h[2];
    const double scale = sqrt(h2 / rot2);
    for (int d = 0; d < 3; d++) ntsd[d] = s->dt * relt[d] + ((rot2 <= 0.0) ? 0.0 : rot[d] * scale);
    /* damping for the restitution coefficient (examples/dem.py:36), or for
       the fixed stiffness: -2 ln(e) sqrt(m_eff k_n) / sqrt(pi^2 + ln^2 e),
       which is the same value at the collision time that gives k_n */
    const double dn = s->kn_fixed > 0.0 ? ((-(2.0) * s->ln_e) * sqrt(meff * kn)) / s->sqrt_pl
                                        : ((-(2.0) * meff) * s->ln_e) / s->t_c;
    const double dt_ = sqrt(s->kappa) * dn;
    double ftls[3];
    for (int d = 0; d < 3; d++) ftls[d] = kt * ntsd[d] + dt_ * relt[d];
    const double ftls_len = sqrt((ftls[0] * ftls[0] + ftls[1] * ftls[1]) + ftls[2] * ftls[2]);
    const int tk = s->type[i] * s->ntypes + s->type[j];
    double fn[3];
    for (int d = 0; d < 3; d++) fn[d] = (kn * delta) * n[d] + dn * reln[d];
    const double fn_len = sqrt((fn[0] * fn[0] + fn[1] * fn[1]) + fn[2] * fn[2]);
    const double f_static = s->mu_s[tk] * fn_len;
    const double f_dynamic = s->mu_d[tk] * fn_len;
    const int cond1 = cond_a && ftls_len < f_static;
    const int cond2 = sticking && ftls_len < f_dynamic;
    const int slide = (!cond1 && !cond2) && kt > 0.0;
    const double f_fric = cond1 ? f_static : f_dynamic;
    const double inv_len = 1.0 / ftls_len;
    double tdir[3];
    for (int d = 0; d < 3; d++) tdir[d] = (ftls_len > 0.0) ? ftls[d] * inv_len : 0.0;
    double *hw = s->tsd + 3 * a;
    for (int d = 0; d < 3; d++) hw[d] = slide ? (f_fric * tdir[d] - dt_ * relt[d]) / kt : ntsd[d];
    const double rel_len = sqrt((rel[0] * rel[0] + rel[1] * rel[1]) + rel[2] * rel[2]);
    s->ivm[a] = (s->ivm[a] > 0.0) ? s->ivm[a] : rel_len;
    s->is_sticking[a] = ((cond1 || cond2) || ftls_len < f_dynamic) ? 1 : 0;
    const double ft_abs = (f_fric < ftls_len) ? f_fric : ftls_len;
    double F[3];
    for (int d = 0; d < 3; d++) F[d] = fn[d] + ft_abs * tdir[d];
    for (int d = 0; d < 3; d++) facc[d] = facc[d] + F[d];
    const double c0 = cp[0] - xi[0], c1 = cp[1] - xi[1], c2 = cp[2] - xi[2];
    tacc[0] = tacc[0] + (c1 * F[2] - c2 * F[1]);
    tacc[1] = tacc[1] + (c2 * F[0] - c0 * F[2]);
    tacc[2] = tacc[2] + (c0 * F[1] - c1 * F[0]);
}

/* the contact module: spheres of cell 0 and the stencil first, then the
   half-spaces (sim/interaction.py:234-264); force += sphere + plane sums,
   torque += plane + sphere sums (ir/apply.py:111-125 reduction order) */
static void contacts(odem_t *s) {
    for (int i = 0; i < s->nlocal; i++) {
        if (s->flags[i] & FLAG_FIXED) continue;
        double fs[3] = {0.0, 0.0, 0.0}, fp[3] = {0.0, 0.0, 0.0}, tp[3] = {0.0, 0.0, 0.0}, ts[3] = {0.0, 0.0, 0.0};
        const int ci = s->particle_cell[i];
        const double *xi = s->x + 3 * i;
        const double ri = s->r[i];
        for (int pass = 0; pass < 2; pass++) {
            for (int k = -1; k < s->nstencil; k++) {
                const int c = k < 0 ? 0 : ci + s->stencil[k];
                if (!(k < 0 || (c > 0 && c < s->ncells))) continue;
                const int *cp_ = s->cell_particles + (size_t) c * s->cellcap;
                const int q0 = pass == 0 ? 0 : s->nshapes[2 * c];
                const int q1 = pass == 0 ? s->nshapes[2 * c] : s->nshapes[2 * c] + s->nshapes[2 * c + 1];
                for (int q = q0; q < q1; q++) {
                    const int j = cp_[q];
                    if (j == i) continue;
                    const double *xj = s->x + 3 * j;
                    double nvec[3], cpt[3], delta;
                    if (pass == 0) {
                        double d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
                        const double rsq = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
                        const double rs = ri + s->r[j];
                        if (!(rsq < rs * (rs + 0.0))) continue;
                        const double dist = sqrt(rsq);
                        const double pen = (dist - ri) - s->r[j];
                        delta = -(pen);
                        if (delta < 0.0) continue;
                        const double inv = 1.0 / dist;
                        for (int e = 0; e < 3; e++) nvec[e] = d[e] * inv;
                        const double reach = s->r[j] + 0.5 * pen;
                        for (int e = 0; e < 3; e++) cpt[e] = xj[e] + nvec[e] * reach;
                    } else {
                        const double *nj = s->nrm + 3 * j;
                        const double pen = (((nj[0] * xi[0] + nj[1] * xi[1]) + nj[2] * xi[2]) - ri) -
                                           ((nj[0] * xj[0] + nj[1] * xj[1]) + nj[2] * xj[2]);
                        if (!(pen < 0.0)) continue;
                        delta = -(pen);
                        if (delta < 0.0) continue;
                        const double reach = ri + pen;
                        for (int e = 0; e < 3; e++) { nvec[e] = nj[e]; cpt[e] = xi[e] - reach * nj[e]; }
                    }
                    const int slot = contact_slot(s, i, s->uid[j]);
                    if (slot < 0) continue;
                    spring_dashpot(s, i, j, slot, nvec, delta, cpt, pass == 0 ? fs : fp, pass == 0 ? ts : tp);
                }
            }
        }
        for (int d = 0; d < 3; d++) {
            s->f[3 * i + d] = s->f[3 * i + d] + (fs[d] + fp[d]);
            s->t[3 * i + d] = s->t[3 * i + d] + (tp[d] + ts[d]);
        }
    }
}

/* gravity (examples/dem.py:88-90): buoyancy-corrected weight on z */
static void gravity(odem_t *s) {
    for (int i = 0; i < s->nlocal; i++) {
        if (s->flags[i] & FLAG_FIXED) continue;
        const double vol = (((4.0 / 3.0) * s->pi) * s->r[i] * s->r[i]) * s->r[i];
        const double drho = (double) (int) (s->rho_p - s->rho_f);
        s->f[3 * i + 2] = s->f[3 * i + 2] - (drho * vol) * s->grav;
    }
}

/* euler (examples/dem.py:77-85) with quaternion(phi, |phi|) and
   quaternion_to_rotation_matrix (src/pairs/mapping/keywords.py:190-252) */
static void euler(odem_t *s) {
    const double dt = s->dt;
    for (int i = 0; i < s->nlocal; i++) {
        if (s->flags[i] & FLAG_FIXED) continue;
        double *x = s->x + 3 * i, *v = s->v + 3 * i, *w = s->w + 3 * i, *q = s->rot_q + 4 * i, *R = s->rot_m + 9 * i;
        const double *f = s->f + 3 * i, *tq = s->t + 3 * i, *I = s->inv_inertia + 9 * i;
        const double im = 1.0 / s->m[i];
        const double hm = 0.5 * im;
        for (int d = 0; d < 3; d++) x[d] = x[d] + (((hm * f[d]) * dt) * dt + v[d] * dt);
        for (int d = 0; d < 3; d++) v[d] = v[d] + (im * f[d]) * dt;
        double a[3], b[3], wd[3];
        for (int r = 0; r < 3; r++) a[r] = (I[3 * r] * tq[0] + I[3 * r + 1] * tq[1]) + I[3 * r + 2] * tq[2];
        for (int r = 0; r < 3; r++) b[r] = (R[3 * r] * a[0] + R[3 * r + 1] * a[1]) + R[3 * r + 2] * a[2];
        /* b * transposed(R): element c = b . (column c of R^T) = b . row c of R... as emitted: b0*R[c] + b1*R[3+c] + b2*R[6+c] */
        for (int c = 0; c < 3; c++) wd[c] = (b[0] * R[c] + b[1] * R[3 + c]) + b[2] * R[6 + c];
        double phi[3];
        for (int d = 0; d < 3; d++) phi[d] = w[d] * dt + ((0.5 * wd[d]) * dt) * dt;
        const double plen = sqrt((phi[0] * phi[0] + phi[1] * phi[1]) + phi[2] * phi[2]);
        double dq[4];
        if (fabs(plen) < 1e-06) {
            dq[0] = 1.0; dq[1] = 0.0; dq[2] = 0.0; dq[3] = 0.0;
        } else {
            const double c = cos(plen * 0.5), sn = sin(plen * 0.5), inv = 1.0 / plen;
            dq[0] = c;
            for (int d = 0; d < 3; d++) dq[1 + d] = sn * (phi[d] * inv);
        }
        const double p0 = ((dq[0] * q[0] - dq[1] * q[1]) - dq[2] * q[2]) - dq[3] * q[3];
        const double p1 = ((dq[0] * q[1] + dq[1] * q[0]) + dq[2] * q[3]) - dq[3] * q[2];
        const double p2 = ((dq[0] * q[2] + dq[2] * q[0]) + dq[3] * q[1]) - dq[1] * q[3];
        const double p3 = ((dq[0] * q[3] + dq[3] * q[0]) + dq[1] * q[2]) - dq[2] * q[1];
        const double n2 = ((p0 * p0 + p1 * p1) + p2 * p2) + p3 * p3;
        const double sc = (fabs(n2 - 1.0) < 1e-08) ? 1.0 : 1.0 / sqrt(n2);
        q[0] = p0 * sc; q[1] = p1 * sc; q[2] = p2 * sc; q[3] = p3 * sc;
        R[0] = (1.0 - (2.0 * q[2]) * q[2]) - (2.0 * q[3]) * q[3];
        R[1] = 2.0 * (q[1] * q[2] - q[0] * q[3]);
        R[2] = 2.0 * (q[1] * q[3] + q[0] * q[2]);
        R[3] = 2.0 * (q[1] * q[2] + q[0] * q[3]);
        R[4] = (1.0 - (2.0 * q[1]) * q[1]) - (2.0 * q[3]) * q[3];
        R[5] = 2.0 * (q[2] * q[3] - q[0] * q[1]);
        R[6] = 2.0 * (q[1] * q[3] - q[0] * q[2]);
        R[7] = 2.0 * (q[2] * q[3] + q[0] * q[1]);
        R[8] = (1.0 - (2.0 * q[1]) * q[1]) - (2.0 * q[2]) * q[2];
        for (int d = 0; d < 3; d++) w[d] = w[d] + wd[d] * dt;
    }
}

/* ---- contact history bookkeeping (src/pairs/sim/contact_history.py:75-127) ---- */

static void reset_usage(odem_t *s) {
    for (int i = 0; i < s->nlocal; i++)
        for (int k = 0; k < s->num_contacts[i]; k++) s->contact_used[(size_t) i * s->ccap + k] = 0;
}

static void clear_unused(odem_t *s) {
    for (int i = 0; i < s->nlocal; i++) {
        int k = 0;
        while (k < s->num_contacts[i]) {
            const size_t a = (size_t) i * s->ccap + k;
            if (s->contact_used[a] == 0) {
                const int last = s->num_contacts[i] - 1;
                if (last > 0) {
                    const size_t b = (size_t) i * s->ccap + last;
                    s->is_sticking[a] = s->is_sticking[b];
                    for (int d = 0; d < 3; d++) s->tsd[3 * a + d] = s->tsd[3 * b + d];
                    s->ivm[a] = s->ivm[b];
                    s->contact_lists[a] = s->contact_lists[b];
                    s->contact_used[a] = s->contact_used[b];
                }
                s->num_contacts[i] = s->num_contacts[i] - 1;
            } else {
                k++;
            }
        }
    }
}

/* one timestep of the generated main loop (reneighbor_every = 1) */
void odem_step(odem_t *s, int ts) {
    (void) ts;
    for (int dim = 0; dim < 3; dim++) exchange_dim(s, dim);
    borders(s);
    build_cells(s);
    reset_usage(s);
    for (int i = 0; i < s->nlocal; i++)
        for (int d = 0; d < 3; d++) { s->f[3 * i + d] = 0.0; s->t[3 * i + d] = 0.0; }
    gravity(s);
    contacts(s);
    euler(s);
    clear_unused(s);
}

void odem_run(odem_t *s, int ts_begin, int ts_end) {
    for (int ts = ts_begin; ts < ts_end; ts++) odem_step(s, ts);
}

/* array accessors for the Python side */
void *odem_ptr(odem_t *s, int which) {
    switch (which) {
    case 0: return s->x;
    case 1: return s->v;
    case 2: return s->w;
    case 3: return s->f;
    case 4: return s->t;
    case 5: return s->r;
    case 6: return s->m;
    case 7: return s->rot_q;
    case 8: return s->rot_m;
    case 9: return s->uid;
    case 10: return s->flags;
    case 11: return s->type;
    case 12: return s->num_contacts;
    case 13: return s->contact_lists;
    case 14: return s->is_sticking;
    case 15: return s->tsd;
    case 16: return s->ivm;
    case 17: return s->shape;
    case 18: return s->nrm;
    case 19: return s->inv_inertia;
    case 20: return s->contact_used;
    default: return 0;
    }
}

int odem_int(const odem_t *s, int which) {
    switch (which) {
    case 0: return s->nlocal;
    case 1: return s->nghost;
    case 2: return s->ccap;
    case 3: return s->overflow;
    case 4: return s->ncells;
    case 5: return s->nsend_all;
    default: return -1;
    }
}

void odem_set_nlocal(odem_t *s, int n) { s->nlocal = n; }
