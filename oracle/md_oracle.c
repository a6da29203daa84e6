/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the reference's generated CPU program for the LJ
 * molecular-dynamics path (examples/md.py lowered by src/pairs).  It exists to
 * check the HIP product path and to serve as the timed CPU baseline; nothing in
 * pairs_amd/ links, imports or calls it.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.
 *
 * Every function names the reference source it follows.  The operation order of
 * each floating-point expression is the one the reference's code generator
 * emits (it prints one C statement per IR node), so with -ffp-contract=off the
 * results are bit-identical to the generated program.
 *
 * Pinning (what checks each part; see tests/test_oracle_*.py):
 *   - omd_fcc_lattice / omd_adjust_thermo : golden init state of
 *     tests/golden/md_fcc4_*.npz (bit-exact) + closed forms (site count, zero
 *     momentum, T == target).
 *   - omd_step (exchange, borders, sync, cells, lists, LJ, VV, thermo): golden
 *     trajectories, neighbour lists, cell membership and send maps of the
 *     reference's generated program (bit-exact), brute-force O(N^2) image
 *     enumeration for the neighbour sets, the closed-form LJ pair force.
 *   - omd_energy (potential energy and virial; the reference has none):
 *     closed-form two-body energy, -dU/dx == F by central differences, brute
 *     force over explicit images.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FLAG_INFINITE 1
#define FLAG_FIXED 4
#define FLAG_GLOBAL 8
#define SHAPE_POINT_MASS 2

typedef struct omd {
    int cap, ncap, ccap, send_cap, ncells_cap;
    int nlocal, nghost;
    double *x, *v, *f, *m;
    int *type, *flags, *uid, *shape;
    double lo[3], hi[3], len[3];
    int periodic[3];
    int pbc[6];
    double spacing;  /* cell spacing == neighbour-list radius (rc + skin) */
    double cutoff;   /* force cutoff rc */
    double dt;
    int half, ntypes;
    double eps[64], sig6[64];
    int reneighbor_every, thermo_every;
    int dim_cells[3], ncells, nstencil, stencil[27];
    int *cell_sizes, *cell_particles, *particle_cell;
    int *numneighs, *neighborlists;
    int nsend_all, nsend[6], nrecv[6], send_offsets[6], recv_offsets[6];
    int *send_map, *send_mult, *exchg_flag, *exchg_copy_to;
    double *send_buf;
    double last_T, last_P;
    int overflow;  /* set when a capacity would be exceeded */
} omd_t;

static void *zalloc(size_t n) { void *p = calloc(1, n ? n : 1); return p; }

omd_t *omd_create(int cap, int ncap, int ccap, int send_cap, int ncells_cap,
                  const double *lo, const double *hi, const int *periodic,
                  double spacing, double cutoff, double dt, int half, int ntypes,
                  const double *eps, const double *sig6, int reneighbor_every, int thermo_every) {
    omd_t *s = (omd_t *) zalloc(sizeof(omd_t));
    s->cap = cap; s->ncap = ncap; s->ccap = ccap; s->send_cap = send_cap; s->ncells_cap = ncells_cap;
    for (int d = 0; d < 3; d++) {
        s->lo[d] = lo[d]; s->hi[d] = hi[d];
        /* the reference prints grid.length(d) = max - min as a literal
           (src/pairs/sim/grid.py:36-37, comm.py:316-317) */
        s->len[d] = hi[d] - lo[d];
        s->periodic[d] = periodic[d];
        /* one rank: pbc_prev = +1, pbc_next = -1 at the global edges
           (runtime/domain/regular_6d_stencil.cpp:72-73) */
        s->pbc[2 * d] = 1; s->pbc[2 * d + 1] = -1;
    }
    s->spacing = spacing; s->cutoff = cutoff; s->dt = dt; s->half = half; s->ntypes = ntypes;
    for (int k = 0; k < ntypes * ntypes; k++) { s->eps[k] = eps[k]; s->sig6[k] = sig6[k]; }
    s->reneighbor_every = reneighbor_every; s->thermo_every = thermo_every;
    s->x = (double *) zalloc(sizeof(double) * 3 * cap);
    s->v = (double *) zalloc(sizeof(double) * 3 * cap);
    s->f = (double *) zalloc(sizeof(double) * 3 * cap);
    s->m = (double *) zalloc(sizeof(double) * cap);
    s->type = (int *) zalloc(sizeof(int) * cap);
    s->flags = (int *) zalloc(sizeof(int) * cap);
    s->uid = (int *) zalloc(sizeof(int) * cap);
    s->shape = (int *) zalloc(sizeof(int) * cap);
    s->cell_sizes = (int *) zalloc(sizeof(int) * ncells_cap);
    s->cell_particles = (int *) zalloc(sizeof(int) * (size_t) ncells_cap * ccap);
    s->particle_cell = (int *) zalloc(sizeof(int) * cap);
    s->numneighs = (int *) zalloc(sizeof(int) * cap);
    s->neighborlists = (int *) zalloc(sizeof(int) * (size_t) cap * ncap);
    s->send_map = (int *) zalloc(sizeof(int) * send_cap);
    s->send_mult = (int *) zalloc(sizeof(int) * 3 * send_cap);
    s->exchg_flag = (int *) zalloc(sizeof(int) * cap);
    s->exchg_copy_to = (int *) zalloc(sizeof(int) * send_cap);
    s->send_buf = (double *) zalloc(sizeof(double) * 11 * (size_t) send_cap);
    return s;
}

void omd_destroy(omd_t *s) {
    if (!s) return;
    free(s->x); free(s->v); free(s->f); free(s->m); free(s->type); free(s->flags);
    free(s->uid); free(s->shape); free(s->cell_sizes); free(s->cell_particles);
    free(s->particle_cell); free(s->numneighs); free(s->neighborlists); free(s->send_map);
    free(s->send_mult); free(s->exchg_flag); free(s->exchg_copy_to); free(s->send_buf);
    free(s);
}

/* ---- initial state ------------------------------------------------------ */

/* Park-Miller minimal standard generator without masking:
   runtime/copper_fcc_lattice.hpp:11-26 */
static double park_miller(int *seed) {
    const int IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836;
    int k = (*seed) / IQ;
    *seed = IA * (*seed - k * IQ) - IR * k;
    if (*seed < 0) *seed += IM;
    return (1.0 / IM) * (*seed);
}

/* runtime/copper_fcc_lattice.hpp:64-145: FCC sites 0.5*a*(i,j,k) with i+j+k
   even, visited in 8x8x8 sub-box order; per-site velocities from the site
   index; types from glibc rand(). */
int omd_fcc_lattice(omd_t *s, int nx, int ny, int nz, double rho, int ntypes) {
    double xprd = s->hi[0] - s->lo[0], yprd = s->hi[1] - s->lo[1], zprd = s->hi[2] - s->lo[2];
    double xlo = 0.0, xhi = xprd, ylo = 0.0, yhi = yprd, zlo = 0.0, zhi = zprd;
    double alat = pow((4.0 / rho), (1.0 / 3.0));
    int ilo = (int) (xlo / (0.5 * alat) - 1), ihi = (int) (xhi / (0.5 * alat) + 1);
    int jlo = (int) (ylo / (0.5 * alat) - 1), jhi = (int) (yhi / (0.5 * alat) + 1);
    int klo = (int) (zlo / (0.5 * alat) - 1), khi = (int) (zhi / (0.5 * alat) + 1);
    if (ilo < 0) ilo = 0;
    if (ihi > 2 * nx - 1) ihi = 2 * nx - 1;
    if (jlo < 0) jlo = 0;
    if (jhi > 2 * ny - 1) jhi = 2 * ny - 1;
    if (klo < 0) klo = 0;
    if (khi > 2 * nz - 1) khi = 2 * nz - 1;
    const double SMALL = 1e-5;  /* runtime/pairs_common.hpp SMALL, isWithinSubdomain */
    int natoms = 0, sx = 0, sy = 0, sz = 0, ox = 0, oy = 0, oz = 0;
    const int sub = 8;
    while (oz * sub <= khi) {
        int k = oz * sub + sz, j = oy * sub + sy, i = ox * sub + sx;
        if (((i + j + k) % 2 == 0) && i >= ilo && i <= ihi && j >= jlo && j <= jhi &&
            k >= klo && k <= khi) {
            double xt = 0.5 * alat * i, yt = 0.5 * alat * j, zt = 0.5 * alat * k;
            if (xt >= s->lo[0] && xt < s->hi[0] - SMALL && yt >= s->lo[1] && yt < s->hi[1] - SMALL &&
                zt >= s->lo[2] && zt < s->hi[2] - SMALL) {
                int n = k * (2 * ny) * (2 * nx) + j * (2 * nx) + i + 1;
                double vx, vy, vz;
                for (int q = 0; q < 5; q++) park_miller(&n);
                vx = park_miller(&n);
                for (int q = 0; q < 5; q++) park_miller(&n);
                vy = park_miller(&n);
                for (int q = 0; q < 5; q++) park_miller(&n);
                vz = park_miller(&n);
                if (natoms < s->cap) {
                    s->m[natoms] = 1.0;
                    s->x[3 * natoms + 0] = xt; s->x[3 * natoms + 1] = yt; s->x[3 * natoms + 2] = zt;
                    s->v[3 * natoms + 0] = vx; s->v[3 * natoms + 1] = vy; s->v[3 * natoms + 2] = vz;
                    s->type[natoms] = rand() % ntypes;
                    s->flags[natoms] = 0;
                    s->shape[natoms] = SHAPE_POINT_MASS;
                }
                natoms++;
            }
        }
        sx++;
        if (sx == sub) { sx = 0; sy++; }
        if (sy == sub) { sy = 0; sz++; }
        if (sz == sub) { sz = 0; ox++; }
        if (ox * sub > ihi) { ox = 0; oy++; }
        if (oy * sub > jhi) { oy = 0; oz++; }
    }
    s->nlocal = natoms;
    s->nghost = 0;
    return natoms;
}

/* runtime/thermo.hpp:11-51 (temperature) */
double omd_temperature(const omd_t *s) {
    const double mvv2e = 1.0;
    const double dof_boltz = (s->nlocal * 3 - 3);
    const double t_scale = mvv2e / dof_boltz;
    double t = 0.0;
    for (int i = 0; i < s->nlocal; i++) {
        const double *v = s->v + 3 * i;
        t += s->m[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    return t * t_scale;
}

/* runtime/thermo.hpp:53-97 */
void omd_adjust_thermo(omd_t *s, double temp) {
    double vx = 0.0, vy = 0.0, vz = 0.0;
    int n = s->nlocal;
    for (int i = 0; i < n; i++) { vx += s->v[3 * i]; vy += s->v[3 * i + 1]; vz += s->v[3 * i + 2]; }
    vx /= n; vy /= n; vz /= n;
    for (int i = 0; i < n; i++) { s->v[3 * i] -= vx; s->v[3 * i + 1] -= vy; s->v[3 * i + 2] -= vz; }
    double t = omd_temperature(s);
    double factor = sqrt(temp / t);
    for (int i = 0; i < n; i++) { s->v[3 * i] *= factor; s->v[3 * i + 1] *= factor; s->v[3 * i + 2] *= factor; }
}

void omd_set_particles(omd_t *s, int n, const double *x, const double *v, const double *m,
                       const int *type, const int *flags) {
    for (int i = 0; i < n && i < s->cap; i++) {
        for (int d = 0; d < 3; d++) { s->x[3 * i + d] = x[3 * i + d]; s->v[3 * i + d] = v[3 * i + d]; }
        s->m[i] = m[i]; s->type[i] = type[i]; s->flags[i] = flags ? flags[i] : 0;
        s->shape[i] = SHAPE_POINT_MASS;
    }
    s->nlocal = n; s->nghost = 0;
}

/* ---- cells ---------------------------------------------------------------- */

/* src/pairs/sim/cell_lists.py:51-87 (build_cell_lists_stencil) */
void omd_setup_cells(omd_t *s) {
    const double sp = s->spacing;
    int total = 1;
    for (int d = 0; d < 3; d++) {
        double dmax = s->hi[d] + sp, dmin = s->lo[d] - sp;
        s->dim_cells[d] = (int) ceil((dmax - dmin) / sp) + 1;
        total *= s->dim_cells[d];
    }
    s->ncells = total + 1;
    if (s->ncells + 1 >= s->ncells_cap) s->overflow |= 1;
    s->nstencil = 0;
    for (int a = -1; a < 2; a++)
        for (int b = -1; b < 2; b++)
            for (int c = -1; c < 2; c++)
                s->stencil[s->nstencil++] = (a * s->dim_cells[1] + b) * s->dim_cells[2] + c;
}

/* cell id of one position: src/pairs/sim/cell_lists.py:114-126 */
static int cell_of(const omd_t *s, const double *x) {
    int idx = 0;
    for (int d = 0; d < 3; d++) {
        double q = (x[d] - (s->lo[d] - s->spacing)) / s->spacing;
        int c = (int) q >= 0 ? (int) q : 0;
        c = c < s->dim_cells[d] ? c : s->dim_cells[d] - 1;
        idx = d == 0 ? c : idx * s->dim_cells[d] + c;
    }
    return idx + 1;
}

/* src/pairs/sim/cell_lists.py:90-134: atomic insertion, on the CPU in index order */
static void build_cells(omd_t *s) {
    for (int c = 0; c < s->ncells; c++) s->cell_sizes[c] = 0;
    int nall = s->nlocal + s->nghost;
    for (int i = 0; i < nall; i++) {
        int flat = 0;
        if (!(s->flags[i] & FLAG_INFINITE)) flat = cell_of(s, s->x + 3 * i);
        if (flat >= 0 && flat < s->ncells) {
            s->particle_cell[i] = flat;
            int slot = s->cell_sizes[flat];
            if (slot + 1 >= s->ccap) { s->overflow |= 2; }
            else s->cell_sizes[flat] = slot + 1;
            s->cell_particles[(size_t) flat * s->ccap + slot] = i;
        }
    }
    /* partition_cell_lists (cell_lists.py:137-171) with one shape: identity */
}

/* src/pairs/sim/neighbor_lists.py:21-48 + interaction.py:90-120 (cell path) */
static void build_neighbors(omd_t *s) {
    const double rc2 = s->spacing * s->spacing;
    for (int i = 0; i < s->nlocal; i++) s->numneighs[i] = 0;
    /* each i fills only its own list: the threads' results equal the serial loop's */
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->nlocal; i++) {
        if (s->flags[i] & FLAG_FIXED) continue;
        int ci = s->particle_cell[i];
        const double *xi = s->x + 3 * i;
        for (int k = -1; k < s->nstencil; k++) {
            int nc = k < 0 ? 0 : ci + s->stencil[k];
            if (!(k < 0 || (nc > 0 && nc < s->ncells))) continue;
            int ncell = s->cell_sizes[nc];
            for (int q = 0; q < ncell; q++) {
                int j = s->cell_particles[(size_t) nc * s->ccap + q];
                int accept = s->half ? (s->shape[j] > SHAPE_POINT_MASS ||
                                        (s->shape[j] == SHAPE_POINT_MASS && i < j))
                                     : (j != i);
                if (!accept) continue;
                const double *xj = s->x + 3 * j;
                double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
                double rsq = dx * dx + dy * dy;
                rsq = rsq + dz * dz;
                if (rsq < rc2) {
                    int n = s->numneighs[i];
                    s->neighborlists[(size_t) i * s->ncap + n] = j;
                    if (n + 1 + 1 >= s->ncap) {
                        #pragma omp atomic
                        s->overflow |= 4;
                    } else {
                        s->numneighs[i] = n + 1;
                    }
                }
            }
        }
    }
}

/* ---- ghosts and exchange --------------------------------------------------- */

/* DetermineGhostParticles (src/pairs/sim/comm.py:223-259) with the predicate of
   DimensionRanges.ghost_particles (src/pairs/sim/domain_partitioning.py:31-65):
   two passes over [0, nlocal+nghost), "< min+offset" then "> max-offset". */
static void determine(omd_t *s, int dim, double offset, int exchange) {
    int nall = s->nlocal + s->nghost;
    if (exchange) for (int i = 0; i < s->nlocal; i++) s->exchg_flag[i] = 0;
    if (!s->periodic[dim]) return;  /* one rank: pbc flags are non-zero only when periodic */
    for (int side = 0; side < 2; side++) {
        int j = dim * 2 + side;
        for (int i = 0; i < nall; i++) {
            if (s->flags[i] & (FLAG_INFINITE | FLAG_GLOBAL)) continue;
            double xd = s->x[3 * i + dim];
            int sel = side == 0 ? (xd < s->lo[dim] + offset) : (xd > s->hi[dim] - offset);
            if (!sel) continue;
            int idx = s->nsend_all++;
            if (idx >= s->send_cap) { s->overflow |= 8; continue; }
            s->send_map[idx] = i;
            if (exchange) s->exchg_flag[i] = 1;
            for (int d = 0; d < 3; d++) s->send_mult[3 * idx + d] = d == dim ? s->pbc[j] : 0;
            s->nsend[j]++;
        }
    }
}

static inline double shifted(const omd_t *s, int idx, int d, double x) {
    /* pack: src += send_mult * length (comm.py:312-318) */
    return x + s->send_mult[3 * idx + d] * s->len[d];
}

/* Comm.exchange for one dimension (src/pairs/sim/comm.py:100-151, 445-520) */
static void exchange_dim(omd_t *s, int dim) {
    s->nsend_all = 0;
    s->nghost = 0;
    for (int j = 0; j < 2 * (dim + 1); j++) { s->nsend[j] = 0; s->nrecv[j] = 0; s->send_offsets[j] = 0; s->recv_offsets[j] = 0; }
    determine(s, dim, 0.0, 1);
    int n = s->nsend_all;
    /* pack (all non-volatile properties) */
    for (int i = 0; i < n; i++) {
        int m = s->send_map[i];
        double *b = s->send_buf + 11 * (size_t) i;
        b[0] = s->uid[m]; b[1] = s->shape[m]; b[2] = s->flags[m];
        for (int d = 0; d < 3; d++) b[3 + d] = shifted(s, i, d, s->x[3 * m + d]);
        b[6] = s->m[m];
        for (int d = 0; d < 3; d++) b[7 + d] = s->v[3 * m + d];
        b[10] = s->type[m];
    }
    /* remove_exchanged_particles_pt1 (host) */
    int send_pos = s->nlocal - 1;
    for (int i = 0; i < n; i++) {
        int pid = s->send_map[i];
        if (pid < s->nlocal - n) {
            while (s->exchg_flag[send_pos] == 1) send_pos--;
            s->exchg_copy_to[i] = send_pos;
            send_pos--;
        } else {
            s->exchg_copy_to[i] = -1;
        }
    }
    /* remove_exchanged_particles_pt2: copy guard is "src > 0" (comm.py:481) */
    for (int i = 0; i < n; i++) {
        int src = s->exchg_copy_to[i];
        if (src > 0) {
            int dst = s->send_map[i];
            s->uid[dst] = s->uid[src]; s->shape[dst] = s->shape[src]; s->flags[dst] = s->flags[src];
            for (int d = 0; d < 3; d++) s->x[3 * dst + d] = s->x[3 * src + d];
            s->m[dst] = s->m[src];
            for (int d = 0; d < 3; d++) s->v[3 * dst + d] = s->v[3 * src + d];
            s->type[dst] = s->type[src];
        }
    }
    s->nlocal -= n;
    /* unpack at nlocal + i (recv offsets of this step are 0) */
    for (int i = 0; i < n; i++) {
        const double *b = s->send_buf + 11 * (size_t) i;
        int dst = s->nlocal + i;
        s->uid[dst] = (int) b[0]; s->shape[dst] = (int) b[1]; s->flags[dst] = (int) b[2];
        for (int d = 0; d < 3; d++) s->x[3 * dst + d] = b[3 + d];
        s->m[dst] = b[6];
        for (int d = 0; d < 3; d++) s->v[3 * dst + d] = b[7 + d];
        s->type[dst] = (int) b[10];
    }
    s->nlocal += n;
}

/* Comm.borders (src/pairs/sim/comm.py:56-98): per dimension, determine, pack
   uid/type/mass/position(+shift)/velocity/shape, unpack at nlocal+offset.
   Ghost flags are not sent. */
static void borders(omd_t *s) {
    s->nsend_all = 0;
    s->nghost = 0;
    for (int dim = 0; dim < 3; dim++) {
        int start = s->nsend_all;
        s->nsend[2 * dim] = s->nsend[2 * dim + 1] = 0;
        determine(s, dim, s->spacing, 0);
        int end = s->nsend_all;
        for (int i = start; i < end; i++) {
            int m = s->send_map[i];
            int g = s->nlocal + i;
            /* pack then unpack: source index m < g always, values read first */
            double xs[3];
            for (int d = 0; d < 3; d++) xs[d] = shifted(s, i, d, s->x[3 * m + d]);
            s->uid[g] = s->uid[m]; s->type[g] = s->type[m]; s->m[g] = s->m[m];
            for (int d = 0; d < 3; d++) { s->x[3 * g + d] = xs[d]; s->v[3 * g + d] = s->v[3 * m + d]; }
            s->shape[g] = s->shape[m];
        }
        s->nghost += end - start;
    }
}

/* Comm.synchronize (src/pairs/sim/comm.py:45-54, 369-442): ALL ghosts are
   packed before ANY is unpacked, so ghosts of ghosts see last step's values. */
static void synchronize(omd_t *s) {
    int n = s->nsend_all;
    for (int i = 0; i < n; i++) {
        int m = s->send_map[i];
        double *b = s->send_buf + 6 * (size_t) i;
        for (int d = 0; d < 3; d++) b[d] = shifted(s, i, d, s->x[3 * m + d]);
        for (int d = 0; d < 3; d++) b[3 + d] = s->v[3 * m + d];
    }
    for (int i = 0; i < n; i++) {
        const double *b = s->send_buf + 6 * (size_t) i;
        int g = s->nlocal + i;
        for (int d = 0; d < 3; d++) { s->x[3 * g + d] = b[d]; s->v[3 * g + d] = b[3 + d]; }
    }
}

/* ---- forces and integration ------------------------------------------------ */

/* lennard_jones (examples/md.py:5-8) lowered by sim/interaction.py:201-292 and
   ir/apply.py:111-125; feature index t_i*ntypes+t_j (ir/features.py:155). */
static void lennard_jones(omd_t *s) {
    const double rc2 = s->cutoff * s->cutoff;
    const int nt = s->ntypes;
    for (int i = 0; i < s->nlocal; i++) {
        s->f[3 * i] = 0.0; s->f[3 * i + 1] = 0.0; s->f[3 * i + 2] = 0.0;  /* reset_volatile_properties */
    }
    /* full lists: each i writes only force[i], so threads may split the loop;
       half lists add mirror forces to other particles in loop order (serial) */
    #pragma omp parallel for schedule(static) if (!s->half)
    for (int i = 0; i < s->nlocal; i++) {
        if (s->flags[i] & FLAG_FIXED) continue;
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
        const double xi0 = s->x[3 * i], xi1 = s->x[3 * i + 1], xi2 = s->x[3 * i + 2];
        const int ti = s->type[i];
        const int nn = s->numneighs[i];
        for (int k = 0; k < nn; k++) {
            int j = s->neighborlists[(size_t) i * s->ncap + k];
            double dx = xi0 - s->x[3 * j], dy = xi1 - s->x[3 * j + 1], dz = xi2 - s->x[3 * j + 2];
            double rsq = dx * dx + dy * dy;
            rsq = rsq + dz * dz;
            if (rsq < rc2) {
                int tj = s->type[j];
                double sr2 = 1.0 / rsq;
                double sr6 = sr2 * sr2;
                sr6 = sr6 * sr2;
                sr6 = sr6 * s->sig6[ti * nt + tj];
                double sc = 48.0 * sr6;
                sc = sc * (sr6 - 0.5);
                sc = sc * sr2;
                sc = sc * s->eps[ti * nt + tj];
                double f0 = dx * sc, f1 = dy * sc, f2 = dz * sc;
                t0 = t0 + f0; t1 = t1 + f1; t2 = t2 + f2;
                if (s->half && j < s->nlocal && !(s->flags[j] & FLAG_FIXED)) {
                    s->f[3 * j] += -f0; s->f[3 * j + 1] += -f1; s->f[3 * j + 2] += -f2;
                }
            }
        }
        s->f[3 * i] = s->f[3 * i] + t0;
        s->f[3 * i + 1] = s->f[3 * i + 1] + t1;
        s->f[3 * i + 2] = s->f[3 * i + 2] + t2;
    }
}

/* initial_integrate (examples/md.py:11-13) */
static void initial_integrate(omd_t *s) {
    const double hdt = s->dt * 0.5;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->nlocal; i++) {
        if (s->flags[i] & FLAG_FIXED) continue;
        for (int d = 0; d < 3; d++) {
            double a = hdt * s->f[3 * i + d];
            a = a / s->m[i];
            s->v[3 * i + d] = s->v[3 * i + d] + a;
        }
        for (int d = 0; d < 3; d++) s->x[3 * i + d] = s->x[3 * i + d] + s->dt * s->v[3 * i + d];
    }
}

/* final_integrate (examples/md.py:16-17) */
static void final_integrate(omd_t *s) {
    const double hdt = s->dt * 0.5;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->nlocal; i++) {
        if (s->flags[i] & FLAG_FIXED) continue;
        for (int d = 0; d < 3; d++) {
            double a = hdt * s->f[3 * i + d];
            a = a / s->m[i];
            s->v[3 * i + d] = s->v[3 * i + d] + a;
        }
    }
}

/* the full neighbour rebuild group of one step (simulation.py:392-409) */
void omd_rebuild(omd_t *s) {
    for (int d = 0; d < 3; d++) exchange_dim(s, d);
    borders(s);
    build_cells(s);
    build_neighbors(s);
}

void omd_force(omd_t *s) { lennard_jones(s); }

/* The older API's modules (examples/lj_onetype.py:5-13), full lists:
   reset_volatile_properties and `lj`, whose pair force
   delta * 48.0 * sr6 * (sr6 - 0.5) * sr2 * epsilon is evaluated left to right
   per component, then `euler`: v += (dt * f) / m, x += dt * v. */
void omd_lj_older(omd_t *s) {
    const double rc2 = s->cutoff * s->cutoff;
    const int nt = s->ntypes;
    for (int i = 0; i < s->nlocal; i++) { s->f[3 * i] = 0.0; s->f[3 * i + 1] = 0.0; s->f[3 * i + 2] = 0.0; }
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->nlocal; i++) {
        if (s->flags[i] & FLAG_FIXED) continue;
        double t[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < s->numneighs[i]; k++) {
            const int j = s->neighborlists[(size_t) i * s->ncap + k];
            double d[3];
            for (int e = 0; e < 3; e++) d[e] = s->x[3 * i + e] - s->x[3 * j + e];
            double rsq = d[0] * d[0] + d[1] * d[1];
            rsq = rsq + d[2] * d[2];
            if (rsq < rc2) {
                const int tk = s->type[i] * nt + s->type[j];
                const double sr2 = 1.0 / rsq;
                double sr6 = sr2 * sr2;
                sr6 = sr6 * sr2;
                sr6 = sr6 * s->sig6[tk];
                for (int e = 0; e < 3; e++) {
                    double c = d[e] * 48.0;
                    c = c * sr6;
                    c = c * (sr6 - 0.5);
                    c = c * sr2;
                    c = c * s->eps[tk];
                    t[e] = t[e] + c;
                }
            }
        }
        for (int e = 0; e < 3; e++) s->f[3 * i + e] = s->f[3 * i + e] + t[e];
    }
}

void omd_euler(omd_t *s) {
    for (int i = 0; i < s->nlocal; i++) {
        if (s->flags[i] & FLAG_FIXED) continue;
        for (int e = 0; e < 3; e++) {
            double a = s->dt * s->f[3 * i + e];
            a = a / s->m[i];
            s->v[3 * i + e] = s->v[3 * i + e] + a;
        }
        for (int e = 0; e < 3; e++) s->x[3 * i + e] = s->x[3 * i + e] + s->dt * s->v[3 * i + e];
    }
}
void omd_sync(omd_t *s) { synchronize(s); }

static int every(int ts, int n) { return n > 0 && (((ts + 1) % n) == 0 || ts == 0); }

/* One iteration ts of the timestep loop (src/pairs/sim/timestep.py:9-72,
   src/pairs/sim/simulation.py:382-452): the loop runs ts = 0..nsteps. */
void omd_step(omd_t *s, int ts) {
    if (ts > 0) initial_integrate(s);
    if (every(ts, s->reneighbor_every)) {
        for (int d = 0; d < 3; d++) exchange_dim(s, d);
        borders(s);
        build_cells(s);
        build_neighbors(s);
    } else {
        synchronize(s);
    }
    lennard_jones(s);
    if (ts > 0) final_integrate(s);
    if (every(ts, s->thermo_every)) {
        double t = omd_temperature(s);
        double V = s->len[0] * s->len[1] * s->len[2];
        double dof = (s->nlocal * 3 - 3);
        s->last_T = t;
        s->last_P = (t * dof) * (1.0 / 3 / s->len[0] / s->len[1] / s->len[2]);
        (void) V;
    }
}

/* threads of the OpenMP loops (1: the reference's serial target_cpu()) */
void omd_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int omd_threads(void) { return omp_get_max_threads(); }

void omd_run(omd_t *s, int ts_begin, int ts_end) {
    for (int ts = ts_begin; ts < ts_end; ts++) omd_step(s, ts);
}

/* Build-defined potential energy and virial over the current neighbour list
   (the reference has none): U = sum over pairs r<rc of 4 eps [(s6/r6)^2 - s6/r6],
   W = sum over pairs of r_ij . F_ij.  A pair seen from both sides (full list, or
   a local-ghost pair of a half list) is weighted 1/2.  Per-atom energies e_i are
   the half-weighted sums over i's list. */
void omd_energy(const omd_t *s, double *pe, double *virial, double *e_atom) {
    const double rc2 = s->cutoff * s->cutoff;
    const int nt = s->ntypes;
    double U = 0.0, W = 0.0;
    for (int i = 0; i < s->nlocal; i++) {
        double ei = 0.0;
        if (s->flags[i] & FLAG_FIXED) { if (e_atom) e_atom[i] = 0.0; continue; }
        for (int k = 0; k < s->numneighs[i]; k++) {
            int j = s->neighborlists[(size_t) i * s->ncap + k];
            double dx = s->x[3 * i] - s->x[3 * j], dy = s->x[3 * i + 1] - s->x[3 * j + 1],
                   dz = s->x[3 * i + 2] - s->x[3 * j + 2];
            double rsq = dx * dx + dy * dy + dz * dz;
            if (rsq >= rc2) continue;
            int t = s->type[i] * nt + s->type[j];
            double sr2 = 1.0 / rsq, sr6 = sr2 * sr2 * sr2 * s->sig6[t];
            double u = 4.0 * s->eps[t] * sr6 * (sr6 - 1.0);
            double w = 48.0 * sr6 * (sr6 - 0.5) * s->eps[t];  /* = rsq * (F.d/rsq) */
            double wt = (!s->half || j >= s->nlocal) ? 0.5 : 1.0;
            U += wt * u;
            W += wt * w;
            ei += 0.5 * u;
        }
        if (e_atom) e_atom[i] = ei;
    }
    *pe = U;
    *virial = W;
}

/* ---- accessors for the test harness ----------------------------------------- */
enum {
    OMD_X = 0, OMD_V, OMD_F, OMD_M, OMD_TYPE, OMD_FLAGS, OMD_NUMNEIGHS, OMD_NEIGHBORLISTS,
    OMD_PARTICLE_CELL, OMD_CELL_SIZES, OMD_CELL_PARTICLES, OMD_SEND_MAP, OMD_SEND_MULT,
    OMD_NSEND, OMD_STENCIL, OMD_DIM_CELLS, OMD_UID
};

void *omd_ptr(omd_t *s, int which) {
    switch (which) {
        case OMD_X: return s->x;
        case OMD_V: return s->v;
        case OMD_F: return s->f;
        case OMD_M: return s->m;
        case OMD_TYPE: return s->type;
        case OMD_FLAGS: return s->flags;
        case OMD_NUMNEIGHS: return s->numneighs;
        case OMD_NEIGHBORLISTS: return s->neighborlists;
        case OMD_PARTICLE_CELL: return s->particle_cell;
        case OMD_CELL_SIZES: return s->cell_sizes;
        case OMD_CELL_PARTICLES: return s->cell_particles;
        case OMD_SEND_MAP: return s->send_map;
        case OMD_SEND_MULT: return s->send_mult;
        case OMD_NSEND: return s->nsend;
        case OMD_STENCIL: return s->stencil;
        case OMD_DIM_CELLS: return s->dim_cells;
        case OMD_UID: return s->uid;
    }
    return 0;
}

enum { OMD_NLOCAL = 0, OMD_NGHOST, OMD_NCELLS, OMD_NSEND_ALL, OMD_OVERFLOW, OMD_NCAP, OMD_CCAP, OMD_CAP };

int omd_int(const omd_t *s, int which) {
    switch (which) {
        case OMD_NLOCAL: return s->nlocal;
        case OMD_NGHOST: return s->nghost;
        case OMD_NCELLS: return s->ncells;
        case OMD_NSEND_ALL: return s->nsend_all;
        case OMD_OVERFLOW: return s->overflow;
        case OMD_NCAP: return s->ncap;
        case OMD_CCAP: return s->ccap;
        case OMD_CAP: return s->cap;
    }
    return -1;
}

double omd_last_temperature(const omd_t *s) { return s->last_T; }
double omd_last_pressure(const omd_t *s) { return s->last_P; }
