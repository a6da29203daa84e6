// Device state and launchers of the LJ molecular-dynamics path.
//
// Storage layout in HBM.  Particles live in "physical" slots: locals first,
// sorted by cell, then ghosts, sorted by cell; every per-particle array is SoA
// (x, y, z, vx, ..., one array per component).  The reference's own particle
// index ("reference index": the slot the generated CPU program would use) is
// kept as a relabeling, ref[physical] and phys[reference], updated exactly as
// the reference's exchange and borders would move particles; every
// order-dependent rule (exchange tail fill, ghost order, cell order, half-list
// i < j, neighbour order) is applied to reference indices.  Positions are
// double-buffered (X[2], Y[2], Z[2]) so that the fused force/integrate kernel
// never overwrites a position another thread gathers.
//
// Work decomposition.  The neighbour build works on "segments": ZC
// consecutive cells of one z-column of the cell grid, whose home particles
// are one contiguous range of physical local slots, and whose candidate cells
// (cell 0 plus the 3x3 columns around it, z range widened by one) are 20
// contiguous physical runs: the segment's tile, staged in LDS once and shared
// by all its home particles.  Neighbour lists hold physical slots (32 bits),
// in the reference's candidate order.  The force kernels work on "waves": 64
// consecutive physical locals, one per lane, gathering neighbour positions
// from a packed (x, y, z, 0) copy of the position buffer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace pk {

enum { FLAG_INFINITE = 1, FLAG_GHOST = 2, FLAG_FIXED = 4, FLAG_GLOBAL = 8 };
enum { SHAPE_SPHERE = 0, SHAPE_HALFSPACE = 1, SHAPE_POINT_MASS = 2 };

// device-resident counters (read by kernels so launches need no host round trip)
enum {
    C_NLOCAL = 0,
    C_NGHOST = 1,
    C_NSEND_ALL = 2,
    C_NSEND = 3,  // 6 entries: per dimension, side 0 then side 1
    C_RANGE = 9,  // particle range of the current compaction
    C_BASE0 = 10,
    C_BASE1 = 11,
    C_OVF = 12,   // first overflow kind (0 = none)
    C_REQ_NEIGH = 13,
    C_REQ_SEND = 14,
    C_REQ_PART = 15,
    C_TILE_MAX = 16,  // largest segment tile of the current lists
    C_UNSUPPORTED = 17,
    C_IMG_OVF = 18,   // a local has more periodic images than IMG_MAX
    C_COUNT = 19
};
enum { OVF_NONE = 0, OVF_PART = 1, OVF_NEIGH = 2, OVF_SEND = 3, OVF_TILE = 4 };

// cell grid and domain, passed by value
struct Geom {
    double lo[3], hi[3], len[3];
    double spacing;
    int dim_cells[3];
    int ncells;
    int nstencil;
    int stencil[27];
    int zc;        // cells per segment along z
    int nchunks;   // segments per z-column
    int nsegs;     // dim_cells[0] * dim_cells[1] * nchunks
};

// Neighbour lists: entry k of physical local i.  Rows come in groups of four
// so that one 16-byte load gives a lane four consecutive entries; a group
// holds, for every particle, its four entries side by side.
__host__ __device__ __forceinline__ size_t nl_at(int k, int i, int nl_stride) {
    return ((size_t) (k >> 2) * nl_stride + i) * 4 + (k & 3);
}

constexpr int IMG_MAX = 8;

// what a position-advancing kernel needs to refresh the ghost images
struct Images {
    const int *cnt, *slot, *src, *mult;
    int cap, on;
};

struct LJTab {
    int nt;
    int unit;  // one type with epsilon = sigma6 = 1
    double eps[16], sig6[16];  // ntypes <= 4
};

struct LJParams {
    double rc2, dt, hdt;
    int ntypes;
    double eps[16], sig6[16];
};

struct DevState {
    int cap, ncap, send_cap, ncells_cap, nl_stride;
    int nlocal;   // host mirror (constant between uploads on one domain)
    int cur;      // which position buffer holds the current state
    int tile_cap; // LDS tile entries the neighbour build may stage
    int pstride;  // row stride of the partials array
    int max_lds;  // shared memory a workgroup may allocate
    double *X[2], *Y[2], *Z[2];
    double4 *P[2];       // (x, y, z, 0) of each position buffer, packed for the force kernels' gathers
    double *vx, *vy, *vz, *fx, *fy, *fz, *mass;
    int *type, *flags, *uid, *shape, *ref;
    int *phys;    // physical slot of each reference index
    // spare arrays the rebuild permutes into (swapped with the live ones)
    double *alt_vx, *alt_vy, *alt_vz, *alt_mass;
    int *alt_type, *alt_flags, *alt_uid, *alt_shape, *alt_ref, *alt_cell;
    int *numneighs;
    uint32_t *nl;  // entry k of physical local i at nl_at(k, i, nl_stride): physical slot
    // half lists: rows [rev0, 2 rev0) hold each local's mirror part (the
    // earlier-index locals whose half list holds it, by reference index);
    // rev0 = 0 for full lists
    int rev0;
    int *nrev;
    int *particle_cell, *slot, *perm;
    int *lcount, *gcount, *lstart, *gstart;  // per-cell counts and CSR starts (locals / ghosts)
    int *send_map, *send_mult, *exchg_flag;
    int *sync_src;       // per ghost (physical order): physical source slot
    int *sync_mult;      // per ghost (physical order): 3 image multipliers
    // periodic images of each local (its ghosts, and ghosts of those): the
    // kernel that advances a local's position also writes its images' next
    // positions, so no separate ghost refresh runs between steps
    int *img_cnt;        // per physical local
    int *img_slot;       // [k][cap]: ghost index (physical - nlocal) of image k
    int img_on;          // host: every local's images fit IMG_MAX
    int *tile_cnt;
    int *counters;
    int *scratch_i;      // 4 * send_cap
    double *stage;       // 10 * send_cap doubles
    double *partials;    // 3 rows of pstride: per force wave (kinetic, energy, virial)
    double *thermo;      // 5 per thermo record
};

void launch_fill_i32(hipStream_t s, int *p, int v, int n);
// P[cur] = (X[cur], Y[cur], Z[cur], 0) for locals and ghosts
void launch_pack_p(hipStream_t s, const DevState &d, int cur);
// waves of the force kernels: 64 locals each
int force_waves(const DevState &d);
// waves of the fused LJ force kernel (two lanes per local: 32 locals per wave)
int lj_waves(const DevState &d);
// the older API's modules (examples/lj_onetype.py): forces with its pair
// expression (full lists), and its Euler update in the current buffer
void launch_lj_older(hipStream_t s, const DevState &d, const Geom &g, const LJParams &prm);
void launch_euler_md(hipStream_t s, const DevState &d, double dt);
void launch_fill_f64(hipStream_t s, double *p, double v, int n);
void launch_set_counters(hipStream_t s, int *c, int first, int count, int v);
// dst[r * C_COUNT + k] = src[r][k] for r < n (device array of counter pointers)
void launch_gather_counters(hipStream_t s, const int *const *src, int n, int *dst);
// the locals' state a rebuild changes in place, and the counters
struct CkptArrays {
    double *X, *Y, *Z, *vx, *vy, *vz, *mass;
    int *type, *flags, *uid, *shape, *ref, *phys, *counters;
};
// copy the first n locals' entries (and all counters) from a to b
void launch_checkpoint(hipStream_t s, const CkptArrays &a, const CkptArrays &b, int n);
void launch_identity(hipStream_t s, int *p, int n);
void launch_exchange(hipStream_t s, const DevState &d, const Geom &g, const int *periodic);
void launch_borders(hipStream_t s, const DevState &d, const Geom &g, const int *periodic);
// bins locals and ghosts into cells and permutes every particle array into
// cell order (swapping d's spare arrays in), then builds the ghost sync maps
void launch_bin_sort(hipStream_t s, DevState &d, const Geom &g);
void launch_build_neighbors(hipStream_t s, const DevState &d, const Geom &g, int half, double rc2);
// largest tile (entries) the build and force kernels can stage in LDS
int build_tile_limit(const DevState &d, const Geom &g, int half);

// one system of a batched cell phase
struct RepDesc {
    DevState d;       // d.cur: the position buffer before the phase
    Geom g;
    int periodic[3];
};
// exchange, borders and cell binning of n systems, each kernel one launch
// (blockIdx.y = system); max_cap / max_ncells: the largest system's
void launch_cells_batch(hipStream_t s, const RepDesc *D, int n, int max_cap, int max_ncells);
// thermo rows (T, P, PE, W, P total) of n systems after a fused step with
// reduction, out[5 r + k]
void launch_thermo_batch(hipStream_t s, const RepDesc *D, int n, double *out);

// the neighbour build's arguments for one system (position buffer d.cur)
struct BuildArgs {
    const int *counters;
    int *counters_rw;
    Geom g;
    const int *lstart, *gstart;
    const double *X, *Y, *Z;
    const int *ref, *flags;
    double rc2;
    uint32_t *nl;
    int nl_stride, ncap, tcap, zt, rev0;
    int *numneighs, *nrev;
};
BuildArgs build_args(const DevState &d, const Geom &g, int half, double rc2);
// dynamic LDS of one system's build workgroup
size_t build_lds_bytes(const DevState &d, const Geom &g, int half);
// the builds of many systems in one launch (block b: system rep[b], its
// segment b - seg_base[rep[b]]); lds = the largest system's need
void launch_build_batch(hipStream_t s, const BuildArgs *desc, const int *rep, const int *seg_base, int nblocks,
                        size_t lds, int half);
void launch_expand_neighbors(hipStream_t s, const DevState &d, const Geom &g, int *out_ref, int ncap_out);
void launch_sync_pp(hipStream_t s, const DevState &d, const Geom &g, int cur, int prev, int grid);
void launch_sync_inplace(hipStream_t s, const DevState &d, const Geom &g, int grid, int with_vel);
// mode: 0 force only, 1 force + integrators (fused), 2 energy only (no writes)
// The force kernel gathers positions from P[cur]: with pack, a pass first
// packs it from X, Y, Z; after_pack, if given, is recorded between that pass
// and the force kernel.  With fuse and do_vv1 the kernel also writes P[next]
// for every local and (images on) every ghost.
void launch_lj_full(hipStream_t s, const DevState &d, const Geom &g, const LJParams &prm, int cur, int next, int fuse,
                    int do_vv2, int do_vv1, int reduce, int pack = 1, hipEvent_t after_pack = nullptr);
void launch_energy(hipStream_t s, const DevState &d, const Geom &g, const LJParams &prm, int half, double *eatom);
void launch_vv_initial(hipStream_t s, const DevState &d, const LJParams &prm);
void launch_vv_post(hipStream_t s, const DevState &d, const Geom &g, const LJParams &prm, int cur, int next,
                    int do_vv2, int do_vv1, int reduce, double4 *Pn = nullptr);
void launch_ke_partial(hipStream_t s, const DevState &d);
// thermo row from the partials: kinetic row over nb_ke entries, energy and
// virial rows over nb_f entries
void launch_thermo_reduce(hipStream_t s, const DevState &d, int nb_ke, int nb_f, int which_mask, const Geom &g,
                          double *out);

// every argument of the full-list force kernel (one system)
struct LJArgs {
    const int *counters;
    Geom g;
    const int *lstart, *gstart;
    const double *X, *Y, *Z;
    const double4 *P;    // packed (x, y, z, 0) of X, Y, Z
    double4 *Pn;         // packed copy of the next positions, or null
    double *Xn, *Yn, *Zn;
    const int *type;
    const uint32_t *nl;
    int nl_stride;
    const int *numneighs, *flags;
    const int *nrev;     // half lists: length of each particle's mirror part
    int rev0;            // half lists: first list row of the mirror part
    int ncap;            // list rows of each part: counts are clamped to it
    double *fx, *fy, *fz, *vx, *vy, *vz;
    const double *mass;
    double rc2, dt, hdt;
    LJTab tab;
    int do_vv2, do_vv1, pstride;
    double *partials, *eatom;
    Images im;
};
// the fused force + integrate arguments of one system (cur -> next buffer)
LJArgs lj_args(const DevState &d, const Geom &g, const LJParams &prm, int cur, int next, int do_vv2, int do_vv1,
               int fused, double *eatom, bool packed = true);
// the packed copy (P) of the buffer each of n descriptors reads (max_cap:
// the largest system's particle capacity)
void launch_pack_p_batch(hipStream_t s, const LJArgs *desc, int n, int max_cap);
// fused force + integrate for many systems: block b is wave b - wave_base[rep[b]] of system rep[b]
// the ghost-image table of a kernel that advances positions (on only when
// it writes the next positions into the other buffer and every local's
// images fit the table)
Images images_of(const DevState &d, bool writes_next);
// one persistent launch over the wave table of many systems: the resident
// workgroups, or at most max_wg of them when max_wg > 0
void launch_lj_full_batch(hipStream_t s, const LJArgs *desc, const int *rep, const int *wave_base, int nblocks,
                          int ntypes1, int unit, int reduce, int max_wg);

}  // namespace pk
