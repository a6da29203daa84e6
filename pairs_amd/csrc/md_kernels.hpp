// Device state and launchers of the LJ molecular-dynamics path.
#pragma once
#include <hip/hip_runtime.h>
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
    C_COUNT = 16
};
enum { OVF_NONE = 0, OVF_PART = 1, OVF_NEIGH = 2, OVF_SEND = 3 };

// cell grid and domain, passed by value
struct Geom {
    double lo[3], hi[3], len[3];
    double spacing;
    int dim_cells[3];
    int ncells;
    int nstencil;
    int stencil[27];
};

struct LJParams {
    double rc2, dt, hdt;
    int ntypes;
    double eps[16], sig6[16];
};

struct DevState {
    int cap, ncap, send_cap, ncells_cap, nl_stride;
    int nlocal;  // host mirror (constant between uploads on one domain)
    int cur;     // which position buffer holds the current state
    double4 *P[2];
    double *vx, *vy, *vz, *fx, *fy, *fz, *mass;
    int *flags, *uid, *shape;
    int *numneighs, *nl;
    int *particle_cell, *slot, *sorted_idx;
    double4 *sorted_pos;
    int *cell_count, *cell_start, *cell_nlocal;
    int *send_map, *send_mult, *exchg_flag;
    int *tile_cnt;
    int *counters;
    int *scratch_i;     // 2 * send_cap
    double *stage;      // max(send_cap * 10, ...) doubles
    double *partials;   // 3 per force block
    double *thermo;     // 5 per thermo record
};

void launch_fill_i32(hipStream_t s, int *p, int v, int n);
void launch_fill_f64(hipStream_t s, double *p, double v, int n);
void launch_set_counters(hipStream_t s, int *c, int first, int count, int v);
void launch_bin(hipStream_t s, const DevState &d, const Geom &g);
void launch_build_neighbors(hipStream_t s, const DevState &d, const Geom &g, int half, double rc2);
void launch_select_dim(hipStream_t s, const DevState &d, const Geom &g, int dim, int exch);
void launch_exchange(hipStream_t s, const DevState &d, const Geom &g, const int *periodic);
void launch_borders(hipStream_t s, const DevState &d, const Geom &g, const int *periodic);
void launch_sync_pp(hipStream_t s, const DevState &d, const Geom &g, int cur, int prev, int grid);
void launch_sync_inplace(hipStream_t s, const DevState &d, const Geom &g, int grid, int with_vel);
void launch_lj_full(hipStream_t s, const DevState &d, const LJParams &prm, int cur, int next, int fuse, int do_vv2,
                    int do_vv1, int reduce);
void launch_lj_half(hipStream_t s, const DevState &d, const LJParams &prm, int cur, int reduce);
void launch_vv_initial(hipStream_t s, const DevState &d, const LJParams &prm);
void launch_vv_post(hipStream_t s, const DevState &d, const LJParams &prm, int cur, int next, int do_vv2, int do_vv1,
                    int reduce);
void launch_ke_partial(hipStream_t s, const DevState &d);
void launch_energy(hipStream_t s, const DevState &d, const LJParams &prm, int half, double *eatom);
void launch_thermo_reduce(hipStream_t s, const DevState &d, int which_mask, const Geom &g, double *out);

}  // namespace pk
