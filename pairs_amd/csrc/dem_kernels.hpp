// Device state and launchers of the DEM path (dem_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace dk {

enum { DFLAG_INFINITE = 1, DFLAG_GHOST = 2, DFLAG_FIXED = 4, DFLAG_GLOBAL = 8 };
enum { DSHAPE_SPHERE = 0, DSHAPE_HALFSPACE = 1 };

// device-resident counters
enum {
    D_NLOCAL = 0,
    D_NGHOST = 1,
    D_NSEND_ALL = 2,
    D_NSEND = 3,   // 6 entries
    D_RANGE = 9,   // nlocal + nghost latched by the selection of a dimension
    D_BASE = 10,   // send entries before this dimension
    D_UNUSED = 11,
    D_OVF = 12,
    D_REQ_SEND = 13,
    D_REQ_CONTACT = 14,
    D_REQ_PART = 15,
    D_COUNT = 16
};
// touching partners recorded per particle by the contact search; a particle
// with more walks its cells again in the force pass
constexpr int DCAND = 16;
// lanes per sphere of the contact search (each a run of the 28 cells)
#ifndef DEM_LPS
#define DEM_LPS 16
#endif

enum { DOVF_NONE = 0, DOVF_PART = 1, DOVF_CONTACT = 2, DOVF_SEND = 3, DOVF_CELLS = 4 };

struct DemGeom {
    double lo[3], hi[3], len[3];
    double width;  // cell width == ghost layer
    int dim_cells[3];
    int ncells;
    int stencil[27];
    int periodic[3];
};

struct DemParams {
    double dt, pi, kappa, sqrt_kappa, ln_e, t_c, drho, grav;
    double kn_fixed;  // > 0: fixed normal stiffness (else from the collision time)
    double sqrt_pl;   // sqrt(pi^2 + ln(e)^2) for the fixed-stiffness damping
    int ntypes;
    double mu_s[16], mu_d[16];
};

// every device array, passed by value to the kernels
struct DemPtrs {
    int cap, ccap, send_cap;
    double *x, *y, *z, *vx, *vy, *vz, *wx, *wy, *wz, *fx, *fy, *fz, *tx, *ty, *tz, *r, *m, *nx, *ny, *nz;
    double *inv_i, *rot_m, *rot_q;  // [k][cap]: 9, 9 and 4 components
    int *uid, *flags, *type, *shape;
    int *num_contacts, *c_uid, *c_used, *c_stick;  // [k][cap]
    double *c_tsd;                                 // [3k + d][cap]
    double *c_ivm;                                 // [k][cap]
    int *cell_count, *cell_start, *cell_particles, *particle_cell, *nshapes;
    int *cell_slot;      // each particle's place in its cell, from the counting pass
    int *send_map, *send_mult, *exchg_flag, *copy_to, *blk, *scan_tmp;
    double *buf, *hbuf;
    int *cand_n, *cand;  // touching partners found by the contact search: [t][cap], t < DCAND
    int *cslot;          // their history slots [t][cap]
    double *cf;          // their force and torque terms [6 t + d][cap]
};

// one strided copy of 32-bit words: rows x words, row r from src + r * src_pitch
// to dst + r * dst_pitch (pitches in words)
struct CopyDesc {
    const void *src;
    void *dst;
    size_t src_pitch, dst_pitch;
    int words, rows;
};
// every copy of a table in one launch (n: the largest row width in particles)
void launch_copy_rows(hipStream_t s, const CopyDesc *table, int ncopies, int n);

void launch_fill(hipStream_t s, int *p, int v, int n);
void launch_mass_inertia(hipStream_t s, const DemPtrs &p, int nl);
// one step = prepare (exchange, borders, cells) + contacts (+ gravity and
// history bookkeeping) + euler
void launch_prepare(hipStream_t s, const DemPtrs &p, int *c, const DemGeom &g);
// between: three events recorded after the search, the slot and the pair
// kernels (timed steps only)
void launch_contacts(hipStream_t s, const DemPtrs &p, int *c, const DemGeom &g, const DemParams &prm,
                     const hipEvent_t *between = nullptr);
void launch_euler(hipStream_t s, const DemPtrs &p, int *c, const DemParams &prm);

}  // namespace dk
