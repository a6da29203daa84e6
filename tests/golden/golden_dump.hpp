// State dump hook spliced into the reference's generated CPU program by
// make_golden.py.  It writes one raw binary record per timestep so that the
// golden fixtures hold the exact FP64 state (the program's own stdout prints
// only 6 significant digits).  Written for this repository; it only reads the
// arrays that the generated main() already owns.
#pragma once
#include <cstdio>
#include <cstdint>

namespace golden {

static inline void put_i(FILE *f, const int *p, long n) { if (n > 0) fwrite(p, sizeof(int), (size_t) n, f); }
static inline void put_d(FILE *f, const double *p, long n) { if (n > 0) fwrite(p, sizeof(double), (size_t) n, f); }

// Record layout (little endian):
//   int32[12] header: magic, ts, nlocal, nghost, nsend_all, ncells,
//                     neighbor_capacity, cell_capacity, has_lists, has_contacts,
//                     contact_capacity, 0
//   f64 position[(nlocal+nghost)*3]
//   f64 linear_velocity[(nlocal+nghost)*3]
//   f64 force[nlocal*3]
//   i32 type[nlocal+nghost]
//   if has_lists: i32 numneighs[nlocal]; i32 neighborlists[nlocal*neighbor_capacity]
//   i32 particle_cell[nlocal+nghost]
//   i32 send_map[nsend_all]; i32 send_mult[nsend_all*3]; i32 nsend[6]
//   i32 cell_sizes[ncells]; i32 cell_particles[ncells*cell_capacity]
inline void dump_md(const char *dir, int ts, int nlocal, int nghost, int nsend_all,
                    int ncells, int neighbor_capacity, int cell_capacity,
                    const double *position, const double *linear_velocity, const double *force,
                    const int *type, const int *numneighs, const int *neighborlists,
                    const int *particle_cell, const int *send_map, const int *send_mult,
                    const int *nsend, const int *cell_sizes, const int *cell_particles) {
    char path[4096];
    snprintf(path, sizeof(path), "%s/dump_%05d.bin", dir, ts + 1);
    FILE *f = fopen(path, "wb");
    if (!f) return;
    const int has_lists = numneighs != nullptr ? 1 : 0;
    int hdr[12] = {0x50414952, ts, nlocal, nghost, nsend_all, ncells, neighbor_capacity,
                   cell_capacity, has_lists, 0, 0, 0};
    fwrite(hdr, sizeof(int), 12, f);
    const long nall = (long) nlocal + nghost;
    put_d(f, position, nall * 3);
    put_d(f, linear_velocity, nall * 3);
    put_d(f, force, (long) nlocal * 3);
    put_i(f, type, nall);
    if (has_lists) {
        put_i(f, numneighs, nlocal);
        put_i(f, neighborlists, (long) nlocal * neighbor_capacity);
    }
    put_i(f, particle_cell, nall);
    put_i(f, send_map, nsend_all);
    put_i(f, send_mult, (long) nsend_all * 3);
    put_i(f, nsend, 6);
    put_i(f, cell_sizes, ncells);
    put_i(f, cell_particles, (long) ncells * cell_capacity);
    fclose(f);
}


// DEM record (magic 0x4D454400 'DEM'):
//   int32[8] header: magic, ts, nlocal, nghost, contact_capacity, 0, 0, 0
//   f64 position[nall*3], linear_velocity[nall*3], angular_velocity[nall*3]
//   f64 force[nlocal*3], torque[nlocal*3], radius[nall], mass[nall]
//   f64 rotation_quat[nlocal*4], rotation_matrix[nlocal*9]
//   i32 uid[nall], flags[nall], type[nall]
//   i32 num_contacts[nlocal], contact_lists[nlocal*cap], is_sticking[nlocal*cap]
//   f64 tangential_spring_displacement[nlocal*cap*3], impact_velocity_magnitude[nlocal*cap]
inline void dump_dem(const char *dir, int ts, int nlocal, int nghost, int cap, const double *position,
                     const double *linear_velocity, const double *angular_velocity, const double *force,
                     const double *torque, const double *radius, const double *mass, const double *rotation_quat,
                     const double *rotation_matrix, const int *uid, const int *flags, const int *type,
                     const int *num_contacts, const int *contact_lists, const int *is_sticking,
                     const double *tangential_spring_displacement, const double *impact_velocity_magnitude) {
    char path[4096];
    snprintf(path, sizeof(path), "%s/dem_%05d.bin", dir, ts + 1);
    FILE *f = fopen(path, "wb");
    if (!f) return;
    int hdr[8] = {0x4D454400, ts, nlocal, nghost, cap, 0, 0, 0};
    fwrite(hdr, sizeof(int), 8, f);
    const long nall = (long) nlocal + nghost, nl = nlocal;
    put_d(f, position, nall * 3);
    put_d(f, linear_velocity, nall * 3);
    put_d(f, angular_velocity, nall * 3);
    put_d(f, force, nl * 3);
    put_d(f, torque, nl * 3);
    put_d(f, radius, nall);
    put_d(f, mass, nall);
    put_d(f, rotation_quat, nl * 4);
    put_d(f, rotation_matrix, nl * 9);
    put_i(f, uid, nall);
    put_i(f, flags, nall);
    put_i(f, type, nall);
    put_i(f, num_contacts, nl);
    put_i(f, contact_lists, nl * cap);
    put_i(f, is_sticking, nl * cap);
    put_d(f, tangential_spring_displacement, nl * cap * 3);
    put_d(f, impact_velocity_magnitude, nl * cap);
    fclose(f);
}

}  // namespace golden
