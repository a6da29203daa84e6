/*
 * pairs_hip_io.h -- particle input (CSV) and output (legacy ASCII VTK).
 *
 * Host-side file formats of the reference's runtime, for the set-up and the
 * visual checks around the hot path:
 *   - CSV input as runtime/read_from_file.hpp:33-115 reads it: one particle
 *     per line, comma-separated, columns in the order of the caller's property
 *     list, a vector property as 3 consecutive columns, a matrix as 9 and a
 *     quaternion as 4; a row is kept if its position lies in the domain
 *     (x >= lo && x < hi - 1e-5 per dimension, Regular6DStencil::
 *     isWithinSubdomain, runtime/domain/regular_6d_stencil.cpp:96-100) or if
 *     its flags have Infinite, Fixed or Global set (read_from_file.hpp:103-105);
 *   - VTK output as runtime/vtk.hpp:11-86 writes it: "<prefix>_<timestep>.vtk",
 *     POINTS, one vertex cell per point, mass as point data, particles flagged
 *     Infinite skipped, 8 decimals, written only when timestep % frequency == 0
 *     (frequency 0: always).
 */
#ifndef PAIRS_HIP_IO_H
#define PAIRS_HIP_IO_H

#include "pairs_hip.h"
#include "pairs_hip_dem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Property kinds of a CSV property list (the value is not the column count
   for INT and REAL, which take one column each). */
enum { PAIRS_HIP_PROP_INT = 0, PAIRS_HIP_PROP_REAL = 1, PAIRS_HIP_PROP_VECTOR = 3, PAIRS_HIP_PROP_QUATERNION = 4,
       PAIRS_HIP_PROP_MATRIX = 9 };

/* Read a particle CSV file with nprops properties of the given kinds.
   position_prop and flags_prop are the indices of the position (a VECTOR)
   and flags (an INT) properties in the list, -1 if absent.  Kept rows are
   written as doubles, the columns of all properties side by side, into
   out[max_rows][ncols] (ncols = total column count).  Returns the number of
   rows kept; PAIRS_HIP_ERR_ARG if the file cannot be opened or a row does not
   parse; PAIRS_HIP_ERR_STATE if more than max_rows rows are kept (*needed,
   if not NULL, receives the count either way). */
int pairs_hip_read_particle_csv(const char *filename, int nprops, const int *kinds, int position_prop,
                                int flags_prop, const double *domain_lo, const double *domain_hi, double *out,
                                int max_rows, int *needed);

/* Write n particles (positions [n][3], masses [n], flags [n]) as the
   reference's VTK file for timestep; returns 1 if written, 0 if skipped by
   the frequency, PAIRS_HIP_ERR_ARG if the file cannot be created. */
int pairs_hip_vtk_write_arrays(const char *prefix, int n, const double *position, const double *mass,
                               const int *flags, int timestep, int frequency);
/* The same for the locals of an LJ context (reference particle order). */
int pairs_hip_vtk_write(pairs_hip_ctx *ctx, const char *prefix, int timestep, int frequency);
/* The same for the locals of a DEM context (spheres and half-spaces). */
int pairs_hip_dem_vtk_write(pairs_hip_dem_ctx *ctx, const char *prefix, int timestep, int frequency);

#ifdef __cplusplus
}
#endif
#endif
