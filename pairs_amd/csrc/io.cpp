// Particle CSV input and legacy-VTK output (include/pairs_hip_io.h).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pairs_hip_io.h"

namespace {

int columns_of(int kind) {
    switch (kind) {
        case PAIRS_HIP_PROP_INT:
        case PAIRS_HIP_PROP_REAL: return 1;
        case PAIRS_HIP_PROP_VECTOR: return 3;
        case PAIRS_HIP_PROP_QUATERNION: return 4;
        case PAIRS_HIP_PROP_MATRIX: return 9;
        default: return -1;
    }
}

// strtod / strtol over the whole field (the reference uses std::stod /
// std::stoi, which accept leading blanks and stop at trailing text)
bool parse_real(const std::string &f, double &v) {
    const char *s = f.c_str();
    char *e = nullptr;
    v = std::strtod(s, &e);
    return e != s;
}
bool parse_int(const std::string &f, int &v) {
    const char *s = f.c_str();
    char *e = nullptr;
    const long x = std::strtol(s, &e, 10);
    v = (int) x;
    return e != s;
}

constexpr int FLAG_INFINITE = 1, FLAG_FIXED = 4, FLAG_GLOBAL = 8;
constexpr double WITHIN_SMALL = 1e-5;  // Regular6DStencil SMALL

}  // namespace

extern "C" {

int pairs_hip_read_particle_csv(const char *filename, int nprops, const int *kinds, int position_prop,
                                int flags_prop, const double *domain_lo, const double *domain_hi, double *out,
                                int max_rows, int *needed) {
    if (!filename || nprops <= 0 || !kinds || max_rows < 0 || (max_rows > 0 && !out)) return PAIRS_HIP_ERR_ARG;
    std::vector<int> col0(nprops + 1, 0);
    for (int p = 0; p < nprops; p++) {
        const int c = columns_of(kinds[p]);
        if (c < 0) return PAIRS_HIP_ERR_ARG;
        col0[p + 1] = col0[p] + c;
    }
    const int ncols = col0[nprops];
    if (position_prop >= 0 && (position_prop >= nprops || kinds[position_prop] != PAIRS_HIP_PROP_VECTOR))
        return PAIRS_HIP_ERR_ARG;
    if (flags_prop >= 0 && (flags_prop >= nprops || kinds[flags_prop] != PAIRS_HIP_PROP_INT)) return PAIRS_HIP_ERR_ARG;
    if (position_prop >= 0 && (!domain_lo || !domain_hi)) return PAIRS_HIP_ERR_ARG;
    std::ifstream in(filename);
    if (!in.is_open()) return PAIRS_HIP_ERR_ARG;
    std::string line;
    std::vector<double> row(ncols);
    int kept = 0;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::stringstream ls(line);
        std::string field;
        int c = 0, fl = 0;
        while (c < ncols && std::getline(ls, field, ',')) {
            // which property does column c belong to
            int p = 0;
            while (col0[p + 1] <= c) p++;
            if (kinds[p] == PAIRS_HIP_PROP_INT) {
                int v;
                if (!parse_int(field, v)) return PAIRS_HIP_ERR_ARG;
                row[c] = v;
                if (p == flags_prop) fl = v;
            } else {
                double v;
                if (!parse_real(field, v)) return PAIRS_HIP_ERR_ARG;
                row[c] = v;
            }
            c++;
        }
        if (c < ncols) return PAIRS_HIP_ERR_ARG;
        int within = 1;
        if (position_prop >= 0) {
            const double *x = &row[col0[position_prop]];
            for (int d = 0; d < 3; d++)
                within &= (x[d] >= domain_lo[d] && x[d] < domain_hi[d] - WITHIN_SMALL) ? 1 : 0;
        }
        if (within || (fl & (FLAG_INFINITE | FLAG_FIXED | FLAG_GLOBAL))) {
            if (kept < max_rows)
                for (int k = 0; k < ncols; k++) out[(size_t) kept * ncols + k] = row[k];
            kept++;
        }
    }
    if (needed) *needed = kept;
    return kept > max_rows ? PAIRS_HIP_ERR_STATE : kept;
}

int pairs_hip_vtk_write_arrays(const char *prefix, int n, const double *position, const double *mass,
                               const int *flags, int timestep, int frequency) {
    if (!prefix || n < 0 || (n > 0 && (!position || !mass || !flags))) return PAIRS_HIP_ERR_ARG;
    if (frequency != 0 && timestep % frequency != 0) return 0;
    const std::string name = std::string(prefix) + "_" + std::to_string(timestep) + ".vtk";
    FILE *f = std::fopen(name.c_str(), "w");
    if (!f) return PAIRS_HIP_ERR_ARG;
    int m = 0;
    for (int i = 0; i < n; i++) m += (flags[i] & FLAG_INFINITE) ? 0 : 1;
    std::fprintf(f, "# vtk DataFile Version 2.0\nParticle data\nASCII\nDATASET UNSTRUCTURED_GRID\n");
    std::fprintf(f, "POINTS %d double\n", m);
    for (int i = 0; i < n; i++)
        if (!(flags[i] & FLAG_INFINITE))
            std::fprintf(f, "%.8f %.8f %.8f\n", position[3 * i], position[3 * i + 1], position[3 * i + 2]);
    std::fprintf(f, "\n\nCELLS %d %d\n", m, 2 * m);
    // the reference numbers each vertex by its particle index (vtk.hpp:63-66)
    for (int i = 0; i < n; i++)
        if (!(flags[i] & FLAG_INFINITE)) std::fprintf(f, "1 %d\n", i);
    std::fprintf(f, "\n\nCELL_TYPES %d\n", m);
    for (int i = 0; i < n; i++)
        if (!(flags[i] & FLAG_INFINITE)) std::fprintf(f, "1\n");
    std::fprintf(f, "\n\nPOINT_DATA %d\nSCALARS mass double\nLOOKUP_TABLE default\n", m);
    for (int i = 0; i < n; i++)
        if (!(flags[i] & FLAG_INFINITE)) std::fprintf(f, "%.8f\n", mass[i]);
    std::fprintf(f, "\n\n");
    std::fclose(f);
    return 1;
}

int pairs_hip_vtk_write(pairs_hip_ctx *ctx, const char *prefix, int timestep, int frequency) {
    if (!ctx || !prefix) return PAIRS_HIP_ERR_ARG;
    if (frequency != 0 && timestep % frequency != 0) return 0;
    const int n = pairs_hip_count(ctx, PAIRS_HIP_NLOCAL);
    if (n < 0) return n;
    std::vector<double> x(3 * (size_t) n), m(n);
    std::vector<int> fl(n);
    // downloads return the element count (or a negative error)
    int r = pairs_hip_download(ctx, PAIRS_HIP_POSITION, x.data(), n);
    if (r >= 0) r = pairs_hip_download(ctx, PAIRS_HIP_MASS, m.data(), n);
    if (r >= 0) r = pairs_hip_download(ctx, PAIRS_HIP_FLAGS, fl.data(), n);
    if (r < 0) return r;
    return pairs_hip_vtk_write_arrays(prefix, n, x.data(), m.data(), fl.data(), timestep, frequency);
}

int pairs_hip_dem_vtk_write(pairs_hip_dem_ctx *ctx, const char *prefix, int timestep, int frequency) {
    if (!ctx || !prefix) return PAIRS_HIP_ERR_ARG;
    if (frequency != 0 && timestep % frequency != 0) return 0;
    const int n = pairs_hip_dem_count(ctx, PAIRS_HIP_DEM_NLOCAL);
    if (n < 0) return n;
    std::vector<double> x(3 * (size_t) n), m(n);
    std::vector<int> fl(n);
    int r = pairs_hip_dem_download(ctx, PAIRS_HIP_DEM_POSITION, x.data(), n);
    if (r >= 0) r = pairs_hip_dem_download(ctx, PAIRS_HIP_DEM_MASS, m.data(), n);
    if (r >= 0) r = pairs_hip_dem_download(ctx, PAIRS_HIP_DEM_FLAGS, fl.data(), n);
    if (r < 0) return r;
    return pairs_hip_vtk_write_arrays(prefix, n, x.data(), m.data(), fl.data(), timestep, frequency);
}

}  // extern "C"
