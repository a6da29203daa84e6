// host program of "dem" for the HIP target (include/pairs_hip_dem.h)
// usage: prog [state.bin]  (state.bin: nlocal, then positions, linear
// velocities and uids of the final state's locals)
#include <cstdio>
#include <vector>
#include "pairs_hip_dem.h"

static int dump(pairs_hip_dem_ctx *d, const char *path) {
    const int n = pairs_hip_dem_count(d, PAIRS_HIP_DEM_NLOCAL);
    if (n < 0) return 1;
    std::vector<double> x(3 * (size_t) n), v(3 * (size_t) n);
    std::vector<int> uid(n);
    if (pairs_hip_dem_download(d, PAIRS_HIP_DEM_POSITION, x.data(), n) != n) return 1;
    if (pairs_hip_dem_download(d, PAIRS_HIP_DEM_LINEAR_VELOCITY, v.data(), n) != n) return 1;
    if (pairs_hip_dem_download(d, PAIRS_HIP_DEM_UID, uid.data(), n) != n) return 1;
    FILE *f = std::fopen(path, "wb");
    if (!f) return 1;
    std::fwrite(&n, sizeof(int), 1, f);
    std::fwrite(x.data(), sizeof(double), x.size(), f);
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fwrite(uid.data(), sizeof(int), uid.size(), f);
    return std::fclose(f) != 0;
}

int main(int argc, char **argv) {
    pairs_hip_dem_config c = {};
    c.particle_capacity = 1000000;
    c.contact_capacity = 20;
    c.send_capacity = 200000;
    c.ncells_capacity = 1 << 22;
    c.box_hi[0] = 0.04; c.box_hi[1] = 0.015; c.box_hi[2] = 0.04;
    c.periodic[0] = 1; c.periodic[1] = 1; c.periodic[2] = 0;
    c.cell_width = 0.0032219;
    c.dt = 5e-05;
    c.density_particle = 2550;
    c.density_fluid = 1000;
    c.gravity = 9.81;
    c.kappa = 0.8764044943820225;
    c.ln_restitution = -2.3025850929940455;
    c.collision_time = 0.0005;
    c.ntypes = 1;
    c.friction_static[0] = 0.0; c.friction_dynamic[0] = 0.5;
    pairs_hip_dem_ctx *d;
    if (pairs_hip_dem_create(&d, &c)) return 1;
    pairs_hip_dem_sc_grid(d, 0.04, 0.015, 0.04, 0.005, 0.0029, 0.00261, 0.00319, 1, 2550, 1);
    { const double p[3] = {0.0, 0.0, 0.0}, n[3] = {0.0, 0.0, 1.0}; pairs_hip_dem_add_halfspace(d, 100000, 0, 1.0, p, n, 13); }
    { const double p[3] = {0.8, 0.015, 0.2}, n[3] = {0.0, 0.0, -1.0}; pairs_hip_dem_add_halfspace(d, 100001, 0, 1.0, p, n, 13); }
    pairs_hip_dem_setup(d);
    int bad = pairs_hip_dem_run(d, 0, 501) != 0;
    if (!bad && argc > 1) bad = dump(d, argv[1]);
    pairs_hip_dem_destroy(d);
    return bad;
}
