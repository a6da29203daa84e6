// host program of "md" for the HIP target (include/pairs_hip.h)
// usage: prog [state.bin]  (state.bin: nlocal, then positions and velocities
// of the final state in the reference's particle order, as doubles)
#include <cstdio>
#include <vector>
#include "pairs_hip.h"

static int dump(pairs_hip_ctx *ctx, const char *path) {
    const int n = pairs_hip_count(ctx, PAIRS_HIP_NLOCAL);
    if (n < 0) return 1;
    std::vector<double> x(3 * (size_t) n), v(3 * (size_t) n);
    if (pairs_hip_download(ctx, PAIRS_HIP_POSITION, x.data(), n) != n) return 1;
    if (pairs_hip_download(ctx, PAIRS_HIP_VELOCITY, v.data(), n) != n) return 1;
    FILE *f = std::fopen(path, "wb");
    if (!f) return 1;
    std::fwrite(&n, sizeof(int), 1, f);
    std::fwrite(x.data(), sizeof(double), x.size(), f);
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    return std::fclose(f) != 0;
}

int main(int argc, char **argv) {
    static double eps[1] = {1.0};
    static double sig6[1] = {1.0};
    pairs_hip_md_config c = {};
    c.particle_capacity = 800000;
    c.neighbor_capacity = 100;
    c.send_capacity = 200000;
    c.ncells_capacity = 100000;
    const double len[3] = {6.718384765530029, 6.718384765530029, 6.718384765530029};
    for (int d = 0; d < 3; d++) { c.box_lo[d] = 0.0; c.box_hi[d] = len[d]; c.periodic[d] = 1; }
    c.cell_spacing = 2.8;
    c.cutoff = 2.5;
    c.dt = 0.005;
    c.half = 0;
    c.ntypes = 1;
    c.epsilon = eps;
    c.sigma6 = sig6;
    c.reneighbor_every = 20;
    c.thermo_every = 100;
    pairs_hip_ctx *ctx;
    if (pairs_hip_create(&ctx, &c)) return 1;
    pairs_hip_copper_fcc_lattice(ctx, 4, 4, 4, 0.8442, 1);
    pairs_hip_adjust_thermo(ctx, 1.44);
    pairs_hip_setup_cells(ctx);
    double th[5 * 4096];
    const int rows = pairs_hip_run(ctx, 0, 101, 100, th, 4096);
    for (int r = 0; r < rows; r++) std::printf("%g\t%g\n", th[5 * r], th[5 * r + 1]);
    int bad = rows < 0;
    if (!bad && argc > 1) bad = dump(ctx, argv[1]);
    pairs_hip_destroy(ctx);
    return bad;
}
