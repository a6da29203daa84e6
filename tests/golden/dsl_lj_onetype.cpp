// host program of "lj" (the older API) for the HIP target (include/pairs_hip.h)
// usage: prog [state.bin [particles.csv]]  (state.bin: nlocal, then positions,
// velocities and forces of the final state in the reference's particle order)
#include <cstdio>
#include <vector>
#include "pairs_hip.h"
#include "pairs_hip_io.h"

// a module call with the resize protocol: grow the capacity it outgrew (x2 the need) and repeat
static int module(pairs_hip_ctx *ctx, int (*fn)(pairs_hip_ctx *)) {
    for (int attempt = 0; attempt < 8; attempt++) {
        const int r = fn(ctx);
        if (r <= 0) return r;
        const int need = pairs_hip_last_required(ctx, r);
        int grow[3] = {0, 0, 0};
        grow[r == PAIRS_HIP_OVERFLOW_PARTICLES ? 0 : r == PAIRS_HIP_OVERFLOW_NEIGHBORS ? 1 : 2] = 2 * need;
        if (pairs_hip_reserve(ctx, grow[0], grow[1], grow[2])) return -1;
    }
    return -1;
}

static int dump(pairs_hip_ctx *ctx, const char *path) {
    const int n = pairs_hip_count(ctx, PAIRS_HIP_NLOCAL);
    if (n < 0) return 1;
    std::vector<double> x(3 * (size_t) n), v(3 * (size_t) n), f(3 * (size_t) n);
    if (pairs_hip_download(ctx, PAIRS_HIP_POSITION, x.data(), n) != n) return 1;
    if (pairs_hip_download(ctx, PAIRS_HIP_VELOCITY, v.data(), n) != n) return 1;
    if (pairs_hip_download(ctx, PAIRS_HIP_FORCE, f.data(), n) != n) return 1;
    FILE *o = std::fopen(path, "wb");
    if (!o) return 1;
    std::fwrite(&n, sizeof(int), 1, o);
    std::fwrite(x.data(), sizeof(double), x.size(), o);
    std::fwrite(v.data(), sizeof(double), v.size(), o);
    std::fwrite(f.data(), sizeof(double), f.size(), o);
    return std::fclose(o) != 0;
}

int main(int argc, char **argv) {
    const char *input = argc > 2 ? argv[2] : "data/minimd_setup_4x4x4_onetype.input";
    const int kinds[3] = {PAIRS_HIP_PROP_REAL, PAIRS_HIP_PROP_VECTOR, PAIRS_HIP_PROP_VECTOR};
    const double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {6.7184, 6.7184, 6.7184};
    int need = 0;
    pairs_hip_read_particle_csv(input, 3, kinds, 1, -1, lo, hi, nullptr, 0, &need);
    std::vector<double> rows((size_t) 7 * (need > 0 ? need : 1));
    const int np = pairs_hip_read_particle_csv(input, 3, kinds, 1, -1, lo, hi,
                                               rows.data(), need, &need);
    if (np < 0) { std::fprintf(stderr, "cannot read %s\n", input); return 1; }
    std::vector<double> x(3 * (size_t) np), v(3 * (size_t) np), m(np);
    for (int i = 0; i < np; i++) {
        for (int d = 0; d < 3; d++) {
            x[3 * i + d] = rows[(size_t) i * 7 + 1 + d];
            v[3 * i + d] = rows[(size_t) i * 7 + 4 + d];
        }
        m[i] = rows[(size_t) i * 7 + 0];
    }
    static double eps[1] = {1.0};
    static double sig6[1] = {1.0};
    pairs_hip_md_config c = {};
    c.particle_capacity = 800000;
    c.neighbor_capacity = 100;
    c.send_capacity = 200000;
    c.ncells_capacity = 100000;
    for (int d = 0; d < 3; d++) { c.box_lo[d] = lo[d]; c.box_hi[d] = hi[d]; }
    c.periodic[0] = 1; c.periodic[1] = 1; c.periodic[2] = 1;
    c.cell_spacing = 2.8;
    c.cutoff = 2.5;
    c.dt = 0.005;
    c.half = 0;
    c.ntypes = 1;
    c.epsilon = eps;
    c.sigma6 = sig6;
    c.reneighbor_every = 1;
    c.thermo_every = 0;
    pairs_hip_ctx *ctx;
    if (pairs_hip_create(&ctx, &c)) return 1;
    int bad = pairs_hip_upload(ctx, np, x.data(), v.data(), m.data(), nullptr, nullptr, nullptr) != 0;
    if (!bad) bad = pairs_hip_setup_cells(ctx) != 0;
    for (int ts = 0; !bad && ts <= 100; ts++) {
        if (1 > 0 && ((ts + 1) % 1 == 0 || ts == 0)) {
            bad = module(ctx, pairs_hip_exchange) || module(ctx, pairs_hip_borders) ||
                  module(ctx, pairs_hip_bin_cells) || module(ctx, pairs_hip_build_neighbors);
        } else {
            bad = pairs_hip_sync_ghosts(ctx) != 0;
        }
        if (!bad) bad = pairs_hip_lj_older(ctx) != 0 || pairs_hip_euler(ctx) != 0;
    }
    if (!bad && argc > 1) bad = dump(ctx, argv[1]);
    pairs_hip_destroy(ctx);
    return bad;
}
