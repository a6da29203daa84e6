// The LJ context behind include/pairs_hip.h (host side), shared by the C API
// translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/pairs_hip.h"
#include "md_kernels.hpp"
#include "mlp_kernels.hpp"

using namespace pk;

struct pairs_hip_ctx {
    pairs_hip_md_config cfg;
    LJParams prm;
    Geom geom;
    DevState d;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = -1;           // the device current at creation
    std::string err;
    int req[5] = {0, 0, 0, 0, 0};
    int h_counters[C_COUNT];
    int nghost = 0, nsend_all = 0;
    int ghost_grid = 1;
    int max_lds = 65536;       // shared memory a workgroup may allocate
    int zc_override = 0;
    int home_target = 40;      // home particles a segment aims at (sets zc; about 3 cells at rho 0.84)
    // graphs of n consecutive non-rebuild steps (slot n - 1, n <=
    // CHAIN_MAX), by parity of the current buffer at their first step: a run
    // of steady steps between two rebuilds is one launch
    static constexpr int CHAIN_MAX = 24;
    bool use_graphs = true;
    hipGraphExec_t step_graph[2][CHAIN_MAX] = {};
    int graph_ghost_grid[2][CHAIN_MAX] = {};
    int graph_img[2][CHAIN_MAX] = {};
    // the first step of a run (initial integrate, in-place ghost refresh,
    // packed copy, fused force step), by parity
    hipGraphExec_t start_graph[2] = {nullptr, nullptr};
    int start_ghost_grid[2] = {0, 0}, start_img[2] = {0, 0};
    // a rebuild (exchange, borders, cell binning, neighbour build), by
    // position buffer
    hipGraphExec_t reb_graph[2] = {nullptr, nullptr};
    // Rebuilds of pairs_hip_run whose counters are checked later: the
    // locals' state is checkpointed, the rebuild and the steps after it are
    // enqueued without waiting for its counters, which are copied to pinned
    // memory behind it and checked at the next rebuild or at the end of the
    // run; an overflow restores the checkpoint and repeats the rebuild with
    // the host loop (and capacity growth).
    bool spec_enabled = true;
    bool spec_pending = false;
    int spec_ts = 0, spec_cur = 0, spec_nthermo = 0, spec_ev_used = 0;
    int *h_spec = nullptr;          // pinned, C_COUNT
    hipEvent_t spec_ev = nullptr;   // behind the counter copy
    // the counter copy runs on a second stream behind the rebuild, so the
    // steps after the rebuild do not queue behind it
    hipStream_t side = nullptr;
    hipEvent_t reb_ev = nullptr;    // behind the rebuild
    double *ckpt = nullptr;         // checkpoint arrays, ckpt_n locals
    int ckpt_n = 0;
    long long spec_count = 0, spec_rollbacks = 0;
    // force-kernel timing
    bool time_force = false;
    int time_every = 10;
    std::vector<hipEvent_t> ev_begin, ev_end;
    int ev_used = 0;
    double force_ms_avg = 0.0;
    int force_launches = 0;
    double *d_eatom = nullptr;
    int eatom_cap = 0;
    int *d_expand = nullptr;
    size_t expand_cap = 0;
    // MLP force field (include/pairs_hip_mlp.h): when on, every force
    // evaluation of the run uses the attached model instead of LJ
    // P[b] (packed x, y, z) matches X[b], Y[b], Z[b] for every particle
    bool xy_valid[2] = {false, false};
    bool mlp_on = false;
    MLPBind mlp{};
    float *mlp_G = nullptr, *mlp_D = nullptr, *mlp_E = nullptr;  // per-slot work arrays
    double *mlp_f = nullptr;                                      // [3][cap] saved forces
    double *mlp_red = nullptr;                                    // reduction partials
    int mlp_cap = 0;
    // training on this context's configurations (pairs_hip_mlp_config_grad)
    float4 *mlp_rho = nullptr;                                    // [cap] dL/dF of every slot
    float *mlp_U = nullptr;                                       // [cap][16] back-projected directions
    int mlp_train_cap = 0;
};

#define HIP_OK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                              \
            return PAIRS_HIP_ERR_HIP;                                                                  \
        }                                                                                              \
    } while (0)

#define CHECK_LAUNCH()                                                                                 \
    do {                                                                                               \
        hipError_t e_ = hipGetLastError();                                                             \
        if (e_ != hipSuccess) {                                                                        \
            ctx->err = std::string("kernel launch: ") + hipGetErrorString(e_);                         \
            return PAIRS_HIP_ERR_HIP;                                                                  \
        }                                                                                              \
    } while (0)

