// Radial-descriptor MLP potential: device layout and launchers.
//
// Descriptor of particle i (16 radial Gaussians under a cosine cutoff):
//   G_k(i) = sum_{j in list(i), r_ij < 2.5} exp(-(r_ij - mu_k)^2 / (2 w^2)) fc(r_ij)
//   mu_k = 0.8 + k * 1.7 / 15 (k = 0..15), w = 0.15, fc(r) = (cos(pi r / 2.5) + 1) / 2
// Energy of i: E_i = e_mean + e_std * y(i), y = 16-64-64-1 MLP with SiLU on the
// standardised descriptor (G - g_mean) * g_istd.  Forces are -dE/dx with
// E = sum_i E_i over the locals:
//   F_i = -sum_{j in list(i), r < 2.5} (c_i(r_ij) + c_j(r_ij)) (x_i - x_j) / r_ij,
//   c_a(r) = sum_k D_ak dphi_k/dr,  D_ak = dE_a / dG_ak,  phi_k = gaussian_k * fc,
// where a ghost's D is its root local's.  The layers run on MFMA
// (v_mfma_f32_16x16x16_bf16: BF16 weights and activations, FP32 accumulation);
// descriptors, gradients and forces are FP32 vector math, positions and the
// integration FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "md_kernels.hpp"

namespace pk {

constexpr int MLP_ND = 16;  // descriptors
constexpr int MLP_H = 64;   // hidden width of both layers
// flat FP32 parameter block
constexpr int MLP_W1 = 0;                         // [64][16]  W1[h][k]
constexpr int MLP_B1 = MLP_W1 + MLP_H * MLP_ND;   // [64]
constexpr int MLP_W2 = MLP_B1 + MLP_H;            // [64][64]  W2[out][in]
constexpr int MLP_B2 = MLP_W2 + MLP_H * MLP_H;    // [64]
constexpr int MLP_W3 = MLP_B2 + MLP_H;            // [64]
constexpr int MLP_B3 = MLP_W3 + MLP_H;            // [1]
constexpr int MLP_NP = MLP_B3 + 1;                // 5313 parameters
// standardisation block
constexpr int MLP_GMEAN = 0, MLP_GISTD = 16, MLP_EMEAN = 32, MLP_ESTD = 33, MLP_NS = 34;

// waves of the list passes over the locals (2 or 4 lanes per local; their
// energy / virial partials are one per wave)
int mlp_waves(const DevState &d);

constexpr double MLP_RC = 2.5;
constexpr float MLP_MU0 = 0.8f, MLP_DMU = 1.7f / 15.0f, MLP_W = 0.15f;

// a model bound to one LJ context: parameters and per-slot work arrays
struct MLPBind {
    const float *p;   // MLP_NP parameters (device)
    const float *sc;  // MLP_NS standardisation constants (device)
    float *G;         // [cap][16] descriptors of the locals
    float *D;         // [cap][16] dE_i/dG_i of locals and (copied) ghosts
    float *E;         // [cap] per-particle energy
    int cap;
};

// descriptors of every local of buffer `cur` into G (physical slots)
void launch_mlp_descriptors(hipStream_t s, const DevState &d, const Geom &g, int cur, float *G, bool pack = true);
// E and dE/dG of n descriptor rows (either output may be null)
void launch_mlp_eval(hipStream_t s, const float *p, const float *sc, const float *G, int n, float *E, float *D);
// full MLP force evaluation on buffer `cur`: descriptors, network, ghost
// gradients, forces (reset + written), and with `reduce` the energy and
// virial partials of the thermo reduction.  With `vv` (and no `reduce`) the
// force pass also integrates, as launch_vv_post(cur, next, vv2, vv1) would:
// final_integrate(ts), initial_integrate(ts+1), next positions and images.
// Pn, when set, receives the packed copy of the next positions.  `pack`:
// the descriptor pass first packs buffer cur (not needed when the previous
// step left its packed copy current).
struct MLPVV {
    const LJParams *prm;
    int next, vv2, vv1;
    double4 *Pn;
};
void launch_mlp_forces(hipStream_t s, const DevState &d, const Geom &g, const MLPBind &b, int cur, int reduce,
                       int ghost_grid, const MLPVV *vv = nullptr, bool pack = true);

// training
struct MLPTrain {
    const float *G;      // [rows][16] descriptors
    const float *label;  // [rows] per-atom energies
    int rows;            // sample from [0, rows)
    int batch;           // rows per step (a multiple of 16)
    int nwaves;          // waves of the gradient launch
    float *partial;      // [nwaves][MLP_NP]
    float *loss;         // [nwaves] partial squared errors
    float gscale = 1.0f; // Adam takes gscale * (sum of the partials)
};

// one configuration's loss gradient (k_mlp_cgrad): rows [0, n) of G, their
// per-atom energy labels and back-projected force directions U
struct MLPCfgGrad {
    const float *G;
    const double *label;
    const float *U;
    int n;
    float eps_scale;     // 2 w_energy / n
    int nwaves;          // waves of the launch (a multiple of 4)
    float *partial;      // [nwaves][MLP_NP]
    float *loss;         // [nwaves] partial sums of the squared standardised energy error
};
// gradient of the mean squared error of the standardised energy over one
// minibatch (rows chosen by a counter hash of seed and step)
void launch_mlp_grad(hipStream_t s, const float *p, const float *sc, const MLPTrain &t, uint64_t seed, int step);
// sum the wave partials and take one Adam step (t = 1, 2, ...)
void launch_mlp_adam(hipStream_t s, float *p, float *m, float *v, const MLPTrain &t, float lr, int step,
                     float *loss_out);
// Gradient of one configuration's loss on buffer `cur` (lists current):
//   w_energy mean_a (y_a - (E*_a - e_mean) / e_std)^2 + w_force sum |F - F*|^2 / sum |F*|^2
// with labels F* = (lx, ly, lz) and E* = eatom (physical slots).  Leaves the
// network's forces in d.f*, the residual/force-norm partials in red ([nb][2])
// and the per-wave gradient and loss partials in t.partial / t.loss.
void launch_mlp_config_grad(hipStream_t s, const DevState &d, const Geom &g, const MLPBind &b, int cur,
                            const double *lx, const double *ly, const double *lz, const double *eatom,
                            double w_energy, double w_force, float4 *rho, float *U, double *red, int nb,
                            int ghost_grid, const MLPCfgGrad &t);
// acc += sum of nwaves partials (acc may be null); out (may be null) = that sum
void launch_mlp_accum(hipStream_t s, const float *partial, int nwaves, float *acc, float *out);
// column sums and sums of squares of rows [0, n): G (16 columns) and labels
void launch_mlp_moments(hipStream_t s, const float *G, const float *label, int n, double *partials, int nblocks);
// sum over rows [0, n) of (E - label)^2 into per-block partials
void launch_mlp_sqerr(hipStream_t s, const float *E, const float *label, int n, double *partials, int nblocks);
// append the locals' descriptors and per-atom energies (FP64 -> FP32)
void launch_mlp_snapshot(hipStream_t s, const float *G, const double *eatom, int nlocal, float *dst_G,
                         float *dst_label);
// per-block partial sums of |fa - fb|^2 and |fb|^2 over n particles
void launch_force_diff(hipStream_t s, const double *ax, const double *ay, const double *az, const double *bx,
                       const double *by, const double *bz, int n, double *partials, int nblocks);

}  // namespace pk
