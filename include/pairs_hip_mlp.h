/*
 * pairs_hip_mlp.h -- radial-descriptor MLP potential for the LJ context.
 *
 * The reference has no machine-learned potential (its force fields are the
 * DSL kernels of examples/md.py and examples/dem.py); this is the build's own
 * force field, evaluated on the same neighbour lists, ghosts and integrator as
 * pairs_hip_lj (include/pairs_hip.h), so it drops into the same timestep loop.
 *
 * Descriptor of particle i, 16 radial Gaussians under a cosine cutoff (r in
 * LJ units, over the particle's full neighbour list, pairs with r < 2.5):
 *   G_k(i) = sum_j exp(-(r_ij - mu_k)^2 / (2 w^2)) * (cos(pi r_ij / 2.5) + 1) / 2
 *   mu_k = 0.8 + k * 1.7 / 15 (k = 0..15),  w = 0.15
 * Energy: E_i = e_mean + e_std * y(i), y a 16-64-64-1 network with SiLU on
 * (G - g_mean) * g_istd; total E = sum over locals; forces F = -dE/dx.
 * Network layers run on MFMA with BF16 weights/activations and FP32
 * accumulation; descriptors and forces in FP32; positions and integration FP64.
 *
 * Parameters (FP32, PAIRS_HIP_MLP_NPARAMS): W1[64][16], b1[64], W2[64][64]
 * (W2[out][in]), b2[64], w3[64], b3.  Scaling (PAIRS_HIP_MLP_NSCALE):
 * g_mean[16], g_istd[16], e_mean, e_std.
 *
 * Training labels come from LJ runs: per-atom LJ energies (pairs_hip_energy's
 * per-atom split, 0.5 * sum_j 4 eps ((s/r)^12 - (s/r)^6) over r < cutoff) and
 * LJ forces.  Two training forms: energy rows in a device-resident set
 * (pairs_hip_mlp_add_snapshot / _add_rows, pairs_hip_mlp_train), and whole
 * configurations of LJ contexts with energy and force terms
 * (pairs_hip_mlp_config_grad, pairs_hip_mlp_apply_grad).  A model is used by
 * a context only if its forces on that context's current state are within a
 * relative RMS tolerance of the LJ forces; otherwise the context keeps LJ
 * (pairs_hip_mlp_attach).
 *
 * Full neighbour lists only (cfg.half == 0).  Return codes as pairs_hip.h.
 */
#ifndef PAIRS_HIP_MLP_H
#define PAIRS_HIP_MLP_H

#include "pairs_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PAIRS_HIP_MLP_NDESC 16
#define PAIRS_HIP_MLP_HIDDEN 64
#define PAIRS_HIP_MLP_NPARAMS 5313
#define PAIRS_HIP_MLP_NSCALE 34

typedef struct pairs_hip_mlp pairs_hip_mlp;

/* A model with Xavier-uniform weights from `seed` (zero biases, identity
   scaling) and room for `dataset_capacity` training rows. */
int pairs_hip_mlp_create(pairs_hip_mlp **model, unsigned long long seed, int dataset_capacity);
/* Detach the model from every context that uses it before destroying it. */
int pairs_hip_mlp_destroy(pairs_hip_mlp *model);
const char *pairs_hip_mlp_last_error(const pairs_hip_mlp *model);
int pairs_hip_mlp_get_params(pairs_hip_mlp *model, float *params, float *scale);
int pairs_hip_mlp_set_params(pairs_hip_mlp *model, const float *params, const float *scale);

/* E_i and dE_i/dG_i (either output may be NULL) of n descriptor rows given
   on the host ([n][16]). */
int pairs_hip_mlp_eval(pairs_hip_mlp *model, int n, const float *G, float *E, float *dEdG);

/* Descriptors of the context's locals on its current state, [nlocal][16] in
   the reference's particle order. */
int pairs_hip_mlp_descriptors(pairs_hip_ctx *ctx, float *G, int count);
/* Append the context's current locals (descriptors, per-atom LJ energy) to
   the model's training set; returns the rows added, or a negative error. */
int pairs_hip_mlp_add_snapshot(pairs_hip_mlp *model, pairs_hip_ctx *ctx);
int pairs_hip_mlp_dataset_size(pairs_hip_mlp *model);
/* Set g_mean, g_istd, e_mean, e_std from rows [0, rows) of the training set. */
int pairs_hip_mlp_fit_scaling(pairs_hip_mlp *model, int rows);
/* Adam on minibatches drawn (with replacement, counter hash of seed and step)
   from rows [0, rows); learning rate decays linearly from lr to lr / 20;
   batch must be a positive multiple of 1024.  loss_out[steps] (may be NULL)
   receives each step's mean squared error of the standardised energy. */
int pairs_hip_mlp_train(pairs_hip_mlp *model, int rows, int steps, int batch, float lr, unsigned long long seed,
                        float *loss_out);
/* Append n rows given on the host (descriptors [n][16], per-atom energy
   labels [n]); returns n, or a negative error. */
int pairs_hip_mlp_add_rows(pairs_hip_mlp *model, int n, const float *G, const float *label);
/* The gradient (grad[PAIRS_HIP_MLP_NPARAMS]) and loss of one minibatch of
   pairs_hip_mlp_train (same row draw for the same seed and step), without
   updating the model. */
int pairs_hip_mlp_energy_grad(pairs_hip_mlp *model, int rows, int batch, unsigned long long seed, int step,
                              float *grad, float *loss);
/* Gradient over the parameters of the loss of the context's current
   configuration, labelled by LJ:
     L = w_energy * mean_i (y_i - (E_lj_i - e_mean) / e_std)^2
       + w_force * sum_i |F_mlp_i - F_lj_i|^2 / sum_i |F_lj_i|^2
   over its locals (F_mlp = -dE/dx of the model, so the force term is
   differentiated through the network's input gradient).  The gradient is
   added to the model's accumulator (pairs_hip_mlp_apply_grad); grad (may be
   NULL) receives this configuration's gradient alone, loss (may be NULL) the
   two unweighted terms {mean squared standardised energy error, relative
   squared force error}.  The context must not have a model attached; its
   force array is left holding the LJ forces. */
int pairs_hip_mlp_config_grad(pairs_hip_ctx *ctx, pairs_hip_mlp *model, double w_energy, double w_force,
                              float *grad, double *loss);
/* One Adam step (as pairs_hip_mlp_train's) on the mean of the gradients
   accumulated by pairs_hip_mlp_config_grad since the last call; clears the
   accumulator.  No-op when nothing was accumulated. */
int pairs_hip_mlp_apply_grad(pairs_hip_mlp *model, float lr);
/* Root mean squared energy error (LJ units) over training rows [begin, end). */
int pairs_hip_mlp_rmse(pairs_hip_mlp *model, int begin, int end, double *rmse);

/* One MLP force evaluation on the context's current state: forces into the
   context's force array (replacing pairs_hip_lj), total energy and virial
   (sum over pairs of r . F_pair) out. */
int pairs_hip_mlp_forces(pairs_hip_ctx *ctx, pairs_hip_mlp *model, double *energy, double *virial);
/* Relative RMS difference |F_mlp - F_lj| / |F_lj| over the locals of the
   current state; leaves the LJ forces in the force array. */
int pairs_hip_mlp_force_error(pairs_hip_ctx *ctx, pairs_hip_mlp *model, double *rel_rms);
/* Use the model as the context's force field for every later step (run,
   step) if its force error on the current state is <= max_rel_rms (a
   negative tolerance skips the check): returns 0 and leaves MLP forces in the
   force array.  Otherwise keeps LJ, leaves LJ forces and returns 1 (the
   fallback).  rel_rms (may be NULL) receives the measured error. */
int pairs_hip_mlp_attach(pairs_hip_ctx *ctx, pairs_hip_mlp *model, double max_rel_rms, double *rel_rms);
int pairs_hip_mlp_detach(pairs_hip_ctx *ctx);
/* 1 if the context's steps use an MLP, else 0. */
int pairs_hip_mlp_active(const pairs_hip_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
