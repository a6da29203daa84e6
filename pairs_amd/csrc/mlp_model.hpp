// The model object behind include/pairs_hip_mlp.h (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "mlp_kernels.hpp"

struct pairs_hip_mlp {
    std::string err;
    hipStream_t stream = nullptr;
    float *p = nullptr, *sc = nullptr;   // parameters, standardisation (device)
    float *m = nullptr, *v = nullptr;    // Adam moments
    int adam_t = 0;                      // Adam steps taken
    float *ds_G = nullptr, *ds_label = nullptr, *ds_E = nullptr;  // training rows
    int ds_cap = 0, ds_rows = 0;
    float *partial = nullptr, *loss = nullptr, *loss_hist = nullptr;
    int nwaves = 64, hist_cap = 0;
    double *red = nullptr;               // reduction partials
    int red_cap = 0;
    float *io_G = nullptr, *io_E = nullptr, *io_D = nullptr;  // pairs_hip_mlp_eval staging
    int io_cap = 0;
    float *gacc = nullptr;               // accumulated configuration gradients
    float *gout = nullptr;               // one gradient for the host
    int gcount = 0;                      // configurations in gacc
};
