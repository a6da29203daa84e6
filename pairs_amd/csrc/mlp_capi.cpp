// C ABI of the MLP model and its training (include/pairs_hip_mlp.h); the
// functions that work on an LJ context live with the context in pairs_hip.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "../../include/pairs_hip_mlp.h"
#include "mlp_model.hpp"

using namespace pk;

static_assert(MLP_NP == PAIRS_HIP_MLP_NPARAMS, "parameter count");
static_assert(MLP_NS == PAIRS_HIP_MLP_NSCALE, "scaling count");

#define MHIP(expr)                                                                                     \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            m->err = std::string(#expr) + ": " + hipGetErrorString(e_);                                \
            return PAIRS_HIP_ERR_HIP;                                                                  \
        }                                                                                              \
    } while (0)

#define MLAUNCH()                                                                                      \
    do {                                                                                               \
        hipError_t e_ = hipGetLastError();                                                             \
        if (e_ != hipSuccess) {                                                                        \
            m->err = std::string("kernel launch: ") + hipGetErrorString(e_);                           \
            return PAIRS_HIP_ERR_HIP;                                                                  \
        }                                                                                              \
    } while (0)

static int ensure_red(pairs_hip_mlp *m, int n) {
    if (m->red_cap >= n) return 0;
    if (m->red) hipFree(m->red);
    m->red = nullptr;
    MHIP(hipMalloc((void **) &m->red, (size_t) n * sizeof(double)));
    m->red_cap = n;
    return 0;
}

extern "C" {

int pairs_hip_mlp_create(pairs_hip_mlp **out, unsigned long long seed, int dataset_capacity) {
    if (!out || dataset_capacity < 0) return PAIRS_HIP_ERR_ARG;
    pairs_hip_mlp *m = new pairs_hip_mlp();
    *out = m;
    MHIP(hipStreamCreate(&m->stream));
    MHIP(hipMalloc((void **) &m->p, MLP_NP * sizeof(float)));
    MHIP(hipMalloc((void **) &m->sc, MLP_NS * sizeof(float)));
    MHIP(hipMalloc((void **) &m->m, MLP_NP * sizeof(float)));
    MHIP(hipMalloc((void **) &m->v, MLP_NP * sizeof(float)));
    MHIP(hipMemset(m->m, 0, MLP_NP * sizeof(float)));
    MHIP(hipMemset(m->v, 0, MLP_NP * sizeof(float)));
    const size_t cap = (size_t) std::max(dataset_capacity, 16);
    MHIP(hipMalloc((void **) &m->ds_G, cap * MLP_ND * sizeof(float)));
    MHIP(hipMalloc((void **) &m->ds_label, cap * sizeof(float)));
    MHIP(hipMalloc((void **) &m->ds_E, cap * sizeof(float)));
    m->ds_cap = (int) cap;
    MHIP(hipMalloc((void **) &m->partial, (size_t) m->nwaves * MLP_NP * sizeof(float)));
    MHIP(hipMalloc((void **) &m->gacc, MLP_NP * sizeof(float)));
    MHIP(hipMalloc((void **) &m->gout, MLP_NP * sizeof(float)));
    MHIP(hipMemset(m->gacc, 0, MLP_NP * sizeof(float)));
    MHIP(hipMalloc((void **) &m->loss, m->nwaves * sizeof(float)));
    // Xavier-uniform weights, zero biases, identity scaling
    std::vector<float> p(MLP_NP, 0.0f), sc(MLP_NS, 0.0f);
    std::mt19937_64 gen(seed);
    auto fill = [&](int off, int fan_out, int fan_in) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        std::uniform_real_distribution<double> u(-a, a);
        for (int k = 0; k < fan_out * fan_in; k++) p[off + k] = (float) u(gen);
    };
    fill(MLP_W1, MLP_H, MLP_ND);
    fill(MLP_W2, MLP_H, MLP_H);
    fill(MLP_W3, 1, MLP_H);
    for (int k = 0; k < MLP_ND; k++) sc[MLP_GISTD + k] = 1.0f;
    sc[MLP_ESTD] = 1.0f;
    MHIP(hipMemcpy(m->p, p.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice));
    MHIP(hipMemcpy(m->sc, sc.data(), MLP_NS * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

int pairs_hip_mlp_destroy(pairs_hip_mlp *m) {
    if (!m) return PAIRS_HIP_ERR_ARG;
    if (m->stream) hipStreamSynchronize(m->stream);
    void *bufs[] = {m->p, m->sc, m->m, m->v, m->ds_G, m->ds_label, m->ds_E, m->partial, m->loss,
                    m->loss_hist, m->red, m->io_G, m->io_E, m->io_D, m->gacc, m->gout};
    for (void *b : bufs)
        if (b) hipFree(b);
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
    return 0;
}

const char *pairs_hip_mlp_last_error(const pairs_hip_mlp *m) { return m ? m->err.c_str() : "null model"; }

int pairs_hip_mlp_get_params(pairs_hip_mlp *m, float *params, float *scale) {
    if (!m) return PAIRS_HIP_ERR_ARG;
    MHIP(hipStreamSynchronize(m->stream));
    if (params) MHIP(hipMemcpy(params, m->p, MLP_NP * sizeof(float), hipMemcpyDeviceToHost));
    if (scale) MHIP(hipMemcpy(scale, m->sc, MLP_NS * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int pairs_hip_mlp_set_params(pairs_hip_mlp *m, const float *params, const float *scale) {
    if (!m) return PAIRS_HIP_ERR_ARG;
    MHIP(hipStreamSynchronize(m->stream));
    if (params) MHIP(hipMemcpy(m->p, params, MLP_NP * sizeof(float), hipMemcpyHostToDevice));
    if (scale) MHIP(hipMemcpy(m->sc, scale, MLP_NS * sizeof(float), hipMemcpyHostToDevice));
    MHIP(hipDeviceSynchronize());
    return 0;
}

int pairs_hip_mlp_eval(pairs_hip_mlp *m, int n, const float *G, float *E, float *dEdG) {
    if (!m || n < 0 || (n > 0 && !G)) return PAIRS_HIP_ERR_ARG;
    if (n == 0) return 0;
    if (m->io_cap < n) {
        for (float **b : {&m->io_G, &m->io_E, &m->io_D})
            if (*b) { hipFree(*b); *b = nullptr; }
        MHIP(hipMalloc((void **) &m->io_G, (size_t) n * MLP_ND * sizeof(float)));
        MHIP(hipMalloc((void **) &m->io_E, (size_t) n * sizeof(float)));
        MHIP(hipMalloc((void **) &m->io_D, (size_t) n * MLP_ND * sizeof(float)));
        m->io_cap = n;
    }
    MHIP(hipMemcpyAsync(m->io_G, G, (size_t) n * MLP_ND * sizeof(float), hipMemcpyHostToDevice, m->stream));
    launch_mlp_eval(m->stream, m->p, m->sc, m->io_G, n, m->io_E, dEdG ? m->io_D : nullptr);
    MLAUNCH();
    if (E) MHIP(hipMemcpyAsync(E, m->io_E, (size_t) n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    if (dEdG)
        MHIP(hipMemcpyAsync(dEdG, m->io_D, (size_t) n * MLP_ND * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP(hipStreamSynchronize(m->stream));
    return 0;
}

int pairs_hip_mlp_dataset_size(pairs_hip_mlp *m) { return m ? m->ds_rows : PAIRS_HIP_ERR_ARG; }

int pairs_hip_mlp_fit_scaling(pairs_hip_mlp *m, int rows) {
    if (!m || rows < 2 || rows > m->ds_rows) return PAIRS_HIP_ERR_ARG;
    const int nb = std::min((rows + 4095) / 4096, 1024);
    if (ensure_red(m, nb * 34)) return PAIRS_HIP_ERR_HIP;
    launch_mlp_moments(m->stream, m->ds_G, m->ds_label, rows, m->red, nb);
    MLAUNCH();
    std::vector<double> part((size_t) nb * 34);
    MHIP(hipMemcpyAsync(part.data(), m->red, part.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MHIP(hipStreamSynchronize(m->stream));
    double s[34] = {0.0};
    for (int b = 0; b < nb; b++)
        for (int k = 0; k < 34; k++) s[k] += part[(size_t) b * 34 + k];
    std::vector<float> sc(MLP_NS);
    auto stdev = [&](double sum, double sum2) {
        const double mean = sum / rows;
        const double var = std::max(sum2 / rows - mean * mean, 0.0);
        return std::max(std::sqrt(var), 1e-6);
    };
    for (int k = 0; k < MLP_ND; k++) {
        sc[MLP_GMEAN + k] = (float) (s[k] / rows);
        sc[MLP_GISTD + k] = (float) (1.0 / stdev(s[k], s[16 + k]));
    }
    sc[MLP_EMEAN] = (float) (s[32] / rows);
    sc[MLP_ESTD] = (float) stdev(s[32], s[33]);
    MHIP(hipMemcpy(m->sc, sc.data(), MLP_NS * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

int pairs_hip_mlp_train(pairs_hip_mlp *m, int rows, int steps, int batch, float lr, unsigned long long seed,
                        float *loss_out) {
    if (!m || rows < 1 || rows > m->ds_rows || steps < 0 || batch <= 0 || batch % (16 * m->nwaves) != 0 ||
        !(lr > 0.0f))
        return PAIRS_HIP_ERR_ARG;
    if (loss_out && m->hist_cap < steps) {
        if (m->loss_hist) hipFree(m->loss_hist);
        m->loss_hist = nullptr;
        MHIP(hipMalloc((void **) &m->loss_hist, (size_t) std::max(steps, 1) * sizeof(float)));
        m->hist_cap = steps;
    }
    MLPTrain t;
    t.G = m->ds_G;
    t.label = m->ds_label;
    t.rows = rows;
    t.batch = batch;
    t.nwaves = m->nwaves;
    t.partial = m->partial;
    t.loss = m->loss;
    for (int k = 0; k < steps; k++) {
        const float frac = steps > 1 ? (float) k / (float) (steps - 1) : 0.0f;
        const float lrk = lr * (1.0f - 0.95f * frac);
        m->adam_t++;
        launch_mlp_grad(m->stream, m->p, m->sc, t, (uint64_t) seed, k);
        launch_mlp_adam(m->stream, m->p, m->m, m->v, t, lrk, m->adam_t, loss_out ? m->loss_hist + k : nullptr);
        MLAUNCH();
    }
    if (loss_out && steps > 0)
        MHIP(hipMemcpyAsync(loss_out, m->loss_hist, (size_t) steps * sizeof(float), hipMemcpyDeviceToHost,
                            m->stream));
    MHIP(hipStreamSynchronize(m->stream));
    return 0;
}

int pairs_hip_mlp_add_rows(pairs_hip_mlp *m, int n, const float *G, const float *label) {
    if (!m || n < 0 || (n > 0 && (!G || !label))) return PAIRS_HIP_ERR_ARG;
    if (m->ds_rows + n > m->ds_cap) {
        m->err = "the model's training set is full";
        return PAIRS_HIP_ERR_STATE;
    }
    if (n == 0) return 0;
    MHIP(hipMemcpyAsync(m->ds_G + (size_t) m->ds_rows * MLP_ND, G, (size_t) n * MLP_ND * sizeof(float),
                        hipMemcpyHostToDevice, m->stream));
    MHIP(hipMemcpyAsync(m->ds_label + m->ds_rows, label, (size_t) n * sizeof(float), hipMemcpyHostToDevice,
                        m->stream));
    MHIP(hipStreamSynchronize(m->stream));
    m->ds_rows += n;
    return n;
}

int pairs_hip_mlp_energy_grad(pairs_hip_mlp *m, int rows, int batch, unsigned long long seed, int step, float *grad,
                              float *loss) {
    if (!m || rows < 1 || rows > m->ds_rows || batch <= 0 || batch % (16 * m->nwaves) != 0 || step < 0 || !grad)
        return PAIRS_HIP_ERR_ARG;
    MLPTrain t;
    t.G = m->ds_G;
    t.label = m->ds_label;
    t.rows = rows;
    t.batch = batch;
    t.nwaves = m->nwaves;
    t.partial = m->partial;
    t.loss = m->loss;
    launch_mlp_grad(m->stream, m->p, m->sc, t, (uint64_t) seed, step);
    launch_mlp_accum(m->stream, m->partial, m->nwaves, nullptr, m->gout);
    MLAUNCH();
    MHIP(hipMemcpyAsync(grad, m->gout, MLP_NP * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    std::vector<float> wl(m->nwaves);
    MHIP(hipMemcpyAsync(wl.data(), m->loss, wl.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP(hipStreamSynchronize(m->stream));
    if (loss) {
        double s = 0.0;
        for (float x : wl) s += x;
        *loss = (float) (s / batch);
    }
    return 0;
}

int pairs_hip_mlp_apply_grad(pairs_hip_mlp *m, float lr) {
    if (!m || !(lr > 0.0f)) return PAIRS_HIP_ERR_ARG;
    if (m->gcount == 0) return 0;
    MLPTrain t{};
    t.partial = m->gacc;
    t.loss = m->loss;
    t.nwaves = 1;
    t.batch = 1;
    t.gscale = 1.0f / (float) m->gcount;
    m->adam_t++;
    launch_mlp_adam(m->stream, m->p, m->m, m->v, t, lr, m->adam_t, nullptr);
    MHIP(hipMemsetAsync(m->gacc, 0, MLP_NP * sizeof(float), m->stream));
    MLAUNCH();
    MHIP(hipStreamSynchronize(m->stream));
    m->gcount = 0;
    return 0;
}

int pairs_hip_mlp_rmse(pairs_hip_mlp *m, int begin, int end, double *rmse) {
    if (!m || begin < 0 || end > m->ds_rows || end <= begin || !rmse) return PAIRS_HIP_ERR_ARG;
    const int n = end - begin;
    launch_mlp_eval(m->stream, m->p, m->sc, m->ds_G + (size_t) begin * MLP_ND, n, m->ds_E + begin, nullptr);
    const int nb = std::min((n + 255) / 256, 1024);
    if (ensure_red(m, nb)) return PAIRS_HIP_ERR_HIP;
    launch_mlp_sqerr(m->stream, m->ds_E + begin, m->ds_label + begin, n, m->red, nb);
    MLAUNCH();
    std::vector<double> part(nb);
    MHIP(hipMemcpyAsync(part.data(), m->red, nb * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MHIP(hipStreamSynchronize(m->stream));
    double s = 0.0;
    for (double x : part) s += x;
    *rmse = std::sqrt(s / n);
    return 0;
}

}  // extern "C"
