// gfx950 kernels of the radial-descriptor MLP potential (mlp_kernels.hpp).
//
// The network runs on MFMA v_mfma_f32_16x16x16_bf16 with the activations kept
// transposed (features x particles): a layer's 16x16 output block in the
// accumulator layout (lane l holds rows 4(l/16)..4(l/16)+3 of column l%16) is
// exactly the B-operand layout of the next layer's K block, so no data moves
// between layers.  Weights are the A operands, converted to BF16 once per wave
// and held in registers while the wave walks its 16-particle tiles.  The
// first layer (and its transpose in the input gradient) carries each operand
// as a pair of BF16 values hi + lo, three MFMAs per block: the energy is a
// nearly linear function of correlated descriptors, and single-BF16 inputs
// cost more accuracy than the network gains over a linear fit.
// The library is compiled with -ffp-contract=off; the FP32 math here uses
// explicit fmaf where a fused multiply-add is wanted.
#include <algorithm>

#include "md_segment.hpp"
#include "mlp_kernels.hpp"

namespace pk {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (inputs are finite)
__device__ __forceinline__ short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short) (u >> 16);
}
__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = f2bf(a); r[1] = f2bf(b); r[2] = f2bf(c); r[3] = f2bf(d);
    return r;
}
__device__ __forceinline__ bf16x4 pack4(const f32x4 &x) { return pack4(x[0], x[1], x[2], x[3]); }
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t) (uint16_t) b) << 16); }
// x = hi + lo with hi = bf16(x), lo = bf16(x - hi): a pair carries 16 mantissa bits
__device__ __forceinline__ void split4(float a, float b, float c, float d, bf16x4 &hi, bf16x4 &lo) {
    hi = pack4(a, b, c, d);
    lo = pack4(a - bf2f(hi[0]), b - bf2f(hi[1]), c - bf2f(hi[2]), d - bf2f(hi[3]));
}
__device__ __forceinline__ void split4(const f32x4 &x, bf16x4 &hi, bf16x4 &lo) { split4(x[0], x[1], x[2], x[3], hi, lo); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// SiLU and its derivative
__device__ __forceinline__ void silu(float z, float &h, float &dh) {
    const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
    h = z * sg;
    dh = sg * fmaf(z, 1.0f - sg, 1.0f);
}

// ---------------------------------------------------------------- network tile

// A operands and per-lane constants of one wave (lane l: q = l / 16, col = l % 16)
struct MLPRegs {
    bf16x4 a1[4];      // W1 rows 16c + col, columns 4q..4q+3 (high BF16 part)
    bf16x4 a1l[4];     // W1 - a1 (low part: the first layer runs at ~16-bit precision)
    bf16x4 a2[4][4];   // W2 [c2][kc]: rows 16c2 + col, columns 16kc + 4q..
    bf16x4 a2t[4][4];  // W2^T [c][kc]: W2[16kc + 4q + v][16c + col]
    bf16x4 a1t[4];     // W1^T [kc]: W1[16kc + 4q + v][col] (high part)
    bf16x4 a1tl[4];    // its low part
    float b1[4][4], b2[4][4], w3[4][4];  // rows 16c + 4q + v
    float b3, gm[4], gi[4], em, es;
};

template <bool BACK>
__device__ __forceinline__ void mlp_load(MLPRegs &R, const float *__restrict__ p, const float *__restrict__ sc,
                                         int lane) {
    const int q = lane >> 4, col = lane & 15;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float *w = p + MLP_W1 + (16 * c + col) * MLP_ND + 4 * q;
        split4(w[0], w[1], w[2], w[3], R.a1[c], R.a1l[c]);
#pragma unroll
        for (int kc = 0; kc < 4; kc++) {
            const float *u = p + MLP_W2 + (16 * c + col) * MLP_H + 16 * kc + 4 * q;
            R.a2[c][kc] = pack4(u[0], u[1], u[2], u[3]);
            if (BACK) {
                const float *t = p + MLP_W2 + (16 * kc + 4 * q) * MLP_H + 16 * c + col;
                R.a2t[c][kc] = pack4(t[0], t[MLP_H], t[2 * MLP_H], t[3 * MLP_H]);
            }
        }
        if (BACK) {
            const float *t = p + MLP_W1 + (16 * c + 4 * q) * MLP_ND + col;
            split4(t[0], t[MLP_ND], t[2 * MLP_ND], t[3 * MLP_ND], R.a1t[c], R.a1tl[c]);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int h = 16 * c + 4 * q + v;
            R.b1[c][v] = p[MLP_B1 + h];
            R.b2[c][v] = p[MLP_B2 + h];
            R.w3[c][v] = p[MLP_W3 + h];
        }
    }
    R.b3 = p[MLP_B3];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        R.gm[v] = sc[MLP_GMEAN + 4 * q + v];
        R.gi[v] = sc[MLP_GISTD + 4 * q + v];
    }
    R.em = sc[MLP_EMEAN];
    R.es = sc[MLP_ESTD];
}

// standardised descriptors of the lane's particle, features 4q..4q+3
__device__ __forceinline__ f32x4 mlp_input(const MLPRegs &R, float4 g) {
    f32x4 s;
    s[0] = (g.x - R.gm[0]) * R.gi[0];
    s[1] = (g.y - R.gm[1]) * R.gi[1];
    s[2] = (g.z - R.gm[2]) * R.gi[2];
    s[3] = (g.w - R.gm[3]) * R.gi[3];
    return s;
}

// forward pass of one 16-particle tile: returns y of particle col (all four
// lane groups hold it); keeps activations and SiLU derivatives
__device__ __forceinline__ float mlp_forward(const MLPRegs &R, bf16x4 b0, bf16x4 b0l, f32x4 (&h1)[4],
                                             f32x4 (&dh1)[4], bf16x4 (&bh1)[4], f32x4 (&h2)[4], f32x4 (&dh2)[4]) {
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        // W1 s with both operands split: hi*hi + lo*hi + hi*lo
        f32x4 z = mfma16(R.a1l[c], b0, zero);
        z = mfma16(R.a1[c], b0l, z);
        z = mfma16(R.a1[c], b0, z);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            float h, dh;
            silu(z[v] + R.b1[c][v], h, dh);
            h1[c][v] = h;
            dh1[c][v] = dh;
        }
        bh1[c] = pack4(h1[c]);
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
        f32x4 z = zero;
#pragma unroll
        for (int kc = 0; kc < 4; kc++) z = mfma16(R.a2[c][kc], bh1[kc], z);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            float h, dh;
            silu(z[v] + R.b2[c][v], h, dh);
            h2[c][v] = h;
            dh2[c][v] = dh;
        }
    }
    float y = 0.f;
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int v = 0; v < 4; v++) y = fmaf(R.w3[c][v], h2[c][v], y);
    y += __shfl_xor(y, 16, 64);
    y += __shfl_xor(y, 32, 64);
    return y + R.b3;
}

// dy/d(standardised descriptor) of particle col, features 4q..4q+3
__device__ __forceinline__ f32x4 mlp_input_grad(const MLPRegs &R, const f32x4 (&dh1)[4], const f32x4 (&dh2)[4]) {
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    bf16x4 bg2[4], bg1[4], bg1l[4];
#pragma unroll
    for (int c = 0; c < 4; c++)
        bg2[c] = pack4(R.w3[c][0] * dh2[c][0], R.w3[c][1] * dh2[c][1], R.w3[c][2] * dh2[c][2],
                       R.w3[c][3] * dh2[c][3]);
#pragma unroll
    for (int c = 0; c < 4; c++) {
        f32x4 a = zero;
#pragma unroll
        for (int kc = 0; kc < 4; kc++) a = mfma16(R.a2t[c][kc], bg2[kc], a);
        split4(a[0] * dh1[c][0], a[1] * dh1[c][1], a[2] * dh1[c][2], a[3] * dh1[c][3], bg1[c], bg1l[c]);
    }
    // W1^T g1, both split like the forward first layer
    f32x4 d = zero;
#pragma unroll
    for (int kc = 0; kc < 4; kc++) {
        d = mfma16(R.a1tl[kc], bg1[kc], d);
        d = mfma16(R.a1t[kc], bg1l[kc], d);
        d = mfma16(R.a1t[kc], bg1[kc], d);
    }
    return d;
}

// E and dE/dG (raw descriptor units) of n rows; one wave per 16-row tile
template <bool WANT_D>
__global__ __launch_bounds__(256) void k_mlp_eval(const float *__restrict__ p, const float *__restrict__ sc,
                                                  const float *__restrict__ G, int n, float *__restrict__ E,
                                                  float *__restrict__ D) {
    const int lane = threadIdx.x & 63, q = lane >> 4, col = lane & 15;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6), nw = gridDim.x * 4;
    const int ntiles = (n + 15) >> 4;
    if (wave >= ntiles) return;  // wave-uniform
    MLPRegs R;
    mlp_load<WANT_D>(R, p, sc, lane);
    for (int t = wave; t < ntiles; t += nw) {
        const int row = t * 16 + col;
        const bool ok = row < n;
        float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ok) g = *(const float4 *) (G + (size_t) row * MLP_ND + 4 * q);
        bf16x4 b0, b0l;
        split4(mlp_input(R, g), b0, b0l);
        f32x4 h1[4], dh1[4], h2[4], dh2[4];
        bf16x4 bh1[4];
        const float y = mlp_forward(R, b0, b0l, h1, dh1, bh1, h2, dh2);
        if (ok && q == 0 && E) E[row] = fmaf(R.es, y, R.em);
        if (WANT_D) {
            const f32x4 d = mlp_input_grad(R, dh1, dh2);
            if (ok) {
                float4 o;
                o.x = R.es * d[0] * R.gi[0];
                o.y = R.es * d[1] * R.gi[1];
                o.z = R.es * d[2] * R.gi[2];
                o.w = R.es * d[3] * R.gi[3];
                *(float4 *) (D + (size_t) row * MLP_ND + 4 * q) = o;
            }
        }
    }
}

// ---------------------------------------------------------------- descriptors and forces

// log2(e) / (2 w^2): exp(-t^2 / (2 w^2)) = exp2(-t^2 * MLP_C)
constexpr float MLP_C = 1.4426950408889634f / (2.0f * MLP_W * MLP_W);
constexpr float MLP_IW2 = 1.0f / (MLP_W * MLP_W);
constexpr float MLP_PIRC = 3.14159265358979323846f / (float) MLP_RC;

struct MLPSegArgs {
    const int *counters;
    const double *X, *Y, *Z;
    const double4 *P;    // (x, y, z, 0) of X, Y, Z packed: one 16-byte and one 8-byte gather
    const uint32_t *nl;
    int nl_stride;
    const int *numneighs;
    int ncap;            // list rows: counts are clamped to it
    int pstride;
    const float *D, *E;
    float *G;
    double *fx, *fy, *fz;
    double *partials;
    const float4 *rho;   // MODE 2: dL/dF of every slot (a ghost holds its root's)
    const float *sc;     // MODE 2: standardisation (e_std, g_istd)
    float *U;            // MODE 2: [cap][16] back-projected directions
    // MODE 1 without RED, vv set: the integration of k_vv_post fused behind
    // the force (final_integrate(ts) if vv2, initial_integrate(ts+1) and the
    // next positions and images if vv1)
    int vv, vv2, vv1;
    double dt, hdt;
    const int *flags;
    const double *mass;
    double *vx, *vy, *vz;
    double *Xn, *Yn, *Zn;
    double4 *Pn;         // packed copy of the next positions, or null
    Images im;
    Geom g;
};

// G of the 32 locals of one wave (MODE 0), or their force, energy and
// virial (MODE 1), or (MODE 2) the back-projection of force residuals onto
// the descriptor gradients: with rho_i = dL/dF_i and F_i = -sum_j w_ij .
// (D_i + D_j), w_ij,k = dphi_k(r_ij) (x_i - x_j) / r_ij,
//   dL/dD_a,k = -sum_{j in list(a)} w_aj,k . (rho_a - rho_j),
// written as U_a,k = e_std g_istd,k dL/dD_a,k (the direction the network's
// input gradient is differentiated along).  L lanes per local (2, or 4 for
// small systems): lane L p + h walks entries 4s + L q + h of local p's list
// (one 16-byte group of four per step), the L partial sums are added across
// the lanes at the end (FP32 / FP64 sums in any order: the potential is not
// bit-matched to a reference), so L times as many waves hide the gathers'
// latency.  Pairs at r >= 2.5 and entries past a lane's count are masked.
template <int MODE, bool RED, int L>
__global__ __launch_bounds__(64) void k_mlp_wave(MLPSegArgs A) {
    constexpr int U = 4 / L;  // entries per lane per step
    const int wv = xcd_block(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x & 63, h = lane & (L - 1);
    const int nlocal = A.counters[C_NLOCAL];
    const uint32_t jmax = (uint32_t) max(nlocal + A.counters[C_NGHOST] - 1, 0);
    const char *__restrict__ P = (const char *) A.P;
    const double rc2 = MLP_RC * MLP_RC;
    double pe = 0.0, vir = 0.0;
    const int i0 = wv * (64 / L);
    if (i0 < nlocal) {  // wave-uniform
        const int i = i0 + lane / L;
        const bool live = i < nlocal;
        const int ii = live ? i : i0;
        const int n = live ? min(A.numneighs[ii], A.ncap) : 0;
        const int nmax = __builtin_amdgcn_readfirstlane(wave_max(n));
        const double xi = A.X[ii], yi = A.Y[ii], zi = A.Z[ii];
        float acc[MLP_ND], di[MLP_ND];
#pragma unroll
        for (int k = 0; k < MLP_ND; k++) acc[k] = 0.f;
        if (MODE == 1) {
            const float4 *dp = (const float4 *) (A.D + (size_t) ii * MLP_ND);
#pragma unroll
            for (int k4 = 0; k4 < 4; k4++) {
                const float4 d = dp[k4];
                di[4 * k4] = d.x; di[4 * k4 + 1] = d.y; di[4 * k4 + 2] = d.z; di[4 * k4 + 3] = d.w;
            }
        }
        const float4 rhoi = MODE == 2 ? A.rho[ii] : make_float4(0.f, 0.f, 0.f, 0.f);
        double f0 = 0.0, f1 = 0.0, f2 = 0.0, w = 0.0;
        const float gb = exp2f(-2.0f * MLP_DMU * MLP_DMU * MLP_C);  // a_{k+1} / a_k
        const uint4 *lp = (const uint4 *) A.nl + ii;
        const int gl = max((n + 3) / 4 - 1, 0);
        const int nst = (nmax + 3) / 4;
        for (int st = 0; st < nst; st++) {
            // groups past the lane's list re-read its last group; entries are
            // clamped to the last particle slot
            const uint4 grp = lp[(size_t) min(st, gl) * A.nl_stride];
            uint32_t jq[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int c = L * q + h;
                const uint32_t e = (c & 2) ? ((c & 1) ? grp.w : grp.z) : ((c & 1) ? grp.y : grp.x);
                jq[q] = min(e, jmax);
            }
            double xj[U], yj[U], zj[U];
            float4 dj[U][4], rj[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                const uint32_t off = jq[q] << 5;
                const double2 pxy = *(const double2 *) (P + off);
                xj[q] = pxy.x; yj[q] = pxy.y; zj[q] = *(const double *) (P + off + 16);
                if (MODE == 1) {
                    const float4 *dp = (const float4 *) (A.D + (size_t) jq[q] * MLP_ND);
#pragma unroll
                    for (int k4 = 0; k4 < 4; k4++) dj[q][k4] = dp[k4];
                }
                if (MODE == 2) rj[q] = A.rho[jq[q]];
            }
#pragma unroll
            for (int q = 0; q < U; q++) {
                const double dx = xi - xj[q], dy = yi - yj[q], dz = zi - zj[q];
                double rsq = dx * dx + dy * dy;
                rsq = rsq + dz * dz;
                const bool in = (4 * st + L * q + h < n) && rsq < rc2;
                // masked entries take an in-range stand-in distance, so the
                // Gaussian recurrence below stays finite; 1/r from the
                // hardware reciprocal square root, r = rsq / r
                const float rsqf = (float) rsq;
                const float rinv = in ? __builtin_amdgcn_rsqf(rsqf) : 1.0f / MLP_MU0;
                const float r = in ? rsqf * rinv : MLP_MU0;
                // g_k by recurrence over the uniform centres,
                // g_{k+1} = g_k a_k, a_{k+1} = a_k b, from two directly
                // computed anchors (k = 0 and k = 8): four exponentials per
                // pair instead of sixteen, relative error below 1e-5
                const float t0 = r - MLP_MU0;
                auto anchor = [&](int k0, float &g, float &a) {
                    const float t = t0 - (float) k0 * MLP_DMU;
                    g = exp2f(-(t * t) * MLP_C);
                    a = exp2f((2.0f * t - MLP_DMU) * MLP_DMU * MLP_C);
                };
                float gk, ak;
                if (MODE == 0) {
                    const float fc = in ? 0.5f * (cosf(MLP_PIRC * r) + 1.0f) : 0.0f;
#pragma unroll
                    for (int kk = 0; kk < MLP_ND; kk++) {
                        if (kk % 8 == 0) anchor(kk, gk, ak);
                        acc[kk] = fmaf(gk, fc, acc[kk]);
                        gk *= ak;
                        ak *= gb;
                    }
                } else if (MODE == 2) {
                    // the cutoff's sine and cosine from the hardware units
                    // (argument in turns: pi r / rc = 2 pi (r / 2 rc), r < rc)
                    const float ang = r * (0.5f / (float) MLP_RC);
                    const float sn = __builtin_amdgcn_sinf(ang), cs = __builtin_amdgcn_cosf(ang);
                    const float fc = 0.5f * (cs + 1.0f), dfc = -0.5f * MLP_PIRC * sn;
                    // (x_a - x_j) / r . (rho_a - rho_j), zero for a masked entry
                    float proj = (float) dx * (rhoi.x - rj[q].x) + (float) dy * (rhoi.y - rj[q].y);
                    proj = proj + (float) dz * (rhoi.z - rj[q].z);
                    proj = in ? -proj * rinv : 0.0f;
#pragma unroll
                    for (int kk = 0; kk < MLP_ND; kk++) {
                        if (kk % 8 == 0) anchor(kk, gk, ak);
                        const float t = t0 - (float) kk * MLP_DMU;
                        const float dphi = gk * fmaf(-t * MLP_IW2, fc, dfc);
                        acc[kk] = fmaf(dphi, proj, acc[kk]);
                        gk *= ak;
                        ak *= gb;
                    }
                } else {
                    const float ang = r * (0.5f / (float) MLP_RC);
                    const float sn = __builtin_amdgcn_sinf(ang), cs = __builtin_amdgcn_cosf(ang);
                    const float fc = 0.5f * (cs + 1.0f), dfc = -0.5f * MLP_PIRC * sn;
                    float ci = 0.f, cj = 0.f;
                    const float *djq = (const float *) dj[q];
#pragma unroll
                    for (int kk = 0; kk < MLP_ND; kk++) {
                        if (kk % 8 == 0) anchor(kk, gk, ak);
                        const float t = t0 - (float) kk * MLP_DMU;
                        const float dphi = gk * fmaf(-t * MLP_IW2, fc, dfc);
                        ci = fmaf(di[kk], dphi, ci);
                        cj = fmaf(djq[kk], dphi, cj);
                        gk *= ak;
                        ak *= gb;
                    }
                    const float cc = ci + cj;
                    const double coef = in ? (double) (-cc * rinv) : 0.0;
                    f0 += coef * dx; f1 += coef * dy; f2 += coef * dz;
                    if (RED) w += in ? (double) (-cc * r) : 0.0;
                }
            }
        }
        // the L lanes of a local: partial sums added
#pragma unroll
        for (int o = 1; o < L; o <<= 1) {
            if (MODE == 1) {
                f0 += __shfl_xor(f0, o, 64);
                f1 += __shfl_xor(f1, o, 64);
                f2 += __shfl_xor(f2, o, 64);
                if (RED) w += __shfl_xor(w, o, 64);
            } else {
#pragma unroll
                for (int k = 0; k < MLP_ND; k++) acc[k] += __shfl_xor(acc[k], o, 64);
            }
        }
        if (live && h == 0) {
            if (MODE == 0) {
                float4 *gp = (float4 *) (A.G + (size_t) i * MLP_ND);
#pragma unroll
                for (int k4 = 0; k4 < 4; k4++)
                    gp[k4] = make_float4(acc[4 * k4], acc[4 * k4 + 1], acc[4 * k4 + 2], acc[4 * k4 + 3]);
            } else if (MODE == 2) {
                const float es = A.sc[MLP_ESTD];
                float4 *up = (float4 *) (A.U + (size_t) i * MLP_ND);
#pragma unroll
                for (int k4 = 0; k4 < 4; k4++) {
                    const float *gi = A.sc + MLP_GISTD + 4 * k4;
                    up[k4] = make_float4(es * gi[0] * acc[4 * k4], es * gi[1] * acc[4 * k4 + 1],
                                         es * gi[2] * acc[4 * k4 + 2], es * gi[3] * acc[4 * k4 + 3]);
                }
            } else {
                A.fx[i] = f0; A.fy[i] = f1; A.fz[i] = f2;
                if (RED) { pe += (double) A.E[i]; vir += 0.5 * w; }
                if (!RED && A.vv) {  // k_vv_post's update, on the forces just summed
                    const double m = A.mass[i];
                    double v0 = A.vx[i], v1 = A.vy[i], v2 = A.vz[i];
                    double px = A.X[i], py = A.Y[i], pz = A.Z[i];
                    if (!(A.flags[i] & FLAG_FIXED)) {
                        const double hdt = A.hdt, dt = A.dt;
                        if (A.vv2) {
                            double a0 = hdt * f0, a1 = hdt * f1, a2 = hdt * f2;
                            a0 = a0 / m; a1 = a1 / m; a2 = a2 / m;
                            v0 = v0 + a0; v1 = v1 + a1; v2 = v2 + a2;
                        }
                        if (A.vv1) {
                            double a0 = hdt * f0, a1 = hdt * f1, a2 = hdt * f2;
                            a0 = a0 / m; a1 = a1 / m; a2 = a2 / m;
                            v0 = v0 + a0; v1 = v1 + a1; v2 = v2 + a2;
                            px = px + dt * v0; py = py + dt * v1; pz = pz + dt * v2;
                        }
                        A.vx[i] = v0; A.vy[i] = v1; A.vz[i] = v2;
                    }
                    if (A.vv1) {
                        A.Xn[i] = px; A.Yn[i] = py; A.Zn[i] = pz;
                        if (A.Pn) A.Pn[i] = make_double4(px, py, pz, 0.0);
                        if (A.im.on)
                            write_images(A.im, i, A.counters[C_NLOCAL], px, py, pz, A.X, A.Y, A.Z, A.Xn, A.Yn, A.Zn,
                                         A.g, A.Pn);
                    }
                }
            }
        }
    }
    if (MODE == 1 && RED) {
        // one partial per wave (energy, virial rows)
        const double a = wave_sum(pe), b = wave_sum(vir);
        if (lane == 0) {
            A.partials[(size_t) A.pstride + wv] = a;
            A.partials[(size_t) 2 * A.pstride + wv] = b;
        }
    }
}

// lanes per local of the list passes: 4 below 32768 locals (a small system
// fills few waves otherwise), 2 above
static int mlp_lanes(const DevState &d) { return d.nlocal < 32768 ? 4 : 2; }
int mlp_waves(const DevState &d) {
    const int per = 64 / mlp_lanes(d);
    return (d.nlocal + per - 1) / per;
}
template <int MODE, bool RED>
static void launch_mlp_wave(hipStream_t s, const DevState &d, const MLPSegArgs &A) {
    if (mlp_lanes(d) == 4) hipLaunchKernelGGL((k_mlp_wave<MODE, RED, 4>), dim3(mlp_waves(d)), dim3(64), 0, s, A);
    else hipLaunchKernelGGL((k_mlp_wave<MODE, RED, 2>), dim3(mlp_waves(d)), dim3(64), 0, s, A);
}

// dE/dG of every ghost = that of the local at the end of its source chain
__global__ __launch_bounds__(256) void k_mlp_ghost_d(const int *__restrict__ counters, const int *__restrict__ sync_src,
                                                     float *__restrict__ D) {
    const int nlocal = counters[C_NLOCAL], ng = counters[C_NGHOST];
    for (int q = blockIdx.x * 256 + threadIdx.x; q < ng; q += gridDim.x * 256) {
        int root = sync_src[q];
        for (int hop = 0; root >= nlocal && hop < 64; hop++) root = sync_src[root - nlocal];
        if (root >= nlocal) root = nlocal + q;  // unresolved chain: leave the slot as it is
        const float4 *s = (const float4 *) (D + (size_t) root * MLP_ND);
        float4 *d = (float4 *) (D + (size_t) (nlocal + q) * MLP_ND);
#pragma unroll
        for (int k4 = 0; k4 < 4; k4++) d[k4] = s[k4];
    }
}

// ---------------------------------------------------------------- training

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// Gradient of L = (1/B) sum_b (y_b - (label_b - e_mean) / e_std)^2 over one
// minibatch.  Every wave takes the same number of 16-row tiles (batch is a
// multiple of 16 * waves); weight gradients are K = 16-particle MFMAs on
// transposed tiles staged in LDS; bias gradients are lane sums.
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ p, const float *__restrict__ sc,
                                                  MLPTrain t, uint64_t seed, int step) {
    __shared__ short lds[4][3 * MLP_H * 16 + MLP_ND * 16];
    const int lane = threadIdx.x & 63, q = lane >> 4, col = lane & 15, wl = threadIdx.x >> 6;
    const int wave = blockIdx.x * 4 + wl, nw = gridDim.x * 4;
    short *sD2 = lds[wl], *sH1 = sD2 + MLP_H * 16, *sD1 = sH1 + MLP_H * 16, *sG = sD1 + MLP_H * 16;
    MLPRegs R;
    mlp_load<true>(R, p, sc, lane);
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 gW2[4][4], gW1[4];
    float gb1[4][4], gb2[4][4], gw3[4][4], gb3 = 0.f, lsum = 0.f;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        gW1[a] = zero;
#pragma unroll
        for (int b = 0; b < 4; b++) { gW2[a][b] = zero; gb1[a][b] = 0.f; gb2[a][b] = 0.f; gw3[a][b] = 0.f; }
    }
    const float inv_es = 1.0f / R.es, scale = 2.0f / (float) t.batch;
    const int ntiles = t.batch >> 4;
    for (int tile = wave; tile < ntiles; tile += nw) {
        const uint64_t j = (uint64_t) (tile * 16 + col);
        const uint64_t h = splitmix64(seed ^ ((uint64_t) step << 32) ^ j);
        const int idx = (int) (h % (uint64_t) t.rows);
        const float4 g = *(const float4 *) (t.G + (size_t) idx * MLP_ND + 4 * q);
        const float lab = t.label[idx];
        bf16x4 b0, b0l;
        split4(mlp_input(R, g), b0, b0l);
#pragma unroll
        for (int v = 0; v < 4; v++) sG[(4 * q + v) * 16 + col] = b0[v];
        f32x4 h1[4], dh1[4], h2[4], dh2[4];
        bf16x4 bh1[4];
        const float y = mlp_forward(R, b0, b0l, h1, dh1, bh1, h2, dh2);
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) sH1[(16 * c + 4 * q + v) * 16 + col] = bh1[c][v];
        const float err = y - (lab - R.em) * inv_es;
        const float e = scale * err;
        if (q == 0) { lsum = fmaf(err, err, lsum); gb3 += e; }
        bf16x4 bd2[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            f32x4 d2;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                gw3[c][v] = fmaf(e, h2[c][v], gw3[c][v]);
                d2[v] = e * R.w3[c][v] * dh2[c][v];
                gb2[c][v] += d2[v];
            }
            bd2[c] = pack4(d2);
#pragma unroll
            for (int v = 0; v < 4; v++) sD2[(16 * c + 4 * q + v) * 16 + col] = bd2[c][v];
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            f32x4 a = zero;
#pragma unroll
            for (int kc = 0; kc < 4; kc++) a = mfma16(R.a2t[c][kc], bd2[kc], a);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const float d1 = a[v] * dh1[c][v];
                gb1[c][v] += d1;
                sD1[(16 * c + 4 * q + v) * 16 + col] = f2bf(d1);
            }
        }
        __syncthreads();
#pragma unroll
        for (int c2 = 0; c2 < 4; c2++) {
            const bf16x4 a = *(const bf16x4 *) (sD2 + (16 * c2 + col) * 16 + 4 * q);
#pragma unroll
            for (int c1 = 0; c1 < 4; c1++) {
                const bf16x4 b = *(const bf16x4 *) (sH1 + (16 * c1 + col) * 16 + 4 * q);
                gW2[c2][c1] = mfma16(a, b, gW2[c2][c1]);
            }
        }
        const bf16x4 bg = *(const bf16x4 *) (sG + col * 16 + 4 * q);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const bf16x4 a = *(const bf16x4 *) (sD1 + (16 * c + col) * 16 + 4 * q);
            gW1[c] = mfma16(a, bg, gW1[c]);
        }
        __syncthreads();
    }
    float *out = t.partial + (size_t) wave * MLP_NP;
#pragma unroll
    for (int c2 = 0; c2 < 4; c2++)
#pragma unroll
        for (int c1 = 0; c1 < 4; c1++)
#pragma unroll
            for (int v = 0; v < 4; v++) out[MLP_W2 + (16 * c2 + 4 * q + v) * MLP_H + 16 * c1 + col] = gW2[c2][c1][v];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int v = 0; v < 4; v++) out[MLP_W1 + (16 * c + 4 * q + v) * MLP_ND + col] = gW1[c][v];
    // bias sums over the 16 particles of a lane group
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            float a = gb1[c][v], b = gb2[c][v], w = gw3[c][v];
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                a += __shfl_xor(a, o, 64); b += __shfl_xor(b, o, 64); w += __shfl_xor(w, o, 64);
            }
            if (col == 0) {
                const int hh = 16 * c + 4 * q + v;
                out[MLP_B1 + hh] = a; out[MLP_B2 + hh] = b; out[MLP_W3 + hh] = w;
            }
        }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) { gb3 += __shfl_xor(gb3, o, 64); lsum += __shfl_xor(lsum, o, 64); }
    if (lane == 0) { out[MLP_B3] = gb3; t.loss[wave] = lsum; }
}

// SiLU with its first and second derivatives:
// h = z s, h' = s (1 + z (1 - s)), h'' = s (1 - s) (2 + z (1 - 2 s)), s = sigmoid(z)
__device__ __forceinline__ void silu3(float z, float &h, float &dh, float &ddh) {
    const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
    h = z * sg;
    dh = sg * fmaf(z, 1.0f - sg, 1.0f);
    ddh = sg * (1.0f - sg) * fmaf(z, 1.0f - 2.0f * sg, 2.0f);
}

// Gradient over the parameters of one configuration's loss (rows [0, n), one
// per local): sum_a [eps_a y_a + (dy_a/ds) . u_a], eps_a = eps_scale *
// (y_a - (E*_a - e_mean) / e_std) and u_a = U row a held fixed.  Per
// 16-particle tile: the forward pass (h1, h2 and both SiLU derivatives), the
// tangent along u (dz1 = W1 u, dh1 = h1' dz1, dz2 = W2 dh1), then reverse mode
// through both:
//   bar dz2 = w3 h2',  bar z2 = w3 (eps h2' + h2'' dz2),
//   bar dz1 = (W2^T bar dz2) h1',  bar z1 = (W2^T bar z2) h1' + (W2^T bar dz2) h1'' dz1,
//   dW2 = bar z2 h1^T + bar dz2 dh1^T,  dW1 = bar z1 s^T + bar dz1 u^T,
//   db1 = bar z1, db2 = bar z2, dw3 = eps h2 + h2' dz2, db3 = eps.
// Layer products and weight gradients are MFMAs in the layout of k_mlp_grad
// (BF16 operands, FP32 accumulation; s and u enter the first layer as hi + lo
// pairs).  Every wave runs the same number of tile iterations (tiles past the
// end have eps = u = 0 and add exact zeros), so the block barriers are uniform.
__global__ __launch_bounds__(256) void k_mlp_cgrad(const float *__restrict__ p, const float *__restrict__ sc,
                                                   MLPCfgGrad t) {
    __shared__ short lds[4][6 * MLP_H * 16 + 2 * MLP_ND * 16];
    const int lane = threadIdx.x & 63, q = lane >> 4, col = lane & 15, wl = threadIdx.x >> 6;
    const int wave = blockIdx.x * 4 + wl, nw = gridDim.x * 4;
    short *sZ2 = lds[wl], *sH1 = sZ2 + MLP_H * 16, *sT2 = sH1 + MLP_H * 16, *sTH1 = sT2 + MLP_H * 16;
    short *sZ1 = sTH1 + MLP_H * 16, *sT1 = sZ1 + MLP_H * 16, *sS = sT1 + MLP_H * 16, *sU = sS + MLP_ND * 16;
    MLPRegs R;
    mlp_load<true>(R, p, sc, lane);
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 gW2[4][4], gW1[4];
    float gb1[4][4], gb2[4][4], gw3[4][4], gb3 = 0.f, lsum = 0.f;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        gW1[a] = zero;
#pragma unroll
        for (int b = 0; b < 4; b++) { gW2[a][b] = zero; gb1[a][b] = 0.f; gb2[a][b] = 0.f; gw3[a][b] = 0.f; }
    }
    const float inv_es = 1.0f / R.es;
    const int ntiles = (t.n + 15) >> 4;
    const int iters = (ntiles + nw - 1) / nw;
    for (int it = 0; it < iters; it++) {
        const int row = (wave + it * nw) * 16 + col;
        const bool ok = row < t.n;
        float4 g = make_float4(0.f, 0.f, 0.f, 0.f), uu = g;
        double lab = 0.0;
        if (ok) {
            g = *(const float4 *) (t.G + (size_t) row * MLP_ND + 4 * q);
            uu = *(const float4 *) (t.U + (size_t) row * MLP_ND + 4 * q);
            lab = t.label[row];
        }
        bf16x4 b0, b0l, u0, u0l;
        split4(mlp_input(R, g), b0, b0l);
        split4(uu.x, uu.y, uu.z, uu.w, u0, u0l);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            sS[(4 * q + v) * 16 + col] = b0[v];
            sU[(4 * q + v) * 16 + col] = u0[v];
        }
        // forward, keeping h, h' and h'' of both layers
        f32x4 h1[4], d1[4], e1[4], h2[4], d2[4], e2[4];
        bf16x4 bh1[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            f32x4 z = mfma16(R.a1l[c], b0, zero);
            z = mfma16(R.a1[c], b0l, z);
            z = mfma16(R.a1[c], b0, z);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                float h, dh, ddh;
                silu3(z[v] + R.b1[c][v], h, dh, ddh);
                h1[c][v] = h; d1[c][v] = dh; e1[c][v] = ddh;
            }
            bh1[c] = pack4(h1[c]);
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            f32x4 z = zero;
#pragma unroll
            for (int kc = 0; kc < 4; kc++) z = mfma16(R.a2[c][kc], bh1[kc], z);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                float h, dh, ddh;
                silu3(z[v] + R.b2[c][v], h, dh, ddh);
                h2[c][v] = h; d2[c][v] = dh; e2[c][v] = ddh;
            }
        }
        float y = 0.f;
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) y = fmaf(R.w3[c][v], h2[c][v], y);
        y += __shfl_xor(y, 16, 64);
        y += __shfl_xor(y, 32, 64);
        y += R.b3;
        const float err = ok ? y - ((float) lab - R.em) * inv_es : 0.0f;
        const float e = t.eps_scale * err;
        if (q == 0) { lsum = fmaf(err, err, lsum); gb3 += e; }
        // tangent along u
        f32x4 t1[4];
        bf16x4 bth1[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            f32x4 z = mfma16(R.a1l[c], u0, zero);
            z = mfma16(R.a1[c], u0l, z);
            t1[c] = mfma16(R.a1[c], u0, z);
            bth1[c] = pack4(d1[c][0] * t1[c][0], d1[c][1] * t1[c][1], d1[c][2] * t1[c][2], d1[c][3] * t1[c][3]);
        }
        bf16x4 bz2[4], bt2[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            f32x4 t2 = zero;
#pragma unroll
            for (int kc = 0; kc < 4; kc++) t2 = mfma16(R.a2[c][kc], bth1[kc], t2);
            f32x4 z2b, t2b;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const float w = R.w3[c][v];
                gw3[c][v] = fmaf(e, h2[c][v], fmaf(d2[c][v], t2[v], gw3[c][v]));
                t2b[v] = w * d2[c][v];
                z2b[v] = w * fmaf(e, d2[c][v], e2[c][v] * t2[v]);
                gb2[c][v] += z2b[v];
            }
            bz2[c] = pack4(z2b);
            bt2[c] = pack4(t2b);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int f = (16 * c + 4 * q + v) * 16 + col;
                sZ2[f] = bz2[c][v];
                sT2[f] = bt2[c][v];
                sH1[f] = bh1[c][v];
                sTH1[f] = bth1[c][v];
            }
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            f32x4 a = zero, at = zero;
#pragma unroll
            for (int kc = 0; kc < 4; kc++) {
                a = mfma16(R.a2t[c][kc], bz2[kc], a);
                at = mfma16(R.a2t[c][kc], bt2[kc], at);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const float z1b = fmaf(at[v] * e1[c][v], t1[c][v], a[v] * d1[c][v]);
                gb1[c][v] += z1b;
                const int f = (16 * c + 4 * q + v) * 16 + col;
                sZ1[f] = f2bf(z1b);
                sT1[f] = f2bf(at[v] * d1[c][v]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int c2 = 0; c2 < 4; c2++) {
            const bf16x4 az = *(const bf16x4 *) (sZ2 + (16 * c2 + col) * 16 + 4 * q);
            const bf16x4 at = *(const bf16x4 *) (sT2 + (16 * c2 + col) * 16 + 4 * q);
#pragma unroll
            for (int c1 = 0; c1 < 4; c1++) {
                const bf16x4 bh = *(const bf16x4 *) (sH1 + (16 * c1 + col) * 16 + 4 * q);
                const bf16x4 bt = *(const bf16x4 *) (sTH1 + (16 * c1 + col) * 16 + 4 * q);
                gW2[c2][c1] = mfma16(az, bh, gW2[c2][c1]);
                gW2[c2][c1] = mfma16(at, bt, gW2[c2][c1]);
            }
        }
        const bf16x4 bs = *(const bf16x4 *) (sS + col * 16 + 4 * q);
        const bf16x4 bu = *(const bf16x4 *) (sU + col * 16 + 4 * q);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const bf16x4 az = *(const bf16x4 *) (sZ1 + (16 * c + col) * 16 + 4 * q);
            const bf16x4 at = *(const bf16x4 *) (sT1 + (16 * c + col) * 16 + 4 * q);
            gW1[c] = mfma16(az, bs, gW1[c]);
            gW1[c] = mfma16(at, bu, gW1[c]);
        }
        __syncthreads();
    }
    float *out = t.partial + (size_t) wave * MLP_NP;
#pragma unroll
    for (int c2 = 0; c2 < 4; c2++)
#pragma unroll
        for (int c1 = 0; c1 < 4; c1++)
#pragma unroll
            for (int v = 0; v < 4; v++) out[MLP_W2 + (16 * c2 + 4 * q + v) * MLP_H + 16 * c1 + col] = gW2[c2][c1][v];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int v = 0; v < 4; v++) out[MLP_W1 + (16 * c + 4 * q + v) * MLP_ND + col] = gW1[c][v];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            float a = gb1[c][v], b = gb2[c][v], w = gw3[c][v];
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                a += __shfl_xor(a, o, 64); b += __shfl_xor(b, o, 64); w += __shfl_xor(w, o, 64);
            }
            if (col == 0) {
                const int hh = 16 * c + 4 * q + v;
                out[MLP_B1 + hh] = a; out[MLP_B2 + hh] = b; out[MLP_W3 + hh] = w;
            }
        }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) { gb3 += __shfl_xor(gb3, o, 64); lsum += __shfl_xor(lsum, o, 64); }
    if (lane == 0) { out[MLP_B3] = gb3; t.loss[wave] = lsum; }
}

// rho_i = 2 w_force (F_i - F*_i) / sum_j |F*_j|^2 for the locals; the sum is
// taken from k_force_diff's per-block partials (rows 2b + 1)
__global__ __launch_bounds__(256) void k_mlp_rho(const int *__restrict__ counters, const double *__restrict__ fx,
                                                 const double *__restrict__ fy, const double *__restrict__ fz,
                                                 const double *__restrict__ lx, const double *__restrict__ ly,
                                                 const double *__restrict__ lz, const double *__restrict__ partials,
                                                 int nb, double w_force, float4 *__restrict__ rho) {
    __shared__ double s_scale;
    if (threadIdx.x == 0) {
        double den = 0.0;
        for (int b = 0; b < nb; b++) den += partials[2 * b + 1];
        s_scale = den > 0.0 ? 2.0 * w_force / den : 0.0;
    }
    __syncthreads();
    const double sc = s_scale;
    const int n = counters[C_NLOCAL];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        rho[i] = make_float4((float) (sc * (fx[i] - lx[i])), (float) (sc * (fy[i] - ly[i])),
                             (float) (sc * (fz[i] - lz[i])), 0.0f);
}

// rho of every ghost = that of the local at the end of its source chain
__global__ __launch_bounds__(256) void k_mlp_ghost_rho(const int *__restrict__ counters,
                                                       const int *__restrict__ sync_src, float4 *__restrict__ rho) {
    const int nlocal = counters[C_NLOCAL], ng = counters[C_NGHOST];
    for (int q = blockIdx.x * 256 + threadIdx.x; q < ng; q += gridDim.x * 256) {
        int root = sync_src[q];
        for (int hop = 0; root >= nlocal && hop < 64; hop++) root = sync_src[root - nlocal];
        rho[nlocal + q] = root < nlocal ? rho[root] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// acc += sum over waves of the gradient partials (in wave order); with out,
// the sum itself is also written there
__global__ __launch_bounds__(256) void k_mlp_accum(const float *__restrict__ partial, int nwaves,
                                                   float *__restrict__ acc, float *__restrict__ out) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= MLP_NP) return;
    float s = 0.f;
    for (int w = 0; w < nwaves; w++) s += partial[(size_t) w * MLP_NP + k];
    if (acc) acc[k] += s;
    if (out) out[k] = s;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the summed wave partials
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                                                  const float *__restrict__ partial, const float *__restrict__ loss,
                                                  int nwaves, int batch, float lr, float bc1, float bc2,
                                                  float gscale, float *__restrict__ loss_out) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < MLP_NP) {
        float gsum = 0.f;
        for (int w = 0; w < nwaves; w++) gsum += partial[(size_t) w * MLP_NP + k];
        gsum *= gscale;
        const float mm = 0.9f * m[k] + 0.1f * gsum;
        const float vv = 0.999f * v[k] + 0.001f * gsum * gsum;
        m[k] = mm;
        v[k] = vv;
        p[k] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
    }
    if (k == 0 && loss_out) {
        float s = 0.f;
        for (int w = 0; w < nwaves; w++) s += loss[w];
        *loss_out = s / (float) batch;
    }
}

// column moments: per block, 16 descriptor sums, 16 sums of squares, label sum and sum of squares
__global__ __launch_bounds__(256) void k_mlp_moments(const float *__restrict__ G, const float *__restrict__ label,
                                                     int n, double *__restrict__ partials) {
    __shared__ double red[16][34];
    const int col = threadIdx.x & 15, sub = threadIdx.x >> 4;
    double s = 0.0, s2 = 0.0, l = 0.0, l2 = 0.0;
    for (int r = blockIdx.x * 16 + sub; r < n; r += gridDim.x * 16) {
        const double x = G[(size_t) r * MLP_ND + col];
        s += x; s2 += x * x;
        if (col == 0) { const double y = label[r]; l += y; l2 += y * y; }
    }
    red[sub][col] = s;
    red[sub][16 + col] = s2;
    if (col == 0) { red[sub][32] = l; red[sub][33] = l2; }
    __syncthreads();
    if (threadIdx.x < 34) {
        double a = 0.0;
        for (int k = 0; k < 16; k++) a += red[k][threadIdx.x];
        partials[(size_t) blockIdx.x * 34 + threadIdx.x] = a;
    }
}

__global__ __launch_bounds__(256) void k_mlp_sqerr(const float *__restrict__ E, const float *__restrict__ label, int n,
                                                   double *__restrict__ partials) {
    double s = 0.0;
    for (int r = blockIdx.x * 256 + threadIdx.x; r < n; r += gridDim.x * 256) {
        const double d = (double) E[r] - (double) label[r];
        s += d * d;
    }
    __shared__ double red[4];
    const double a = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ __launch_bounds__(256) void k_mlp_snapshot(const float *__restrict__ G, const double *__restrict__ eatom,
                                                      int nlocal, float *__restrict__ dG, float *__restrict__ dl) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nlocal; i += gridDim.x * 256) {
        const float4 *s = (const float4 *) (G + (size_t) i * MLP_ND);
        float4 *d = (float4 *) (dG + (size_t) i * MLP_ND);
#pragma unroll
        for (int k4 = 0; k4 < 4; k4++) d[k4] = s[k4];
        dl[i] = (float) eatom[i];
    }
}

__global__ __launch_bounds__(256) void k_force_diff(const double *__restrict__ ax, const double *__restrict__ ay,
                                                    const double *__restrict__ az, const double *__restrict__ bx,
                                                    const double *__restrict__ by, const double *__restrict__ bz,
                                                    int n, double *__restrict__ partials) {
    double s = 0.0, r = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double d0 = ax[i] - bx[i], d1 = ay[i] - by[i], d2 = az[i] - bz[i];
        s += d0 * d0 + d1 * d1 + d2 * d2;
        r += bx[i] * bx[i] + by[i] * by[i] + bz[i] * bz[i];
    }
    __shared__ double red[2][4];
    const double a = wave_sum(s), b = wave_sum(r);
    if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = a; red[1][threadIdx.x >> 6] = b; }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int k = threadIdx.x;
        partials[2 * blockIdx.x + k] = ((red[k][0] + red[k][1]) + red[k][2]) + red[k][3];
    }
}

// ---------------------------------------------------------------- launchers

static MLPSegArgs seg_args(const DevState &d, int cur) {
    MLPSegArgs A{};
    A.counters = d.counters;
    A.X = d.X[cur]; A.Y = d.Y[cur]; A.Z = d.Z[cur];
    A.P = d.P[cur];
    A.nl = d.nl;
    A.nl_stride = d.nl_stride;
    A.numneighs = d.numneighs;
    A.ncap = d.ncap;
    A.pstride = d.pstride;
    return A;
}

void launch_mlp_descriptors(hipStream_t s, const DevState &d, const Geom &g, int cur, float *G, bool pack) {
    if (d.nlocal <= 0) return;
    (void) g;
    if (pack) launch_pack_p(s, d, cur);  // both passes gather packed positions
    MLPSegArgs A = seg_args(d, cur);
    A.G = G;
    launch_mlp_wave<0, false>(s, d, A);
}

// each wave converts the weights to BF16 registers once (about 21 KB read)
// and then strides over 16-row tiles: 512 workgroups (two waves per SIMD)
// amortise that over several tiles at 131072 rows
static int eval_grid(int n) {
    const int ntiles = (n + 15) / 16;
    return std::max(1, std::min((ntiles + 3) / 4, 512));
}

void launch_mlp_eval(hipStream_t s, const float *p, const float *sc, const float *G, int n, float *E, float *D) {
    if (n <= 0) return;
    if (D) hipLaunchKernelGGL(k_mlp_eval<true>, dim3(eval_grid(n)), dim3(256), 0, s, p, sc, G, n, E, D);
    else hipLaunchKernelGGL(k_mlp_eval<false>, dim3(eval_grid(n)), dim3(256), 0, s, p, sc, G, n, E, D);
}

void launch_mlp_forces(hipStream_t s, const DevState &d, const Geom &g, const MLPBind &b, int cur, int reduce,
                       int ghost_grid, const MLPVV *vv, bool pack) {
    if (d.nlocal <= 0) return;
    launch_mlp_descriptors(s, d, g, cur, b.G, pack);
    launch_mlp_eval(s, b.p, b.sc, b.G, d.nlocal, b.E, b.D);
    hipLaunchKernelGGL(k_mlp_ghost_d, dim3(std::max(ghost_grid, 1)), dim3(256), 0, s, d.counters, d.sync_src, b.D);
    MLPSegArgs A = seg_args(d, cur);
    A.D = b.D;
    A.E = b.E;
    A.fx = d.fx; A.fy = d.fy; A.fz = d.fz;
    A.partials = d.partials;
    if (vv && !reduce) {
        A.vv = 1;
        A.vv2 = vv->vv2; A.vv1 = vv->vv1;
        A.dt = vv->prm->dt; A.hdt = vv->prm->hdt;
        A.flags = d.flags; A.mass = d.mass;
        A.vx = d.vx; A.vy = d.vy; A.vz = d.vz;
        A.Xn = d.X[vv->next]; A.Yn = d.Y[vv->next]; A.Zn = d.Z[vv->next];
        A.im = images_of(d, vv->vv1 && vv->next != cur);
        A.g = g;
        A.Pn = vv->Pn;
    }
    if (reduce) launch_mlp_wave<1, true>(s, d, A);
    else launch_mlp_wave<1, false>(s, d, A);
}

void launch_mlp_grad(hipStream_t s, const float *p, const float *sc, const MLPTrain &t, uint64_t seed, int step) {
    hipLaunchKernelGGL(k_mlp_grad, dim3(t.nwaves / 4), dim3(256), 0, s, p, sc, t, seed, step);
}

void launch_mlp_adam(hipStream_t s, float *p, float *m, float *v, const MLPTrain &t, float lr, int step,
                     float *loss_out) {
    const float bc1 = 1.0f - powf(0.9f, (float) step), bc2 = 1.0f - powf(0.999f, (float) step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NP + 255) / 256), dim3(256), 0, s, p, m, v, t.partial, t.loss,
                       t.nwaves, t.batch, lr, bc1, bc2, t.gscale, loss_out);
}

void launch_mlp_config_grad(hipStream_t s, const DevState &d, const Geom &g, const MLPBind &b, int cur,
                            const double *lx, const double *ly, const double *lz, const double *eatom,
                            double w_energy, double w_force, float4 *rho, float *U, double *red, int nb,
                            int ghost_grid, const MLPCfgGrad &t) {
    const int n = d.nlocal;
    if (n <= 0) return;
    // the network's forces (descriptors, E, dE/dG of locals and ghosts) into d.f*
    launch_mlp_forces(s, d, g, b, cur, 0, ghost_grid);
    launch_force_diff(s, d.fx, d.fy, d.fz, lx, ly, lz, n, red, nb);
    hipLaunchKernelGGL(k_mlp_rho, dim3(std::min((n + 255) / 256, 1024)), dim3(256), 0, s, d.counters, d.fx, d.fy,
                       d.fz, lx, ly, lz, red, nb, w_force, rho);
    hipLaunchKernelGGL(k_mlp_ghost_rho, dim3(std::max(ghost_grid, 1)), dim3(256), 0, s, d.counters, d.sync_src,
                       rho);
    MLPSegArgs A = seg_args(d, cur);
    A.rho = rho;
    A.sc = b.sc;
    A.U = U;
    launch_mlp_wave<2, false>(s, d, A);
    MLPCfgGrad tt = t;
    tt.G = b.G;
    tt.label = eatom;
    tt.U = U;
    tt.n = n;
    tt.eps_scale = (float) (2.0 * w_energy / n);
    hipLaunchKernelGGL(k_mlp_cgrad, dim3(tt.nwaves / 4), dim3(256), 0, s, b.p, b.sc, tt);
}

void launch_mlp_accum(hipStream_t s, const float *partial, int nwaves, float *acc, float *out) {
    hipLaunchKernelGGL(k_mlp_accum, dim3((MLP_NP + 255) / 256), dim3(256), 0, s, partial, nwaves, acc, out);
}

void launch_mlp_moments(hipStream_t s, const float *G, const float *label, int n, double *partials, int nblocks) {
    hipLaunchKernelGGL(k_mlp_moments, dim3(nblocks), dim3(256), 0, s, G, label, n, partials);
}

void launch_mlp_sqerr(hipStream_t s, const float *E, const float *label, int n, double *partials, int nblocks) {
    hipLaunchKernelGGL(k_mlp_sqerr, dim3(nblocks), dim3(256), 0, s, E, label, n, partials);
}

void launch_mlp_snapshot(hipStream_t s, const float *G, const double *eatom, int nlocal, float *dst_G,
                         float *dst_label) {
    if (nlocal <= 0) return;
    hipLaunchKernelGGL(k_mlp_snapshot, dim3(std::min((nlocal + 255) / 256, 2048)), dim3(256), 0, s, G, eatom, nlocal,
                       dst_G, dst_label);
}

void launch_force_diff(hipStream_t s, const double *ax, const double *ay, const double *az, const double *bx,
                       const double *by, const double *bz, int n, double *partials, int nblocks) {
    hipLaunchKernelGGL(k_force_diff, dim3(nblocks), dim3(256), 0, s, ax, ay, az, bx, by, bz, n, partials);
}

}  // namespace pk
