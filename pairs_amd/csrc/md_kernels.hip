// gfx950 kernels of the LJ molecular-dynamics hot path.
//
// Every floating-point expression follows the operation order of the module
// the reference's code generator emits for it (cited per kernel), and the
// library is compiled with -ffp-contract=off, so each value is rounded exactly
// as in the reference's generated CPU program.  Integer/index work (cells,
// ghost maps, neighbour lists) is deterministic and reproduces the CPU
// program's order: stable compactions by ballot/prefix scan instead of the
// atomic insertion the generated CUDA code uses.
#include "md_kernels.hpp"

namespace pk {

// ---------------------------------------------------------------- utilities

__global__ void k_fill_i32(int *__restrict__ p, int v, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double *__restrict__ p, double v, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_set_counters(int *__restrict__ c, int first, int count, int v) {
    int t = threadIdx.x;
    if (t < count) c[first + t] = v;
}

// exclusive scan of one value per thread over a block of BS threads (BS = 64 * nwaves)
template <int BS>
__device__ __forceinline__ int block_exclusive_scan(int v, int *lds_waves, int *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds_waves[wave] = x;
    __syncthreads();
    if (wave == 0) {
        int w = lane < BS / 64 ? lds_waves[lane] : 0;
        int s = w;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(s, off, 64);
            if (lane >= off) s += y;
        }
        if (lane < BS / 64) lds_waves[lane] = s - w;
        if (lane == BS / 64 - 1) lds_waves[BS / 64] = s;
    }
    __syncthreads();
    int r = x - v + lds_waves[wave];
    *total = lds_waves[BS / 64];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------- cells

// cell of a position: src/pairs/sim/cell_lists.py:114-126 (generated as
// (int)((x - (min - s)) / s), clamped to [0, n-1], flattened z fastest, + 1)
__device__ __forceinline__ int cell_of(double x, double y, double z, const Geom &g) {
    const double xs[3] = {x, y, z};
    int idx = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double q = (xs[d] - (g.lo[d] - g.spacing)) / g.spacing;
        int c = (int) q >= 0 ? (int) q : 0;
        c = c < g.dim_cells[d] ? c : g.dim_cells[d] - 1;
        idx = d == 0 ? c : idx * g.dim_cells[d] + c;
    }
    return idx + 1;
}

// BuildCellLists (src/pairs/sim/cell_lists.py:90-134), histogram half:
// cell id of every local and ghost particle, per-cell counts.
__global__ __launch_bounds__(256) void k_bin_count(const double4 *__restrict__ P, const int *__restrict__ flags,
                                                   const int *__restrict__ counters, Geom g,
                                                   int *__restrict__ particle_cell, int *__restrict__ cell_count,
                                                   int *__restrict__ slot) {
    if (counters[C_OVF]) return;
    const int n = counters[C_NLOCAL] + counters[C_NGHOST];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        int c = 0;
        if (!(flags[i] & FLAG_INFINITE)) {
            double4 p = P[i];
            c = cell_of(p.x, p.y, p.z, g);
        }
        particle_cell[i] = c;
        slot[i] = atomicAdd(&cell_count[c], 1);
    }
}

// exclusive scan of the cell counts -> cell_start[0..ncells]; one block
__global__ __launch_bounds__(1024) void k_scan_cells(const int *__restrict__ counters, const int *__restrict__ cell_count,
                                                     int *__restrict__ cell_start, int ncells) {
    __shared__ int lds[20];
    if (counters[C_OVF]) return;
    int carry = 0;
    for (int base = 0; base < ncells; base += 1024) {
        int c = base + threadIdx.x;
        int v = c < ncells ? cell_count[c] : 0;
        int total;
        int ex = block_exclusive_scan<1024>(v, lds, &total);
        if (c < ncells) cell_start[c] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) cell_start[ncells] = carry;
}

__global__ __launch_bounds__(256) void k_bin_scatter(const int *__restrict__ counters, const int *__restrict__ particle_cell,
                                                     const int *__restrict__ slot, const int *__restrict__ cell_start,
                                                     int *__restrict__ sorted_idx) {
    if (counters[C_OVF]) return;
    const int n = counters[C_NLOCAL] + counters[C_NGHOST];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        sorted_idx[cell_start[particle_cell[i]] + slot[i]] = i;
}

// Order each cell by particle index -- the order in which the CPU program's
// sequential "atomic" insertion fills cell_particles -- then gather the
// cell-sorted copy of the positions and count the locals of each cell.
__global__ __launch_bounds__(256) void k_sort_cells(const int *__restrict__ counters, const int *__restrict__ cell_start,
                                                    int *__restrict__ sorted_idx, const double4 *__restrict__ P,
                                                    double4 *__restrict__ sorted_pos, int *__restrict__ cell_nlocal,
                                                    int ncells) {
    if (counters[C_OVF]) return;
    const int nlocal = counters[C_NLOCAL];
    for (int c = blockIdx.x * 256 + threadIdx.x; c < ncells; c += gridDim.x * 256) {
        const int b = cell_start[c], e = cell_start[c + 1];
        for (int k = b + 1; k < e; k++) {  // insertion sort: ~18 entries per cell
            int v = sorted_idx[k];
            int q = k - 1;
            while (q >= b && sorted_idx[q] > v) { sorted_idx[q + 1] = sorted_idx[q]; q--; }
            sorted_idx[q + 1] = v;
        }
        int nl = 0;
        for (int k = b; k < e; k++) {
            int i = sorted_idx[k];
            sorted_pos[k] = P[i];
            nl += i < nlocal;
        }
        cell_nlocal[c] = nl;
    }
}

// ---------------------------------------------------------------- neighbour lists

// BuildNeighborLists (src/pairs/sim/neighbor_lists.py:21-48) over the cell
// stencil (src/pairs/sim/interaction.py:90-120): candidates are cell 0 then
// the 27 stencil cells in order, each cell in index order; accepted when
// j != i (full) or i < j (half, point masses) and rsq < (rc+skin)^2.
//
// One workgroup per home cell: the candidate tile (positions + indices) is
// staged in LDS once and reused by every home particle; each wave takes one
// home particle at a time, its 64 lanes test 64 candidates, and a ballot with
// a lane prefix count compacts the accepted ones in candidate order.
constexpr int NB_TILE = 1536;
constexpr int NB_HOME = 256;

__global__ __launch_bounds__(256) void k_build_neighbors(const int *__restrict__ counters, Geom g,
                                                         const int *__restrict__ cell_start,
                                                         const int *__restrict__ cell_nlocal,
                                                         const int *__restrict__ sorted_idx,
                                                         const double4 *__restrict__ sorted_pos,
                                                         const int *__restrict__ flags, int half, double rc2,
                                                         int *__restrict__ nl, int nl_stride, int ncap,
                                                         int *__restrict__ numneighs, int *__restrict__ ovf) {
    __shared__ double tx[NB_TILE], ty[NB_TILE], tz[NB_TILE];
    __shared__ int tj[NB_TILE];
    __shared__ int seg_begin[28], seg_off[29];
    __shared__ int home_cnt[NB_HOME];
    if (counters[C_OVF]) return;
    const int c = xcd_block(blockIdx.x, gridDim.x);
    if (c >= g.ncells) return;
    const int nhome = cell_nlocal[c];
    if (nhome == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) {
        int off = 0;
        for (int s = 0; s < 28; s++) {
            int nc = s == 0 ? 0 : c + g.stencil[s - 1];
            int ok = s == 0 || (s - 1 < g.nstencil && nc > 0 && nc < g.ncells);
            int len = ok ? cell_start[nc + 1] - cell_start[nc] : 0;
            seg_begin[s] = ok ? cell_start[nc] : 0;
            seg_off[s] = off;
            off += len;
        }
        seg_off[28] = off;
    }
    __syncthreads();
    const int total = seg_off[28];
    const int home0 = cell_start[c];
    for (int hb = 0; hb < nhome; hb += NB_HOME) {
        const int nh = min(NB_HOME, nhome - hb);
        for (int a = threadIdx.x; a < nh; a += 256) home_cnt[a] = 0;
        for (int c0 = 0; c0 < total; c0 += NB_TILE) {
            const int len = min(NB_TILE, total - c0);
            __syncthreads();
            for (int p = threadIdx.x; p < len; p += 256) {
                const int q = c0 + p;
                int s = 0;
                while (seg_off[s + 1] <= q) s++;
                const int gi = seg_begin[s] + (q - seg_off[s]);
                const double4 pp = sorted_pos[gi];
                tx[p] = pp.x; ty[p] = pp.y; tz[p] = pp.z;
                tj[p] = sorted_idx[gi];
            }
            __syncthreads();
            for (int a = wave; a < nh; a += 4) {
                const int i = sorted_idx[home0 + hb + a];
                if (flags[i] & FLAG_FIXED) continue;
                const double4 pi = sorted_pos[home0 + hb + a];
                int cnt = home_cnt[a];
                for (int q0 = 0; q0 < len; q0 += 64) {
                    const int q = q0 + lane;
                    bool acc = false;
                    int j = 0;
                    if (q < len) {
                        j = tj[q];
                        const bool keep = half ? (i < j) : (j != i);
                        const double dx = pi.x - tx[q], dy = pi.y - ty[q], dz = pi.z - tz[q];
                        double rsq = dx * dx + dy * dy;
                        rsq = rsq + dz * dz;
                        acc = keep && rsq < rc2;
                    }
                    const unsigned long long m = __ballot(acc);
                    const int k = cnt + lane_rank(m);
                    if (acc && k < ncap) nl[(size_t) k * nl_stride + i] = j;
                    cnt += __popcll(m);
                }
                if (lane == 0) home_cnt[a] = cnt;
            }
        }
        __syncthreads();
        for (int a = threadIdx.x; a < nh; a += 256) {
            const int i = sorted_idx[home0 + hb + a];
            int K = home_cnt[a];
            if (flags[i] & FLAG_FIXED) K = 0;
            // reference overflow rule (generated resize check): count+1 >= capacity
            if (K + 1 >= ncap) {
                atomicMax(ovf, K + 2);
                K = ncap - 2;
            }
            numneighs[i] = K;
        }
    }
}

// ---------------------------------------------------------------- Lennard-Jones

struct LJTab {
    int nt;
    double eps[16], sig6[16];  // ntypes <= 4
};

__device__ __forceinline__ void lj_pair(const double4 &pi, const double4 &pj, double rc2, const LJTab &tab, int ti,
                                        double &t0, double &t1, double &t2, double &u, double &w, bool energy,
                                        double &g0, double &g1, double &g2, bool &inside) {
    // examples/md.py:5-8 as emitted: delta, rsq = dx*dx + dy*dy + dz*dz,
    // sr2 = 1/rsq, sr6 = sr2*sr2*sr2*sigma6, s = 48*sr6*(sr6-0.5)*sr2*eps
    const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    double rsq = dx * dx + dy * dy;
    rsq = rsq + dz * dz;
    inside = rsq < rc2;
    if (inside) {
        double sg6, ep;
        if (tab.nt == 1) { sg6 = tab.sig6[0]; ep = tab.eps[0]; }
        else { const int k = ti * tab.nt + pos_type(pj); sg6 = tab.sig6[k]; ep = tab.eps[k]; }
        const double sr2 = 1.0 / rsq;
        double sr6 = sr2 * sr2;
        sr6 = sr6 * sr2;
        sr6 = sr6 * sg6;
        double s = 48.0 * sr6;
        s = s * (sr6 - 0.5);
        s = s * sr2;
        s = s * ep;
        g0 = dx * s; g1 = dy * s; g2 = dz * s;
        t0 = t0 + g0; t1 = t1 + g1; t2 = t2 + g2;
        if (energy) {
            u += 4.0 * ep * sr6 * (sr6 - 1.0);
            w += 48.0 * sr6 * (sr6 - 0.5) * ep;
        }
    }
}

// Full-list LJ force, optionally fused with the integrator around it:
//   F(ts); [final_integrate(ts)]; [thermo partial]; [initial_integrate(ts+1)]
// Positions are double-buffered: the kernel reads P (locals + ghosts of ts)
// and writes the locals of ts+1 into Pn, so no thread ever overwrites a
// position another thread may still gather.
// One lane per particle; neighbour indices stream from the SoA list
// (nl[k*stride + i]: 64 consecutive lanes read 256 contiguous bytes) and the
// loop is unrolled U-deep so that U index loads and U position gathers are in
// flight per lane while the partial sums still add up in list order.
template <int U, bool FUSE, bool RED>
__global__ __launch_bounds__(256) void k_lj_full(const double4 *__restrict__ P, double4 *__restrict__ Pn,
                                                 const int *__restrict__ nl, int nl_stride,
                                                 const int *__restrict__ numneighs, const int *__restrict__ flags,
                                                 double *__restrict__ fx, double *__restrict__ fy,
                                                 double *__restrict__ fz, double *__restrict__ vx,
                                                 double *__restrict__ vy, double *__restrict__ vz,
                                                 const double *__restrict__ mass, int nlocal, double rc2, LJTab tab,
                                                 double dt, double hdt, int do_vv2, int do_vv1,
                                                 double *__restrict__ partials) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int i = b * 256 + threadIdx.x;
    double ke = 0.0, pe = 0.0, vir = 0.0;
    if (i < nlocal) {
        const double4 pi = P[i];
        const int fl = flags[i];
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
        double u = 0.0, w = 0.0, g0, g1, g2;
        bool ins;
        if (!(fl & FLAG_FIXED)) {
            const int ti = pos_type(pi);
            const int n = numneighs[i];
            const int *lp = nl + i;
            int k = 0;
            for (; k + U <= n; k += U) {
                int jj[U];
                double4 pj[U];
#pragma unroll
                for (int q = 0; q < U; q++) jj[q] = __builtin_nontemporal_load(lp + (size_t) (k + q) * nl_stride);
#pragma unroll
                for (int q = 0; q < U; q++) pj[q] = P[jj[q]];
#pragma unroll
                for (int q = 0; q < U; q++) lj_pair(pi, pj[q], rc2, tab, ti, t0, t1, t2, u, w, RED, g0, g1, g2, ins);
            }
            for (; k < n; k++) {
                const int j = lp[(size_t) k * nl_stride];
                lj_pair(pi, P[j], rc2, tab, ti, t0, t1, t2, u, w, RED, g0, g1, g2, ins);
            }
        }
        // reset_volatile_properties then force[i] = force[i] + reduction
        const double f0 = 0.0 + t0, f1 = 0.0 + t1, f2 = 0.0 + t2;
        fx[i] = f0; fy[i] = f1; fz[i] = f2;
        if (RED) { pe = 0.5 * u; vir = 0.5 * w; }
        if (FUSE) {
            double4 pn = pi;
            if (!(fl & FLAG_FIXED)) {
                const double m = mass[i];
                double v0 = vx[i], v1 = vy[i], v2 = vz[i];
                if (do_vv2) {  // final_integrate(ts)
                    double a0 = hdt * f0, a1 = hdt * f1, a2 = hdt * f2;
                    a0 = a0 / m; a1 = a1 / m; a2 = a2 / m;
                    v0 = v0 + a0; v1 = v1 + a1; v2 = v2 + a2;
                }
                if (RED) ke = m * (v0 * v0 + v1 * v1 + v2 * v2);
                if (do_vv1) {  // initial_integrate(ts+1)
                    double a0 = hdt * f0, a1 = hdt * f1, a2 = hdt * f2;
                    a0 = a0 / m; a1 = a1 / m; a2 = a2 / m;
                    v0 = v0 + a0; v1 = v1 + a1; v2 = v2 + a2;
                    pn.x = pi.x + dt * v0; pn.y = pi.y + dt * v1; pn.z = pi.z + dt * v2;
                }
                vx[i] = v0; vy[i] = v1; vz[i] = v2;
            } else if (RED) {
                const double v0 = vx[i], v1 = vy[i], v2 = vz[i];
                ke = mass[i] * (v0 * v0 + v1 * v1 + v2 * v2);
            }
            if (do_vv1) Pn[i] = pn;
        }
    }
    if (RED) {
        __shared__ double red[3][4];
        const double a = wave_sum(ke), bb = wave_sum(pe), cc = wave_sum(vir);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) { red[0][wave] = a; red[1][wave] = bb; red[2][wave] = cc; }
        __syncthreads();
        if (threadIdx.x < 3) {
            const int r = threadIdx.x;
            partials[(size_t) b * 3 + r] = ((red[r][0] + red[r][1]) + red[r][2]) + red[r][3];
        }
    }
}

// Half-list LJ (compute_half): i accumulates its own sum; the mirror force on a
// local, non-fixed j is an FP64 atomic subtract (src/pairs/ir/apply.py:111-125).
// Forces must be zeroed before the launch (reset_volatile_properties).
template <int U, bool RED>
__global__ __launch_bounds__(256) void k_lj_half(const double4 *__restrict__ P, const int *__restrict__ nl,
                                                 int nl_stride, const int *__restrict__ numneighs,
                                                 const int *__restrict__ flags, double *__restrict__ fx,
                                                 double *__restrict__ fy, double *__restrict__ fz, int nlocal,
                                                 double rc2, LJTab tab, double *__restrict__ partials) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int i = b * 256 + threadIdx.x;
    double pe = 0.0, vir = 0.0;
    if (i < nlocal && !(flags[i] & FLAG_FIXED)) {
        const double4 pi = P[i];
        const int ti = pos_type(pi);
        const int n = numneighs[i];
        const int *lp = nl + i;
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, u = 0.0, w = 0.0, g0, g1, g2;
        bool ins;
        int k = 0;
        for (; k + U <= n; k += U) {
            int jj[U];
            double4 pj[U];
#pragma unroll
            for (int q = 0; q < U; q++) jj[q] = __builtin_nontemporal_load(lp + (size_t) (k + q) * nl_stride);
#pragma unroll
            for (int q = 0; q < U; q++) pj[q] = P[jj[q]];
#pragma unroll
            for (int q = 0; q < U; q++) {
                double uu = 0.0, ww = 0.0;
                lj_pair(pi, pj[q], rc2, tab, ti, t0, t1, t2, uu, ww, RED, g0, g1, g2, ins);
                if (ins) {
                    const int j = jj[q];
                    const bool loc = j < nlocal;
                    if (RED) { const double wt = loc ? 1.0 : 0.5; u += wt * uu; w += wt * ww; }
                    if (loc && !(flags[j] & FLAG_FIXED)) {
                        atomic_add_f64(fx + j, -g0); atomic_add_f64(fy + j, -g1); atomic_add_f64(fz + j, -g2);
                    }
                }
            }
        }
        for (; k < n; k++) {
            const int j = lp[(size_t) k * nl_stride];
            double uu = 0.0, ww = 0.0;
            lj_pair(pi, P[j], rc2, tab, ti, t0, t1, t2, uu, ww, RED, g0, g1, g2, ins);
            if (ins) {
                const bool loc = j < nlocal;
                if (RED) { const double wt = loc ? 1.0 : 0.5; u += wt * uu; w += wt * ww; }
                if (loc && !(flags[j] & FLAG_FIXED)) {
                    atomic_add_f64(fx + j, -g0); atomic_add_f64(fy + j, -g1); atomic_add_f64(fz + j, -g2);
                }
            }
        }
        atomic_add_f64(fx + i, t0); atomic_add_f64(fy + i, t1); atomic_add_f64(fz + i, t2);
        pe = u; vir = w;
    }
    if (RED) {
        __shared__ double red[2][4];
        const double a = wave_sum(pe), bb = wave_sum(vir);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) { red[0][wave] = a; red[1][wave] = bb; }
        __syncthreads();
        if (threadIdx.x < 2) {
            const int r = threadIdx.x;
            partials[(size_t) b * 3 + 1 + r] = ((red[r][0] + red[r][1]) + red[r][2]) + red[r][3];
        }
    }
}

// Potential energy and virial over the current list (no reference counterpart):
// U = sum 4 eps sr6 (sr6 - 1), W = sum r.F over pairs r < rc, a pair seen from
// both sides weighted 1/2; e_i = 1/2 sum_j u_ij (per-atom, full lists).
__global__ __launch_bounds__(256) void k_energy(const double4 *__restrict__ P, const int *__restrict__ nl,
                                                int nl_stride, const int *__restrict__ numneighs,
                                                const int *__restrict__ flags, int nlocal, int half, double rc2,
                                                LJTab tab, double *__restrict__ partials, double *__restrict__ eatom) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    double pe = 0.0, vir = 0.0;
    if (i < nlocal) {
        double ei = 0.0;
        if (!(flags[i] & FLAG_FIXED)) {
            const double4 pi = P[i];
            const int ti = pos_type(pi);
            const int n = numneighs[i];
            for (int k = 0; k < n; k++) {
                const int j = nl[(size_t) k * nl_stride + i];
                const double4 pj = P[j];
                const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                const double rsq = dx * dx + dy * dy + dz * dz;
                if (rsq >= rc2) continue;
                const int t = tab.nt == 1 ? 0 : ti * tab.nt + pos_type(pj);
                const double sr2 = 1.0 / rsq, sr6 = sr2 * sr2 * sr2 * tab.sig6[t];
                const double u = 4.0 * tab.eps[t] * sr6 * (sr6 - 1.0);
                const double w = 48.0 * sr6 * (sr6 - 0.5) * tab.eps[t];
                const double wt = (!half || j >= nlocal) ? 0.5 : 1.0;
                pe += wt * u;
                vir += wt * w;
                ei += 0.5 * u;
            }
        }
        if (eatom) eatom[i] = ei;
    }
    __shared__ double red[2][4];
    const double a = wave_sum(pe), bb = wave_sum(vir);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { red[0][wave] = a; red[1][wave] = bb; }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int r = threadIdx.x;
        partials[(size_t) blockIdx.x * 3 + 1 + r] = ((red[r][0] + red[r][1]) + red[r][2]) + red[r][3];
    }
}

// ---------------------------------------------------------------- integration

// initial_integrate (examples/md.py:11-13): v += (dt*0.5)*f/m; x += dt*v
__global__ __launch_bounds__(256) void k_vv_initial(double4 *__restrict__ P, double *__restrict__ vx,
                                                    double *__restrict__ vy, double *__restrict__ vz,
                                                    const double *__restrict__ fx, const double *__restrict__ fy,
                                                    const double *__restrict__ fz, const double *__restrict__ mass,
                                                    const int *__restrict__ flags, int nlocal, double dt, double hdt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nlocal || (flags[i] & FLAG_FIXED)) return;
    const double m = mass[i];
    double a0 = hdt * fx[i], a1 = hdt * fy[i], a2 = hdt * fz[i];
    a0 = a0 / m; a1 = a1 / m; a2 = a2 / m;
    const double v0 = vx[i] + a0, v1 = vy[i] + a1, v2 = vz[i] + a2;
    vx[i] = v0; vy[i] = v1; vz[i] = v2;
    double4 p = P[i];
    p.x = p.x + dt * v0; p.y = p.y + dt * v1; p.z = p.z + dt * v2;
    P[i] = p;
}

// final_integrate (examples/md.py:16-17), optionally followed by the thermo
// partial and the next step's initial_integrate written into Pn.
template <bool RED>
__global__ __launch_bounds__(256) void k_vv_post(const double4 *__restrict__ P, double4 *__restrict__ Pn,
                                                 double *__restrict__ vx, double *__restrict__ vy,
                                                 double *__restrict__ vz, const double *__restrict__ fx,
                                                 const double *__restrict__ fy, const double *__restrict__ fz,
                                                 const double *__restrict__ mass, const int *__restrict__ flags,
                                                 int nlocal, double dt, double hdt, int do_vv2, int do_vv1,
                                                 double *__restrict__ partials) {
    const int b = blockIdx.x;
    const int i = b * 256 + threadIdx.x;
    double ke = 0.0;
    if (i < nlocal) {
        const int fl = flags[i];
        const double m = mass[i];
        double v0 = vx[i], v1 = vy[i], v2 = vz[i];
        double4 p = P[i];
        if (!(fl & FLAG_FIXED)) {
            const double f0 = fx[i], f1 = fy[i], f2 = fz[i];
            if (do_vv2) {
                double a0 = hdt * f0, a1 = hdt * f1, a2 = hdt * f2;
                a0 = a0 / m; a1 = a1 / m; a2 = a2 / m;
                v0 = v0 + a0; v1 = v1 + a1; v2 = v2 + a2;
            }
            if (RED) ke = m * (v0 * v0 + v1 * v1 + v2 * v2);
            if (do_vv1) {
                double a0 = hdt * f0, a1 = hdt * f1, a2 = hdt * f2;
                a0 = a0 / m; a1 = a1 / m; a2 = a2 / m;
                v0 = v0 + a0; v1 = v1 + a1; v2 = v2 + a2;
                p.x = p.x + dt * v0; p.y = p.y + dt * v1; p.z = p.z + dt * v2;
            }
            vx[i] = v0; vy[i] = v1; vz[i] = v2;
        } else if (RED) {
            ke = m * (v0 * v0 + v1 * v1 + v2 * v2);
        }
        if (do_vv1) Pn[i] = p;
    }
    if (RED) {
        __shared__ double red[4];
        const double a = wave_sum(ke);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
        __syncthreads();
        if (threadIdx.x == 0) partials[(size_t) b * 3] = ((red[0] + red[1]) + red[2]) + red[3];
    }
}

// kinetic partial only (stand-alone compute_thermo)
__global__ __launch_bounds__(256) void k_ke_partial(const double *__restrict__ vx, const double *__restrict__ vy,
                                                    const double *__restrict__ vz, const double *__restrict__ mass,
                                                    int nlocal, double *__restrict__ partials) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    double ke = 0.0;
    if (i < nlocal) {
        const double v0 = vx[i], v1 = vy[i], v2 = vz[i];
        ke = mass[i] * (v0 * v0 + v1 * v1 + v2 * v2);
    }
    __shared__ double red[4];
    const double a = wave_sum(ke);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) partials[(size_t) blockIdx.x * 3] = ((red[0] + red[1]) + red[2]) + red[3];
}

// fixed-order sum of the per-block partials; thermo as runtime/thermo.hpp:11-51
__global__ __launch_bounds__(256) void k_thermo_reduce(const double *__restrict__ partials, int nblocks,
                                                       int which_mask, int natoms, double lx, double ly, double lz,
                                                       double *__restrict__ out) {
    __shared__ double red[3][4];
    double s[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += 256)
        for (int r = 0; r < 3; r++) s[r] += partials[(size_t) b * 3 + r];
    for (int r = 0; r < 3; r++) {
        const double a = wave_sum(s[r]);
        if ((threadIdx.x & 63) == 0) red[r][threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double ke = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
        const double pe = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
        const double w = ((red[2][0] + red[2][1]) + red[2][2]) + red[2][3];
        const double dof = (natoms * 3 - 3);
        const double t_scale = 1.0 / dof;
        const double t = ke * t_scale;
        const double p_scale = 1.0 / 3 / lx / ly / lz;
        const double p = (t * dof) * p_scale;
        if (which_mask & 1) { out[0] = t; out[1] = p; }
        if (which_mask & 2) { out[2] = pe; out[3] = w; out[4] = p + w / (3.0 * lx * ly * lz); }
    }
}

// ---------------------------------------------------------------- ghosts / exchange

// Selection predicate of DimensionRanges.ghost_particles
// (src/pairs/sim/domain_partitioning.py:31-65): side 0 "x_d < min_d + offset"
// (shift +L), side 1 "x_d > max_d - offset" (shift -L), skipping
// Infinite|Global particles.  offset 0 selects leavers (exchange).
__device__ __forceinline__ void ghost_sel(const double4 &p, int fl, int dim, const Geom &g, double offset, bool exch,
                                          bool &s0, bool &s1) {
    const double x = dim == 0 ? p.x : (dim == 1 ? p.y : p.z);
    const bool ok = !(fl & (FLAG_INFINITE | FLAG_GLOBAL));
    if (exch) { s0 = ok && x < g.lo[dim]; s1 = ok && x > g.hi[dim]; }
    else { s0 = ok && x < g.lo[dim] + offset; s1 = ok && x > g.hi[dim] - offset; }
}

// pass 1 of the stable two-sided compaction: per-tile counts
__global__ __launch_bounds__(256) void k_sel_count(const double4 *__restrict__ P, const int *__restrict__ flags,
                                                   int *__restrict__ counters, Geom g, int dim, int exch,
                                                   int *__restrict__ tile_cnt, int *__restrict__ exchg_flag) {
    __shared__ int wc[2][4];
    if (counters[C_OVF]) return;
    const int n = exch ? counters[C_NLOCAL] : counters[C_NLOCAL] + counters[C_NGHOST];
    const int ntiles = (n + 255) / 256;
    if (blockIdx.x == 0 && threadIdx.x == 0) counters[C_RANGE] = n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int i = t * 256 + threadIdx.x;
        bool s0 = false, s1 = false;
        if (i < n) ghost_sel(P[i], flags[i], dim, g, g.spacing, exch, s0, s1);
        if (exch && i < n) exchg_flag[i] = (s0 || s1) ? 1 : 0;
        const unsigned long long m0 = __ballot(s0), m1 = __ballot(s1);
        if (lane == 0) { wc[0][wave] = __popcll(m0); wc[1][wave] = __popcll(m1); }
        __syncthreads();
        if (threadIdx.x == 0) {
            tile_cnt[2 * t] = wc[0][0] + wc[0][1] + wc[0][2] + wc[0][3];
            tile_cnt[2 * t + 1] = wc[1][0] + wc[1][1] + wc[1][2] + wc[1][3];
        }
        __syncthreads();
    }
}

// pass 2: exclusive scan of the tile counts (side 0 entries precede side 1),
// counts per direction, capacity checks.
__global__ __launch_bounds__(1024) void k_sel_scan(int *__restrict__ counters, int *__restrict__ tile_cnt, int dim,
                                                   int exch, int send_cap, int cap) {
    __shared__ int lds[20];
    if (counters[C_OVF]) return;
    const int n = counters[C_RANGE];
    const int ntiles = (n + 255) / 256;
    const int base = exch ? 0 : counters[C_NSEND_ALL];
    int carry[2] = {0, 0};
    for (int side = 0; side < 2; side++) {
        for (int t0 = 0; t0 < ntiles; t0 += 1024) {
            const int t = t0 + threadIdx.x;
            const int v = t < ntiles ? tile_cnt[2 * t + side] : 0;
            int total;
            const int ex = block_exclusive_scan<1024>(v, lds, &total);
            if (t < ntiles) tile_cnt[2 * t + side] = carry[side] + ex;
            carry[side] += total;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int n0 = carry[0], n1 = carry[1];
        counters[C_BASE0] = base;
        counters[C_BASE1] = base + n0;
        counters[C_NSEND + 2 * dim] = n0;
        counters[C_NSEND + 2 * dim + 1] = n1;
        const int nall = base + n0 + n1;
        counters[C_NSEND_ALL] = nall;
        if (nall > send_cap) {
            counters[C_OVF] = OVF_SEND;
            counters[C_REQ_SEND] = max(counters[C_REQ_SEND], nall);
        }
        if (!exch) {
            const int nghost = counters[C_NGHOST] + n0 + n1;
            if (counters[C_NLOCAL] + nghost > cap) {
                counters[C_OVF] = OVF_PART;
                counters[C_REQ_PART] = max(counters[C_REQ_PART], counters[C_NLOCAL] + nghost);
            } else {
                counters[C_NGHOST] = nghost;
            }
        }
    }
}

// pass 3: stable scatter into the send map; for borders the ghost is created
// in the same pass (pack + communicateData self-copy + unpack of comm.py:291-366:
// uid, type, mass, position + shift, velocity, shape; flags are not sent).
__global__ __launch_bounds__(256) void k_sel_scatter(double4 *__restrict__ P, double *__restrict__ vx,
                                                     double *__restrict__ vy, double *__restrict__ vz,
                                                     double *__restrict__ mass, int *__restrict__ uid,
                                                     int *__restrict__ shape, const int *__restrict__ flags,
                                                     const int *__restrict__ counters, Geom g, int dim, int exch,
                                                     const int *__restrict__ tile_off, int *__restrict__ send_map,
                                                     int *__restrict__ send_mult) {
    __shared__ int wc[2][5];
    if (counters[C_OVF]) return;
    const int n = counters[C_RANGE];
    const int ntiles = (n + 255) / 256;
    const int nlocal = counters[C_NLOCAL];
    const int base1 = counters[C_BASE1] - counters[C_BASE0];  // side-1 block offsets are relative
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int i = t * 256 + threadIdx.x;
        bool s0 = false, s1 = false;
        double4 p = make_double4(0, 0, 0, 0);
        if (i < n) { p = P[i]; ghost_sel(p, flags[i], dim, g, g.spacing, exch, s0, s1); }
        const unsigned long long m0 = __ballot(s0), m1 = __ballot(s1);
        if (lane == 0) { wc[0][wave] = __popcll(m0); wc[1][wave] = __popcll(m1); }
        __syncthreads();
        if (threadIdx.x == 0) {
            int a = 0, b = 0;
            for (int w = 0; w < 4; w++) { int x = wc[0][w]; wc[0][w] = a; a += x; int y = wc[1][w]; wc[1][w] = b; b += y; }
        }
        __syncthreads();
        if (s0 || s1) {
            const int side = s0 ? 0 : 1;
            const int r = side == 0 ? lane_rank(m0) : lane_rank(m1);
            const int pos = counters[C_BASE0] + (side == 0 ? tile_off[2 * t] : base1 + tile_off[2 * t + 1]) +
                            wc[side][wave] + r;
            const int mult = side == 0 ? 1 : -1;  // pbc_prev = +1, pbc_next = -1 on one rank
            send_map[pos] = i;
            send_mult[3 * pos + 0] = dim == 0 ? mult : 0;
            send_mult[3 * pos + 1] = dim == 1 ? mult : 0;
            send_mult[3 * pos + 2] = dim == 2 ? mult : 0;
            if (!exch) {
                const int gslot = nlocal + pos;
                double4 q = p;
                q.x = p.x + (double) send_mult[3 * pos + 0] * g.len[0];
                q.y = p.y + (double) send_mult[3 * pos + 1] * g.len[1];
                q.z = p.z + (double) send_mult[3 * pos + 2] * g.len[2];
                P[gslot] = q;
                vx[gslot] = vx[i]; vy[gslot] = vy[i]; vz[gslot] = vz[i];
                mass[gslot] = mass[i];
                uid[gslot] = uid[i];
                shape[gslot] = shape[i];
            }
        }
        __syncthreads();
    }
}

// Exchange of one dimension after its send map is built: pack (with shift),
// remove_exchanged_particles_pt1/pt2 (tail fill, copy guard "src > 0"), unpack
// appended at nlocal - n (src/pairs/sim/comm.py:100-151, 445-520).  The
// sequential tail walk of pt1 is the r-th non-exchanged index below nlocal
// (descending) for the r-th exchanged particle that lies below nlocal - n.
// One workgroup: a rebuild moves few particles.
__global__ __launch_bounds__(1024) void k_exchange_fix(double4 *__restrict__ P, double *__restrict__ vx,
                                                       double *__restrict__ vy, double *__restrict__ vz,
                                                       double *__restrict__ mass, int *__restrict__ uid,
                                                       int *__restrict__ shape, int *__restrict__ flags,
                                                       const int *__restrict__ counters, Geom g,
                                                       const int *__restrict__ send_map,
                                                       const int *__restrict__ send_mult,
                                                       const int *__restrict__ exchg_flag, double *__restrict__ buf,
                                                       int *__restrict__ tail, int *__restrict__ copy_to) {
    __shared__ int lds[20];
    if (counters[C_OVF]) return;
    const int n = counters[C_NSEND_ALL];
    if (n == 0) return;
    const int nlocal = counters[C_NLOCAL];
    // pack
    for (int e = threadIdx.x; e < n; e += 1024) {
        const int m = send_map[e];
        double4 p = P[m];
        p.x = p.x + (double) send_mult[3 * e + 0] * g.len[0];
        p.y = p.y + (double) send_mult[3 * e + 1] * g.len[1];
        p.z = p.z + (double) send_mult[3 * e + 2] * g.len[2];
        double *b = buf + (size_t) e * 10;
        b[0] = p.x; b[1] = p.y; b[2] = p.z; b[3] = p.w;
        b[4] = vx[m]; b[5] = vy[m]; b[6] = vz[m]; b[7] = mass[m];
        int *bi = (int *) (b + 8);
        bi[0] = uid[m]; bi[1] = shape[m]; bi[2] = flags[m]; bi[3] = 0;
    }
    // tail list: non-exchanged indices nlocal-1, nlocal-2, ... >= nlocal-n
    int carry = 0;
    for (int t0 = 0; t0 < n; t0 += 1024) {
        const int t = t0 + threadIdx.x;
        const int idx = nlocal - 1 - t;
        const int keep = (t < n && exchg_flag[idx] == 0) ? 1 : 0;
        int total;
        const int ex = block_exclusive_scan<1024>(keep, lds, &total);
        if (keep) tail[carry + ex] = idx;
        carry += total;
    }
    __syncthreads();
    carry = 0;
    for (int e0 = 0; e0 < n; e0 += 1024) {
        const int e = e0 + threadIdx.x;
        const int q = (e < n && send_map[e] < nlocal - n) ? 1 : 0;
        int total;
        const int ex = block_exclusive_scan<1024>(q, lds, &total);
        if (e < n) copy_to[e] = q ? tail[carry + ex] : -1;
        carry += total;
    }
    __syncthreads();
    // pt2: fill the holes from the tail
    for (int e = threadIdx.x; e < n; e += 1024) {
        const int src = copy_to[e];
        if (src > 0) {
            const int dst = send_map[e];
            P[dst] = P[src];
            vx[dst] = vx[src]; vy[dst] = vy[src]; vz[dst] = vz[src];
            mass[dst] = mass[src]; uid[dst] = uid[src]; shape[dst] = shape[src]; flags[dst] = flags[src];
        }
    }
    __syncthreads();
    // unpack at the end of the locals
    for (int e = threadIdx.x; e < n; e += 1024) {
        const int dst = nlocal - n + e;
        const double *b = buf + (size_t) e * 10;
        P[dst] = make_double4(b[0], b[1], b[2], b[3]);
        vx[dst] = b[4]; vy[dst] = b[5]; vz[dst] = b[6]; mass[dst] = b[7];
        const int *bi = (const int *) (b + 8);
        uid[dst] = bi[0]; shape[dst] = bi[1]; flags[dst] = bi[2];
    }
}

// Comm.synchronize with the stale multi-hop semantics of the reference (all
// ghosts are packed before any is unpacked, comm.py:45-54): a ghost whose
// source is local takes the source's current position; a ghost whose source
// is another ghost takes that ghost's previous value.  Double-buffered form:
// Pcur holds this step's locals, Pprev the previous step's ghosts.
__global__ __launch_bounds__(256) void k_sync_pp(double4 *__restrict__ Pcur, const double4 *__restrict__ Pprev,
                                                 const int *__restrict__ counters, const int *__restrict__ send_map,
                                                 const int *__restrict__ send_mult, Geom g) {
    const int n = counters[C_NSEND_ALL];
    const int nlocal = counters[C_NLOCAL];
    for (int e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) {
        const int src = send_map[e];
        double4 p = src < nlocal ? Pcur[src] : Pprev[src];
        p.x = p.x + (double) send_mult[3 * e + 0] * g.len[0];
        p.y = p.y + (double) send_mult[3 * e + 1] * g.len[1];
        p.z = p.z + (double) send_mult[3 * e + 2] * g.len[2];
        Pcur[nlocal + e] = p;
    }
}

// single-buffer form: pack everything into a staging buffer, then unpack
__global__ __launch_bounds__(256) void k_sync_pack(const double4 *__restrict__ P, const double *__restrict__ vx,
                                                   const double *__restrict__ vy, const double *__restrict__ vz,
                                                   const int *__restrict__ counters, const int *__restrict__ send_map,
                                                   const int *__restrict__ send_mult, Geom g,
                                                   double *__restrict__ stage, int with_vel) {
    const int n = counters[C_NSEND_ALL];
    for (int e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) {
        const int src = send_map[e];
        double4 p = P[src];
        double *s = stage + (size_t) e * 8;
        s[0] = p.x + (double) send_mult[3 * e + 0] * g.len[0];
        s[1] = p.y + (double) send_mult[3 * e + 1] * g.len[1];
        s[2] = p.z + (double) send_mult[3 * e + 2] * g.len[2];
        s[3] = p.w;
        if (with_vel) { s[4] = vx[src]; s[5] = vy[src]; s[6] = vz[src]; }
    }
}
__global__ __launch_bounds__(256) void k_sync_unpack(double4 *__restrict__ P, double *__restrict__ vx,
                                                     double *__restrict__ vy, double *__restrict__ vz,
                                                     const int *__restrict__ counters, const double *__restrict__ stage,
                                                     int with_vel) {
    const int n = counters[C_NSEND_ALL];
    const int nlocal = counters[C_NLOCAL];
    for (int e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) {
        const double *s = stage + (size_t) e * 8;
        const int gs = nlocal + e;
        P[gs] = make_double4(s[0], s[1], s[2], s[3]);
        if (with_vel) { vx[gs] = s[4]; vy[gs] = s[5]; vz[gs] = s[6]; }
    }
}

// ---------------------------------------------------------------- launchers

void launch_fill_i32(hipStream_t s, int *p, int v, int n) {
    if (n <= 0) return;
    int blocks = min((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, s, p, v, n);
}
void launch_fill_f64(hipStream_t s, double *p, double v, int n) {
    if (n <= 0) return;
    int blocks = min((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_f64, dim3(blocks), dim3(256), 0, s, p, v, n);
}
void launch_set_counters(hipStream_t s, int *c, int first, int count, int v) {
    hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, s, c, first, count, v);
}

static int tile_grid(int cap) { return max(1, min((cap + 255) / 256, 2048)); }

void launch_bin(hipStream_t s, const DevState &d, const Geom &g) {
    launch_fill_i32(s, d.cell_count, 0, g.ncells + 1);
    const int grid = tile_grid(d.cap);
    hipLaunchKernelGGL(k_bin_count, dim3(grid), dim3(256), 0, s, d.P[d.cur], d.flags, d.counters, g,
                       d.particle_cell, d.cell_count, d.slot);
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, s, d.counters, d.cell_count, d.cell_start, g.ncells);
    hipLaunchKernelGGL(k_bin_scatter, dim3(grid), dim3(256), 0, s, d.counters, d.particle_cell, d.slot,
                       d.cell_start, d.sorted_idx);
    hipLaunchKernelGGL(k_sort_cells, dim3((g.ncells + 255) / 256), dim3(256), 0, s, d.counters, d.cell_start,
                       d.sorted_idx, d.P[d.cur], d.sorted_pos, d.cell_nlocal, g.ncells);
}

void launch_build_neighbors(hipStream_t s, const DevState &d, const Geom &g, int half, double rc2) {
    hipLaunchKernelGGL(k_build_neighbors, dim3(g.ncells), dim3(256), 0, s, d.counters, g, d.cell_start,
                       d.cell_nlocal, d.sorted_idx, d.sorted_pos, d.flags, half, rc2, d.nl, d.nl_stride, d.ncap,
                       d.numneighs, d.counters + C_REQ_NEIGH);
}

void launch_select_dim(hipStream_t s, const DevState &d, const Geom &g, int dim, int exch) {
    const int grid = tile_grid(d.cap);
    double4 *P = d.P[d.cur];
    hipLaunchKernelGGL(k_sel_count, dim3(grid), dim3(256), 0, s, P, d.flags, d.counters, g, dim, exch, d.tile_cnt,
                       d.exchg_flag);
    hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(1024), 0, s, d.counters, d.tile_cnt, dim, exch, d.send_cap, d.cap);
    hipLaunchKernelGGL(k_sel_scatter, dim3(grid), dim3(256), 0, s, P, d.vx, d.vy, d.vz, d.mass, d.uid, d.shape,
                       d.flags, d.counters, g, dim, exch, d.tile_cnt, d.send_map, d.send_mult);
}

void launch_exchange(hipStream_t s, const DevState &d, const Geom &g, const int *periodic) {
    for (int dim = 0; dim < 3; dim++) {
        if (!periodic[dim]) continue;
        // nsend_all = 0; nghost = 0 (comm.py:105-118)
        launch_set_counters(s, d.counters, C_NSEND_ALL, 1, 0);
        launch_set_counters(s, d.counters, C_NGHOST, 1, 0);
        launch_select_dim(s, d, g, dim, 1);
        hipLaunchKernelGGL(k_exchange_fix, dim3(1), dim3(1024), 0, s, d.P[d.cur], d.vx, d.vy, d.vz, d.mass, d.uid,
                           d.shape, d.flags, d.counters, g, d.send_map, d.send_mult, d.exchg_flag, d.stage,
                           d.scratch_i, d.scratch_i + d.send_cap);
    }
}

void launch_borders(hipStream_t s, const DevState &d, const Geom &g, const int *periodic) {
    launch_set_counters(s, d.counters, C_NSEND_ALL, 1, 0);
    launch_set_counters(s, d.counters, C_NGHOST, 1, 0);
    launch_set_counters(s, d.counters, C_NSEND, 6, 0);
    for (int dim = 0; dim < 3; dim++) {
        if (!periodic[dim]) continue;
        launch_select_dim(s, d, g, dim, 0);
    }
}

void launch_sync_pp(hipStream_t s, const DevState &d, const Geom &g, int cur, int prev, int grid) {
    hipLaunchKernelGGL(k_sync_pp, dim3(grid), dim3(256), 0, s, d.P[cur], d.P[prev], d.counters, d.send_map,
                       d.send_mult, g);
}

void launch_sync_inplace(hipStream_t s, const DevState &d, const Geom &g, int grid, int with_vel) {
    hipLaunchKernelGGL(k_sync_pack, dim3(grid), dim3(256), 0, s, d.P[d.cur], d.vx, d.vy, d.vz, d.counters,
                       d.send_map, d.send_mult, g, d.stage, with_vel);
    hipLaunchKernelGGL(k_sync_unpack, dim3(grid), dim3(256), 0, s, d.P[d.cur], d.vx, d.vy, d.vz, d.counters,
                       d.stage, with_vel);
}

static LJTab make_tab(const LJParams &prm) {
    LJTab t;
    t.nt = prm.ntypes;
    for (int k = 0; k < 16; k++) { t.eps[k] = k < prm.ntypes * prm.ntypes ? prm.eps[k] : 0.0; t.sig6[k] = k < prm.ntypes * prm.ntypes ? prm.sig6[k] : 0.0; }
    return t;
}

constexpr int LJ_UNROLL = 8;

void launch_lj_full(hipStream_t s, const DevState &d, const LJParams &prm, int cur, int next, int fuse, int do_vv2,
                    int do_vv1, int reduce) {
    const int nlocal = d.nlocal;
    if (nlocal <= 0) return;
    const int grid = (nlocal + 255) / 256;
    LJTab tab = make_tab(prm);
    const double4 *P = d.P[cur];
    double4 *Pn = d.P[next];
    if (fuse) {
        if (reduce)
            hipLaunchKernelGGL((k_lj_full<LJ_UNROLL, true, true>), dim3(grid), dim3(256), 0, s, P, Pn, d.nl, d.nl_stride,
                               d.numneighs, d.flags, d.fx, d.fy, d.fz, d.vx, d.vy, d.vz, d.mass, nlocal, prm.rc2, tab,
                               prm.dt, prm.hdt, do_vv2, do_vv1, d.partials);
        else
            hipLaunchKernelGGL((k_lj_full<LJ_UNROLL, true, false>), dim3(grid), dim3(256), 0, s, P, Pn, d.nl,
                               d.nl_stride, d.numneighs, d.flags, d.fx, d.fy, d.fz, d.vx, d.vy, d.vz, d.mass, nlocal,
                               prm.rc2, tab, prm.dt, prm.hdt, do_vv2, do_vv1, d.partials);
    } else {
        if (reduce)
            hipLaunchKernelGGL((k_lj_full<LJ_UNROLL, false, true>), dim3(grid), dim3(256), 0, s, P, Pn, d.nl,
                               d.nl_stride, d.numneighs, d.flags, d.fx, d.fy, d.fz, d.vx, d.vy, d.vz, d.mass, nlocal,
                               prm.rc2, tab, prm.dt, prm.hdt, 0, 0, d.partials);
        else
            hipLaunchKernelGGL((k_lj_full<LJ_UNROLL, false, false>), dim3(grid), dim3(256), 0, s, P, Pn, d.nl,
                               d.nl_stride, d.numneighs, d.flags, d.fx, d.fy, d.fz, d.vx, d.vy, d.vz, d.mass, nlocal,
                               prm.rc2, tab, prm.dt, prm.hdt, 0, 0, d.partials);
    }
}

void launch_lj_half(hipStream_t s, const DevState &d, const LJParams &prm, int cur, int reduce) {
    const int nlocal = d.nlocal;
    if (nlocal <= 0) return;
    const int grid = (nlocal + 255) / 256;
    LJTab tab = make_tab(prm);
    launch_fill_f64(s, d.fx, 0.0, nlocal);
    launch_fill_f64(s, d.fy, 0.0, nlocal);
    launch_fill_f64(s, d.fz, 0.0, nlocal);
    if (reduce)
        hipLaunchKernelGGL((k_lj_half<LJ_UNROLL, true>), dim3(grid), dim3(256), 0, s, d.P[cur], d.nl, d.nl_stride,
                           d.numneighs, d.flags, d.fx, d.fy, d.fz, nlocal, prm.rc2, tab, d.partials);
    else
        hipLaunchKernelGGL((k_lj_half<LJ_UNROLL, false>), dim3(grid), dim3(256), 0, s, d.P[cur], d.nl, d.nl_stride,
                           d.numneighs, d.flags, d.fx, d.fy, d.fz, nlocal, prm.rc2, tab, d.partials);
}

void launch_vv_initial(hipStream_t s, const DevState &d, const LJParams &prm) {
    if (d.nlocal <= 0) return;
    hipLaunchKernelGGL(k_vv_initial, dim3((d.nlocal + 255) / 256), dim3(256), 0, s, d.P[d.cur], d.vx, d.vy, d.vz,
                       d.fx, d.fy, d.fz, d.mass, d.flags, d.nlocal, prm.dt, prm.hdt);
}

void launch_vv_post(hipStream_t s, const DevState &d, const LJParams &prm, int cur, int next, int do_vv2, int do_vv1,
                    int reduce) {
    if (d.nlocal <= 0) return;
    const int grid = (d.nlocal + 255) / 256;
    if (reduce)
        hipLaunchKernelGGL(k_vv_post<true>, dim3(grid), dim3(256), 0, s, d.P[cur], d.P[next], d.vx, d.vy, d.vz, d.fx,
                           d.fy, d.fz, d.mass, d.flags, d.nlocal, prm.dt, prm.hdt, do_vv2, do_vv1, d.partials);
    else
        hipLaunchKernelGGL(k_vv_post<false>, dim3(grid), dim3(256), 0, s, d.P[cur], d.P[next], d.vx, d.vy, d.vz, d.fx,
                           d.fy, d.fz, d.mass, d.flags, d.nlocal, prm.dt, prm.hdt, do_vv2, do_vv1, d.partials);
}

void launch_ke_partial(hipStream_t s, const DevState &d) {
    if (d.nlocal <= 0) return;
    hipLaunchKernelGGL(k_ke_partial, dim3((d.nlocal + 255) / 256), dim3(256), 0, s, d.vx, d.vy, d.vz, d.mass,
                       d.nlocal, d.partials);
}

void launch_energy(hipStream_t s, const DevState &d, const LJParams &prm, int half, double *eatom) {
    if (d.nlocal <= 0) return;
    LJTab tab = make_tab(prm);
    hipLaunchKernelGGL(k_energy, dim3((d.nlocal + 255) / 256), dim3(256), 0, s, d.P[d.cur], d.nl, d.nl_stride,
                       d.numneighs, d.flags, d.nlocal, half, prm.rc2, tab, d.partials, eatom);
}

void launch_thermo_reduce(hipStream_t s, const DevState &d, int which_mask, const Geom &g, double *out) {
    const int nb = (d.nlocal + 255) / 256;
    hipLaunchKernelGGL(k_thermo_reduce, dim3(1), dim3(256), 0, s, d.partials, nb, which_mask, d.nlocal, g.len[0],
                       g.len[1], g.len[2], out);
}

}  // namespace pk
