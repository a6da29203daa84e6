// DEM hot path on gfx950: spheres and half-spaces, linear spring-dashpot
// contacts with per-contact tangential history (examples/dem.py).
//
// Storage: structure of arrays in the reference's particle order (slot i is
// the reference's particle i: exchange and borders move particles exactly as
// the generated program does, so every order-dependent rule -- exchange tail
// fill, ghost order, cell order, contact slot order -- is the reference's).
// Contact history is [slot k][particle i] (k-major) so that the lanes of a
// wave, which hold consecutive particles, touch consecutive words.
//
// One step is a fixed sequence of kernels whose grids depend on capacities
// only; every count lives in device memory (DemCounters), so the step is
// captured once as a hipGraph and replayed.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "dem_kernels.hpp"

namespace dk {

// copy table entry blockIdx.y, grid-strided over its rows x words
__global__ __launch_bounds__(256) void k_copy_rows(const CopyDesc *__restrict__ table) {
    const CopyDesc c = table[blockIdx.y];
    const uint32_t *__restrict__ src = (const uint32_t *) c.src;
    uint32_t *__restrict__ dst = (uint32_t *) c.dst;
    const size_t total = (size_t) c.words * (size_t) c.rows;
    for (size_t e = (size_t) blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t) gridDim.x * 256) {
        const size_t r = e / (size_t) c.words, w = e - r * (size_t) c.words;
        dst[r * c.dst_pitch + w] = src[r * c.src_pitch + w];
    }
}

void launch_copy_rows(hipStream_t s, const CopyDesc *table, int ncopies, int n) {
    if (ncopies <= 0 || n <= 0) return;
    // the widest entry: a double array of 9 rows of n particles
    const size_t most = (size_t) 2 * n * 9;
    const int gx = (int) std::min<size_t>((most + 255) / 256, 1024);
    hipLaunchKernelGGL(k_copy_rows, dim3(gx, ncopies), dim3(256), 0, s, table);
}

__global__ void k_fill_i32(int *__restrict__ p, int v, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------- scans

template <int BS>
__device__ __forceinline__ int block_scan_excl(int v, int *lds, int *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < BS / 64; w++) { const int t = lds[w]; lds[w] = acc; acc += t; }
        lds[BS / 64] = acc;
    }
    __syncthreads();
    const int r = lds[wave] + x - v;
    *total = lds[BS / 64];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------- selection (exchange / borders)

// DimensionRanges predicates (src/pairs/sim/domain_partitioning.py:31-65):
// side 0 "x_d < min + offset", side 1 "x_d > max - offset", Infinite and
// Global particles excluded (flags & 9).
__device__ __forceinline__ void dsel(double xd, int fl, double lo, double hi, double off, bool &s0, bool &s1) {
    const bool ok = (fl & (DFLAG_INFINITE | DFLAG_GLOBAL)) == 0;
    s0 = ok && xd < lo + off;
    s1 = ok && xd > hi - off;
}

// per-block counts of side 0 and side 1 over [0, nlocal + nghost); an
// exchange, and the borders of the first periodic dimension, start from no
// ghosts and an empty send map (comm.py:105-118: the counts are reset before
// them), a later dimension selects among the locals and the ghosts so far and
// appends behind them; workgroup 0 latches the range and the base for the
// scatter (nothing writes them while this pass runs); an exchange clears
// exchg_flag of locals
__global__ __launch_bounds__(256) void k_sel_count(const double *__restrict__ Xd, const int *__restrict__ flags,
                                                   int *__restrict__ c, double lo, double hi, double off, int exch,
                                                   int first, int *__restrict__ blk, int *__restrict__ exchg_flag) {
    __shared__ int s0c, s1c;
    const bool fresh = exch || first;
    const int nl = c[D_NLOCAL];
    const int n = nl + (fresh ? 0 : c[D_NGHOST]);
    if (threadIdx.x == 0) { s0c = 0; s1c = 0; }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c[D_RANGE] = n;
        c[D_BASE] = fresh ? 0 : c[D_NSEND_ALL];
    }
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    bool a = false, b = false;
    if (i < n) dsel(Xd[i], flags[i], lo, hi, off, a, b);
    if (exch && i < nl) exchg_flag[i] = 0;
    const unsigned long long ma = __ballot(a), mb = __ballot(b);
    if ((threadIdx.x & 63) == 0) { atomicAdd(&s0c, __popcll(ma)); atomicAdd(&s1c, __popcll(mb)); }
    __syncthreads();
    if (threadIdx.x == 0) { blk[2 * blockIdx.x] = s0c; blk[2 * blockIdx.x + 1] = s1c; }
}

// sums of two per-thread values over a workgroup of 256 threads
__device__ __forceinline__ void block_sum2_256(int a, int b, int *lds, int &sa, int &sb) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    __syncthreads();
    if (lane == 0) { lds[2 * wave] = a; lds[2 * wave + 1] = b; }
    __syncthreads();
    sa = lds[0] + lds[2] + lds[4] + lds[6];
    sb = lds[1] + lds[3] + lds[5] + lds[7];
}

// ordered scatter into the send map (index order within each side, side 0
// entries before side 1); borders also create the ghosts at nlocal + send
// index (uid, type, mass, radius, position + shift, linear and angular
// velocity, shape; flags are not sent, src/pairs/sim/comm.py:61-72).  Every
// workgroup sums the per-block counts of the count pass itself (totals, and
// the counts of the blocks before its own), so the offsets need no pass of
// their own; all workgroups take the same capacity decision, and workgroup 0
// updates the counts (a border dimension's ghosts join the ghost range), which
// no other workgroup of this pass reads.
__global__ __launch_bounds__(256) void k_sel_scatter(DemPtrs p, int *__restrict__ c, int dim, double lo, double hi,
                                                     double off, double len, int exch, const int *__restrict__ blk,
                                                     int pbc0, int pbc1, int send_cap, int cap) {
    __shared__ int wbase[2][4];
    __shared__ int red[8];
    if (c[D_OVF]) return;
    const int n = c[D_RANGE], nl = c[D_NLOCAL];
    const int nblk = (n + 255) / 256;
    if (blockIdx.x > 0 && (int) blockIdx.x >= nblk) return;  // workgroup 0 always writes the counts
    const int base = c[D_BASE];
    int t0 = 0, t1 = 0, q0 = 0, q1 = 0;
    for (int u = threadIdx.x; u < nblk; u += 256) {
        const int x = blk[2 * u], y = blk[2 * u + 1];
        t0 += x; t1 += y;
        if (u < (int) blockIdx.x) { q0 += x; q1 += y; }
    }
    int T0, T1, P0, P1;
    block_sum2_256(t0, t1, red, T0, T1);
    block_sum2_256(q0, q1, red, P0, P1);
    const int need = base + T0 + T1;
    if (need > send_cap || (!exch && nl + need > cap)) {
        // the counts stay within the maps and arrays
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            if (need > send_cap) {
                atomicMax(&c[D_REQ_SEND], need);
                atomicMax(&c[D_OVF], DOVF_SEND);
            } else {
                atomicMax(&c[D_REQ_PART], nl + need);
                atomicMax(&c[D_OVF], DOVF_PART);
            }
        }
        return;
    }
    const double *Xd = dim == 0 ? p.x : (dim == 1 ? p.y : p.z);
    const int i = blockIdx.x * 256 + threadIdx.x;
    bool a = false, b = false;
    if (i < n) dsel(Xd[i], p.flags[i], lo, hi, off, a, b);
    const unsigned long long ma = __ballot(a), mb = __ballot(b);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { wbase[0][wave] = __popcll(ma); wbase[1][wave] = __popcll(mb); }
    __syncthreads();
    int pa = 0, pb = 0;
    for (int w = 0; w < wave; w++) { pa += wbase[0][w]; pb += wbase[1][w]; }
    for (int s = 0; s < 2; s++) {
        const bool sel = s == 0 ? a : b;
        if (!sel) continue;
        const int idx = base + (s == 0 ? P0 + pa + lane_rank(ma) : T0 + P1 + pb + lane_rank(mb));
        const int mult = s == 0 ? pbc0 : pbc1;
        p.send_map[idx] = i;
        p.send_mult[3 * idx + 0] = dim == 0 ? mult : 0;
        p.send_mult[3 * idx + 1] = dim == 1 ? mult : 0;
        p.send_mult[3 * idx + 2] = dim == 2 ? mult : 0;
        if (exch) {
            p.exchg_flag[i] = 1;
        } else {
            const int g = nl + idx;
            // pack: src += send_mult * length (comm.py:312-318)
            const double sh = (double) mult * len;
            p.uid[g] = p.uid[i]; p.type[g] = p.type[i]; p.m[g] = p.m[i]; p.r[g] = p.r[i];
            p.x[g] = dim == 0 ? p.x[i] + sh : p.x[i] + 0 * len;
            p.y[g] = dim == 1 ? p.y[i] + sh : p.y[i] + 0 * len;
            p.z[g] = dim == 2 ? p.z[i] + sh : p.z[i] + 0 * len;
            p.vx[g] = p.vx[i]; p.vy[g] = p.vy[i]; p.vz[g] = p.vz[i];
            p.wx[g] = p.wx[i]; p.wy[g] = p.wy[i]; p.wz[g] = p.wz[i];
            p.shape[g] = p.shape[i];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c[D_NSEND + 2 * dim] = T0;
        c[D_NSEND + 2 * dim + 1] = T1;
        c[D_NSEND_ALL] = need;
        if (!exch) c[D_NGHOST] = need;
    }
}

__global__ void k_reset_comm(int *__restrict__ c) {
    if (threadIdx.x == 0) {
        c[D_NSEND_ALL] = 0;
        c[D_NGHOST] = 0;
        for (int k = 0; k < 6; k++) c[D_NSEND + k] = 0;
    }
}

// ---------------------------------------------------------------- exchange with contact history

// remove_exchanged_particles_pt1 (src/pairs/sim/comm.py:445-470): holes of
// leavers below nlocal - n are filled from the tail, skipping leavers
__device__ __forceinline__ void exch_pt1(const int *__restrict__ c, const int *__restrict__ send_map,
                                         const int *__restrict__ exchg_flag, int *__restrict__ copy_to) {
    const int n = c[D_NSEND_ALL], nl = c[D_NLOCAL];
    int pos = nl - 1;
    for (int i = 0; i < n; i++) {
        if (send_map[i] < nl - n) {
            while (exchg_flag[pos] == 1) pos--;
            copy_to[i] = pos;
            pos--;
        } else {
            copy_to[i] = -1;
        }
    }
}

constexpr int EXB = 40;  // doubles per packed particle (comm.py:61-72 property list of the exchange)

__device__ __forceinline__ void copy_particle(DemPtrs &p, int dst, int src) {
    p.uid[dst] = p.uid[src]; p.shape[dst] = p.shape[src]; p.flags[dst] = p.flags[src];
    p.x[dst] = p.x[src]; p.y[dst] = p.y[src]; p.z[dst] = p.z[src];
    p.m[dst] = p.m[src];
    p.vx[dst] = p.vx[src]; p.vy[dst] = p.vy[src]; p.vz[dst] = p.vz[src];
    p.wx[dst] = p.wx[src]; p.wy[dst] = p.wy[src]; p.wz[dst] = p.wz[src];
    p.r[dst] = p.r[src];
    p.nx[dst] = p.nx[src]; p.ny[dst] = p.ny[src]; p.nz[dst] = p.nz[src];
    for (int k = 0; k < 9; k++) {
        p.inv_i[(size_t) k * p.cap + dst] = p.inv_i[(size_t) k * p.cap + src];
        p.rot_m[(size_t) k * p.cap + dst] = p.rot_m[(size_t) k * p.cap + src];
    }
    for (int k = 0; k < 4; k++) p.rot_q[(size_t) k * p.cap + dst] = p.rot_q[(size_t) k * p.cap + src];
    p.type[dst] = p.type[src];
}

// pack the leavers (all non-volatile properties, position shifted)
__device__ __forceinline__ void exch_pack(DemPtrs &p, const int *__restrict__ c, double lx, double ly, double lz) {
    const int n = c[D_NSEND_ALL];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int m = p.send_map[i];
        double *b = p.buf + (size_t) EXB * i;
        const int *mu = p.send_mult + 3 * i;
        b[0] = p.uid[m]; b[1] = p.shape[m]; b[2] = p.flags[m];
        b[3] = p.x[m] + mu[0] * lx; b[4] = p.y[m] + mu[1] * ly; b[5] = p.z[m] + mu[2] * lz;
        b[6] = p.m[m];
        b[7] = p.vx[m]; b[8] = p.vy[m]; b[9] = p.vz[m];
        b[10] = p.wx[m]; b[11] = p.wy[m]; b[12] = p.wz[m];
        b[13] = p.r[m];
        b[14] = p.nx[m]; b[15] = p.ny[m]; b[16] = p.nz[m];
        for (int k = 0; k < 9; k++) { b[17 + k] = p.inv_i[(size_t) k * p.cap + m]; b[26 + k] = p.rot_m[(size_t) k * p.cap + m]; }
        for (int k = 0; k < 4; k++) b[35 + k] = p.rot_q[(size_t) k * p.cap + m];
        b[39] = p.type[m];
    }
}

// remove_exchanged_particles_pt2 (comm.py:473-506): copy guard "src > 0",
// properties and contact history move from the tail into the hole
__device__ __forceinline__ void exch_pt2(DemPtrs &p, const int *__restrict__ c) {
    const int n = c[D_NSEND_ALL];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int src = p.copy_to[i];
        if (src <= 0) continue;
        const int dst = p.send_map[i];
        copy_particle(p, dst, src);
        const int nc = p.num_contacts[src];
        for (int k = 0; k < nc; k++) {
            const size_t a = (size_t) k * p.cap + dst, b = (size_t) k * p.cap + src;
            p.c_uid[a] = p.c_uid[b];
            p.c_used[a] = p.c_used[b];
            p.c_stick[a] = p.c_stick[b];
            p.c_tsd[3 * (size_t) k * p.cap + dst] = p.c_tsd[3 * (size_t) k * p.cap + src];
            p.c_tsd[(3 * (size_t) k + 1) * p.cap + dst] = p.c_tsd[(3 * (size_t) k + 1) * p.cap + src];
            p.c_tsd[(3 * (size_t) k + 2) * p.cap + dst] = p.c_tsd[(3 * (size_t) k + 2) * p.cap + src];
            p.c_ivm[a] = p.c_ivm[b];
        }
        p.num_contacts[dst] = nc;
    }
}

// unpack the leavers at nlocal - n + i; then pack their contact history from
// the leaver's slot as it is after the hole filling (the reference packs the
// history after pt2); every contact's uid lands in one buffer word, so the
// last contact's uid is the one that travels
__device__ __forceinline__ void exch_unpack(DemPtrs &p, const int *__restrict__ c) {
    const int n = c[D_NSEND_ALL], base = c[D_NLOCAL] - n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double *b = p.buf + (size_t) EXB * i;
        const int g = base + i;
        p.uid[g] = (int) b[0]; p.shape[g] = (int) b[1]; p.flags[g] = (int) b[2];
        p.x[g] = b[3]; p.y[g] = b[4]; p.z[g] = b[5];
        p.m[g] = b[6];
        p.vx[g] = b[7]; p.vy[g] = b[8]; p.vz[g] = b[9];
        p.wx[g] = b[10]; p.wy[g] = b[11]; p.wz[g] = b[12];
        p.r[g] = b[13];
        p.nx[g] = b[14]; p.ny[g] = b[15]; p.nz[g] = b[16];
        for (int k = 0; k < 9; k++) { p.inv_i[(size_t) k * p.cap + g] = b[17 + k]; p.rot_m[(size_t) k * p.cap + g] = b[26 + k]; }
        for (int k = 0; k < 4; k++) p.rot_q[(size_t) k * p.cap + g] = b[35 + k];
        p.type[g] = (int) b[39];
    }
}

__device__ __forceinline__ void hist_pack(DemPtrs &p, const int *__restrict__ c) {
    const int n = c[D_NSEND_ALL];
    const int hb = 2 + 5 * p.ccap;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int m = p.send_map[i];
        double *b = p.hbuf + (size_t) hb * i;
        const int nc = p.num_contacts[m];
        b[0] = nc;
        double last = 0.0;
        for (int k = 0; k < nc; k++) {
            const size_t a = (size_t) k * p.cap + m;
            last = p.c_uid[a];
            b[2 + 5 * k] = p.c_stick[a];
            b[3 + 5 * k] = p.c_tsd[3 * (size_t) k * p.cap + m];
            b[4 + 5 * k] = p.c_tsd[(3 * (size_t) k + 1) * p.cap + m];
            b[5 + 5 * k] = p.c_tsd[(3 * (size_t) k + 2) * p.cap + m];
            b[6 + 5 * k] = p.c_ivm[a];
        }
        b[1] = last;
    }
}

__device__ __forceinline__ void hist_unpack(DemPtrs &p, const int *__restrict__ c) {
    const int n = c[D_NSEND_ALL], base = c[D_NLOCAL] - n;
    const int hb = 2 + 5 * p.ccap;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double *b = p.hbuf + (size_t) hb * i;
        const int g = base + i;
        const int nc = (int) b[0];
        p.num_contacts[g] = nc;
        for (int k = 0; k < nc; k++) {
            const size_t a = (size_t) k * p.cap + g;
            p.c_uid[a] = (int) b[1];
            p.c_used[a] = 1;
            p.c_stick[a] = (int) b[2 + 5 * k];
            p.c_tsd[3 * (size_t) k * p.cap + g] = b[3 + 5 * k];
            p.c_tsd[(3 * (size_t) k + 1) * p.cap + g] = b[4 + 5 * k];
            p.c_tsd[(3 * (size_t) k + 2) * p.cap + g] = b[5 + 5 * k];
            p.c_ivm[a] = b[6 + 5 * k];
        }
    }
}

// The exchange of one dimension after its send map is built, in one
// workgroup (a step moves few particles across a face): the tail fill plan
// (pt1, sequential as the reference's), the leavers packed, the holes filled
// from the tail with their contact history (pt2), the leavers unpacked at
// nlocal - n + i, their history packed after the hole filling and unpacked --
// the reference's order, each phase finished before the next begins.
__global__ __launch_bounds__(1024) void k_exchange(DemPtrs p, const int *__restrict__ c, double lx, double ly,
                                                   double lz) {
    if (c[D_OVF] || c[D_NSEND_ALL] == 0) return;
    if (threadIdx.x == 0) exch_pt1(c, p.send_map, p.exchg_flag, p.copy_to);
    exch_pack(p, c, lx, ly, lz);
    __syncthreads();
    exch_pt2(p, c);
    __syncthreads();
    exch_unpack(p, c);
    __syncthreads();
    hist_pack(p, c);
    __syncthreads();
    hist_unpack(p, c);
}

// ---------------------------------------------------------------- cells

__device__ __forceinline__ int dcell_of(double x, double y, double z, const DemGeom &g) {
    const double xs[3] = {x, y, z};
    int idx = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double q = (xs[d] - (g.lo[d] - g.width)) / g.width;
        int cc = (int) q >= 0 ? (int) q : 0;
        cc = cc < g.dim_cells[d] ? cc : g.dim_cells[d] - 1;
        idx = d == 0 ? cc : idx * g.dim_cells[d] + cc;
    }
    return idx + 1;
}

// BuildCellLists (src/pairs/sim/cell_lists.py:90-134): Infinite -> cell 0
__global__ __launch_bounds__(256) void k_cell_count(DemPtrs p, const int *__restrict__ c, DemGeom g) {
    const int n = c[D_NLOCAL] + c[D_NGHOST];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int cell = (p.flags[i] & DFLAG_INFINITE) ? 0 : dcell_of(p.x[i], p.y[i], p.z[i], g);
    p.particle_cell[i] = cell;
    p.cell_slot[i] = atomicAdd(&p.cell_count[cell], 1);  // its place among the cell's entries
}

// exclusive scan of cell counts: block partials, then each block adds its offset
__global__ __launch_bounds__(1024) void k_scan_partial(const int *__restrict__ in, int *__restrict__ out,
                                                       int *__restrict__ bsum, int n) {
    __shared__ int lds[20];
    const int i = blockIdx.x * 1024 + threadIdx.x;
    int t;
    const int e = block_scan_excl<1024>(i < n ? in[i] : 0, lds, &t);
    if (i < n) out[i] = e;
    if (threadIdx.x == 0) bsum[blockIdx.x] = t;
}

// each block adds the totals of the blocks before it (it sums them itself:
// a few hundred words, so the block offsets need no pass of their own)
__global__ __launch_bounds__(1024) void k_scan_add(int *__restrict__ out, const int *__restrict__ bsum, int n) {
    __shared__ int red[16];
    int q = 0;
    for (int b = threadIdx.x; b < (int) blockIdx.x; b += 1024) q += bsum[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) q += __shfl_xor(q, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = q;
    __syncthreads();
    int off = 0;
#pragma unroll
    for (int w = 0; w < 16; w++) off += red[w];
    const int i = blockIdx.x * 1024 + threadIdx.x;
    if (i < n) out[i] += off;
}

__global__ __launch_bounds__(256) void k_cell_fill(DemPtrs p, const int *__restrict__ c) {
    const int n = c[D_NLOCAL] + c[D_NGHOST];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int cell = p.particle_cell[i];
    p.cell_particles[p.cell_start[cell] + p.cell_slot[i]] = i;
}

// per cell: index order (the CPU's insertion order), then PartitionCellLists
// (cell_lists.py:137-171): in-place two-pointer partition by shape,
// spheres then half-spaces
__global__ __launch_bounds__(256) void k_cell_order(DemPtrs p, int ncells) {
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= ncells) return;
    const int s0 = p.cell_start[cell], cnt = p.cell_count[cell];
    // the counts are not used after this pass: cleared for the next binning
    p.cell_count[cell] = 0;
    int *cp = p.cell_particles + s0;
    for (int a = 1; a < cnt; a++) {
        const int v = cp[a];
        int b = a - 1;
        while (b >= 0 && cp[b] > v) { cp[b + 1] = cp[b]; b--; }
        cp[b + 1] = v;
    }
    int lo = 0;
    for (int sh = 0; sh < 2; sh++) {
        const int first = lo;
        int hi = cnt - 1, ns = 0;
        while (lo <= hi) {
            const int q = cp[lo];
            if (p.shape[q] != sh) {
                if (lo != hi) { cp[lo] = cp[hi]; cp[hi] = q; }
                hi--;
            } else {
                lo++;
                ns = lo - first;
            }
        }
        p.nshapes[2 * cell + sh] = ns;
    }
}

// ---------------------------------------------------------------- contacts

struct Acc3 { double x, y, z; };

// linear_spring_dashpot (examples/dem.py:18-74) for one contact: normal n,
// penetration delta, contact point cp, history slot k; updates the slot and
// returns the contact's force F and torque term T = (cp - x_i) x F
__device__ __forceinline__ void spring_dashpot(const DemPtrs &p, const DemParams &prm, int i, int j, int k,
                                               const double n[3], double delta, const double cp[3], double F[3],
                                               double T[3]) {
    const size_t a = (size_t) k * p.cap + i;
    const bool sticking = p.c_stick[a] == 1;
    const double xi[3] = {p.x[i], p.y[i], p.z[i]}, xj[3] = {p.x[j], p.y[j], p.z[j]};
    const double vi[3] = {p.vx[i], p.vy[i], p.vz[i]}, vj[3] = {p.vx[j], p.vy[j], p.vz[j]};
    const double wi[3] = {p.wx[i], p.wy[i], p.wz[i]}, wj[3] = {p.wx[j], p.wy[j], p.wz[j]};
    double ai[3], aj[3], rel[3], reln[3], relt[3];
    for (int d = 0; d < 3; d++) { ai[d] = cp[d] - xi[d]; aj[d] = cp[d] - xj[d]; }
    double vwi[3], vwj[3];
    vwi[0] = vi[0] + (wi[1] * ai[2] - wi[2] * ai[1]);
    vwi[1] = vi[1] + (wi[2] * ai[0] - wi[0] * ai[2]);
    vwi[2] = vi[2] + (wi[0] * ai[1] - wi[1] * ai[0]);
    vwj[0] = vj[0] + (wj[1] * aj[2] - wj[2] * aj[1]);
    vwj[1] = vj[1] + (wj[2] * aj[0] - wj[0] * aj[2]);
    vwj[2] = vj[2] + (wj[0] * aj[1] - wj[1] * aj[0]);
    for (int d = 0; d < 3; d++) rel[d] = -(vwi[d] - vwj[d]);
    const double rn = (rel[0] * n[0] + rel[1] * n[1]) + rel[2] * n[2];
    for (int d = 0; d < 3; d++) { reln[d] = rn * n[d]; relt[d] = rel[d] - reln[d]; }
    const double relt_len = sqrt((relt[0] * relt[0] + relt[1] * relt[1]) + relt[2] * relt[2]);
    const bool cond_a = sticking && relt_len < 1e-08;
    const double meff = 1.0 / (1.0 / p.m[i] + 1.0 / p.m[j]);
    // stiffness from the collision time (examples/dem.py:34-35), or fixed
    const double kn = prm.kn_fixed > 0.0 ? prm.kn_fixed
                                         : (meff * (prm.pi * prm.pi + prm.ln_e * prm.ln_e)) / (prm.t_c * prm.t_c);
    const double kt = prm.kappa * kn;
    const double h[3] = {p.c_tsd[3 * (size_t) k * p.cap + i], p.c_tsd[(3 * (size_t) k + 1) * p.cap + i],
                         p.c_tsd[(3 * (size_t) k + 2) * p.cap + i]};
    const double hn = (h[0] * n[0] + h[1] * n[1]) + h[2] * n[2];
    double rot[3], ntsd[3];
    for (int d = 0; d < 3; d++) rot[d] = h[d] - n[d] * hn;
    const double rot2 = (rot[0] * rot[0] + rot[1] * rot[1]) + rot[2] * rot[2];
    const double h2 = (h[0] * h[0] + h[1] * h[1]) + h[2] * h[2];
    const double scale = sqrt(h2 / rot2);
    for (int d = 0; d < 3; d++) ntsd[d] = prm.dt * relt[d] + ((rot2 <= 0.0) ? 0.0 : rot[d] * scale);
    // damping for the restitution coefficient: from the collision time
    // (examples/dem.py:36), or for the fixed stiffness
    const double dn = prm.kn_fixed > 0.0 ? ((-(2.0) * prm.ln_e) * sqrt(meff * kn)) / prm.sqrt_pl
                                         : ((-(2.0) * meff) * prm.ln_e) / prm.t_c;
    const double dtan = prm.sqrt_kappa * dn;
    double ftls[3];
    for (int d = 0; d < 3; d++) ftls[d] = kt * ntsd[d] + dtan * relt[d];
    const double ftls_len = sqrt((ftls[0] * ftls[0] + ftls[1] * ftls[1]) + ftls[2] * ftls[2]);
    const int tk = p.type[i] * prm.ntypes + p.type[j];
    double fn[3];
    for (int d = 0; d < 3; d++) fn[d] = (kn * delta) * n[d] + dn * reln[d];
    const double fn_len = sqrt((fn[0] * fn[0] + fn[1] * fn[1]) + fn[2] * fn[2]);
    const double f_static = prm.mu_s[tk] * fn_len;
    const double f_dynamic = prm.mu_d[tk] * fn_len;
    const bool cond1 = cond_a && ftls_len < f_static;
    const bool cond2 = sticking && ftls_len < f_dynamic;
    const bool slide = (!cond1 && !cond2) && kt > 0.0;
    const double f_fric = cond1 ? f_static : f_dynamic;
    const double inv_len = 1.0 / ftls_len;
    double tdir[3];
    for (int d = 0; d < 3; d++) tdir[d] = (ftls_len > 0.0) ? ftls[d] * inv_len : 0.0;
    for (int d = 0; d < 3; d++)
        p.c_tsd[(3 * (size_t) k + d) * p.cap + i] = slide ? (f_fric * tdir[d] - dtan * relt[d]) / kt : ntsd[d];
    const double rel_len = sqrt((rel[0] * rel[0] + rel[1] * rel[1]) + rel[2] * rel[2]);
    const double iv = p.c_ivm[a];
    p.c_ivm[a] = (iv > 0.0) ? iv : rel_len;
    p.c_stick[a] = ((cond1 || cond2) || ftls_len < f_dynamic) ? 1 : 0;
    const double ft_abs = (f_fric < ftls_len) ? f_fric : ftls_len;
    for (int d = 0; d < 3; d++) F[d] = fn[d] + ft_abs * tdir[d];
    const double c0 = cp[0] - xi[0], c1 = cp[1] - xi[1], c2 = cp[2] - xi[2];
    T[0] = c1 * F[2] - c2 * F[1];
    T[1] = c2 * F[0] - c0 * F[2];
    T[2] = c0 * F[1] - c1 * F[0];
}

// force += F, torque += T of one contact, in the caller's contact order
__device__ __forceinline__ void accumulate(Acc3 &fa, Acc3 &ta, const double F[3], const double T[3]) {
    fa.x = fa.x + F[0]; fa.y = fa.y + F[1]; fa.z = fa.z + F[2];
    ta.x = ta.x + T[0]; ta.y = ta.y + T[1]; ta.z = ta.z + T[2];
}

// contact slot of uid_j in i's history: the last match, or a new slot with
// zeroed history (src/pairs/mapping/funcs.py:230-270); marks it used
__device__ __forceinline__ int contact_slot(const DemPtrs &p, int *c, int i, int uid_j) {
    const int nc = p.num_contacts[i];
    int slot = -1;
    for (int k = 0; k < nc; k++)
        if (p.c_uid[(size_t) k * p.cap + i] == uid_j) slot = k;
    if (slot == -1) {
        if (nc >= p.ccap) {
            atomicMax(&c[D_REQ_CONTACT], nc + 1);
            atomicMax(&c[D_OVF], DOVF_CONTACT);
            return -1;
        }
        slot = nc;
        p.num_contacts[i] = nc + 1;
        const size_t a = (size_t) slot * p.cap + i;
        p.c_uid[a] = uid_j;
        p.c_stick[a] = 0;
        p.c_tsd[3 * (size_t) slot * p.cap + i] = 0.0;
        p.c_tsd[(3 * (size_t) slot + 1) * p.cap + i] = 0.0;
        p.c_tsd[(3 * (size_t) slot + 2) * p.cap + i] = 0.0;
        p.c_ivm[a] = 0.0;
    }
    p.c_used[(size_t) slot * p.cap + i] = 1;
    return slot;
}

// contact test of i (position xi, radius ri) against candidate j of the given
// pass (0: sphere, 1: half-space), as linear_spring_dashpot's guard
// (examples/dem.py:18-30 with the sphere/half-space geometry of
// src/pairs/sim/interaction.py); on contact fills the normal, the contact
// point and the penetration depth
__device__ __forceinline__ bool contact_geom(const DemPtrs &p, const double xi[3], double ri, int j, int pass,
                                             double nv[3], double cpt[3], double &delta) {
    const double xj[3] = {p.x[j], p.y[j], p.z[j]};
    if (pass == 0) {
        const double d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
        const double rsq = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        const double rj = p.r[j];
        const double rs = ri + rj;
        if (!(rsq < rs * (rs + 0.0))) return false;
        const double dist = sqrt(rsq);
        const double pen = (dist - ri) - rj;
        delta = -(pen);
        if (delta < 0.0) return false;
        const double inv = 1.0 / dist;
        for (int e = 0; e < 3; e++) nv[e] = d[e] * inv;
        const double reach = rj + 0.5 * pen;
        for (int e = 0; e < 3; e++) cpt[e] = xj[e] + nv[e] * reach;
    } else {
        const double nj[3] = {p.nx[j], p.ny[j], p.nz[j]};
        const double pen = (((nj[0] * xi[0] + nj[1] * xi[1]) + nj[2] * xi[2]) - ri) -
                           ((nj[0] * xj[0] + nj[1] * xj[1]) + nj[2] * xj[2]);
        if (!(pen < 0.0)) return false;
        delta = -(pen);
        if (delta < 0.0) return false;
        const double reach = ri + pen;
        for (int e = 0; e < 3; e++) { nv[e] = nj[e]; cpt[e] = xi[e] - reach * nj[e]; }
    }
    return true;
}

// Candidates of i in the reference's visiting order: the spheres of cell 0
// and of the 27 stencil cells, then the half-spaces of the same cells
// (cell_lists.py partitions each cell spheres-first).  fn(j, pass) is called
// for each candidate j != i and returns false to stop the walk.
template <typename F>
__device__ __forceinline__ void walk_candidates(const DemPtrs &p, const DemGeom &g, int i, int ci, F &&fn) {
    for (int pass = 0; pass < 2; pass++) {
        for (int s = -1; s < 27; s++) {
            const int cell = s < 0 ? 0 : ci + g.stencil[s];
            if (!(s < 0 || (cell > 0 && cell < g.ncells))) continue;
            const int st = p.cell_start[cell], nsph = p.nshapes[2 * cell];
            const int q0 = pass == 0 ? 0 : nsph;
            const int q1 = pass == 0 ? nsph : nsph + p.nshapes[2 * cell + 1];
            for (int q = q0; q < q1; q++) {
                const int j = p.cell_particles[st + q];
                if (j == i) continue;
                if (!fn(j, pass)) return;
            }
        }
    }
}

// inclusive sum over each group of W lanes
template <int W>
__device__ __forceinline__ int group_scan(int x) {
    const int l = threadIdx.x & (W - 1);
#pragma unroll
    for (int off = 1; off < W; off <<= 1) {
        const int y = __shfl_up(x, off, W);
        if (l >= off) x += y;
    }
    return x;
}

// Contact search, LPS lanes per local particle.  The visiting order's 28
// cells (cell 0, then the 27 stencil cells) are dealt to the lanes in
// contiguous runs of CPL = ceil(28 / LPS).  Each lane counts the touching
// spheres and half-spaces of its cells; prefix sums over the lanes give every
// hit its place in the reference's visiting order (the spheres of cell 0 and
// of the stencil cells, then the half-spaces of the same cells;
// cell_lists.py partitions each cell spheres-first), and a second walk stores
// them as cand[t][i] (half-spaces as ~j), from a bit mask of the lane's
// first 32 candidates (beyond them the geometry test is repeated).
// cand_n[i] = DCAND + 1 when more than DCAND touch (the force pass then walks
// the cells itself).  Only geometry is read: no history is touched.
template <int LPS>
__global__ __launch_bounds__(256) void k_contact_find(DemPtrs p, const int *__restrict__ c, DemGeom g) {
    constexpr int CPL = (28 + LPS - 1) / LPS;
    if (c[D_OVF]) return;
    const int nl = c[D_NLOCAL];
    const int i = (blockIdx.x * 256 + threadIdx.x) / LPS;
    const int l = threadIdx.x & (LPS - 1);
    const bool act = i < nl;
    const int ii = act ? i : 0;
    const bool fixed = !act || (p.flags[ii] & DFLAG_FIXED);
    // this lane's cells: start, sphere count, candidate count, and the
    // offset of each cell's candidates in the lane's sequence
    int cst[CPL], csph[CPL], cn[CPL], cbase[CPL + 1];
    const int ci = fixed ? 0 : p.particle_cell[ii];
    cbase[0] = 0;
#pragma unroll
    for (int u = 0; u < CPL; u++) {
        const int k = l * CPL + u;
        cst[u] = 0; csph[u] = 0; cn[u] = 0;
        if (!fixed && k < 28) {
            const int cell = k == 0 ? 0 : ci + g.stencil[k - 1];
            if (k == 0 || (cell > 0 && cell < g.ncells)) {
                cst[u] = p.cell_start[cell];
                const int2 ns = reinterpret_cast<const int2 *>(p.nshapes)[cell];
                csph[u] = ns.x;
                cn[u] = ns.x + ns.y;
            }
        }
        cbase[u + 1] = cbase[u] + cn[u];
    }
    const double ri = p.r[ii];
    const double xi[3] = {p.x[ii], p.y[ii], p.z[ii]};
    int hs = 0, hh = 0;
    unsigned hit = 0u;
#pragma unroll
    for (int u = 0; u < CPL; u++)
        for (int q = 0; q < cn[u]; q++) {
            const int j = p.cell_particles[cst[u] + q];
            const bool sph = q < csph[u];
            double nv[3], cpt[3], delta;
            if (j != ii && contact_geom(p, xi, ri, j, sph ? 0 : 1, nv, cpt, delta)) {
                if (sph) hs++; else hh++;
                const int b = cbase[u] + q;
                if (b < 32) hit |= 1u << b;
            }
        }
    const int ps = group_scan<LPS>(hs), ph = group_scan<LPS>(hh);
    const int tot_s = __shfl(ps, LPS - 1, LPS), tot_h = __shfl(ph, LPS - 1, LPS);
    const int n = tot_s + tot_h;
    if (n <= DCAND) {
        int os = ps - hs, oh = tot_s + ph - hh;
        while (hit) {
            const int b = __builtin_ctz(hit);
            hit &= hit - 1u;
            int u = 0;
#pragma unroll
            for (int v = 1; v < CPL; v++) u += b >= cbase[v] ? 1 : 0;
            int st = cst[0], q = b, sp = csph[0];
#pragma unroll
            for (int v = 1; v < CPL; v++)
                if (u == v) { st = cst[v]; q = b - cbase[v]; sp = csph[v]; }
            const int j = p.cell_particles[st + q];
            const int t = q < sp ? os++ : oh++;
            p.cand[(size_t) t * p.cap + ii] = q < sp ? j : ~j;
        }
        if (cbase[CPL] > 32) {
#pragma unroll
            for (int u = 0; u < CPL; u++)
                for (int q = max(0, 32 - cbase[u]); q < cn[u]; q++) {
                    const int j = p.cell_particles[cst[u] + q];
                    const bool sph = q < csph[u];
                    double nv[3], cpt[3], delta;
                    if (j != ii && contact_geom(p, xi, ri, j, sph ? 0 : 1, nv, cpt, delta)) {
                        const int t = sph ? os++ : oh++;
                        p.cand[(size_t) t * p.cap + ii] = sph ? j : ~j;
                    }
                }
        }
    }
    if (act && l == 0) p.cand_n[i] = n <= DCAND ? n : DCAND + 1;
}

// One thread per local particle: reset_contact_history_usage_status, then the
// history slot of each touching partner in order (the last match, or a new
// slot; src/pairs/mapping/funcs.py:230-270), marked used.
__global__ __launch_bounds__(256) void k_contact_slots(DemPtrs p, int *__restrict__ c) {
    const int nl = c[D_NLOCAL];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nl || c[D_OVF]) return;
    const int nc0 = p.num_contacts[i];
    for (int k = 0; k < nc0; k++) p.c_used[(size_t) k * p.cap + i] = 0;
    const int nt = p.cand_n[i];
    if (nt > DCAND) return;
    for (int t = 0; t < nt; t++) {
        const int e = p.cand[(size_t) t * p.cap + i];
        p.cslot[(size_t) t * p.cap + i] = contact_slot(p, c, i, p.uid[e < 0 ? ~e : e]);
    }
}

// One thread per (partner t, local particle i): linear_spring_dashpot of the
// contact, its history slot updated, its force and torque term stored as
// cf[6 t + 0..5][i] for the in-order sums of k_contacts.
__global__ __launch_bounds__(256) void k_contact_pairs(DemPtrs p, const int *__restrict__ c, DemParams prm) {
    const int nl = c[D_NLOCAL];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int t = blockIdx.y;
    if (i >= nl || c[D_OVF]) return;
    const int nt = p.cand_n[i];
    if (t >= nt || nt > DCAND) return;
    const size_t a = (size_t) t * p.cap + i;
    const int slot = p.cslot[a];
    if (slot < 0) return;
    const int e = p.cand[a];
    const int j = e < 0 ? ~e : e;
    const double ri = p.r[i];
    const double xi[3] = {p.x[i], p.y[i], p.z[i]};
    double nv[3], cpt[3], delta, F[3], T[3];
    if (!contact_geom(p, xi, ri, j, e < 0 ? 1 : 0, nv, cpt, delta)) return;
    spring_dashpot(p, prm, i, j, slot, nv, delta, cpt, F, T);
    for (int d = 0; d < 3; d++) {
        p.cf[(6 * (size_t) t + d) * p.cap + i] = F[d];
        p.cf[(6 * (size_t) t + 3 + d) * p.cap + i] = T[d];
    }
}

// One thread per local particle: reset_volatile_properties, gravity, the
// contacts' forces and torques summed in contact order (force += sphere +
// plane sums, torque += plane + sphere sums) and
// clear_unused_contact_history -- every one of them touches only the
// thread's own particle and history rows.  A particle with more than DCAND
// touching partners walks its cells here, slot search and spring-dashpot
// inline.
__global__ __launch_bounds__(128) void k_contacts(DemPtrs p, int *__restrict__ c, DemGeom g, DemParams prm) {
    const int nl = c[D_NLOCAL];
    const int i = blockIdx.x * 128 + threadIdx.x;
    if (i >= nl || c[D_OVF]) return;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0;
    if (!(p.flags[i] & DFLAG_FIXED)) {
        // gravity (examples/dem.py:88-90)
        const double ri = p.r[i];
        const double vol = (((4.0 / 3.0) * prm.pi) * ri * ri) * ri;
        f2 = f2 - (prm.drho * vol) * prm.grav;
        Acc3 fs = {0.0, 0.0, 0.0}, fp = {0.0, 0.0, 0.0}, ts = {0.0, 0.0, 0.0}, tp = {0.0, 0.0, 0.0};
        const int nt = p.cand_n[i];
        if (nt <= DCAND) {
            for (int t = 0; t < nt; t++) {
                const size_t a = (size_t) t * p.cap + i;
                if (p.cslot[a] < 0) continue;
                double F[3], T[3];
                for (int d = 0; d < 3; d++) {
                    F[d] = p.cf[(6 * (size_t) t + d) * p.cap + i];
                    T[d] = p.cf[(6 * (size_t) t + 3 + d) * p.cap + i];
                }
                if (p.cand[a] >= 0) accumulate(fs, ts, F, T);
                else accumulate(fp, tp, F, T);
            }
        } else {
            const double xi[3] = {p.x[i], p.y[i], p.z[i]};
            walk_candidates(p, g, i, p.particle_cell[i], [&](int j, int pass) {
                double nv[3], cpt[3], delta, F[3], T[3];
                if (!contact_geom(p, xi, ri, j, pass, nv, cpt, delta)) return true;
                const int slot = contact_slot(p, c, i, p.uid[j]);
                if (slot < 0) return true;
                spring_dashpot(p, prm, i, j, slot, nv, delta, cpt, F, T);
                if (pass == 0) accumulate(fs, ts, F, T);
                else accumulate(fp, tp, F, T);
                return true;
            });
        }
        f0 = f0 + (fs.x + fp.x); f1 = f1 + (fs.y + fp.y); f2 = f2 + (fs.z + fp.z);
        t0 = t0 + (tp.x + ts.x); t1 = t1 + (tp.y + ts.y); t2 = t2 + (tp.z + ts.z);
    }
    p.fx[i] = f0; p.fy[i] = f1; p.fz[i] = f2;
    p.tx[i] = t0; p.ty[i] = t1; p.tz[i] = t2;
    // clear_unused_contact_history (src/pairs/sim/contact_history.py:91-127)
    int k = 0;
    int nc = p.num_contacts[i];
    while (k < nc) {
        const size_t a = (size_t) k * p.cap + i;
        if (p.c_used[a] == 0) {
            const int last = nc - 1;
            if (last > 0) {
                const size_t b = (size_t) last * p.cap + i;
                p.c_stick[a] = p.c_stick[b];
                for (int d = 0; d < 3; d++)
                    p.c_tsd[(3 * (size_t) k + d) * p.cap + i] = p.c_tsd[(3 * (size_t) last + d) * p.cap + i];
                p.c_ivm[a] = p.c_ivm[b];
                p.c_uid[a] = p.c_uid[b];
                p.c_used[a] = p.c_used[b];
            }
            nc = nc - 1;
        } else {
            k++;
        }
    }
    p.num_contacts[i] = nc;
}

// ---------------------------------------------------------------- integration

// euler (examples/dem.py:77-85) with quaternion(phi, |phi|) and
// quaternion_to_rotation_matrix (src/pairs/mapping/keywords.py:190-252)
__global__ __launch_bounds__(256) void k_euler(DemPtrs p, const int *__restrict__ c, double dt) {
    const int nl = c[D_NLOCAL];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nl || c[D_OVF] || (p.flags[i] & DFLAG_FIXED)) return;
    const size_t C = p.cap;
    const double im = 1.0 / p.m[i];
    const double hm = 0.5 * im;
    const double f[3] = {p.fx[i], p.fy[i], p.fz[i]}, tq[3] = {p.tx[i], p.ty[i], p.tz[i]};
    double x[3] = {p.x[i], p.y[i], p.z[i]}, v[3] = {p.vx[i], p.vy[i], p.vz[i]}, w[3] = {p.wx[i], p.wy[i], p.wz[i]};
    for (int d = 0; d < 3; d++) x[d] = x[d] + (((hm * f[d]) * dt) * dt + v[d] * dt);
    for (int d = 0; d < 3; d++) v[d] = v[d] + (im * f[d]) * dt;
    double I[9], R[9], q[4];
    for (int k = 0; k < 9; k++) { I[k] = p.inv_i[k * C + i]; R[k] = p.rot_m[k * C + i]; }
    for (int k = 0; k < 4; k++) q[k] = p.rot_q[k * C + i];
    double a[3], b[3], wd[3];
    for (int r = 0; r < 3; r++) a[r] = (I[3 * r] * tq[0] + I[3 * r + 1] * tq[1]) + I[3 * r + 2] * tq[2];
    for (int r = 0; r < 3; r++) b[r] = (R[3 * r] * a[0] + R[3 * r + 1] * a[1]) + R[3 * r + 2] * a[2];
    for (int cc = 0; cc < 3; cc++) wd[cc] = (b[0] * R[cc] + b[1] * R[3 + cc]) + b[2] * R[6 + cc];
    double phi[3];
    for (int d = 0; d < 3; d++) phi[d] = w[d] * dt + ((0.5 * wd[d]) * dt) * dt;
    const double plen = sqrt((phi[0] * phi[0] + phi[1] * phi[1]) + phi[2] * phi[2]);
    double dq[4];
    if (fabs(plen) < 1e-06) {
        dq[0] = 1.0; dq[1] = 0.0; dq[2] = 0.0; dq[3] = 0.0;
    } else {
        const double cs = cos(plen * 0.5), sn = sin(plen * 0.5), inv = 1.0 / plen;
        dq[0] = cs;
        for (int d = 0; d < 3; d++) dq[1 + d] = sn * (phi[d] * inv);
    }
    const double p0 = ((dq[0] * q[0] - dq[1] * q[1]) - dq[2] * q[2]) - dq[3] * q[3];
    const double p1 = ((dq[0] * q[1] + dq[1] * q[0]) + dq[2] * q[3]) - dq[3] * q[2];
    const double p2 = ((dq[0] * q[2] + dq[2] * q[0]) + dq[3] * q[1]) - dq[1] * q[3];
    const double p3 = ((dq[0] * q[3] + dq[3] * q[0]) + dq[1] * q[2]) - dq[2] * q[1];
    const double n2 = ((p0 * p0 + p1 * p1) + p2 * p2) + p3 * p3;
    const double sc = (fabs(n2 - 1.0) < 1e-08) ? 1.0 : 1.0 / sqrt(n2);
    q[0] = p0 * sc; q[1] = p1 * sc; q[2] = p2 * sc; q[3] = p3 * sc;
    R[0] = (1.0 - (2.0 * q[2]) * q[2]) - (2.0 * q[3]) * q[3];
    R[1] = 2.0 * (q[1] * q[2] - q[0] * q[3]);
    R[2] = 2.0 * (q[1] * q[3] + q[0] * q[2]);
    R[3] = 2.0 * (q[1] * q[2] + q[0] * q[3]);
    R[4] = (1.0 - (2.0 * q[1]) * q[1]) - (2.0 * q[3]) * q[3];
    R[5] = 2.0 * (q[2] * q[3] - q[0] * q[1]);
    R[6] = 2.0 * (q[1] * q[3] - q[0] * q[2]);
    R[7] = 2.0 * (q[2] * q[3] + q[0] * q[1]);
    R[8] = (1.0 - (2.0 * q[1]) * q[1]) - (2.0 * q[2]) * q[2];
    for (int d = 0; d < 3; d++) w[d] = w[d] + wd[d] * dt;
    p.x[i] = x[0]; p.y[i] = x[1]; p.z[i] = x[2];
    p.vx[i] = v[0]; p.vy[i] = v[1]; p.vz[i] = v[2];
    p.wx[i] = w[0]; p.wy[i] = w[1]; p.wz[i] = w[2];
    for (int k = 0; k < 9; k++) p.rot_m[k * C + i] = R[k];
    for (int k = 0; k < 4; k++) p.rot_q[k * C + i] = q[k];
}

// update_mass_and_inertia (examples/dem.py:6-15)
__global__ __launch_bounds__(256) void k_mass_inertia(DemPtrs p, int nl) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nl) return;
    const size_t C = p.cap;
    for (int k = 0; k < 9; k++) p.rot_m[k * C + i] = (k % 4 == 0) ? 1.0 : 0.0;
    p.rot_q[i] = 1.0; p.rot_q[C + i] = 0.0; p.rot_q[2 * C + i] = 0.0; p.rot_q[3 * C + i] = 0.0;
    if (p.shape[i] == DSHAPE_SPHERE) {
        double a = 0.4 * p.m[i];
        a = a * p.r[i];
        a = a * p.r[i];
        const double M[9] = {a, 0.0, 0.0, 0.0, a, 0.0, 0.0, 0.0, a};
        double det = M[0] * (M[4] * M[8] - M[7] * M[5]);
        det = det + M[1] * (M[5] * M[6] - M[8] * M[3]);
        det = det + M[2] * (M[3] * M[7] - M[4] * M[6]);
        const double id = 1.0 / det;
        const double inv[9] = {id * (M[4] * M[8] - M[5] * M[7]), id * (M[7] * M[2] - M[8] * M[1]),
                               id * (M[1] * M[5] - M[2] * M[4]), id * (M[5] * M[6] - M[3] * M[8]),
                               id * (M[8] * M[0] - M[6] * M[2]), id * (M[2] * M[3] - M[0] * M[5]),
                               id * (M[3] * M[7] - M[4] * M[6]), id * (M[6] * M[1] - M[7] * M[0]),
                               id * (M[0] * M[4] - M[1] * M[3])};
        for (int k = 0; k < 9; k++) p.inv_i[k * C + i] = inv[k];
    } else {
        p.m[i] = __longlong_as_double(0x7ff0000000000000LL);  // +infinity
        for (int k = 0; k < 9; k++) p.inv_i[k * C + i] = 0.0;
    }
}

// ---------------------------------------------------------------- launchers

static int grid_of(int n, int bs) { return n <= 0 ? 1 : (n + bs - 1) / bs; }

void launch_fill(hipStream_t s, int *p, int v, int n) {
    hipLaunchKernelGGL(k_fill_i32, dim3(std::max(1, std::min(grid_of(n, 256), 4096))), dim3(256), 0, s, p, v, n);
}

void launch_mass_inertia(hipStream_t s, const DemPtrs &p, int nl) {
    hipLaunchKernelGGL(k_mass_inertia, dim3(grid_of(nl, 256)), dim3(256), 0, s, p, nl);
}

static void select_dim(hipStream_t s, const DemPtrs &p, int *c, const DemGeom &g, int dim, int exch, int first) {
    const double *Xd = dim == 0 ? p.x : (dim == 1 ? p.y : p.z);
    const int nb = grid_of(p.cap, 256);
    const double off = exch ? 0.0 : g.width;
    hipLaunchKernelGGL(k_sel_count, dim3(nb), dim3(256), 0, s, Xd, p.flags, c, g.lo[dim], g.hi[dim], off, exch, first,
                       p.blk, p.exchg_flag);
    hipLaunchKernelGGL(k_sel_scatter, dim3(nb), dim3(256), 0, s, p, c, dim, g.lo[dim], g.hi[dim], off,
                       g.hi[dim] - g.lo[dim], exch, p.blk, 1, -1, p.send_cap, p.cap);
}

// exchange, borders, cells of one step (the cell counts are zero on entry:
// allocated zeroed, cleared by every cell ordering pass)
void launch_prepare(hipStream_t s, const DemPtrs &p, int *c, const DemGeom &g) {
    // exchange, dimension by dimension (comm.py:100-151); one rank: only the
    // periodic dimensions send
    for (int dim = 0; dim < 3; dim++) {
        if (!g.periodic[dim]) continue;
        select_dim(s, p, c, g, dim, 1, 0);
        hipLaunchKernelGGL(k_exchange, dim3(1), dim3(1024), 0, s, p, c, g.len[0], g.len[1], g.len[2]);
    }
    // borders (comm.py:56-98)
    bool first = true;
    for (int dim = 0; dim < 3; dim++) {
        if (!g.periodic[dim]) continue;
        select_dim(s, p, c, g, dim, 0, first ? 1 : 0);
        first = false;
    }
    if (first) hipLaunchKernelGGL(k_reset_comm, dim3(1), dim3(64), 0, s, c);  // no periodic dimension: no ghosts
    // cells
    const int pg = grid_of(p.cap, 256);
    hipLaunchKernelGGL(k_cell_count, dim3(pg), dim3(256), 0, s, p, c, g);
    const int nb = grid_of(g.ncells, 1024);
    hipLaunchKernelGGL(k_scan_partial, dim3(nb), dim3(1024), 0, s, p.cell_count, p.cell_start, p.scan_tmp, g.ncells);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(1024), 0, s, p.cell_start, p.scan_tmp, g.ncells);
    hipLaunchKernelGGL(k_cell_fill, dim3(pg), dim3(256), 0, s, p, c);
    hipLaunchKernelGGL(k_cell_order, dim3(grid_of(g.ncells, 256)), dim3(256), 0, s, p, g.ncells);
}

// forces + history bookkeeping, then integration
void launch_contacts(hipStream_t s, const DemPtrs &p, int *c, const DemGeom &g, const DemParams &prm,
                     const hipEvent_t *between) {
    hipLaunchKernelGGL(k_contact_find<DEM_LPS>, dim3(grid_of(p.cap, 256 / DEM_LPS)), dim3(256), 0, s, p, c, g);
    if (between) hipEventRecord(between[0], s);
    hipLaunchKernelGGL(k_contact_slots, dim3(grid_of(p.cap, 256)), dim3(256), 0, s, p, c);
    if (between) hipEventRecord(between[1], s);
    hipLaunchKernelGGL(k_contact_pairs, dim3(grid_of(p.cap, 256), DCAND), dim3(256), 0, s, p, c, prm);
    if (between) hipEventRecord(between[2], s);
    hipLaunchKernelGGL(k_contacts, dim3(grid_of(p.cap, 128)), dim3(128), 0, s, p, c, g, prm);
}

void launch_euler(hipStream_t s, const DemPtrs &p, int *c, const DemParams &prm) {
    hipLaunchKernelGGL(k_euler, dim3(grid_of(p.cap, 256)), dim3(256), 0, s, p, c, prm.dt);
}

}  // namespace dk
