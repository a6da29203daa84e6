// Device helpers shared by the kernels that work on segments of the cell
// grid (neighbour build, LJ force, MLP descriptors and forces): a segment's
// tile of physical runs, its staging into LDS, and the walk over a particle's
// packed 16-bit neighbour list.
#pragma once
#include "md_kernels.hpp"

namespace pk {

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

struct SegInfo {
    int cx, cy, za, zb;  // column and z range [za, zb) of the home cells
    int ca, cb;          // flat cell ids of the first home cell and one past the last
};

__device__ __forceinline__ SegInfo seg_info(int seg, const Geom &g) {
    SegInfo s;
    const int chunk = seg % g.nchunks, col = seg / g.nchunks;
    s.cy = col % g.dim_cells[1];
    s.cx = col / g.dim_cells[1];
    s.za = chunk == 0 ? 0 : 1 + chunk * g.zc;
    s.zb = chunk == g.nchunks - 1 ? g.dim_cells[2] : 1 + (chunk + 1) * g.zc;
    const int base = (s.cx * g.dim_cells[1] + s.cy) * g.dim_cells[2] + 1;
    s.ca = base + s.za;
    s.cb = base + s.zb;
    return s;
}

constexpr int NRUNS = 20;

// The 20 physical runs of a segment's tile, in the reference's candidate
// order of cells: cell 0 (locals, ghosts), then the columns (dx, dy) of the
// stencil (dx outer, dy inner; src/pairs/sim/cell_lists.py:70-87), each over
// flat cells [B + za - 1, B + zb] clipped to the grid (a stencil cell of the
// reference is valid iff 0 < nc < ncells, interaction.py:92-98), locals then
// ghosts.  Written by threads 0..19; r_off is the exclusive prefix (r_off[20]
// = tile size).  lo of each column is kept for candidate enumeration.
__device__ __forceinline__ void seg_runs(const SegInfo &si, const Geom &g, const int *__restrict__ lstart,
                                         const int *__restrict__ gstart, int nlocal, int *r_start, int *r_len,
                                         int *r_off, int *col_lo) {
    const int t = threadIdx.x;
    if (t < NRUNS) {
        int start = 0, len = 0;
        if (t < 2) {
            if (t == 0) { start = lstart[0]; len = lstart[1] - lstart[0]; }
            else { start = nlocal + gstart[0]; len = gstart[1] - gstart[0]; }
        } else {
            const int col = (t - 2) >> 1, loc = ((t - 2) & 1) == 0;
            const int dx = col / 3 - 1, dy = col % 3 - 1;
            const int B = ((si.cx + dx) * g.dim_cells[1] + (si.cy + dy)) * g.dim_cells[2] + 1;
            int lo = B + si.za - 1, hi = B + si.zb;
            lo = max(lo, 1);
            hi = min(hi, g.ncells - 1);
            if (lo <= hi) {
                if (loc) { start = lstart[lo]; len = lstart[hi + 1] - lstart[lo]; }
                else { start = nlocal + gstart[lo]; len = gstart[hi + 1] - gstart[lo]; }
            }
            if (loc) col_lo[col] = lo;
        }
        r_start[t] = start;
        r_len[t] = len;
    }
    __syncthreads();
    if (t == 0) {
        int off = 0;
        for (int r = 0; r < NRUNS; r++) { r_off[r] = off; off += r_len[r]; }
        r_off[NRUNS] = off;
    }
    __syncthreads();
}

__device__ __forceinline__ int run_of(const int *r_off, int p) {
    int r = 0;
#pragma unroll
    for (int k = 1; k < NRUNS; k++) r += (p >= r_off[k]) ? 1 : 0;
    return r;
}

constexpr int FB = 64;   // threads per force workgroup (one wave)

__device__ __forceinline__ void stage_slots(const int *r_start, const int *r_off, int T, int *tslot) {
    for (int p = threadIdx.x; p < T; p += blockDim.x) {
        const int r = run_of(r_off, p);
        tslot[p] = r_start[r] + (p - r_off[r]);
    }
}

// Walk rows [row0, row0 + cnt) of a lane's list to the wave-uniform depth
// cmax, U (4 or 8) entries per step: body(k, j) gets entries k .. k+U-1 as
// physical slots.  A step takes U/4 packed groups of four entries (one
// 16-byte load each); the groups two steps ahead are loaded first, in a ring
// of three register sets written out (no register moves), so a list load
// never stalls the step that issues it.  Groups past the list re-read its
// last group.  Entries past cnt are masked by the body; every entry is
// clamped to jmax (the last particle slot), so no gather leaves the particle
// arrays whatever a row past a particle's count holds.
template <int U, typename Body>
__device__ __forceinline__ void list_walk32(const uint32_t *__restrict__ nl, int row0, int i, int stride, int cnt,
                                            int cmax, uint32_t jmax, Body &&body) {
    static_assert(U == 4 || U == 8, "a step takes one or two packed groups");
    constexpr int G = U / 4;  // groups per step
    const uint4 *lp = (const uint4 *) nl + (size_t) (row0 >> 2) * stride + i;
    const int gl = max((cnt + 3) / 4 - 1, 0);
    // a relaxed atomic load is an ordered memory access to the compiler: it
    // stays where it is issued instead of being sunk next to its use
    auto ld1 = [&](int gq) {
        const uint32_t *p = (const uint32_t *) (lp + (size_t) min(gq, gl) * stride);
        uint4 v;
        v.x = __hip_atomic_load(p + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        v.y = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        v.z = __hip_atomic_load(p + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        v.w = __hip_atomic_load(p + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        return v;
    };
    struct Grp { uint4 v[G]; };
    auto ld = [&](int step) {
        Grp r;
#pragma unroll
        for (int q = 0; q < G; q++) r.v[q] = ld1(step * G + q);
        return r;
    };
    auto run = [&](int k, const Grp &r) {
        uint32_t j[U];
#pragma unroll
        for (int q = 0; q < G; q++) {
            j[4 * q + 0] = min(r.v[q].x, jmax);
            j[4 * q + 1] = min(r.v[q].y, jmax);
            j[4 * q + 2] = min(r.v[q].z, jmax);
            j[4 * q + 3] = min(r.v[q].w, jmax);
        }
        body(k, j);
    };
    Grp r0 = ld(0), r1 = ld(1), r2;
    for (int k = 0; k < cmax; k += 3 * U) {
        const int st = k / U;
        r2 = ld(st + 2);
        run(k, r0);
        if (k + U >= cmax) break;
        r0 = ld(st + 3);
        run(k + U, r1);
        if (k + 2 * U >= cmax) break;
        r1 = ld(st + 4);
        run(k + 2 * U, r2);
    }
}

// Next positions of local i's images (the stale multi-hop refresh of
// src/pairs/sim/comm.py:45-54,369-442 one step early): an image of i itself
// is i's new position + shift, an image of a ghost is that ghost's current
// position + shift -- exactly what the next step's synchronize packs.
__device__ __forceinline__ void write_images(const Images &im, int i, int nlocal, double px, double py, double pz,
                                             const double *__restrict__ X, const double *__restrict__ Y,
                                             const double *__restrict__ Z, double *__restrict__ Xn,
                                             double *__restrict__ Yn, double *__restrict__ Zn, const Geom &g,
                                             double4 *__restrict__ Pn = nullptr) {
    const int n = min(im.cnt[i], IMG_MAX);
    for (int k = 0; k < n; k++) {
        const int q = im.slot[(size_t) k * im.cap + i];
        const int src = im.src[q];
        const bool own = src == i;
        const double bx = own ? px : X[src], by = own ? py : Y[src], bz = own ? pz : Z[src];
        const double x = bx + (double) im.mult[3 * q + 0] * g.len[0];
        const double y = by + (double) im.mult[3 * q + 1] * g.len[1];
        const double z = bz + (double) im.mult[3 * q + 2] * g.len[2];
        Xn[nlocal + q] = x;
        Yn[nlocal + q] = y;
        Zn[nlocal + q] = z;
        if (Pn) Pn[nlocal + q] = make_double4(x, y, z, 0.0);
    }
}

}  // namespace pk
