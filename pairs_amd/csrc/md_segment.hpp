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

// Walk a lane's neighbour list eight tile offsets per step, to the
// wave-uniform depth nmax: body(k, on) gets entries k .. k+7.  A step
// consumes two packed groups of four entries; the loads of the groups two
// steps ahead are issued first, in a ring of three register pairs written
// out (no register moves), so a list load never stalls the step that issues
// it.  Groups past the list re-read its last group (in bounds, no branch);
// offsets are clamped into the tile.
template <int U, typename Body>
__device__ __forceinline__ void list_walk(const uint16_t *nl, int i, int nl_stride, int n, int nmax, int T,
                                          Body &&body) {
    static_assert(U == 4 || U == 8, "a step takes one or two packed groups");
    const uint64_t *lp = (const uint64_t *) nl + i;
    const int gl = max((n + 3) / 4 - 1, 0), tlast = T - 1;
    // a relaxed atomic load is an ordered memory access to the compiler: it
    // stays where it is issued instead of being sunk next to its use
    auto ld = [&](int gq) {
        return __hip_atomic_load(lp + (size_t) min(gq, gl) * nl_stride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    };
    auto step = [&](int k, uint64_t a, uint64_t b) {
        int on[8];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            on[q] = min((int) ((a >> (16 * q)) & 0xffff), tlast);
            on[4 + q] = min((int) ((b >> (16 * q)) & 0xffff), tlast);
        }
        body(k, on);
    };
    if (U == 4) {
        // one group per step, same ring of three
        auto step4 = [&](int k, uint64_t a) {
            int on[8];
#pragma unroll
            for (int q = 0; q < 4; q++) on[q] = min((int) ((a >> (16 * q)) & 0xffff), tlast);
            body(k, on);
        };
        uint64_t r0 = ld(0), r1 = ld(1), r2;
        for (int k = 0; k < nmax; k += 12) {
            const int g = k >> 2;
            r2 = ld(g + 2);
            step4(k, r0);
            if (k + 4 >= nmax) break;
            r0 = ld(g + 3);
            step4(k + 4, r1);
            if (k + 8 >= nmax) break;
            r1 = ld(g + 4);
            step4(k + 8, r2);
        }
        return;
    }
    uint64_t r0a = ld(0), r0b = ld(1), r1a = ld(2), r1b = ld(3), r2a, r2b;
    for (int k = 0; k < nmax; k += 24) {
        const int g = k >> 2;
        r2a = ld(g + 4); r2b = ld(g + 5);
        step(k, r0a, r0b);
        if (k + 8 >= nmax) break;
        r0a = ld(g + 6); r0b = ld(g + 7);
        step(k + 8, r1a, r1b);
        if (k + 16 >= nmax) break;
        r1a = ld(g + 8); r1b = ld(g + 9);
        step(k + 16, r2a, r2b);
    }
}


}  // namespace pk
