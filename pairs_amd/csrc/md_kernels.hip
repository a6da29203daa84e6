// gfx950 kernels of the LJ molecular-dynamics hot path.
//
// Every floating-point expression follows the operation order of the module
// the reference's code generator emits for it (cited per kernel), and the
// library is compiled with -ffp-contract=off, so each value is rounded exactly
// as in the reference's generated CPU program.  Integer/index work (cells,
// ghost maps, neighbour lists) is deterministic and reproduces the CPU
// program's order on reference indices: stable compactions by ballot/prefix
// scan instead of the atomic insertion the generated CUDA code uses.
// The storage layout (cell-ordered physical slots, reference-index
// relabeling, segments and their LDS tiles) is described in md_kernels.hpp.
#include "md_kernels.hpp"
#include "md_segment.hpp"

namespace pk {

// ---------------------------------------------------------------- utilities

__global__ void k_fill_i32(int *__restrict__ p, int v, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double *__restrict__ p, double v, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_identity(int *__restrict__ p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}
__global__ void k_set_counters(int *__restrict__ c, int first, int count, int v) {
    const int t = threadIdx.x;
    if (t < count) c[first + t] = v;
}

// the counters of many systems into one array, [r][C_COUNT]
__global__ __launch_bounds__(256) void k_gather_counters(const int *const *__restrict__ src, int n,
                                                         int *__restrict__ dst) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n * C_COUNT) dst[t] = src[t / C_COUNT][t % C_COUNT];
}

// The per-particle state of the first n locals that a rebuild changes in
// place (positions of the buffer it starts from, velocities, mass, type,
// flags, uid, shape, reference index by slot and slot by reference index)
// and the counters, copied from a to b: a checkpoint before a rebuild whose
// counters are checked later, and the restore when they show an overflow.
__global__ __launch_bounds__(256) void k_checkpoint(CkptArrays a, CkptArrays b, int n) {
    if (blockIdx.x == 0 && threadIdx.x < C_COUNT) b.counters[threadIdx.x] = a.counters[threadIdx.x];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        b.X[i] = a.X[i]; b.Y[i] = a.Y[i]; b.Z[i] = a.Z[i];
        b.vx[i] = a.vx[i]; b.vy[i] = a.vy[i]; b.vz[i] = a.vz[i];
        b.mass[i] = a.mass[i];
        b.type[i] = a.type[i]; b.flags[i] = a.flags[i]; b.uid[i] = a.uid[i]; b.shape[i] = a.shape[i];
        b.ref[i] = a.ref[i]; b.phys[i] = a.phys[i];
    }
}

// exclusive scan of one value per thread over a block of BS threads
template <int BS>
__device__ __forceinline__ int block_exclusive_scan(int v, int *lds_waves, int *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds_waves[wave] = x;
    __syncthreads();
    if (wave == 0) {
        const int w = lane < BS / 64 ? lds_waves[lane] : 0;
        int s = w;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(s, off, 64);
            if (lane >= off) s += y;
        }
        if (lane < BS / 64) lds_waves[lane] = s - w;
        if (lane == BS / 64 - 1) lds_waves[BS / 64] = s;
    }
    __syncthreads();
    const int r = x - v + lds_waves[wave];
    *total = lds_waves[BS / 64];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------- geometry

// cell of a position: src/pairs/sim/cell_lists.py:114-126 (generated as
// (int)((x - (min - s)) / s), clamped to [0, n-1], flattened z fastest, + 1)
__device__ __forceinline__ int cell_of(double x, double y, double z, const Geom &g) {
    const double xs[3] = {x, y, z};
    int idx = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double q = (xs[d] - (g.lo[d] - g.spacing)) / g.spacing;
        int c = (int) q >= 0 ? (int) q : 0;
        c = c < g.dim_cells[d] ? c : g.dim_cells[d] - 1;
        idx = d == 0 ? c : idx * g.dim_cells[d] + c;
    }
    return idx + 1;
}

// ---------------------------------------------------------------- cells (binning by counting sort)

// BuildCellLists (src/pairs/sim/cell_lists.py:90-134), histogram half: cell
// of every local and ghost, per-cell counts of locals and of ghosts, and a
// slot per particle within its cell's locals or ghosts.  The locals are in
// cell order from the last rebuild, so the lanes of a wave share few cells:
// each group of lanes with one key takes consecutive slots with one atomic
// (up to 8 groups per wave, single atomics for the rest).  Slot order is
// arrival order; the cell sort orders each cell by reference index after.
__device__ __forceinline__ void k_bin_count_body(const double *__restrict__ X, const double *__restrict__ Y,
                                                   const double *__restrict__ Z, const int *__restrict__ flags,
                                                   const int *__restrict__ counters, Geom g,
                                                   int *__restrict__ cell, int *__restrict__ lcount,
                                                   int *__restrict__ gcount, int *__restrict__ slot) {
    if (counters[C_OVF]) return;
    const int nlocal = counters[C_NLOCAL];
    const int n = nlocal + counters[C_NGHOST];
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int p0 = blockIdx.x * 256 + (threadIdx.x & ~63); p0 < n; p0 += gridDim.x * 256) {
        const int p = p0 + lane;
        const bool live = p < n;
        int c = 0;
        if (live) {
            c = (flags[p] & FLAG_INFINITE) ? 0 : cell_of(X[p], Y[p], Z[p], g);
            cell[p] = c;
        }
        const int key = live ? 2 * c + (p < nlocal ? 0 : 1) : -1;
        unsigned long long todo = __ballot(live);
        int my = 0;
        for (int it = 0; it < 8 && todo; it++) {
            const int leader = __ffsll((long long) todo) - 1;
            const int kl = __shfl(key, leader, 64);
            const unsigned long long m = __ballot(key == kl) & todo;
            int base = 0;
            if (lane == leader) base = atomicAdd(((kl & 1) ? gcount : lcount) + (kl >> 1), __popcll(m));
            base = __shfl(base, leader, 64);
            if ((m >> lane) & 1ull) my = base + __popcll(m & below);
            todo &= ~m;
        }
        if ((todo >> lane) & 1ull) my = atomicAdd(((key & 1) ? gcount : lcount) + (key >> 1), 1);
        if (live) slot[p] = my;
    }
}

__global__ __launch_bounds__(256) void k_bin_count(const double *__restrict__ X, const double *__restrict__ Y,
                                                   const double *__restrict__ Z, const int *__restrict__ flags,
                                                   const int *__restrict__ counters, Geom g,
                                                   int *__restrict__ cell, int *__restrict__ lcount,
                                                   int *__restrict__ gcount, int *__restrict__ slot) {
    k_bin_count_body(X, Y, Z, flags, counters, g, cell, lcount, gcount, slot);
}

// exclusive scans of the local and ghost cell counts -> lstart, gstart
// [0..ncells], one workgroup of 16 waves: each wave takes a contiguous run
// of rows of 64 cells (coalesced loads), the wave totals give each run's
// base, and each row is scanned across the lanes.  The counts are cleared
// behind the reads (nothing else uses them, so the next binning starts from
// zeros).  The flags the rest of the rebuild raises by atomics (images,
// largest tile) are cleared here too.
__device__ __forceinline__ void k_scan_cells_body(int *__restrict__ counters, int *__restrict__ lcount,
                                                     int *__restrict__ gcount, int *__restrict__ lstart,
                                                     int *__restrict__ gstart, int ncells) {
    __shared__ int tot[2][16];
    if (threadIdx.x == 0) { counters[C_IMG_OVF] = 0; counters[C_TILE_MAX] = 0; }
    if (counters[C_OVF]) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int RW = 16;
    constexpr int ROUND = 16 * RW * 64;
    int carry_l = 0, carry_g = 0;
    for (int r0 = 0; r0 < ncells; r0 += ROUND) {
        const int rn = min(ncells - r0, ROUND);
        const int rows = (rn + 63) / 64;
        const int per = (rows + 15) / 16;
        const int row0 = wave * per;
        int vl[RW], vg[RW];
        int sl = 0, sg = 0;
#pragma unroll
        for (int k = 0; k < RW; k++) {
            const int cr = (row0 + k) * 64 + lane;
            const bool in = k < per && cr < rn;
            vl[k] = in ? lcount[r0 + cr] : 0;
            vg[k] = in ? gcount[r0 + cr] : 0;
            sl += vl[k];
            sg += vg[k];
        }
#pragma unroll
        for (int k = 0; k < RW; k++) {
            const int cr = (row0 + k) * 64 + lane;
            if (k < per && cr < rn) { lcount[r0 + cr] = 0; gcount[r0 + cr] = 0; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            sl += __shfl_xor(sl, off, 64);
            sg += __shfl_xor(sg, off, 64);
        }
        if (lane == 0) { tot[0][wave] = sl; tot[1][wave] = sg; }
        __syncthreads();
        int bl = carry_l, bg = carry_g, al = carry_l, ag = carry_g;
        for (int w = 0; w < 16; w++) {
            const int a = tot[0][w], b = tot[1][w];
            if (w < wave) { bl += a; bg += b; }
            al += a; ag += b;
        }
#pragma unroll
        for (int k = 0; k < RW; k++) {
            int xl = vl[k], xg = vg[k];
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int yl = __shfl_up(xl, off, 64), yg = __shfl_up(xg, off, 64);
                if (lane >= off) { xl += yl; xg += yg; }
            }
            const int cr = (row0 + k) * 64 + lane;
            if (k < per && cr < rn) { lstart[r0 + cr] = bl + xl - vl[k]; gstart[r0 + cr] = bg + xg - vg[k]; }
            bl += __shfl(xl, 63, 64);
            bg += __shfl(xg, 63, 64);
        }
        carry_l = al;
        carry_g = ag;
        __syncthreads();
    }
    if (threadIdx.x == 0) { lstart[ncells] = carry_l; gstart[ncells] = carry_g; }
}

__global__ __launch_bounds__(1024) void k_scan_cells(int *__restrict__ counters, int *__restrict__ lcount,
                                                     int *__restrict__ gcount, int *__restrict__ lstart,
                                                     int *__restrict__ gstart, int ncells) {
    k_scan_cells_body(counters, lcount, gcount, lstart, gstart, ncells);
}

__device__ __forceinline__ void k_bin_scatter_body(const int *__restrict__ counters, const int *__restrict__ cell,
                                                     const int *__restrict__ slot, const int *__restrict__ lstart,
                                                     const int *__restrict__ gstart, int *__restrict__ perm) {
    if (counters[C_OVF]) return;
    const int nlocal = counters[C_NLOCAL];
    const int n = nlocal + counters[C_NGHOST];
    for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
        const int c = cell[p];
        const int np = p < nlocal ? lstart[c] + slot[p] : nlocal + gstart[c] + slot[p];
        perm[np] = p;
    }
}

__global__ __launch_bounds__(256) void k_bin_scatter(const int *__restrict__ counters, const int *__restrict__ cell,
                                                     const int *__restrict__ slot, const int *__restrict__ lstart,
                                                     const int *__restrict__ gstart, int *__restrict__ perm) {
    k_bin_scatter_body(counters, cell, slot, lstart, gstart, perm);
}

// Order the locals and the ghosts of every cell by reference index -- the
// order in which the CPU program's sequential insertion fills cell_particles.
// One wave per (cell, locals|ghosts) run: a rank sort of up to 64 keys in
// registers, a serial insertion sort beyond that.
// one wave per item (cell c = item / 2; its locals, or its ghosts)
__device__ __forceinline__ void k_sort_cells_body(const int *__restrict__ counters, const int *__restrict__ lstart,
                                                    const int *__restrict__ gstart, const int *__restrict__ ref,
                                                    int *__restrict__ perm, int ncells, int item) {
    if (counters[C_OVF]) return;
    const int nlocal = counters[C_NLOCAL];
    const int lane = threadIdx.x & 63;
    if (item >= 2 * ncells) return;
    const int c = item >> 1, ghosts = item & 1;
    const int b = ghosts ? nlocal + gstart[c] : lstart[c];
    const int e = ghosts ? nlocal + gstart[c + 1] : lstart[c + 1];
    const int len = e - b;
    if (len <= 1) return;
    if (len <= 64) {
        const int p = lane < len ? perm[b + lane] : 0;
        const int key = lane < len ? ref[p] : 0x7fffffff;
        int rank = 0;
        for (int q = 0; q < len; q++) {
            const int kq = __shfl(key, q, 64);
            rank += (kq < key) ? 1 : 0;
        }
        if (lane < len) perm[b + rank] = p;
    } else if (lane == 0) {
        for (int k = b + 1; k < e; k++) {
            const int v = perm[k], kv = ref[v];
            int q = k - 1;
            while (q >= b && ref[perm[q]] > kv) { perm[q + 1] = perm[q]; q--; }
            perm[q + 1] = v;
        }
    }
}

__global__ __launch_bounds__(256) void k_sort_cells(const int *__restrict__ counters, const int *__restrict__ lstart,
                                                    const int *__restrict__ gstart, const int *__restrict__ ref,
                                                    int *__restrict__ perm, int ncells) {
    k_sort_cells_body(counters, lstart, gstart, ref, perm, ncells, blockIdx.x * 4 + (threadIdx.x >> 6));
}

// Move every particle array into cell order (into the spare arrays) and
// update phys[reference index].
__device__ __forceinline__ void k_permute_body(const int *__restrict__ counters, const int *__restrict__ perm,
                                                 const double *__restrict__ X, const double *__restrict__ Y,
                                                 const double *__restrict__ Z, double *__restrict__ Xn,
                                                 double *__restrict__ Yn, double *__restrict__ Zn,
                                                 const double *__restrict__ vx, const double *__restrict__ vy,
                                                 const double *__restrict__ vz, const double *__restrict__ mass,
                                                 const int *__restrict__ type, const int *__restrict__ flags,
                                                 const int *__restrict__ uid, const int *__restrict__ shape,
                                                 const int *__restrict__ ref, const int *__restrict__ cell,
                                                 double *__restrict__ avx, double *__restrict__ avy,
                                                 double *__restrict__ avz, double *__restrict__ amass,
                                                 int *__restrict__ atype, int *__restrict__ aflags,
                                                 int *__restrict__ auid, int *__restrict__ ashape,
                                                 int *__restrict__ aref, int *__restrict__ acell,
                                                 int *__restrict__ phys) {
    if (counters[C_OVF]) return;
    const int n = counters[C_NLOCAL] + counters[C_NGHOST];
    for (int np = blockIdx.x * 256 + threadIdx.x; np < n; np += gridDim.x * 256) {
        const int p = perm[np];
        Xn[np] = X[p]; Yn[np] = Y[p]; Zn[np] = Z[p];
        avx[np] = vx[p]; avy[np] = vy[p]; avz[np] = vz[p]; amass[np] = mass[p];
        atype[np] = type[p]; aflags[np] = flags[p]; auid[np] = uid[p]; ashape[np] = shape[p];
        const int r = ref[p];
        aref[np] = r;
        acell[np] = cell[p];
        phys[r] = np;
    }
}

__global__ __launch_bounds__(256) void k_permute(const int *__restrict__ counters, const int *__restrict__ perm,
                                                 const double *__restrict__ X, const double *__restrict__ Y,
                                                 const double *__restrict__ Z, double *__restrict__ Xn,
                                                 double *__restrict__ Yn, double *__restrict__ Zn,
                                                 const double *__restrict__ vx, const double *__restrict__ vy,
                                                 const double *__restrict__ vz, const double *__restrict__ mass,
                                                 const int *__restrict__ type, const int *__restrict__ flags,
                                                 const int *__restrict__ uid, const int *__restrict__ shape,
                                                 const int *__restrict__ ref, const int *__restrict__ cell,
                                                 double *__restrict__ avx, double *__restrict__ avy,
                                                 double *__restrict__ avz, double *__restrict__ amass,
                                                 int *__restrict__ atype, int *__restrict__ aflags,
                                                 int *__restrict__ auid, int *__restrict__ ashape,
                                                 int *__restrict__ aref, int *__restrict__ acell,
                                                 int *__restrict__ phys) {
    k_permute_body(counters, perm, X, Y, Z, Xn, Yn, Zn, vx, vy, vz, mass, type, flags, uid, shape, ref, cell, avx, avy, avz, amass, atype, aflags, auid, ashape, aref, acell, phys);
}

// Copy the permuted arrays back (array addresses stay fixed, so captured
// graphs remain valid across rebuilds), clear the locals' image counts, and
// build the ghost refresh maps in physical order: the physical slot of each
// ghost's source and its image multipliers (send_map / send_mult are in
// reference ghost order: ghost g = reference index nlocal + g, read from the
// permuted copy of ref).
__device__ __forceinline__ void k_copy_back_maps_body(
    const int *__restrict__ counters, const double *__restrict__ avx, const double *__restrict__ avy,
    const double *__restrict__ avz, const double *__restrict__ amass, const int *__restrict__ atype,
    const int *__restrict__ aflags, const int *__restrict__ auid, const int *__restrict__ ashape,
    const int *__restrict__ aref, const int *__restrict__ acell, double *__restrict__ vx, double *__restrict__ vy,
    double *__restrict__ vz, double *__restrict__ mass, int *__restrict__ type, int *__restrict__ flags,
    int *__restrict__ uid, int *__restrict__ shape, int *__restrict__ ref, int *__restrict__ cell,
    const int *__restrict__ phys, const int *__restrict__ send_map, const int *__restrict__ send_mult,
    int *__restrict__ sync_src, int *__restrict__ sync_mult, int *__restrict__ img_cnt) {
    if (counters[C_OVF]) return;
    const int nlocal = counters[C_NLOCAL], ng = counters[C_NGHOST];
    const int n = nlocal + ng;
    for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
        vx[p] = avx[p]; vy[p] = avy[p]; vz[p] = avz[p]; mass[p] = amass[p];
        type[p] = atype[p]; flags[p] = aflags[p]; uid[p] = auid[p]; shape[p] = ashape[p];
        ref[p] = aref[p]; cell[p] = acell[p];
        if (p < nlocal) img_cnt[p] = 0;
    }
    for (int q = blockIdx.x * 256 + threadIdx.x; q < ng; q += gridDim.x * 256) {
        const int g = aref[nlocal + q] - nlocal;
        sync_src[q] = phys[send_map[g]];
        sync_mult[3 * q + 0] = send_mult[3 * g + 0];
        sync_mult[3 * q + 1] = send_mult[3 * g + 1];
        sync_mult[3 * q + 2] = send_mult[3 * g + 2];
    }
}

__global__ __launch_bounds__(256) void k_copy_back_maps(
    const int *__restrict__ counters, const double *__restrict__ avx, const double *__restrict__ avy,
    const double *__restrict__ avz, const double *__restrict__ amass, const int *__restrict__ atype,
    const int *__restrict__ aflags, const int *__restrict__ auid, const int *__restrict__ ashape,
    const int *__restrict__ aref, const int *__restrict__ acell, double *__restrict__ vx, double *__restrict__ vy,
    double *__restrict__ vz, double *__restrict__ mass, int *__restrict__ type, int *__restrict__ flags,
    int *__restrict__ uid, int *__restrict__ shape, int *__restrict__ ref, int *__restrict__ cell,
    const int *__restrict__ phys, const int *__restrict__ send_map, const int *__restrict__ send_mult,
    int *__restrict__ sync_src, int *__restrict__ sync_mult, int *__restrict__ img_cnt) {
    k_copy_back_maps_body(counters, avx, avy, avz, amass, atype, aflags, auid, ashape, aref, acell, vx, vy, vz, mass,
                          type, flags, uid, shape, ref, cell, phys, send_map, send_mult, sync_src, sync_mult, img_cnt);
}

// A segment's tile as physical slots: entry p of the tile is particle
// tslot[p] (the force kernels gather positions from HBM through it, so the
// tile costs 4 B of LDS per entry and many workgroups share a CU).
// Images of each local: every ghost's chain of sources ends at its root
// local; a root gets a slot per image (order irrelevant: positions are
// written to the next buffer from values of the current one).
__device__ __forceinline__ void k_img_build_body(int *__restrict__ counters, const int *__restrict__ sync_src,
                                                   int *__restrict__ img_cnt, int *__restrict__ img_slot, int cap) {
    if (counters[C_OVF]) return;
    const int nlocal = counters[C_NLOCAL], ng = counters[C_NGHOST];
    for (int q = blockIdx.x * 256 + threadIdx.x; q < ng; q += gridDim.x * 256) {
        int root = sync_src[q];
        for (int hop = 0; root >= nlocal && hop < 64; hop++) root = sync_src[root - nlocal];
        if (root >= nlocal) { atomicMax(&counters[C_IMG_OVF], 1); continue; }
        const int k = atomicAdd(&img_cnt[root], 1);
        if (k < IMG_MAX) img_slot[(size_t) k * cap + root] = q;
        else atomicMax(&counters[C_IMG_OVF], 1);
    }
}

__global__ __launch_bounds__(256) void k_img_build(int *__restrict__ counters, const int *__restrict__ sync_src,
                                                   int *__restrict__ img_cnt, int *__restrict__ img_slot, int cap) {
    k_img_build_body(counters, sync_src, img_cnt, img_slot, cap);
}


// ---------------------------------------------------------------- neighbour lists

// BuildNeighborLists (src/pairs/sim/neighbor_lists.py:21-48) over the cell
// stencil (src/pairs/sim/interaction.py:90-120): candidates of a particle in
// cell c are cell 0, then the 27 stencil cells in order, each cell in
// reference-index order; accepted when j != i (full) or i < j (half, point
// masses) and rsq < (rc + skin)^2.
//
// One workgroup of BB threads (BW waves) per segment; each wave takes home
// cells of the segment in turn.  The tile (positions, + reference indices
// for half lists) is staged in LDS and a table gives, per column of the tile,
// each flat cell's local and ghost sub-range (tile offset, length).  For a
// home cell the wave lays out its candidates in the reference's order (cell
// 0; then per stencil column and dz: that cell's locals, then its ghosts) as
// tile offsets in LDS; then each chunk of 64 candidates is held one per lane
// while the home particles of the cell are taken one at a time (broadcast
// from the lane that holds them): a ballot over the chunk and a lane prefix
// count append the accepted candidates in candidate order, written to the
// list directly (the four entries of a group share 16 bytes, so the L2
// merges the appends of a cell's home particles before they reach HBM; an
// LDS row buffer that wrote whole groups cost more in occupancy than it saved).
constexpr int BB = 128;        // threads per build workgroup
constexpr int BW = BB / 64;    // waves per build workgroup
constexpr int NB_CAND = 2048;  // candidates of one home cell
// half lists: mirror entries of one home particle collected in the top
// RMAX slots of the wave's candidate buffer before they are sorted
constexpr int RMAX = 256;
constexpr int REF_MASK = (1 << 30) - 1;  // tref: reference index; bit 30: FIXED

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned) __builtin_amdgcn_readlane((int) b, l);
    const int hi = __builtin_amdgcn_readlane((int) (b >> 32), l);
    return __longlong_as_double(((long long) hi << 32) | lo);
}

template <bool HALF>
__device__ __forceinline__ void build_segment(const BuildArgs &A, int seg) {
    const int *__restrict__ counters = A.counters;
    const Geom &g = A.g;
    const int *__restrict__ lstart = A.lstart, *__restrict__ gstart = A.gstart;
    const double *__restrict__ X = A.X, *__restrict__ Y = A.Y, *__restrict__ Z = A.Z;
    const int *__restrict__ ref = A.ref, *__restrict__ flags = A.flags;
    const double rc2 = A.rc2;
    uint32_t *__restrict__ nl = A.nl;
    const int nl_stride = A.nl_stride, ncap = A.ncap, tcap = A.tcap, zt = A.zt, rev0 = A.rev0;
    int *__restrict__ numneighs = A.numneighs, *__restrict__ ctr = A.counters_rw, *__restrict__ nrev = A.nrev;
    extern __shared__ char smem[];
    int *tslot = (int *) smem;                                  // [tcap] physical slot of each tile entry
    int *ctab = tslot + tcap;                                   // [9][zt][4]: local st, len, ghost st, len
    int *tref = ctab + 9 * zt * 4;                              // [tcap] half lists only
    uint16_t *cand_all = (uint16_t *) (tref + (HALF ? tcap : 0));  // [BW][NB_CAND]
    __shared__ int r_start[NRUNS], r_len[NRUNS], r_off[NRUNS + 1], col_lo[9];
    if (counters[C_OVF]) return;
    const int nlocal = counters[C_NLOCAL];
    if (seg >= g.nsegs) return;
    const SegInfo si = seg_info(seg, g);
    const int h0 = lstart[si.ca], h1 = lstart[si.cb];
    if (h1 <= h0) return;
    seg_runs(si, g, lstart, gstart, nlocal, r_start, r_len, r_off, col_lo);
    const int T = r_off[NRUNS];
    const int nz = si.zb - si.za + 2;  // flat cells B+za-1 .. B+zb of each column
    if (T > tcap || nz > zt) {
        if (threadIdx.x == 0) { atomicMax(&ctr[C_TILE_MAX], T); atomicMax(&ctr[C_OVF], OVF_TILE); }
        return;
    }
    if (threadIdx.x == 0) atomicMax(&ctr[C_TILE_MAX], T);
    stage_slots(r_start, r_off, T, tslot);
    __syncthreads();
    if (HALF)
        for (int p = threadIdx.x; p < T; p += BB)
            tref[p] = ref[tslot[p]] | ((flags[tslot[p]] & FLAG_FIXED) ? (1 << 30) : 0);
    // sub-range table: a flat cell cc of column col is a stencil cell of the
    // reference iff 0 < cc < ncells (interaction.py:92-98)
    for (int e = threadIdx.x; e < 9 * nz; e += BB) {
        const int col = e / nz, zz = e - col * nz;
        const int dx = col / 3 - 1, dy = col % 3 - 1;
        const int cc = ((si.cx + dx) * g.dim_cells[1] + (si.cy + dy)) * g.dim_cells[2] + 1 + si.za - 1 + zz;
        int v0 = 0, v1 = 0, v2 = 0, v3 = 0;
        if (cc > 0 && cc < g.ncells) {
            const int lo = col_lo[col];
            v0 = r_off[2 + 2 * col] + (lstart[cc] - lstart[lo]);
            v1 = lstart[cc + 1] - lstart[cc];
            v2 = r_off[3 + 2 * col] + (gstart[cc] - gstart[lo]);
            v3 = gstart[cc + 1] - gstart[cc];
        }
        int *t = ctab + (col * zt + zz) * 4;
        t[0] = v0; t[1] = v1; t[2] = v2; t[3] = v3;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint16_t *cand = cand_all + wave * NB_CAND;
    // home particles are the locals of the centre column (col 4), table rows
    // 1 .. nz-2; a home particle's index in the segment is its tile offset
    // minus own_base (= its physical slot minus h0)
    const int own_base = r_off[2 + 2 * 4] + (h0 - lstart[col_lo[4]]);
    for (int zz = 1 + wave; zz <= nz - 2; zz += BW) {
        const int *tc = ctab + (4 * zt + zz) * 4;
        const int hst = tc[0], hlen = tc[1];
        if (hlen == 0) continue;
        // candidate sub-ranges, one per lane: cell 0 locals, ghosts; then per
        // column (dx outer, dy inner) and dz = -1, 0, 1: locals, ghosts
        int st = 0, ln = 0;
        if (lane < 2) {
            st = r_off[lane];
            ln = r_len[lane];
        } else if (lane < 56) {
            const int u = lane - 2, col = u / 6, dz = (u % 6) / 2 - 1, gh = u & 1;
            const int *t = ctab + (col * zt + zz + dz) * 4;
            st = t[2 * gh];
            ln = t[2 * gh + 1];
        }
        int incl = ln;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        const int ncand = __shfl(incl, 63, 64);
        if (ncand > (HALF ? NB_CAND - RMAX : NB_CAND)) {
            if (lane == 0) { atomicMax(&ctr[C_TILE_MAX], tcap + 1); atomicMax(&ctr[C_OVF], OVF_TILE); }
            continue;
        }
        for (int k = 0, e0 = incl - ln; k < ln; k++) cand[e0 + k] = (uint16_t) (st + k);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        // the cell's candidates, chunk c on lane c * 64 + lane; the first
        // CREG chunks stay in registers while every home particle tests them
        // (half lists: 6, which leaves room for the mirror bookkeeping at the
        // same occupancy; 272 against 289 us at 131072 atoms)
        constexpr int CREG = HALF ? 6 : 8;
        const int nchunk = (ncand + 63) >> 6;
        double rcx[CREG], rcy[CREG], rcz[CREG];
        int rco[CREG], rcr[CREG];
#pragma unroll
        for (int c = 0; c < CREG; c++) {
            rco[c] = 0; rcr[c] = 0;
            rcx[c] = 0.0; rcy[c] = 0.0; rcz[c] = 0.0;
            if (c < nchunk) {
                const int q = c * 64 + lane;
                const int o = q < ncand ? cand[q] : 0;
                const int pc = tslot[o];
                rco[c] = o;
                rcx[c] = X[pc]; rcy[c] = Y[pc]; rcz[c] = Z[pc];
                if (HALF) rcr[c] = tref[o];
            }
        }
        for (int hb = 0; hb < hlen; hb += 64) {
            const int nh = min(64, hlen - hb);
            // lane h holds home particle h of this pass
            const int oh = hst + hb + min(lane, nh - 1);
            const int ph = tslot[oh];
            const double hx = X[ph], hy = Y[ph], hz = Z[ph];
            const int href = HALF ? (tref[oh] & REF_MASK) : 0;
            const int hfix = flags[h0 + (oh - own_base)] & FLAG_FIXED;
            for (int h = 0; h < nh; h++) {
                const int ownh = hst + hb + h;
                const int idx = ownh - own_base;
                int cnt = 0;   // wave-uniform: entries of home particle h so far
                int cntr = 0;  // half lists: its mirror entries so far
                uint16_t *rbuf = cand + (NB_CAND - RMAX);
                if (!__builtin_amdgcn_readlane(hfix, h)) {
                    const double xi = readlane_f64(hx, h), yi = readlane_f64(hy, h), zi = readlane_f64(hz, h);
                    const int hr = HALF ? __builtin_amdgcn_readlane(href, h) : 0;
                    // one chunk: accepted candidates append in lane (= candidate) order
                    auto test = [&](int c0, int o, double cx, double cy, double cz, int cr) {
                        const bool qv = c0 + lane < ncand;
                        const double dx = xi - cx, dy = yi - cy, dz = zi - cz;
                        double rsq = dx * dx + dy * dy;
                        rsq = rsq + dz * dz;
                        const int crf = cr & REF_MASK;
                        const bool keep = HALF ? (hr < crf) : (o != ownh);
                        const bool acc = qv && keep && rsq < rc2;
                        const unsigned long long m = __ballot(acc);
                        const int k = cnt + lane_rank(m);
                        if (acc && k < ncap) nl[nl_at(k, h0 + idx, nl_stride)] = (uint32_t) tslot[o];
                        cnt += __popcll(m);
                        if (HALF) {
                            // a mirror entry: an earlier-index local, not fixed
                            const bool mir = qv && crf < hr && !(cr >> 30) && rsq < rc2;
                            const unsigned long long mr = __ballot(mir);
                            const int kr = cntr + lane_rank(mr);
                            if (mir && kr < RMAX) rbuf[kr] = (uint16_t) o;
                            cntr += __popcll(mr);
                        }
                    };
#pragma unroll
                    for (int c = 0; c < CREG; c++)
                        if (c < nchunk) test(c * 64, rco[c], rcx[c], rcy[c], rcz[c], rcr[c]);
                    for (int c = CREG; c < nchunk; c++) {  // more candidates than the registers hold
                        const int q = c * 64 + lane;
                        const int o = q < ncand ? cand[q] : 0;
                        const int pc = tslot[o];
                        test(c * 64, o, X[pc], Y[pc], Z[pc], HALF ? tref[o] : 0);
                    }
                }
                if (HALF) {
                    // the mirror entries go to rows rev0.. sorted by reference
                    // index: an entry's rank is the number of entries whose
                    // index is smaller (indices are distinct)
                    if (cntr > RMAX) {
                        if (lane == 0) { atomicMax(&ctr[C_TILE_MAX], tcap + 1); atomicMax(&ctr[C_OVF], OVF_TILE); }
                        cntr = RMAX;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    int KR = cntr;
                    if (KR + 1 >= rev0) {
                        if (lane == 0) atomicMax(&ctr[C_REQ_NEIGH], KR + 2);
                        KR = rev0 - 2;
                    }
                    if (cntr <= 64) {
                        // one entry per lane, its index compared with the
                        // others' through readlane (no memory round trips)
                        const int oe = lane < cntr ? rbuf[lane] : 0;
                        const int re = lane < cntr ? (tref[oe] & REF_MASK) : 0;
                        int rank = 0;
                        for (int q = 0; q < cntr; q++) rank += __builtin_amdgcn_readlane(re, q) < re ? 1 : 0;
                        if (lane < cntr && rank < KR) nl[nl_at(rev0 + rank, h0 + idx, nl_stride)] = (uint32_t) tslot[oe];
                    } else {
                        for (int e = lane; e < cntr; e += 64) {
                            const int oe = rbuf[e];
                            const int re = tref[oe] & REF_MASK;
                            int rank = 0;
                            for (int q = 0; q < cntr; q++) rank += (tref[rbuf[q]] & REF_MASK) < re ? 1 : 0;
                            if (rank < KR) nl[nl_at(rev0 + rank, h0 + idx, nl_stride)] = (uint32_t) tslot[oe];
                        }
                    }
                    if (lane == 0) nrev[h0 + idx] = KR;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                }
                if (lane == 0) {
                    // reference overflow rule (generated resize check): count+1 >= capacity
                    int K = cnt;
                    if (K + 1 >= ncap) {
                        atomicMax(&ctr[C_REQ_NEIGH], K + 2);
                        K = ncap - 2;
                    }
                    numneighs[h0 + idx] = K;
                }
            }
        }
    }
}

template <bool HALF>
__global__ __launch_bounds__(BB) void k_build_neighbors(BuildArgs A) {
    build_segment<HALF>(A, xcd_block(blockIdx.x, gridDim.x));
}

// the segments of many independent systems in one launch: block b belongs to
// system rep[b] and is its segment b - seg_base[rep[b]]
template <bool HALF>
__global__ __launch_bounds__(BB) void k_build_neighbors_batch(const BuildArgs *__restrict__ desc,
                                                              const int *__restrict__ rep,
                                                              const int *__restrict__ seg_base) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r = rep[b];
    build_segment<HALF>(desc[r], b - seg_base[r]);
}

// Expand the lists into reference indices (for download):
// out[physical local][k] = reference index of neighbour k.
__global__ __launch_bounds__(256) void k_expand_neighbors(const int *__restrict__ counters,
                                                          const int *__restrict__ ref,
                                                          const uint32_t *__restrict__ nl, int nl_stride,
                                                          const int *__restrict__ numneighs,
                                                          int *__restrict__ out, int ncap_out) {
    const int nlocal = counters[C_NLOCAL];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nlocal; i += gridDim.x * 256) {
        const int n = numneighs[i];
        for (int k = 0; k < n && k < ncap_out; k++) out[(size_t) i * ncap_out + k] = ref[nl[nl_at(k, i, nl_stride)]];
    }
}

// ---------------------------------------------------------------- Lennard-Jones

// 1 / b rounded to nearest, for a finite b of magnitude in [2^-1000, 2^1000]
// (squared pair distances): Newton-Raphson from the hardware estimate, two
// iterations, then Markstein's final correction q = y + (1 - b y) y with
// both steps fused.  This is the division sequence the compiler emits for
// 1.0 / b with its operand-scaling and special-case fix-up steps left out,
// which are identities in this range, so the bits are the same (and those
// of the correctly rounded quotient the reference's CPU division gives).
__device__ __forceinline__ double recip_rn(double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    const double r = __builtin_fma(-b, y, 1.0);
    return __builtin_fma(r, y, y);
}

// One pair of examples/md.py:5-8 as the reference emits it: delta,
// rsq = dx*dx + dy*dy + dz*dz, sr2 = 1/rsq, sr6 = sr2*sr2*sr2*sigma6,
// s = 48*sr6*(sr6-0.5)*sr2*eps, force contribution delta*s.
// Branch-free: every lane evaluates the pair and a pair that is not counted
// (past the lane's list, or rsq >= rc2) gets s = 0, so its contribution is a
// signed zero.  A running sum that starts at +0.0 is never -0.0, and t + (+-0)
// == t exactly for every other t, so the masked pairs leave the sum
// bit-identical to the reference's conditional add; without branches the
// reciprocals of a step overlap.
// UNIT: sigma6 = epsilon = 1 (one type with the reference's defaults); the
// multiplications by them are then exact identities and are left out.
// OLDER: the pair module of the older API (examples/lj_onetype.py:5-8),
// force[i] += delta * 48.0 * sr6 * (sr6 - 0.5) * sr2 * epsilon evaluated left
// to right, so each component is ((((d * 48) * sr6) * (sr6 - 0.5)) * sr2) * eps.
template <bool ENERGY, bool OLDER = false, bool UNIT = false>
__device__ __forceinline__ bool lj_pair(double xi, double yi, double zi, double xj, double yj, double zj, bool valid,
                                        double rc2, double sg6, double ep, double &g0, double &g1, double &g2,
                                        double &u, double &w) {
    const double dx = xi - xj, dy = yi - yj, dz = zi - zj;
    double rsq = dx * dx + dy * dy;
    rsq = rsq + dz * dz;
    const bool in = valid && rsq < rc2;
    const double sr2 = recip_rn(rsq);
    double sr6 = sr2 * sr2;
    sr6 = sr6 * sr2;
    if (!UNIT) sr6 = sr6 * sg6;
    if (OLDER) {
        const double h = sr6 - 0.5;
        auto term = [&](double d) { return in ? (((d * 48.0) * sr6) * h) * sr2 * ep : 0.0; };
        g0 = term(dx); g1 = term(dy); g2 = term(dz);
    } else {
        double s = 48.0 * sr6;
        s = s * (sr6 - 0.5);
        s = s * sr2;
        if (!UNIT) s = s * ep;
        s = in ? s : 0.0;
        g0 = dx * s; g1 = dy * s; g2 = dz * s;
    }
    if (ENERGY) {
        u = in ? 4.0 * ep * sr6 * (sr6 - 1.0) : 0.0;
        w = in ? 48.0 * sr6 * (sr6 - 0.5) * ep : 0.0;
    }
    return in;
}

// LJ force of the locals of one wave: lanes take 64 consecutive physical
// locals (cell-sorted, so their neighbours are close in memory), each lane
// walks its own list in order and adds its pair forces in list order, so the
// sum is bit-identical to the reference's.  Neighbour positions are gathered
// from the packed copy P (x, y, z, 0 at 32 bytes per particle): one 16-byte
// and one 8-byte load per entry at a 32-bit byte offset.
// MODE 0: forces only (lennard_jones with reset_volatile_properties);
// MODE 1: fused F(ts) + final_integrate(ts) + [thermo] + initial_integrate(ts+1)
//         writing the next positions into the other buffer;
// MODE 2: energy and virial only (no writes but the partials).
// HALF: the system uses half lists.  lennard_jones then adds, for local i,
// each pair of its half list into its own sum T_i and the mirror force of
// every earlier-index local j whose half list holds i into force[i] before
// i's turn, so force[i] = R_i + T_i with R_i the mirror contributions summed
// in reference-index order of j (src/pairs/ir/apply.py:111-125 with the
// sequential loop of the generated CPU code).  Both sums are taken here from
// i's side -- the mirror pair's force recomputed from i is exactly the
// negated force j computes -- so the result is the reference's bit for bit,
// without atomics.
template <int NT, int MODE, bool RED, bool HALF = false, bool OLDER = false, bool UNIT = false>
__device__ __forceinline__ void lj_wave(const LJArgs &A, int wv) {
    const int lane = threadIdx.x & 63;
    const int *__restrict__ counters = A.counters;
    const int nlocal = counters[C_NLOCAL];
    const uint32_t jmax = (uint32_t) max(nlocal + counters[C_NGHOST] - 1, 0);
    const double *__restrict__ X = A.X, *__restrict__ Y = A.Y, *__restrict__ Z = A.Z;
    double *__restrict__ Xn = A.Xn, *__restrict__ Yn = A.Yn, *__restrict__ Zn = A.Zn;
    const int *__restrict__ type = A.type;
    const int nl_stride = A.nl_stride;
    const int *__restrict__ numneighs = A.numneighs, *__restrict__ flags = A.flags;
    double *__restrict__ fx = A.fx, *__restrict__ fy = A.fy, *__restrict__ fz = A.fz;
    double *__restrict__ vx = A.vx, *__restrict__ vy = A.vy, *__restrict__ vz = A.vz;
    const double *__restrict__ mass = A.mass;
    const char *__restrict__ P = (const char *) A.P;
    const double rc2 = A.rc2, dt = A.dt, hdt = A.hdt;
    const LJTab &tab = A.tab;
    const int do_vv2 = A.do_vv2, do_vv1 = A.do_vv1, pstride = A.pstride;
    double *__restrict__ partials = A.partials;
    double *__restrict__ eatom = A.eatom;
    const Images &im = A.im;
    __shared__ double s_eps[16], s_sig6[16];
    if (NT != 1) {
        if (lane < 16) { s_eps[lane] = tab.eps[lane]; s_sig6[lane] = tab.sig6[lane]; }
        __syncthreads();
    }
    const int i0 = wv * 64;
    double ke = 0.0, pe = 0.0, vir = 0.0;
    if (i0 < nlocal) {  // wave-uniform
        const int i = i0 + lane;
        const bool live = i < nlocal;
        const int ii = live ? i : i0;
        const int fl = live ? flags[ii] : FLAG_FIXED;
        const bool moves = !(fl & FLAG_FIXED);
        // counts past the list rows (a build never writes them) are clamped,
        // so no list load leaves the allocation
        const int n = moves ? min(numneighs[ii], A.ncap) : 0;
        const int nmax = __builtin_amdgcn_readfirstlane(wave_max(n));
        const double xi = X[ii], yi = Y[ii], zi = Z[ii];
        const int ti = NT != 1 ? type[ii] : 0;
        const double eps0 = tab.eps[0], sig0 = tab.sig6[0];
        const int nt = tab.nt;
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, r0 = 0.0, r1 = 0.0, r2 = 0.0, u = 0.0, w = 0.0, ue = 0.0;
        // one list part walked in order into (a0, a1, a2).  Energy and
        // virial: a full list weighs every pair 1/2 (u, w are halved below);
        // a half list weighs a local pair 1 and a local-ghost pair 1/2 over
        // its own part only, and its per-atom energy is half the own part's
        // sum (ue)
        auto walk = [&](int row0, int cnt, int cmax, bool mirror, double &a0, double &a1, double &a2) {
            list_walk32<8>(A.nl, row0, ii, nl_stride, cnt, cmax, jmax, [&](int k, const uint32_t *j) {
                double xj[8], yj[8], zj[8];
                int tj[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const uint32_t off = j[q] << 5;
                    const double2 p = *(const double2 *) (P + off);
                    xj[q] = p.x;
                    yj[q] = p.y;
                    zj[q] = *(const double *) (P + off + 16);
                    if (NT != 1) tj[q] = type[j[q]];
                }
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    double sg6 = sig0, ep = eps0;
                    if (NT != 1) { const int kt = ti * nt + tj[q]; sg6 = s_sig6[kt]; ep = s_eps[kt]; }
                    double g0, g1, g2, uu, ww;
                    lj_pair<RED, OLDER, UNIT>(xi, yi, zi, xj[q], yj[q], zj[q], k + q < cnt, rc2, sg6, ep, g0, g1,
                                              g2, uu, ww);
                    a0 = a0 + g0; a1 = a1 + g1; a2 = a2 + g2;
                    if (RED && !HALF) { u += uu; w += ww; }
                    if (RED && HALF && !mirror) {
                        const double wt = j[q] < (uint32_t) nlocal ? 2.0 : 1.0;
                        u += wt * uu; w += wt * ww; ue += uu;
                    }
                }
            });
        };
        walk(0, n, nmax, false, t0, t1, t2);
        if (HALF) {
            const int nr = moves ? min(A.nrev[ii], A.ncap) : 0;
            const int nrmax = __builtin_amdgcn_readfirstlane(wave_max(nr));
            walk(A.rev0, nr, nrmax, true, r0, r1, r2);
        }
        if (live) {
            if (RED) { pe = 0.5 * u; vir = 0.5 * w; }
            if (MODE == 2) {
                if (eatom) eatom[i] = 0.5 * (HALF ? ue : u);
            } else {
                // reset_volatile_properties, the mirror contributions (half
                // lists), then force[i] = force[i] + reduction
                const double f0 = r0 + t0, f1 = r1 + t1, f2 = r2 + t2;
                fx[i] = f0; fy[i] = f1; fz[i] = f2;
                if (MODE == 1) {
                    const double m = mass[i];
                    double v0 = vx[i], v1 = vy[i], v2 = vz[i];
                    double px = xi, py = yi, pz = zi;
                    if (moves && do_vv2) {  // final_integrate(ts)
                        double a0 = hdt * f0, a1 = hdt * f1, a2 = hdt * f2;
                        a0 = a0 / m; a1 = a1 / m; a2 = a2 / m;
                        v0 = v0 + a0; v1 = v1 + a1; v2 = v2 + a2;
                    }
                    if (RED) ke = m * (v0 * v0 + v1 * v1 + v2 * v2);
                    if (moves) {
                        if (do_vv1) {  // initial_integrate(ts+1)
                            double a0 = hdt * f0, a1 = hdt * f1, a2 = hdt * f2;
                            a0 = a0 / m; a1 = a1 / m; a2 = a2 / m;
                            v0 = v0 + a0; v1 = v1 + a1; v2 = v2 + a2;
                            px = xi + dt * v0; py = yi + dt * v1; pz = zi + dt * v2;
                        }
                        vx[i] = v0; vy[i] = v1; vz[i] = v2;
                    }
                    if (do_vv1) {
                        Xn[i] = px; Yn[i] = py; Zn[i] = pz;
                        if (A.Pn) A.Pn[i] = make_double4(px, py, pz, 0.0);
                        if (im.on) write_images(im, i, nlocal, px, py, pz, X, Y, Z, Xn, Yn, Zn, A.g, A.Pn);
                    }
                }
            }
        }
    }
    if (RED) {
        // one partial per wave (kinetic row only from the fused kernel)
        const double a = wave_sum(ke), bb = wave_sum(pe), cc = wave_sum(vir);
        if (lane == 0) {
            if (MODE == 1) partials[wv] = a;
            partials[(size_t) pstride + wv] = bb;
            partials[(size_t) 2 * pstride + wv] = cc;
        }
    }
}

// one wave per workgroup, waves of one system
template <int NT, int MODE, bool RED, bool HALF = false, bool OLDER = false, bool UNIT = false>
__global__ __launch_bounds__(64) void k_lj(LJArgs A) {
    lj_wave<NT, MODE, RED, HALF, OLDER, UNIT>(A, xcd_block(blockIdx.x, gridDim.x));
}

// The 64-bit pattern of lane 2p + 1 in lane 2p (DPP quad_perm 1,1,3,3 on
// both 32-bit halves; the odd lanes receive lane 2p + 1's own value).  Every
// lane is written by a quad permutation, so the move needs no old value (no
// register to clear first).
__device__ __forceinline__ double odd_lane(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int) b, 0xF5, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int) (b >> 32), 0xF5, 0xF, 0xF, true);
    return __longlong_as_double(((long long) hi << 32) | (unsigned int) lo);
}

// the 64-bit value of the other lane of each lane pair (DPP quad_perm 1,0,3,2)
__device__ __forceinline__ double pair_swap(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int) b, 0xB1, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int) (b >> 32), 0xB1, 0xF, 0xF, true);
    return __longlong_as_double(((long long) hi << 32) | (unsigned int) lo);
}

// The force kernel of the timestep (MODE 0: forces; MODE 1: fused with
// final/initial integrate and the next positions) on TWO lanes per local: a
// wave takes 32 consecutive cell-sorted locals, lane 2p + h of local p
// evaluates list entries 4s + h and 4s + 2 + h of step s (one 16-byte group
// of four rows per step), and lane 2p adds the pair terms in list order --
// its own entry, then lane 2p + 1's, received by DPP -- so every sum is the
// reference's, bit for bit, as in lj_wave.  Twice the waves of lj_wave hide
// the gather latency better; the positions of step s + 1 are gathered while
// step s computes (two register sets, a loop body without branches: a last
// odd step is computed fully masked, which adds +0).  The two lanes of a
// local load each neighbour record together: for the entry pair (e, o) one
// 16-byte load has lane 2p read record e's (x, y) and lane 2p + 1 record e's
// (z, 0), the next has lane 2p read record o's (z, 0) and lane 2p + 1 record
// o's (x, y), and the z halves are swapped by DPP, so a record is touched by
// one load instruction (one cache-line lookup) instead of two.
// HALF: the own part (rows [0, n)) into T and the mirror part (rows
// [rev0, rev0 + nrev)) into R, force = R + T, as lj_wave.
// RED: kinetic energy of the locals (lane 2p) and energy / virial of every
// lane's pairs (full lists: weight 1/2; half lists: 1 for local pairs and 1/2
// for local-ghost pairs, own part only) into one partial per wave; the sums
// over a wave's lanes are taken in lane order, so these totals may differ
// from lj_wave's in the last bits.
template <int NT, int MODE, bool RED, bool HALF, bool UNIT>
__device__ __forceinline__ void lj_wave2(const LJArgs &A, int wv) {
    constexpr int U = 2;  // entries per lane per step (4 per local)
    const int lane = threadIdx.x & 63, h = lane & 1;
    const int *__restrict__ counters = A.counters;
    const int nlocal = counters[C_NLOCAL];
    const uint32_t jmax = (uint32_t) max(nlocal + counters[C_NGHOST] - 1, 0);
    const double *__restrict__ X = A.X, *__restrict__ Y = A.Y, *__restrict__ Z = A.Z;
    const int *__restrict__ type = A.type;
    const char *__restrict__ P = (const char *) A.P;
    const double rc2 = A.rc2, dt = A.dt, hdt = A.hdt;
    const LJTab &tab = A.tab;
    __shared__ double s_eps[16], s_sig6[16];
    if (NT != 1) {
        if (lane < 16) { s_eps[lane] = tab.eps[lane]; s_sig6[lane] = tab.sig6[lane]; }
        __syncthreads();
    }
    const int i0 = wv * 32;
    double ke = 0.0, pe = 0.0, vir = 0.0;
    if (i0 < nlocal) {  // wave-uniform
        const int i = i0 + (lane >> 1);
        const bool live = i < nlocal;
        const int ii = live ? i : i0;
        const int fl = live ? A.flags[ii] : FLAG_FIXED;
        const bool moves = !(fl & FLAG_FIXED);
        const double xi = X[ii], yi = Y[ii], zi = Z[ii];
        const int ti = NT != 1 ? type[ii] : 0;
        const double eps0 = tab.eps[0], sig0 = tab.sig6[0];
        const int nt = tab.nt;
        const int stride = A.nl_stride;
        double u = 0.0, w = 0.0;
        struct Pos { double2 a[U], b[U]; int t[U]; uint32_t j[U]; };
        // j[2q] / j[2q + 1]: the even / odd entry of pair q (both lanes)
        auto gather = [&](const uint32_t (&j)[2 * U], Pos &p) {
#pragma unroll
            for (int q = 0; q < U; q++) {
                p.a[q] = *(const double2 *) (P + (j[2 * q] << 5) + (h << 4));
                p.b[q] = *(const double2 *) (P + (j[2 * q + 1] << 5) + ((h ^ 1) << 4));
                const uint32_t jo = h ? j[2 * q + 1] : j[2 * q];
                if (NT != 1) p.t[q] = type[jo];
                p.j[q] = jo;
            }
        };
        // one list part, rows [row0, row0 + cnt), summed in order into (a0, a1, a2) of lane 2p
        auto walk = [&](int row0, int cnt, bool mirror, double &a0, double &a1, double &a2) {
            const int nmax = __builtin_amdgcn_readfirstlane(wave_max(cnt));
            const int nst = (nmax + 2 * U - 1) / (2 * U);
            if (nst <= 0) return;  // wave-uniform
            const uint4 *lp = (const uint4 *) A.nl + (size_t) (row0 >> 2) * stride + ii;
            const int gl = max((cnt + 3) / 4 - 1, 0);
            auto ldl = [&](int st, uint32_t (&j)[2 * U]) {
                const uint4 g = lp[(size_t) min(st, gl) * stride];
                j[0] = min(g.x, jmax);
                j[1] = min(g.y, jmax);
                j[2] = min(g.z, jmax);
                j[3] = min(g.w, jmax);
            };
            auto compute = [&](int st, const Pos &p) {
#pragma unroll
                for (int q = 0; q < U; q++) {
                    double sg6 = sig0, ep = eps0;
                    if (NT != 1) { const int kt = ti * nt + p.t[q]; sg6 = s_sig6[kt]; ep = s_eps[kt]; }
                    double g0, g1, g2, uu, ww;
                    const double2 xy = h ? p.b[q] : p.a[q];
                    const double zj = pair_swap(h ? p.a[q].x : p.b[q].x);
                    lj_pair<RED, false, UNIT>(xi, yi, zi, xy.x, xy.y, zj, 4 * st + 2 * q + h < cnt, rc2, sg6,
                                              ep, g0, g1, g2, uu, ww);
                    // lane 2p: entry 4s + 2q (its own), then 4s + 2q + 1 (lane 2p + 1's)
                    a0 = a0 + g0; a1 = a1 + g1; a2 = a2 + g2;
                    a0 = a0 + odd_lane(g0); a1 = a1 + odd_lane(g1); a2 = a2 + odd_lane(g2);
                    if (RED && !HALF) { u += uu; w += ww; }
                    if (RED && HALF && !mirror) {
                        const double wt = p.j[q] < (uint32_t) nlocal ? 2.0 : 1.0;
                        u += wt * uu; w += wt * ww;
                    }
                }
            };
            uint32_t LA[2 * U], LB[2 * U];
            Pos PA, PB;
            ldl(0, LA);
            ldl(1, LB);
            gather(LA, PA);
            ldl(2, LA);
            for (int st = 0; st < nst; st += 2) {
                gather(LB, PB);
                ldl(st + 3, LB);
                compute(st, PA);
                gather(LA, PA);
                ldl(st + 4, LA);
                compute(st + 1, PB);
            }
        };
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, r0 = 0.0, r1 = 0.0, r2 = 0.0;
        // counts past the list rows (a build never writes them) are clamped
        walk(0, moves ? min(A.numneighs[ii], A.ncap) : 0, false, t0, t1, t2);
        if (HALF) walk(A.rev0, moves ? min(A.nrev[ii], A.ncap) : 0, true, r0, r1, r2);
        if (RED) { pe = 0.5 * u; vir = 0.5 * w; }
        if (live && h == 0) {
            const double f0 = r0 + t0, f1 = r1 + t1, f2 = r2 + t2;
            A.fx[i] = f0; A.fy[i] = f1; A.fz[i] = f2;
            if (MODE == 1) {
                const double m = A.mass[i];
                double v0 = A.vx[i], v1 = A.vy[i], v2 = A.vz[i];
                double px = xi, py = yi, pz = zi;
                if (moves && A.do_vv2) {  // final_integrate(ts)
                    double b0 = hdt * f0, b1 = hdt * f1, b2 = hdt * f2;
                    b0 = b0 / m; b1 = b1 / m; b2 = b2 / m;
                    v0 = v0 + b0; v1 = v1 + b1; v2 = v2 + b2;
                }
                if (RED) ke = m * (v0 * v0 + v1 * v1 + v2 * v2);
                if (moves) {
                    if (A.do_vv1) {  // initial_integrate(ts+1)
                        double b0 = hdt * f0, b1 = hdt * f1, b2 = hdt * f2;
                        b0 = b0 / m; b1 = b1 / m; b2 = b2 / m;
                        v0 = v0 + b0; v1 = v1 + b1; v2 = v2 + b2;
                        px = xi + dt * v0; py = yi + dt * v1; pz = zi + dt * v2;
                    }
                    A.vx[i] = v0; A.vy[i] = v1; A.vz[i] = v2;
                }
                if (A.do_vv1) {
                    A.Xn[i] = px; A.Yn[i] = py; A.Zn[i] = pz;
                    if (A.Pn) A.Pn[i] = make_double4(px, py, pz, 0.0);
                    if (A.im.on) write_images(A.im, i, nlocal, px, py, pz, X, Y, Z, A.Xn, A.Yn, A.Zn, A.g, A.Pn);
                }
            }
        }
    }
    if (RED) {
        // one partial per wave (kinetic row from the fused kernel only)
        const double a = wave_sum(ke), bb = wave_sum(pe), cc = wave_sum(vir);
        if (lane == 0) {
            if (MODE == 1) A.partials[wv] = a;
            A.partials[(size_t) A.pstride + wv] = bb;
            A.partials[(size_t) 2 * A.pstride + wv] = cc;
        }
    }
}

template <int NT, int MODE, bool RED, bool HALF, bool UNIT>
__global__ __launch_bounds__(64) void k_lj2(LJArgs A) {
    lj_wave2<NT, MODE, RED, HALF, UNIT>(A, xcd_block(blockIdx.x, gridDim.x));
}

// the waves of many independent systems in one persistent launch: the grid
// is what the device holds resident at once, and workgroup w takes the
// consecutive table entries [w * per, (w + 1) * per); entry b belongs to
// system rep[b] and is its wave b - wave_base[rep[b]] (one lane per local, 64
// locals per wave).  The XCD-aware order of w gives each XCD one contiguous
// eighth of the table.
template <int NT, int MODE, bool RED, bool UNIT>
__global__ __launch_bounds__(64) void k_lj_batch(const LJArgs *__restrict__ desc, const int *__restrict__ rep,
                                                 const int *__restrict__ wave_base, int nblocks) {
    const int G = gridDim.x;
    const int per = (nblocks + G - 1) / G;
    const int b0 = xcd_block(blockIdx.x, G) * per, b1 = min(b0 + per, nblocks);
    for (int b = b0; b < b1; b++) {
        const int r = rep[b];
        lj_wave<NT, MODE, RED, false, false, UNIT>(desc[r], b - wave_base[r]);
    }
}

// euler of the older API (examples/lj_onetype.py:11-13) for the non-fixed
// locals: velocity += dt * force / mass, then position += dt * velocity, in
// the current position buffer
__global__ __launch_bounds__(256) void k_euler_md(double *__restrict__ X, double *__restrict__ Y,
                                                  double *__restrict__ Z, double *__restrict__ vx,
                                                  double *__restrict__ vy, double *__restrict__ vz,
                                                  const double *__restrict__ fx, const double *__restrict__ fy,
                                                  const double *__restrict__ fz, const double *__restrict__ mass,
                                                  const int *__restrict__ flags, int nlocal, double dt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nlocal || (flags[i] & FLAG_FIXED)) return;
    const double m = mass[i];
    const double v0 = vx[i] + (dt * fx[i]) / m, v1 = vy[i] + (dt * fy[i]) / m, v2 = vz[i] + (dt * fz[i]) / m;
    vx[i] = v0; vy[i] = v1; vz[i] = v2;
    X[i] = X[i] + dt * v0; Y[i] = Y[i] + dt * v1; Z[i] = Z[i] + dt * v2;
}

void launch_euler_md(hipStream_t s, const DevState &d, double dt) {
    if (d.nlocal <= 0) return;
    const int c = d.cur;
    hipLaunchKernelGGL(k_euler_md, dim3((d.nlocal + 255) / 256), dim3(256), 0, s, d.X[c], d.Y[c], d.Z[c], d.vx, d.vy,
                       d.vz, d.fx, d.fy, d.fz, d.mass, d.flags, d.nlocal, dt);
}

// forces only with the older API's pair expression (full lists)
void launch_lj_older(hipStream_t s, const DevState &d, const Geom &g, const LJParams &prm) {
    if (d.nlocal <= 0) return;
    launch_pack_p(s, d, d.cur);
    const LJArgs A = lj_args(d, g, prm, d.cur, d.cur, 0, 0, 0, nullptr);
    const dim3 grid(force_waves(d));
    if (prm.ntypes == 1) hipLaunchKernelGGL((k_lj<1, 0, false, false, true>), grid, dim3(64), 0, s, A);
    else hipLaunchKernelGGL((k_lj<0, 0, false, false, true>), grid, dim3(64), 0, s, A);
}

// the packed (x, y, z, 0) copy the force kernels gather from
__global__ __launch_bounds__(256) void k_pack_p(const int *__restrict__ counters, const double *__restrict__ X,
                                                const double *__restrict__ Y, const double *__restrict__ Z,
                                                double4 *__restrict__ P) {
    const int n = counters[C_NLOCAL] + counters[C_NGHOST];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) P[i] = make_double4(X[i], Y[i], Z[i], 0.0);
}

// ---------------------------------------------------------------- integration

// initial_integrate (examples/md.py:11-13): v += (dt*0.5)*f/m; x += dt*v
__global__ __launch_bounds__(256) void k_vv_initial(double *__restrict__ X, double *__restrict__ Y,
                                                    double *__restrict__ Z, double *__restrict__ vx,
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
    X[i] = X[i] + dt * v0; Y[i] = Y[i] + dt * v1; Z[i] = Z[i] + dt * v2;
}

// final_integrate (examples/md.py:16-17), optionally followed by the thermo
// partial and the next step's initial_integrate written into (Xn, Yn, Zn).
template <bool RED>
__global__ __launch_bounds__(256) void k_vv_post(const double *__restrict__ X, const double *__restrict__ Y,
                                                 const double *__restrict__ Z, double *__restrict__ Xn,
                                                 double *__restrict__ Yn, double *__restrict__ Zn,
                                                 double *__restrict__ vx, double *__restrict__ vy,
                                                 double *__restrict__ vz, const double *__restrict__ fx,
                                                 const double *__restrict__ fy, const double *__restrict__ fz,
                                                 const double *__restrict__ mass, const int *__restrict__ flags,
                                                 int nlocal, double dt, double hdt, int do_vv2, int do_vv1,
                                                 double *__restrict__ partials, Images im, Geom g,
                                                 double4 *__restrict__ Pn) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    double ke = 0.0;
    if (i < nlocal) {
        const int fl = flags[i];
        const double m = mass[i];
        double v0 = vx[i], v1 = vy[i], v2 = vz[i];
        double px = X[i], py = Y[i], pz = Z[i];
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
                px = px + dt * v0; py = py + dt * v1; pz = pz + dt * v2;
            }
            vx[i] = v0; vy[i] = v1; vz[i] = v2;
        } else if (RED) {
            ke = m * (v0 * v0 + v1 * v1 + v2 * v2);
        }
        if (do_vv1) {
            Xn[i] = px; Yn[i] = py; Zn[i] = pz;
            if (Pn) Pn[i] = make_double4(px, py, pz, 0.0);
            if (im.on) write_images(im, i, nlocal, px, py, pz, X, Y, Z, Xn, Yn, Zn, g, Pn);
        }
    }
    if (RED) {
        __shared__ double red[4];
        const double a = wave_sum(ke);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
        __syncthreads();
        if (threadIdx.x == 0) partials[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
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
    if (threadIdx.x == 0) partials[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Fixed-order sums of the per-block partials (row 0 kinetic over nb_ke blocks,
// rows 1-2 energy/virial over nb_f blocks); thermo as runtime/thermo.hpp:11-51.
__device__ __forceinline__ void thermo_reduce_body(const double *__restrict__ partials, int pstride, int nb_ke,
                                                   int nb_f, int which_mask, int natoms, double lx, double ly,
                                                   double lz, double *__restrict__ out) {
    __shared__ double red[3][4];
    double s[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb_ke; b += 256) s[0] += partials[b];
    for (int b = threadIdx.x; b < nb_f; b += 256) { s[1] += partials[pstride + b]; s[2] += partials[2 * pstride + b]; }
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
        if (which_mask & 2) { out[2] = pe; out[3] = w; }
        if (which_mask == 3) out[4] = p + w / (3.0 * lx * ly * lz);
    }
}

__global__ __launch_bounds__(256) void k_thermo_reduce(const double *__restrict__ partials, int pstride, int nb_ke,
                                                       int nb_f, int which_mask, int natoms, double lx, double ly,
                                                       double lz, double *__restrict__ out) {
    thermo_reduce_body(partials, pstride, nb_ke, nb_f, which_mask, natoms, lx, ly, lz, out);
}

// the thermo rows of many systems after a fused step with reduction
// (kinetic energy per segment): out[5 r + k], one workgroup per system
__global__ __launch_bounds__(256) void k_thermo_reduce_batch(const RepDesc *__restrict__ D, double *__restrict__ out) {
    const RepDesc &R = D[blockIdx.x];
    const int nw = (R.d.nlocal + 63) / 64;  // waves of the batched force kernel (lj_wave)
    thermo_reduce_body(R.d.partials, R.d.pstride, nw, nw, 3, R.d.nlocal, R.g.len[0], R.g.len[1], R.g.len[2],
                       out + 5 * (size_t) blockIdx.x);
}

void launch_thermo_batch(hipStream_t s, const RepDesc *D, int n, double *out) {
    if (n > 0) hipLaunchKernelGGL(k_thermo_reduce_batch, dim3(n), dim3(256), 0, s, D, out);
}

// ---------------------------------------------------------------- ghosts / exchange

// Selection predicate of DimensionRanges.ghost_particles
// (src/pairs/sim/domain_partitioning.py:31-65): side 0 "x_d < min_d + offset"
// (shift +L), side 1 "x_d > max_d - offset" (shift -L), skipping
// Infinite|Global particles.  offset 0 selects leavers (exchange).
__device__ __forceinline__ void ghost_sel(double x, int fl, int dim, const Geom &g, bool exch, bool &s0, bool &s1) {
    const bool ok = !(fl & (FLAG_INFINITE | FLAG_GLOBAL));
    if (exch) { s0 = ok && x < g.lo[dim]; s1 = ok && x > g.hi[dim]; }
    else { s0 = ok && x < g.lo[dim] + g.spacing; s1 = ok && x > g.hi[dim] - g.spacing; }
}

// pass 1 of the stable two-sided compaction over reference indices: per-tile counts
// The exchange of a dimension selects among the locals and appends at send
// map entry 0; the borders of the first periodic dimension start with no
// ghosts (comm.py:105-118), those of a later one select among locals and the
// ghosts so far and append behind them.  The range and the base are kept for
// pass 2 (nothing writes them while pass 1 runs).
__device__ __forceinline__ void k_sel_count_body(const double *__restrict__ Xd, const int *__restrict__ flags,
                                                   const int *__restrict__ phys, int *__restrict__ counters, Geom g,
                                                   int dim, int exch, int first, int *__restrict__ tile_cnt,
                                                   int *__restrict__ exchg_flag) {
    __shared__ int wc[2][4];
    if (counters[C_OVF]) return;
    const bool fresh = exch || first;
    const int n = counters[C_NLOCAL] + (fresh ? 0 : counters[C_NGHOST]);
    const int ntiles = (n + 255) / 256;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[C_RANGE] = n;
        counters[C_BASE0] = fresh ? 0 : counters[C_NSEND_ALL];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int r = t * 256 + threadIdx.x;
        bool s0 = false, s1 = false;
        if (r < n) {
            const int p = phys[r];
            ghost_sel(Xd[p], flags[p], dim, g, exch, s0, s1);
            if (exch) exchg_flag[r] = (s0 || s1) ? 1 : 0;
        }
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

__global__ __launch_bounds__(256) void k_sel_count(const double *__restrict__ Xd, const int *__restrict__ flags,
                                                   const int *__restrict__ phys, int *__restrict__ counters, Geom g,
                                                   int dim, int exch, int first, int *__restrict__ tile_cnt,
                                                   int *__restrict__ exchg_flag) {
    k_sel_count_body(Xd, flags, phys, counters, g, dim, exch, first, tile_cnt, exchg_flag);
}

// sums of two per-thread values over a workgroup of 256 threads
__device__ __forceinline__ void block_sum2_256(int a, int b, int *lds, int &sa, int &sb) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    __syncthreads();
    if (lane == 0) { lds[2 * wave] = a; lds[2 * wave + 1] = b; }
    __syncthreads();
    sa = lds[0] + lds[2] + lds[4] + lds[6];
    sb = lds[1] + lds[3] + lds[5] + lds[7];
}

// pass 2: stable scatter into the send map (reference indices); for borders
// the ghost is created in the same pass (pack + communicateData self-copy +
// unpack of comm.py:291-366: uid, type, mass, position + shift, velocity,
// shape) at physical slot = reference index nlocal + g.  Ghost flags are not
// sent by the reference; a ghost slot's flags read 0 (zeroed memory).
// Side 0 entries precede side 1.  Every workgroup sums the tile counts of
// pass 1 itself (totals, and its tiles' offsets: the counts of the tiles
// before them), so the offsets need no pass of their own; all workgroups
// see the same totals and take the same capacity decision, and workgroup 0
// updates the counts, which no other workgroup of this pass reads.
__device__ __forceinline__ void k_sel_scatter_body(double *__restrict__ X, double *__restrict__ Y,
                                                     double *__restrict__ Z, double *__restrict__ vx,
                                                     double *__restrict__ vy, double *__restrict__ vz,
                                                     double *__restrict__ mass, int *__restrict__ type,
                                                     int *__restrict__ uid, int *__restrict__ shape,
                                                     int *__restrict__ flags, int *__restrict__ ref,
                                                     int *__restrict__ phys, int *__restrict__ counters, Geom g,
                                                     int dim, int exch, const int *__restrict__ tile_cnt,
                                                     int *__restrict__ send_map, int *__restrict__ send_mult,
                                                     int send_cap, int cap) {
    __shared__ int wc[2][5];
    __shared__ int red[8];
    if (counters[C_OVF]) return;
    const int n = counters[C_RANGE];
    const int ntiles = (n + 255) / 256;
    if (blockIdx.x > 0 && (int) blockIdx.x >= ntiles) return;  // workgroup 0 always writes the counts
    const int nlocal = counters[C_NLOCAL];
    const int base = counters[C_BASE0];
    int t0 = 0, t1 = 0;
    for (int u = threadIdx.x; u < ntiles; u += 256) { t0 += tile_cnt[2 * u]; t1 += tile_cnt[2 * u + 1]; }
    int T0, T1;
    block_sum2_256(t0, t1, red, T0, T1);
    const int nall = base + T0 + T1;
    const bool ovf_send = nall > send_cap, ovf_part = !exch && nlocal + nall > cap;
    if (ovf_send || ovf_part) {
        // the counts stay within the maps (steps may still run on them
        // before the host sees the overflow)
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            if (ovf_send) {
                counters[C_OVF] = OVF_SEND;
                counters[C_REQ_SEND] = max(counters[C_REQ_SEND], nall);
            }
            if (ovf_part) {
                counters[C_OVF] = OVF_PART;
                counters[C_REQ_PART] = max(counters[C_REQ_PART], nlocal + nall);
            }
        }
        return;
    }
    const double *Xd = dim == 0 ? X : (dim == 1 ? Y : Z);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        int q0 = 0, q1 = 0;
        for (int u = threadIdx.x; u < t; u += 256) { q0 += tile_cnt[2 * u]; q1 += tile_cnt[2 * u + 1]; }
        int P0, P1;
        block_sum2_256(q0, q1, red, P0, P1);
        const int r = t * 256 + threadIdx.x;
        bool s0 = false, s1 = false;
        int p = 0;
        if (r < n) { p = phys[r]; ghost_sel(Xd[p], flags[p], dim, g, exch, s0, s1); }
        const unsigned long long m0 = __ballot(s0), m1 = __ballot(s1);
        if (lane == 0) { wc[0][wave] = __popcll(m0); wc[1][wave] = __popcll(m1); }
        __syncthreads();
        if (threadIdx.x == 0) {
            int a = 0, b = 0;
            for (int w = 0; w < 4; w++) { int x = wc[0][w]; wc[0][w] = a; a += x; int y = wc[1][w]; wc[1][w] = b; b += y; }
        }
        __syncthreads();
        // each side on its own: in a box shorter than two cell spacings a
        // particle near both faces of the dimension is sent to both sides
        // (the reference's two selection loops, domain_partitioning.py:31-65)
        for (int side = 0; side < 2; side++) {
            if (!(side == 0 ? s0 : s1)) continue;
            const int rk = side == 0 ? lane_rank(m0) : lane_rank(m1);
            const int pos = base + (side == 0 ? P0 : T0 + P1) + wc[side][wave] + rk;
            const int mult = side == 0 ? 1 : -1;  // pbc_prev = +1, pbc_next = -1 on one rank
            const int mx = dim == 0 ? mult : 0, my = dim == 1 ? mult : 0, mz = dim == 2 ? mult : 0;
            send_map[pos] = r;
            send_mult[3 * pos + 0] = mx;
            send_mult[3 * pos + 1] = my;
            send_mult[3 * pos + 2] = mz;
            if (!exch) {
                const int gp = nlocal + pos;
                X[gp] = X[p] + (double) mx * g.len[0];
                Y[gp] = Y[p] + (double) my * g.len[1];
                Z[gp] = Z[p] + (double) mz * g.len[2];
                vx[gp] = vx[p]; vy[gp] = vy[p]; vz[gp] = vz[p];
                mass[gp] = mass[p]; type[gp] = type[p]; uid[gp] = uid[p]; shape[gp] = shape[p];
                flags[gp] = 0;
                ref[gp] = gp;
                phys[gp] = gp;
            }
        }
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[C_NSEND + 2 * dim] = T0;
        counters[C_NSEND + 2 * dim + 1] = T1;
        counters[C_NSEND_ALL] = nall;
        if (!exch) counters[C_NGHOST] = nall;
    }
}

__global__ __launch_bounds__(256) void k_sel_scatter(double *__restrict__ X, double *__restrict__ Y,
                                                     double *__restrict__ Z, double *__restrict__ vx,
                                                     double *__restrict__ vy, double *__restrict__ vz,
                                                     double *__restrict__ mass, int *__restrict__ type,
                                                     int *__restrict__ uid, int *__restrict__ shape,
                                                     int *__restrict__ flags, int *__restrict__ ref,
                                                     int *__restrict__ phys, int *__restrict__ counters, Geom g,
                                                     int dim, int exch, const int *__restrict__ tile_cnt,
                                                     int *__restrict__ send_map, int *__restrict__ send_mult,
                                                     int send_cap, int cap) {
    k_sel_scatter_body(X, Y, Z, vx, vy, vz, mass, type, uid, shape, flags, ref, phys, counters, g, dim, exch,
                       tile_cnt, send_map, send_mult, send_cap, cap);
}

// Exchange of one dimension after its send map is built, as a relabeling:
// the reference packs the leavers (position + shift), fills their holes from
// the tail (remove_exchanged_particles_pt1/pt2, copy guard "src > 0") and
// appends them at nlocal - n (src/pairs/sim/comm.py:100-151, 445-520).  Here
// the particles stay in their physical slots: leavers get their shift and
// the reference index the reference would give them, tail particles moved
// into a hole get that hole's index.  pt1's sequential tail walk gives the
// r-th exchanged particle below nlocal - n the r-th non-exchanged reference
// index below nlocal (descending).  One workgroup: a rebuild moves few particles.
__device__ __forceinline__ void k_exchange_relabel_body(double *__restrict__ X, double *__restrict__ Y,
                                                           double *__restrict__ Z, int *__restrict__ ref,
                                                           int *__restrict__ phys, const int *__restrict__ counters,
                                                           Geom g, const int *__restrict__ send_map,
                                                           const int *__restrict__ send_mult,
                                                           const int *__restrict__ exchg_flag, int *__restrict__ tail,
                                                           int *__restrict__ copy_to, int *__restrict__ p_leaver,
                                                           int *__restrict__ p_src) {
    __shared__ int lds[20];
    __shared__ int s_tail[1024];
    if (counters[C_OVF]) return;
    const int n = counters[C_NSEND_ALL];
    if (n == 0) return;
    const int nlocal = counters[C_NLOCAL];
    if (n <= 1024) {
        // one leaver per thread (the usual case): the tail list in LDS and
        // each leaver's entries in registers, the same steps as below
        const int t = threadIdx.x;
        const bool in = t < n;
        const int keep = (in && exchg_flag[nlocal - 1 - t] == 0) ? 1 : 0;
        int total;
        const int ex = block_exclusive_scan<1024>(keep, lds, &total);
        if (keep) s_tail[ex] = nlocal - 1 - t;
        const int sm = in ? send_map[t] : 0;
        const int q = (in && sm < nlocal - n) ? 1 : 0;
        const int eq = block_exclusive_scan<1024>(q, lds, &total);  // (its barriers publish s_tail)
        const int src = q ? s_tail[eq] : -1;
        const int pl = in ? phys[sm] : 0;
        const int ps = src > 0 ? phys[src] : -1;
        __syncthreads();
        if (ps >= 0) { ref[ps] = sm; phys[sm] = ps; }
        __syncthreads();
        if (in) {
            X[pl] = X[pl] + (double) send_mult[3 * t + 0] * g.len[0];
            Y[pl] = Y[pl] + (double) send_mult[3 * t + 1] * g.len[1];
            Z[pl] = Z[pl] + (double) send_mult[3 * t + 2] * g.len[2];
            const int nr = nlocal - n + t;
            ref[pl] = nr;
            phys[nr] = pl;
        }
        return;
    }
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
    // read the current physical slots before relabeling anything
    for (int e = threadIdx.x; e < n; e += 1024) {
        p_leaver[e] = phys[send_map[e]];
        const int src = copy_to[e];
        p_src[e] = src > 0 ? phys[src] : -1;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += 1024) {
        const int ps = p_src[e];
        if (ps >= 0) { const int dst = send_map[e]; ref[ps] = dst; phys[dst] = ps; }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += 1024) {
        const int p = p_leaver[e];
        X[p] = X[p] + (double) send_mult[3 * e + 0] * g.len[0];
        Y[p] = Y[p] + (double) send_mult[3 * e + 1] * g.len[1];
        Z[p] = Z[p] + (double) send_mult[3 * e + 2] * g.len[2];
        const int nr = nlocal - n + e;
        ref[p] = nr;
        phys[nr] = p;
    }
}

__global__ __launch_bounds__(1024) void k_exchange_relabel(double *__restrict__ X, double *__restrict__ Y,
                                                           double *__restrict__ Z, int *__restrict__ ref,
                                                           int *__restrict__ phys, const int *__restrict__ counters,
                                                           Geom g, const int *__restrict__ send_map,
                                                           const int *__restrict__ send_mult,
                                                           const int *__restrict__ exchg_flag, int *__restrict__ tail,
                                                           int *__restrict__ copy_to, int *__restrict__ p_leaver,
                                                           int *__restrict__ p_src) {
    k_exchange_relabel_body(X, Y, Z, ref, phys, counters, g, send_map, send_mult, exchg_flag, tail, copy_to, p_leaver, p_src);
}

// Comm.synchronize with the stale multi-hop semantics of the reference (all
// ghosts are packed before any is unpacked, comm.py:45-54): a ghost whose
// source is local takes the source's current position; a ghost whose source
// is another ghost takes that ghost's previous value.  Double-buffered form:
// (X, Y, Z)[cur] holds this step's locals, [prev] the previous ghosts.
__global__ __launch_bounds__(256) void k_sync_pp(double *__restrict__ Xc, double *__restrict__ Yc,
                                                 double *__restrict__ Zc, const double *__restrict__ Xp,
                                                 const double *__restrict__ Yp, const double *__restrict__ Zp,
                                                 const int *__restrict__ counters, const int *__restrict__ sync_src,
                                                 const int *__restrict__ sync_mult, Geom g,
                                                 double4 *__restrict__ Pc) {
    const int n = counters[C_NSEND_ALL];
    const int nlocal = counters[C_NLOCAL];
    for (int q = blockIdx.x * 256 + threadIdx.x; q < n; q += gridDim.x * 256) {
        const int src = sync_src[q];
        const bool loc = src < nlocal;
        const double x = loc ? Xc[src] : Xp[src];
        const double y = loc ? Yc[src] : Yp[src];
        const double z = loc ? Zc[src] : Zp[src];
        const double xg = x + (double) sync_mult[3 * q + 0] * g.len[0];
        const double yg = y + (double) sync_mult[3 * q + 1] * g.len[1];
        const double zg = z + (double) sync_mult[3 * q + 2] * g.len[2];
        Xc[nlocal + q] = xg;
        Yc[nlocal + q] = yg;
        Zc[nlocal + q] = zg;
        Pc[nlocal + q] = make_double4(xg, yg, zg, 0.0);
    }
}

// single-buffer form: pack every ghost into a staging buffer, then unpack
__global__ __launch_bounds__(256) void k_sync_pack(const double *__restrict__ X, const double *__restrict__ Y,
                                                   const double *__restrict__ Z, const double *__restrict__ vx,
                                                   const double *__restrict__ vy, const double *__restrict__ vz,
                                                   const int *__restrict__ counters, const int *__restrict__ sync_src,
                                                   const int *__restrict__ sync_mult, Geom g,
                                                   double *__restrict__ stage, int with_vel) {
    const int n = counters[C_NSEND_ALL];
    for (int q = blockIdx.x * 256 + threadIdx.x; q < n; q += gridDim.x * 256) {
        const int src = sync_src[q];
        double *s = stage + (size_t) q * 6;
        s[0] = X[src] + (double) sync_mult[3 * q + 0] * g.len[0];
        s[1] = Y[src] + (double) sync_mult[3 * q + 1] * g.len[1];
        s[2] = Z[src] + (double) sync_mult[3 * q + 2] * g.len[2];
        if (with_vel) { s[3] = vx[src]; s[4] = vy[src]; s[5] = vz[src]; }
    }
}
__global__ __launch_bounds__(256) void k_sync_unpack(double *__restrict__ X, double *__restrict__ Y,
                                                     double *__restrict__ Z, double *__restrict__ vx,
                                                     double *__restrict__ vy, double *__restrict__ vz,
                                                     const int *__restrict__ counters, const double *__restrict__ stage,
                                                     int with_vel) {
    const int n = counters[C_NSEND_ALL];
    const int nlocal = counters[C_NLOCAL];
    for (int q = blockIdx.x * 256 + threadIdx.x; q < n; q += gridDim.x * 256) {
        const double *s = stage + (size_t) q * 6;
        const int gs = nlocal + q;
        X[gs] = s[0]; Y[gs] = s[1]; Z[gs] = s[2];
        if (with_vel) { vx[gs] = s[3]; vy[gs] = s[4]; vz[gs] = s[5]; }
    }
}

// ---------------------------------------------------------------- launchers

void launch_fill_i32(hipStream_t s, int *p, int v, int n) {
    if (n <= 0) return;
    const int blocks = min((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, s, p, v, n);
}
void launch_fill_f64(hipStream_t s, double *p, double v, int n) {
    if (n <= 0) return;
    const int blocks = min((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_f64, dim3(blocks), dim3(256), 0, s, p, v, n);
}
void launch_identity(hipStream_t s, int *p, int n) {
    if (n <= 0) return;
    const int blocks = min((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_identity, dim3(blocks), dim3(256), 0, s, p, n);
}
void launch_set_counters(hipStream_t s, int *c, int first, int count, int v) {
    hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, s, c, first, count, v);
}

void launch_checkpoint(hipStream_t s, const CkptArrays &a, const CkptArrays &b, int n) {
    hipLaunchKernelGGL(k_checkpoint, dim3(std::max(1, std::min((n + 255) / 256, 2048))), dim3(256), 0, s, a, b, n);
}

void launch_gather_counters(hipStream_t s, const int *const *src, int n, int *dst) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_counters, dim3((n * C_COUNT + 255) / 256), dim3(256), 0, s, src, n, dst);
}

static int tile_grid(int cap) { return max(1, min((cap + 255) / 256, 2048)); }

// (the cell counts are zero on entry: allocated zeroed, cleared by every scan)
void launch_bin_sort(hipStream_t s, DevState &d, const Geom &g) {
    const int grid = tile_grid(d.cap);
    const int c = d.cur, o = 1 - d.cur;
    hipLaunchKernelGGL(k_bin_count, dim3(grid), dim3(256), 0, s, d.X[c], d.Y[c], d.Z[c], d.flags, d.counters, g,
                       d.particle_cell, d.lcount, d.gcount, d.slot);
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, s, d.counters, d.lcount, d.gcount, d.lstart, d.gstart,
                       g.ncells);
    hipLaunchKernelGGL(k_bin_scatter, dim3(grid), dim3(256), 0, s, d.counters, d.particle_cell, d.slot, d.lstart,
                       d.gstart, d.perm);
    hipLaunchKernelGGL(k_sort_cells, dim3((2 * g.ncells + 3) / 4), dim3(256), 0, s, d.counters, d.lstart, d.gstart,
                       d.ref, d.perm, g.ncells);
    hipLaunchKernelGGL(k_permute, dim3(grid), dim3(256), 0, s, d.counters, d.perm, d.X[c], d.Y[c], d.Z[c], d.X[o],
                       d.Y[o], d.Z[o], d.vx, d.vy, d.vz, d.mass, d.type, d.flags, d.uid, d.shape, d.ref,
                       d.particle_cell, d.alt_vx, d.alt_vy, d.alt_vz, d.alt_mass, d.alt_type, d.alt_flags,
                       d.alt_uid, d.alt_shape, d.alt_ref, d.alt_cell, d.phys);
    d.cur = o;
    hipLaunchKernelGGL(k_copy_back_maps, dim3(grid), dim3(256), 0, s, d.counters, d.alt_vx, d.alt_vy, d.alt_vz,
                       d.alt_mass, d.alt_type, d.alt_flags, d.alt_uid, d.alt_shape, d.alt_ref, d.alt_cell, d.vx,
                       d.vy, d.vz, d.mass, d.type, d.flags, d.uid, d.shape, d.ref, d.particle_cell, d.phys,
                       d.send_map, d.send_mult, d.sync_src, d.sync_mult, d.img_cnt);
    hipLaunchKernelGGL(k_img_build, dim3(grid), dim3(256), 0, s, d.counters, d.sync_src, d.img_cnt, d.img_slot, d.cap);
}

// ---------------------------------------------------------------- batched cell phase
//
// The cell phase of a rebuild (exchange, borders, binning) of many systems,
// each kernel one launch over all of them: blockIdx.y is the system, whose
// arguments come from its descriptor; the x grid is the largest system's and
// every kernel loops or bounds-checks by the system's own device counts.

#define REP_ARGS                                     \
    const RepDesc &R = D[blockIdx.y];                \
    const DevState &d = R.d;                         \
    const Geom &g = R.g;                             \
    const int c = d.cur;                             \
    (void) g;                                        \
    (void) c

// dim < 0: every system; else only systems periodic in dim
__global__ void k_set_counters_batch(const RepDesc *__restrict__ D, int dim, int first, int count, int v) {
    REP_ARGS;
    if (dim >= 0 && !R.periodic[dim]) return;
    if ((int) threadIdx.x < count) d.counters[first + threadIdx.x] = v;
}

// dim is the system's first periodic dimension
__device__ __forceinline__ bool first_periodic(const RepDesc &R, int dim) {
    for (int k = 0; k < dim; k++)
        if (R.periodic[k]) return false;
    return true;
}

__global__ __launch_bounds__(256) void k_sel_count_batch(const RepDesc *__restrict__ D, int dim, int exch) {
    REP_ARGS;
    if (!R.periodic[dim]) return;
    const double *Xd = dim == 0 ? d.X[c] : (dim == 1 ? d.Y[c] : d.Z[c]);
    k_sel_count_body(Xd, d.flags, d.phys, d.counters, g, dim, exch, first_periodic(R, dim) ? 1 : 0, d.tile_cnt,
                     d.exchg_flag);
}

__global__ __launch_bounds__(256) void k_sel_scatter_batch(const RepDesc *__restrict__ D, int dim, int exch) {
    REP_ARGS;
    if (!R.periodic[dim]) return;
    k_sel_scatter_body(d.X[c], d.Y[c], d.Z[c], d.vx, d.vy, d.vz, d.mass, d.type, d.uid, d.shape, d.flags, d.ref,
                       d.phys, d.counters, g, dim, exch, d.tile_cnt, d.send_map, d.send_mult, d.send_cap, d.cap);
}

__global__ __launch_bounds__(1024) void k_exchange_relabel_batch(const RepDesc *__restrict__ D, int dim) {
    REP_ARGS;
    if (!R.periodic[dim]) return;
    int *sc = d.scratch_i;
    k_exchange_relabel_body(d.X[c], d.Y[c], d.Z[c], d.ref, d.phys, d.counters, g, d.send_map, d.send_mult,
                            d.exchg_flag, sc, sc + d.send_cap, sc + 2 * d.send_cap, sc + 3 * d.send_cap);
}

// which: 0 lcount, 1 gcount (ncells + 1 entries), 2 img_cnt (cap entries)
__global__ __launch_bounds__(256) void k_fill_batch(const RepDesc *__restrict__ D, int which, int v) {
    REP_ARGS;
    int *p = which == 0 ? d.lcount : (which == 1 ? d.gcount : d.img_cnt);
    const int n = which == 2 ? d.cap : g.ncells + 1;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = v;
}

__global__ __launch_bounds__(256) void k_bin_count_batch(const RepDesc *__restrict__ D) {
    REP_ARGS;
    k_bin_count_body(d.X[c], d.Y[c], d.Z[c], d.flags, d.counters, g, d.particle_cell, d.lcount, d.gcount, d.slot);
}

__global__ __launch_bounds__(1024) void k_scan_cells_batch(const RepDesc *__restrict__ D) {
    REP_ARGS;
    k_scan_cells_body(d.counters, d.lcount, d.gcount, d.lstart, d.gstart, g.ncells);
}

__global__ __launch_bounds__(256) void k_bin_scatter_batch(const RepDesc *__restrict__ D) {
    REP_ARGS;
    k_bin_scatter_body(d.counters, d.particle_cell, d.slot, d.lstart, d.gstart, d.perm);
}

__global__ __launch_bounds__(256) void k_sort_cells_batch(const RepDesc *__restrict__ D) {
    REP_ARGS;
    for (int item = blockIdx.x * 4 + (threadIdx.x >> 6); item < 2 * g.ncells; item += gridDim.x * 4)
        k_sort_cells_body(d.counters, d.lstart, d.gstart, d.ref, d.perm, g.ncells, item);
}

__global__ __launch_bounds__(256) void k_permute_batch(const RepDesc *__restrict__ D) {
    REP_ARGS;
    const int o = 1 - c;
    k_permute_body(d.counters, d.perm, d.X[c], d.Y[c], d.Z[c], d.X[o], d.Y[o], d.Z[o], d.vx, d.vy, d.vz, d.mass,
                   d.type, d.flags, d.uid, d.shape, d.ref, d.particle_cell, d.alt_vx, d.alt_vy, d.alt_vz, d.alt_mass,
                   d.alt_type, d.alt_flags, d.alt_uid, d.alt_shape, d.alt_ref, d.alt_cell, d.phys);
}

__global__ __launch_bounds__(256) void k_copy_back_maps_batch(const RepDesc *__restrict__ D) {
    REP_ARGS;
    k_copy_back_maps_body(d.counters, d.alt_vx, d.alt_vy, d.alt_vz, d.alt_mass, d.alt_type, d.alt_flags, d.alt_uid,
                          d.alt_shape, d.alt_ref, d.alt_cell, d.vx, d.vy, d.vz, d.mass, d.type, d.flags, d.uid,
                          d.shape, d.ref, d.particle_cell, d.phys, d.send_map, d.send_mult, d.sync_src, d.sync_mult,
                          d.img_cnt);
}

__global__ __launch_bounds__(256) void k_img_build_batch(const RepDesc *__restrict__ D) {
    REP_ARGS;
    k_img_build_body(d.counters, d.sync_src, d.img_cnt, d.img_slot, d.cap);
}

#undef REP_ARGS

void launch_cells_batch(hipStream_t s, const RepDesc *D, int n, int max_cap, int max_ncells) {
    if (n <= 0) return;
    // every kernel but the single-workgroup ones strides over its system's
    // range, so a few workgroups per system fill the device
    const int per = std::max(1, std::min(tile_grid(max_cap), 32));
    const dim3 one(1, n), tiles(per, n), cells(per, n), fills(per, n);
    (void) max_ncells;
    // exchange (comm.py:105-151), one dimension at a time
    for (int dim = 0; dim < 3; dim++) {
        hipLaunchKernelGGL(k_sel_count_batch, tiles, dim3(256), 0, s, D, dim, 1);
        hipLaunchKernelGGL(k_sel_scatter_batch, tiles, dim3(256), 0, s, D, dim, 1);
        hipLaunchKernelGGL(k_exchange_relabel_batch, one, dim3(1024), 0, s, D, dim);
    }
    // borders
    for (int dim = 0; dim < 3; dim++) {
        hipLaunchKernelGGL(k_sel_count_batch, tiles, dim3(256), 0, s, D, dim, 0);
        hipLaunchKernelGGL(k_sel_scatter_batch, tiles, dim3(256), 0, s, D, dim, 0);
    }
    // cell binning and the permutation into cell order (flips each system's
    // position buffer: the caller flips d.cur on the host); the cell counts
    // are zero on entry, the scan clears the image and tile flags
    hipLaunchKernelGGL(k_bin_count_batch, tiles, dim3(256), 0, s, D);
    hipLaunchKernelGGL(k_scan_cells_batch, one, dim3(1024), 0, s, D);
    hipLaunchKernelGGL(k_bin_scatter_batch, tiles, dim3(256), 0, s, D);
    hipLaunchKernelGGL(k_sort_cells_batch, cells, dim3(256), 0, s, D);
    hipLaunchKernelGGL(k_permute_batch, tiles, dim3(256), 0, s, D);
    hipLaunchKernelGGL(k_copy_back_maps_batch, tiles, dim3(256), 0, s, D);
    hipLaunchKernelGGL(k_img_build_batch, tiles, dim3(256), 0, s, D);
}

int build_zt(const Geom &g) { return g.nchunks == 1 ? g.dim_cells[2] + 2 : g.zc + 3; }

static size_t build_lds(const DevState &d, const Geom &g, int half, int tcap) {
    (void) d;
    return (size_t) tcap * (sizeof(int) + (half ? sizeof(int) : 0)) + (size_t) BW * NB_CAND * sizeof(uint16_t) +
           (size_t) 9 * build_zt(g) * 4 * sizeof(int);
}

int build_tile_limit(const DevState &d, const Geom &g, int half) {
    const size_t fixed = build_lds(d, g, half, 0) + 2048;  // + static LDS
    const size_t per = sizeof(int) + (half ? sizeof(int) : 0);
    const int lim = d.max_lds > (int) fixed ? (int) ((d.max_lds - fixed) / per) : 0;
    return min(lim, 65536);  // tile offsets are 16-bit
}

BuildArgs build_args(const DevState &d, const Geom &g, int half, double rc2) {
    BuildArgs A;
    const int c = d.cur;
    A.counters = d.counters; A.counters_rw = d.counters;
    A.g = g;
    A.lstart = d.lstart; A.gstart = d.gstart;
    A.X = d.X[c]; A.Y = d.Y[c]; A.Z = d.Z[c];
    A.ref = d.ref; A.flags = d.flags;
    A.rc2 = rc2;
    A.nl = d.nl; A.nl_stride = d.nl_stride; A.ncap = d.ncap;
    A.tcap = min(d.tile_cap, build_tile_limit(d, g, half));
    A.zt = build_zt(g);
    A.numneighs = d.numneighs; A.nrev = d.nrev; A.rev0 = d.rev0;
    return A;
}

size_t build_lds_bytes(const DevState &d, const Geom &g, int half) {
    return build_lds(d, g, half, min(d.tile_cap, build_tile_limit(d, g, half)));
}

void launch_build_neighbors(hipStream_t s, const DevState &d, const Geom &g, int half, double rc2) {
    const BuildArgs A = build_args(d, g, half, rc2);
    const size_t lds = build_lds(d, g, half, A.tcap);
    if (half) hipLaunchKernelGGL(k_build_neighbors<true>, dim3(g.nsegs), dim3(BB), lds, s, A);
    else hipLaunchKernelGGL(k_build_neighbors<false>, dim3(g.nsegs), dim3(BB), lds, s, A);
}

void launch_build_batch(hipStream_t s, const BuildArgs *desc, const int *rep, const int *seg_base, int nblocks,
                        size_t lds, int half) {
    if (nblocks <= 0) return;
    if (half) hipLaunchKernelGGL(k_build_neighbors_batch<true>, dim3(nblocks), dim3(BB), lds, s, desc, rep, seg_base);
    else hipLaunchKernelGGL(k_build_neighbors_batch<false>, dim3(nblocks), dim3(BB), lds, s, desc, rep, seg_base);
}

void launch_expand_neighbors(hipStream_t s, const DevState &d, const Geom &g, int *out_ref, int ncap_out) {
    (void) g;
    hipLaunchKernelGGL(k_expand_neighbors, dim3(std::max(1, std::min((d.nlocal + 255) / 256, 2048))), dim3(256), 0,
                       s, d.counters, d.ref, d.nl, d.nl_stride, d.numneighs, out_ref, ncap_out);
}

static void launch_select_dim(hipStream_t s, const DevState &d, const Geom &g, int dim, int exch, int first) {
    const int grid = tile_grid(d.cap);
    const int c = d.cur;
    const double *Xd = dim == 0 ? d.X[c] : (dim == 1 ? d.Y[c] : d.Z[c]);
    hipLaunchKernelGGL(k_sel_count, dim3(grid), dim3(256), 0, s, Xd, d.flags, d.phys, d.counters, g, dim, exch, first,
                       d.tile_cnt, d.exchg_flag);
    hipLaunchKernelGGL(k_sel_scatter, dim3(grid), dim3(256), 0, s, d.X[c], d.Y[c], d.Z[c], d.vx, d.vy, d.vz, d.mass,
                       d.type, d.uid, d.shape, d.flags, d.ref, d.phys, d.counters, g, dim, exch, d.tile_cnt,
                       d.send_map, d.send_mult, d.send_cap, d.cap);
}

void launch_exchange(hipStream_t s, const DevState &d, const Geom &g, const int *periodic) {
    const int c = d.cur;
    for (int dim = 0; dim < 3; dim++) {
        if (!periodic[dim]) continue;
        launch_select_dim(s, d, g, dim, 1, 0);
        int *sc = d.scratch_i;
        hipLaunchKernelGGL(k_exchange_relabel, dim3(1), dim3(1024), 0, s, d.X[c], d.Y[c], d.Z[c], d.ref, d.phys,
                           d.counters, g, d.send_map, d.send_mult, d.exchg_flag, sc, sc + d.send_cap,
                           sc + 2 * d.send_cap, sc + 3 * d.send_cap);
    }
}

void launch_borders(hipStream_t s, const DevState &d, const Geom &g, const int *periodic) {
    // the first periodic dimension starts from no ghosts and an empty send map
    bool first = true;
    for (int dim = 0; dim < 3; dim++) {
        if (!periodic[dim]) continue;
        launch_select_dim(s, d, g, dim, 0, first ? 1 : 0);
        first = false;
    }
    // no periodic dimension: no ghosts, nothing sent (adjacent counters)
    static_assert(C_NSEND_ALL == C_NGHOST + 1 && C_NSEND == C_NSEND_ALL + 1, "counter layout");
    if (first) launch_set_counters(s, d.counters, C_NGHOST, 8, 0);
}

void launch_sync_pp(hipStream_t s, const DevState &d, const Geom &g, int cur, int prev, int grid) {
    hipLaunchKernelGGL(k_sync_pp, dim3(grid), dim3(256), 0, s, d.X[cur], d.Y[cur], d.Z[cur], d.X[prev], d.Y[prev],
                       d.Z[prev], d.counters, d.sync_src, d.sync_mult, g, d.P[cur]);
}

void launch_sync_inplace(hipStream_t s, const DevState &d, const Geom &g, int grid, int with_vel) {
    const int c = d.cur;
    hipLaunchKernelGGL(k_sync_pack, dim3(grid), dim3(256), 0, s, d.X[c], d.Y[c], d.Z[c], d.vx, d.vy, d.vz,
                       d.counters, d.sync_src, d.sync_mult, g, d.stage, with_vel);
    hipLaunchKernelGGL(k_sync_unpack, dim3(grid), dim3(256), 0, s, d.X[c], d.Y[c], d.Z[c], d.vx, d.vy, d.vz,
                       d.counters, d.stage, with_vel);
}

static LJTab make_tab(const LJParams &prm) {
    LJTab t;
    t.nt = prm.ntypes;
    for (int k = 0; k < 16; k++) {
        t.eps[k] = k < prm.ntypes * prm.ntypes ? prm.eps[k] : 0.0;
        t.sig6[k] = k < prm.ntypes * prm.ntypes ? prm.sig6[k] : 0.0;
    }
    t.unit = (prm.ntypes == 1 && prm.eps[0] == 1.0 && prm.sig6[0] == 1.0) ? 1 : 0;
    return t;
}

int force_waves(const DevState &d) { return (d.nlocal + 63) / 64; }
int lj_waves(const DevState &d) { return (d.nlocal + 31) / 32; }

// image refresh only when the kernel writes next positions into the other
// buffer and every local's images fit the table
Images images_of(const DevState &d, bool writes_next) {
    Images im;
    im.cnt = d.img_cnt; im.slot = d.img_slot; im.src = d.sync_src; im.mult = d.sync_mult;
    im.cap = d.cap;
    im.on = (writes_next && d.img_on) ? 1 : 0;
    return im;
}

LJArgs lj_args(const DevState &d, const Geom &g, const LJParams &prm, int cur, int next, int do_vv2, int do_vv1,
               int fused, double *eatom, bool packed) {
    LJArgs A;
    A.counters = d.counters;
    A.g = g;
    A.lstart = d.lstart; A.gstart = d.gstart;
    A.X = d.X[cur]; A.Y = d.Y[cur]; A.Z = d.Z[cur];
    A.P = d.P[cur];
    A.Pn = (packed && fused && do_vv1 && next != cur) ? d.P[next] : nullptr;
    A.Xn = d.X[next]; A.Yn = d.Y[next]; A.Zn = d.Z[next];
    A.type = d.type; A.nl = d.nl; A.nl_stride = d.nl_stride; A.numneighs = d.numneighs; A.flags = d.flags;
    A.nrev = d.nrev; A.rev0 = d.rev0; A.ncap = d.ncap;
    A.fx = d.fx; A.fy = d.fy; A.fz = d.fz; A.vx = d.vx; A.vy = d.vy; A.vz = d.vz; A.mass = d.mass;
    A.rc2 = prm.rc2; A.tab = make_tab(prm); A.dt = prm.dt; A.hdt = prm.hdt;
    A.do_vv2 = do_vv2; A.do_vv1 = do_vv1;
    A.partials = d.partials; A.pstride = d.pstride; A.eatom = eatom;
    A.im = images_of(d, fused && do_vv1 && next != cur);
    return A;
}

template <int NT, int MODE, bool RED, bool UNIT>
static void lj_launch(hipStream_t s, const DevState &d, const Geom &g, const LJParams &prm, int cur, int next,
                      int do_vv2, int do_vv1, double *eatom) {
    const LJArgs A = lj_args(d, g, prm, cur, next, do_vv2, do_vv1, MODE == 1, eatom);
    if (MODE == 2) {  // energy and per-atom energies: one lane per local
        const dim3 grid(force_waves(d));
        if (d.rev0 > 0) hipLaunchKernelGGL((k_lj<NT, MODE, RED, true, false, UNIT>), grid, dim3(64), 0, s, A);
        else hipLaunchKernelGGL((k_lj<NT, MODE, RED, false, false, UNIT>), grid, dim3(64), 0, s, A);
    } else {
        const dim3 grid(lj_waves(d));
        if (d.rev0 > 0) hipLaunchKernelGGL((k_lj2<NT, MODE, RED, true, UNIT>), grid, dim3(64), 0, s, A);
        else hipLaunchKernelGGL((k_lj2<NT, MODE, RED, false, UNIT>), grid, dim3(64), 0, s, A);
    }
}

// workgroups of a kernel that the current device holds resident at once
template <typename K>
static int resident_grid(K kernel, int block) {
    int dev = 0, occ = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, block, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 1024;
    return std::max(1, occ * cus);
}

void launch_lj_full_batch(hipStream_t s, const LJArgs *desc, const int *rep, const int *wave_base, int nblocks,
                          int ntypes1, int unit, int reduce, int max_wg) {
    if (nblocks <= 0) return;
    const dim3 blk(64);
    // the resident workgroups, or at most max_wg of them
    auto grid = [&](int resident) { return dim3(std::min(nblocks, max_wg > 0 ? std::min(max_wg, resident) : resident)); };
#define LJB(NT, RED, UNIT)                                                                                     \
    hipLaunchKernelGGL((k_lj_batch<NT, 1, RED, UNIT>), grid(resident_grid(k_lj_batch<NT, 1, RED, UNIT>, 64)), blk, 0, \
                       s, desc, rep, wave_base, nblocks)
    if (ntypes1 && unit) { if (reduce) LJB(1, true, true); else LJB(1, false, true); }
    else if (ntypes1) { if (reduce) LJB(1, true, false); else LJB(1, false, false); }
    else { if (reduce) LJB(0, true, false); else LJB(0, false, false); }
#undef LJB
}

template <int NT, bool UNIT>
static void lj_nt(hipStream_t s, const DevState &d, const Geom &g, const LJParams &prm, int cur, int next, int fuse,
                  int do_vv2, int do_vv1, int reduce) {
    if (fuse) {
        if (reduce) lj_launch<NT, 1, true, UNIT>(s, d, g, prm, cur, next, do_vv2, do_vv1, nullptr);
        else lj_launch<NT, 1, false, UNIT>(s, d, g, prm, cur, next, do_vv2, do_vv1, nullptr);
    } else {
        if (reduce) lj_launch<NT, 0, true, UNIT>(s, d, g, prm, cur, next, 0, 0, nullptr);
        else lj_launch<NT, 0, false, UNIT>(s, d, g, prm, cur, next, 0, 0, nullptr);
    }
}

// the packed copies of many systems, one launch: system blockIdx.y packs
// the position buffer its force descriptor reads
__global__ __launch_bounds__(256) void k_pack_p_batch(const LJArgs *__restrict__ desc) {
    const LJArgs &A = desc[blockIdx.y];
    const int n = A.counters[C_NLOCAL] + A.counters[C_NGHOST];
    double4 *P = (double4 *) A.P;  // the system's own packed buffer
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        P[i] = make_double4(A.X[i], A.Y[i], A.Z[i], 0.0);
}

void launch_pack_p_batch(hipStream_t s, const LJArgs *desc, int n, int max_cap) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pack_p_batch, dim3(std::max(1, std::min((max_cap + 255) / 256, 256)), n), dim3(256), 0, s,
                       desc);
}

void launch_pack_p(hipStream_t s, const DevState &d, int cur) {
    hipLaunchKernelGGL(k_pack_p, dim3(std::max(1, std::min((d.cap + 255) / 256, 1024))), dim3(256), 0, s,
                       d.counters, d.X[cur], d.Y[cur], d.Z[cur], d.P[cur]);
}

void launch_lj_full(hipStream_t s, const DevState &d, const Geom &g, const LJParams &prm, int cur, int next, int fuse,
                    int do_vv2, int do_vv1, int reduce, int pack, hipEvent_t after_pack) {
    if (d.nlocal <= 0) return;
    if (pack) launch_pack_p(s, d, cur);
    if (after_pack) hipEventRecord(after_pack, s);
    if (prm.ntypes != 1) lj_nt<0, false>(s, d, g, prm, cur, next, fuse, do_vv2, do_vv1, reduce);
    else if (make_tab(prm).unit) lj_nt<1, true>(s, d, g, prm, cur, next, fuse, do_vv2, do_vv1, reduce);
    else lj_nt<1, false>(s, d, g, prm, cur, next, fuse, do_vv2, do_vv1, reduce);
}

void launch_energy(hipStream_t s, const DevState &d, const Geom &g, const LJParams &prm, int half, double *eatom) {
    if (d.nlocal <= 0) return;
    (void) half;  // half lists: the kernel adds the mirror part (d.rev0 > 0)
    launch_pack_p(s, d, d.cur);
    if (prm.ntypes == 1) lj_launch<1, 2, true, false>(s, d, g, prm, d.cur, d.cur, 0, 0, eatom);
    else lj_launch<0, 2, true, false>(s, d, g, prm, d.cur, d.cur, 0, 0, eatom);
}

void launch_vv_initial(hipStream_t s, const DevState &d, const LJParams &prm) {
    if (d.nlocal <= 0) return;
    const int c = d.cur;
    hipLaunchKernelGGL(k_vv_initial, dim3((d.nlocal + 255) / 256), dim3(256), 0, s, d.X[c], d.Y[c], d.Z[c], d.vx,
                       d.vy, d.vz, d.fx, d.fy, d.fz, d.mass, d.flags, d.nlocal, prm.dt, prm.hdt);
}

void launch_vv_post(hipStream_t s, const DevState &d, const Geom &g, const LJParams &prm, int cur, int next,
                    int do_vv2, int do_vv1, int reduce, double4 *Pn) {
    if (d.nlocal <= 0) return;
    const int grid = (d.nlocal + 255) / 256;
    if (reduce)
        hipLaunchKernelGGL(k_vv_post<true>, dim3(grid), dim3(256), 0, s, d.X[cur], d.Y[cur], d.Z[cur], d.X[next],
                           d.Y[next], d.Z[next], d.vx, d.vy, d.vz, d.fx, d.fy, d.fz, d.mass, d.flags, d.nlocal, prm.dt,
                           prm.hdt, do_vv2, do_vv1, d.partials, images_of(d, do_vv1 && next != cur), g, Pn);
    else
        hipLaunchKernelGGL(k_vv_post<false>, dim3(grid), dim3(256), 0, s, d.X[cur], d.Y[cur], d.Z[cur], d.X[next],
                           d.Y[next], d.Z[next], d.vx, d.vy, d.vz, d.fx, d.fy, d.fz, d.mass, d.flags, d.nlocal, prm.dt,
                           prm.hdt, do_vv2, do_vv1, d.partials, images_of(d, do_vv1 && next != cur), g, Pn);
}

void launch_ke_partial(hipStream_t s, const DevState &d) {
    if (d.nlocal <= 0) return;
    hipLaunchKernelGGL(k_ke_partial, dim3((d.nlocal + 255) / 256), dim3(256), 0, s, d.vx, d.vy, d.vz, d.mass,
                       d.nlocal, d.partials);
}

void launch_thermo_reduce(hipStream_t s, const DevState &d, int nb_ke, int nb_f, int which_mask, const Geom &g,
                          double *out) {
    hipLaunchKernelGGL(k_thermo_reduce, dim3(1), dim3(256), 0, s, d.partials, d.pstride, nb_ke, nb_f,
                       which_mask, d.nlocal, g.len[0], g.len[1], g.len[2], out);
}

}  // namespace pk
