// neighbor.hip — cell binning by counting sort and the Verlet list build.
//
// Reference semantics (single rank, orthorhombic):
//   grid:     CellList::compute_grid / build_stencil (src/cpu/neighbor/cell_list.cpp:83-232)
//   binning:  CellList::bin_atoms (:234-314) — wrapped copy for the cell index only,
//             floor(x/cs) clamped, z-fastest cell index
//   pairs:    the half-stencil predicate of refresh_half (:539-738): a pair whose
//             cells differ is tested in the orientation of the FORWARD stencil
//             offset as  d = x_i - (x_j + shift),  a home-cell pair as d = x_i - x_j,
//             accepted iff (dx*dx + dy*dy) + dz*dz < (rc+skin)^2. With the small-grid
//             fallback (fewer than 2*reach+1 cells on an axis) the full stencil with
//             while-loop min-image is used instead, as the reference does.
//   cadence:  NeighborList::decide_reneighbor (include/atomiverse/cpu/neighbor/neighbor_list.h:128-141)
//
// MI355X layout: atoms are physically re-sorted at every build by (cell
// rank, tag), where the rank orders cells brick by brick (brick.hpp), so a
// cell owns the contiguous slot range [cell_start[r], cell_start[r+1]) and a
// brick owns a contiguous range of cells. The list is a FULL list (both
// directions of every accepted pair) so each force row has a single writer;
// every unordered pair is tested once per side in the canonical orientation
// above, so both sides reach the same bit-identical decision. In brick mode
// one workgroup per brick stages its halo in LDS and stores 16-bit halo rows,
// column-major inside the brick (a wave's k-th loads are one 128 B line). The
// small-grid fallback keeps a 32-bit column-major list over global slots.
// The whole rebuild is predicated on a device flag, so the host never waits
// for the decision. All expressions here are compiled with -ffp-contract=off.

#include "atvh_internal.hpp"
#include "batch.hpp"
#include "brick.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

using namespace atvh;

namespace {

constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems; // elements per scan block

// Arguments of the counting-sort kernels (one system).
struct SortArgs {
    const int* pred;
    int64_t n;
    int ncells;
    const double *x, *y, *z;
    double lx, ly, lz, csx, csy, csz;
    int ncx, ncy, ncz;
    const int* cell_rank;
    int* cell_of; // cell of every slot in the current (pre-sort) order
    int* count;
    int* start;
    int* tile_sums;
    int ntiles;
    int* fill;
    int* perm;
    const int* tag;
    // OpBin's snapshot of the per-atom arrays in the current slot order
    // (positions wrapped); OpPermute gathers from it into the system's arrays
    const double *vx, *vy, *vz, *fx, *fy, *fz, *q;
    const int* type;
    double *ox, *oy, *oz, *ovx, *ovy, *ovz, *ofx, *ofy, *ofz, *oq;
    int *otype, *otag;
    // arrays whose permutation is a no-op or unread are skipped: forces
    // outside the first build, charges of an uncharged frame (all zero),
    // types of a one-type frame (all zero)
    int with_forces, with_q, with_type;
};

// Runs of equal keys in consecutive lanes (slots are in cell order from the
// previous build, and most atoms keep their cell, so runs average the atoms
// per cell). For every lane: the first lane of its run and the run length
// (meaningful on the first lane). Every lane of the wave must call it;
// inactive lanes pass a key no other lane holds.
struct LaneRun {
    int head;   // first lane of this lane's run
    int length; // lanes in the run
};

__device__ __forceinline__ LaneRun lane_run(int key) {
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1, 64);
    const bool first = lane == 0 || prev != key;
    const unsigned long long heads = __ballot(first);
    const unsigned long long below = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1)); // heads <= lane
    const unsigned long long above = heads & ~((lane == 63) ? ~0ull : ((2ull << lane) - 1)); // heads > lane
    LaneRun r;
    r.head = 63 - __clzll(below);
    r.length = (above ? __ffsll(static_cast<long long>(above)) - 1 : 64) - lane;
    return r;
}

// CellList::bin_atoms: cell of every slot (current order) + histogram (one
// atomic per run of equal cells in a wave). The slot's arrays are copied to
// the snapshot on the way (positions wrapped), so the permutation can write
// the system's arrays in place.
struct OpBin {
    using Args = SortArgs;
    static __device__ void run(const SortArgs& a, int bid, int nbk) {
        if (!*a.pred)
            return;
        const int lane = threadIdx.x & 63;
        for (int64_t w = bid * (int64_t)blockDim.x + (threadIdx.x & ~63); w < a.n;
             w += (int64_t)nbk * blockDim.x) { // wave-uniform
            const int64_t i = w + lane;
            const bool act = i < a.n;
            int key = -1 - lane;
            if (act) {
                // wrap_orthorhombic (pbc.h:37-51) is applied here, at the rebuild,
                // rather than every step: between rebuilds positions stay unwrapped so
                // the periodic image of every listed pair is the one fixed at the build.
                double xi = a.x[i], yi = a.y[i], zi = a.z[i];
                xi -= a.lx * floor(xi / a.lx);
                yi -= a.ly * floor(yi / a.ly);
                zi -= a.lz * floor(zi / a.lz);
                a.ox[i] = xi;
                a.oy[i] = yi;
                a.oz[i] = zi;
                a.ovx[i] = a.vx[i];
                a.ovy[i] = a.vy[i];
                a.ovz[i] = a.vz[i];
                a.otag[i] = a.tag[i];
                if (a.with_q)
                    a.oq[i] = a.q[i];
                if (a.with_type)
                    a.otype[i] = a.type[i];
                if (a.with_forces) {
                    a.ofx[i] = a.fx[i];
                    a.ofy[i] = a.fy[i];
                    a.ofz[i] = a.fz[i];
                }
                // bin_atoms wraps its own copy once more (a no-op except at x == L)
                xi -= a.lx * floor(xi / a.lx);
                yi -= a.ly * floor(yi / a.ly);
                zi -= a.lz * floor(zi / a.lz);
                int cx = static_cast<int>(floor((xi - 0.0) / a.csx));
                int cy = static_cast<int>(floor((yi - 0.0) / a.csy));
                int cz = static_cast<int>(floor((zi - 0.0) / a.csz));
                cx = max(0, min(cx, a.ncx - 1));
                cy = max(0, min(cy, a.ncy - 1));
                cz = max(0, min(cz, a.ncz - 1));
                const int c = cx * a.ncy * a.ncz + cy * a.ncz + cz;
                a.cell_of[i] = c;
                key = a.cell_rank[c];
            }
            const LaneRun r = lane_run(key);
            if (act && r.head == lane)
                atomicAdd(&a.count[key], r.length);
        }
    }
};

// Block-wide exclusive scan of one int per thread (LDS Hillis-Steele).
__device__ int block_exclusive_scan(int v, int* sh, int& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        int add = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const int incl = sh[t];
    total = sh[kBlock - 1];
    __syncthreads();
    return incl - v;
}

// Scan phase 1: per-tile exclusive scan (tile = kScanTile cells) + tile total.
struct OpScanTiles {
    using Args = SortArgs;
    static __device__ void run(const SortArgs& a, int bid, int) {
        if (!*a.pred)
            return;
        __shared__ int sh[kBlock];
        const int base = bid * kScanTile + threadIdx.x * kScanItems;
        int v[kScanItems];
        int sum = 0;
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = (base + k < a.ncells) ? a.count[base + k] : 0;
            sum += v[k];
        }
        int total;
        int ex = block_exclusive_scan(sum, sh, total);
        for (int k = 0; k < kScanItems; ++k) {
            if (base + k < a.ncells)
                a.start[base + k] = ex;
            ex += v[k];
        }
        if (threadIdx.x == 0)
            a.tile_sums[bid] = total;
    }
};

// Scan phase 2: one block scans the tile totals (any count, chunked) into
// the tiles' offsets. A cell's start is its tile-local start plus its
// tile's offset: the scatter adds it on the fly, and the in-cell order pass
// writes the sum back as the cell's start.
struct OpScanSums {
    using Args = SortArgs;
    static __device__ void run(const SortArgs& a, int, int) {
        if (!*a.pred)
            return;
        __shared__ int sh[kBlock];
        int carry = 0;
        for (int b0 = 0; b0 < a.ntiles; b0 += kBlock) {
            const int idx = b0 + threadIdx.x;
            const int v = idx < a.ntiles ? a.tile_sums[idx] : 0;
            int total;
            const int ex = block_exclusive_scan(v, sh, total);
            if (idx < a.ntiles)
                a.tile_sums[idx] = ex + carry;
            carry += total;
        }
        if (threadIdx.x == 0)
            a.start[a.ncells] = carry;
    }
};

// Counting-sort scatter: slot of every atom inside its cell (order fixed
// afterwards by the in-cell order pass). Also clears the histogram.
struct OpScatter {
    using Args = SortArgs;
    static __device__ void run(const SortArgs& a, int bid, int nbk) {
        if (!*a.pred)
            return;
        const int lane = threadIdx.x & 63;
        for (int64_t w = bid * (int64_t)blockDim.x + (threadIdx.x & ~63); w < a.n;
             w += (int64_t)nbk * blockDim.x) { // wave-uniform
            const int64_t i = w + lane;
            const bool act = i < a.n;
            const int c = act ? a.cell_rank[a.cell_of[i]] : -1 - lane;
            // one atomic per run of equal cells; the run's lanes take
            // consecutive places from its base
            const LaneRun r = lane_run(c);
            int base = 0;
            if (act && r.head == lane)
                base = atomicAdd(&a.fill[c], r.length);
            base = __shfl(base, r.head, 64);
            if (act) {
                a.perm[a.start[c] + a.tile_sums[c / kScanTile] + base + (lane - r.head)] = static_cast<int>(i);
                a.count[c] = 0;
            }
        }
    }
};

// Deterministic order inside a cell: ascending tag (insertion sort; cells
// hold a handful of atoms). The cell's start becomes final here (tile-local
// start plus tile offset; its end from its own fill count, so no thread reads
// a start another one rewrites). Also clears the fill counters.
struct OpCellOrder {
    using Args = SortArgs;
    static __device__ void run(const SortArgs& a, int bid, int nbk) {
        if (!*a.pred)
            return;
        for (int c = bid * blockDim.x + threadIdx.x; c < a.ncells; c += nbk * blockDim.x) {
            const int b = a.start[c] + a.tile_sums[c / kScanTile], e = b + a.fill[c];
            a.start[c] = b;
            for (int k = b + 1; k < e; ++k) {
                const int p = a.perm[k];
                const int tp = a.tag[p];
                int m = k - 1;
                while (m >= b && a.tag[a.perm[m]] > tp) {
                    a.perm[m + 1] = a.perm[m];
                    --m;
                }
                a.perm[m + 1] = p;
            }
            a.fill[c] = 0;
        }
    }
};

struct PermArgs {
    const int* pred;
    int64_t n;
    const int* perm;
    // OpBin's snapshot (previous slot order)
    const double *ox, *oy, *oz, *ovx, *ovy, *ovz, *ofx, *ofy, *ofz, *oq;
    const int *otype, *otag, *ocell_of;
    // the system's arrays, written in the new slot order, and x_save
    double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz, *q, *xs, *ys, *zs;
    int *type, *tag, *cell_of;
    int with_forces;
    // arrays whose permutation is a no-op or unread are skipped: charges of an
    // uncharged frame (all zero), types of a one-type frame (all zero), and
    // the cell index in brick mode (only the small-grid build reads it)
    int with_q, with_type, with_cell_of;
};

// The sorted order, gathered from the snapshot into the system's arrays,
// and x_save (CellList::save_positions) at the same positions.
struct OpPermute {
    using Args = PermArgs;
    static __device__ void run(const PermArgs& a, int bid, int nbk) {
    if (!*a.pred)
        return;
    for (int64_t s = bid * (int64_t)blockDim.x + threadIdx.x; s < a.n; s += (int64_t)nbk * blockDim.x) {
        const int p = a.perm[s];
        const double xx = a.ox[p], yy = a.oy[p], zz = a.oz[p];
        a.x[s] = xx;
        a.y[s] = yy;
        a.z[s] = zz;
        a.xs[s] = xx;
        a.ys[s] = yy;
        a.zs[s] = zz;
        a.vx[s] = a.ovx[p];
        a.vy[s] = a.ovy[p];
        a.vz[s] = a.ovz[p];
        if (a.with_q)
            a.q[s] = a.oq[p];
        if (a.with_type)
            a.type[s] = a.otype[p];
        a.tag[s] = a.otag[p];
        if (a.with_cell_of)
            a.cell_of[s] = a.ocell_of[p];
        if (a.with_forces) {
            a.fx[s] = a.ofx[p];
            a.fy[s] = a.ofy[p];
            a.fz[s] = a.ofz[p];
        }
    }
    }
};

struct BuildArgs {
    const double *x, *y, *z;
    const int* cell_of;
    const int* cell_rank;
    const int* start;
    const int4* stencil; // {di, dj, dk, kind}: kind 0 home, 1 forward, 2 backward
    int nstencil;
    int ncx, ncy, ncz;
    double lx, ly, lz;
    double rc_sq;
    int use_half;
    int* nbr;
    int* numneigh;
    int cap;
    int64_t npad;
    DevCtl* ctl;
};

// One thread per slot i: scan the full stencil, test each candidate with the
// canonical orientation, append accepted j (full list). HALF selects the
// reference's half-stencil arithmetic (every axis has >= 2*reach+1 cells);
// otherwise the small-grid min-image fallback.
template <bool HALF>
__global__ __launch_bounds__(kBlock) void k_nl_build(const int* __restrict__ pred, int64_t n,
                                                     BuildArgs a) {
    if (!*pred)
        return;
    __shared__ long long sh_tot[kBlock / 64];
    __shared__ int sh_max[kBlock / 64];
    long long my_cnt_sum = 0;
    int my_max = 0;
    int my_over = 0;
    const double half_lx = 0.5 * a.lx, half_ly = 0.5 * a.ly, half_lz = 0.5 * a.lz;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ci = a.cell_of[i];
        const int cz0 = ci % a.ncz;
        const int cy0 = (ci / a.ncz) % a.ncy;
        const int cx0 = ci / (a.ncy * a.ncz);
        const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
        int cnt = 0;
        int visited[HALF ? 1 : 128];
        int nvis = 0;
        for (int s = 0; s < a.nstencil; ++s) {
            const int4 o = a.stencil[s];
            int cx = cx0 + o.x, cy = cy0 + o.y, cz = cz0 + o.z;
            double shx = 0.0, shy = 0.0, shz = 0.0;
            if (cx < 0) { cx += a.ncx; shx = -a.lx; } else if (cx >= a.ncx) { cx -= a.ncx; shx = a.lx; }
            if (cy < 0) { cy += a.ncy; shy = -a.ly; } else if (cy >= a.ncy) { cy -= a.ncy; shy = a.ly; }
            if (cz < 0) { cz += a.ncz; shz = -a.lz; } else if (cz >= a.ncz) { cz -= a.ncz; shz = a.lz; }
            if (!HALF) {
                // small grids: fold residues, dedupe aliased cells
                if (cx < 0 || cx >= a.ncx) cx = ((cx % a.ncx) + a.ncx) % a.ncx;
                if (cy < 0 || cy >= a.ncy) cy = ((cy % a.ncy) + a.ncy) % a.ncy;
                if (cz < 0 || cz >= a.ncz) cz = ((cz % a.ncz) + a.ncz) % a.ncz;
            }
            const int cj = cx * a.ncy * a.ncz + cy * a.ncz + cz;
            if (!HALF) {
                bool seen = false;
                for (int v = 0; v < nvis; ++v)
                    if (visited[v] == cj) { seen = true; break; }
                if (seen)
                    continue;
                if (nvis < 128)
                    visited[nvis++] = cj;
            }
            const int rj = a.cell_rank[cj];
            const int jb = a.start[rj], je = a.start[rj + 1];
            for (int j = jb; j < je; ++j) {
                if (j == i)
                    continue;
                const double xj = a.x[j], yj = a.y[j], zj = a.z[j];
                double dx, dy, dz;
                if (!HALF) {
                    dx = xi - xj;
                    dy = yi - yj;
                    dz = zi - zj;
                    while (dx > half_lx) dx -= a.lx;
                    while (dx < -half_lx) dx += a.lx;
                    while (dy > half_ly) dy -= a.ly;
                    while (dy < -half_ly) dy += a.ly;
                    while (dz > half_lz) dz -= a.lz;
                    while (dz < -half_lz) dz += a.lz;
                } else if (o.w == 2) {
                    // backward offset: the pair belongs to j's forward scan
                    dx = xj - (xi - shx);
                    dy = yj - (yi - shy);
                    dz = zj - (zi - shz);
                } else {
                    dx = xi - (xj + shx);
                    dy = yi - (yj + shy);
                    dz = zi - (zj + shz);
                }
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 < a.rc_sq) {
                    if (cnt < a.cap)
                        a.nbr[(int64_t)cnt * a.npad + i] = j;
                    ++cnt;
                }
            }
        }
        a.numneigh[i] = cnt < a.cap ? cnt : a.cap;
        my_cnt_sum += cnt;
        my_max = max(my_max, cnt);
        if (cnt > a.cap)
            ++my_over;
    }
    // block statistics: total entries, max, overflow
    for (int off = 32; off > 0; off >>= 1) {
        my_cnt_sum += __shfl_down(my_cnt_sum, off, 64);
        my_max = max(my_max, __shfl_down(my_max, off, 64));
        my_over += __shfl_down(my_over, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __shared__ int sh_over[kBlock / 64];
    if (lane == 0) {
        sh_tot[wid] = my_cnt_sum;
        sh_max[wid] = my_max;
        sh_over[wid] = my_over;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        int m = 0, ov = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            t += sh_tot[w];
            m = max(m, sh_max[w]);
            ov += sh_over[w];
        }
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.ctl->nbr_total),
                  static_cast<unsigned long long>(t));
        atomicMax(&a.ctl->max_nbr, m);
        if (ov)
            atomicAdd(&a.ctl->overflow, ov);
    }
}

struct BrickBuildArgs {
    const int* pred;
    BrickGeom g;
    const int* halo_cells; // static per-brick halo cell table (rank | code << 26)
    int* halo_off;         // [nbricks*(kHaloCellCap+1)] halo row offset of every halo cell
    int nbricks;
    const double *x, *y, *z;
    const int* cell_of;
    const int* start;
    const int4* rows;   // stencil rows {di, dj, kmax, 0}: offsets (di, dj, -kmax..kmax)
    int nrows;
    double rc_sq;
    double rcs;         // cutoff + skin (for the conservative culling only)
    uint16_t* nbr16;
    uint16_t* self16;
    int* numneigh;
    int* halo_rows;
    int* halo_n;
    int2* brick_range;
    int cap;
    DevCtl* ctl;
};

constexpr int kMaxCols = 25;     // stencil columns: (2*reach+1)^2 with reach <= 2

// Rebuild step 1 (one workgroup per brick, small LDS, high occupancy): the
// brick's halo table from the static halo-cell table and the new cell_start —
// halo rows (global slot | image code), per-halo-cell row offsets, the halo
// size and the brick's slot range.
struct OpHalo {
    using Args = BrickBuildArgs;
    static __device__ void run(const BrickBuildArgs& a, int bid, int nbk);
};

__device__ void OpHalo::run(const BrickBuildArgs& a, int bid, int nbk) {
    if (!*a.pred)
        return;
    __shared__ int h_start[kHaloCellCap + 1];
    __shared__ int h_off[kHaloCellCap + 1];
    __shared__ int h_shift[kHaloCellCap];
    __shared__ int scan_tmp[kBlock];
    for (int t = bid; t < a.nbricks; t += nbk) {
        const int b = xcd_brick(t, a.nbricks);
        HaloTable ht;
        build_halo_table(a.g, a.halo_cells, a.start, b, ht, h_start, h_off, h_shift, scan_tmp);
        const int a0 = a.start[a.g.brick_cell_base[b]];
        const int nat = a.start[a.g.brick_cell_base[b + 1]] - a0;
        if (threadIdx.x == 0) {
            a.halo_n[b] = ht.natoms;
            a.brick_range[b] = make_int2(a0, nat);
            if (brick_skipped(ht.natoms, nat, a.g.max_atoms))
                atomicAdd(&a.ctl->halo_overflow, 1);
        }
        int* __restrict__ hoff = a.halo_off + (int64_t)b * (kHaloCellCap + 1);
        for (int h = threadIdx.x; h <= ht.ncell; h += blockDim.x)
            hoff[h] = h_off[h];
        if (ht.natoms <= kHaloRows) {
            int* __restrict__ rows = a.halo_rows + (int64_t)b * kHaloCap;
            for (int h = threadIdx.x; h < ht.ncell; h += blockDim.x) {
                const int q0 = h_off[h], nq = h_off[h + 1] - q0, g0 = h_start[h];
                const int tag = h_shift[h] << kRowSlotBits;
                for (int k = 0; k < nq; ++k)
                    rows[q0 + k] = (g0 + k) | tag;
            }
        }
        __syncthreads(); // LDS reused by the next brick
    }
}

// The permutation (blocks [0, gp)) and the halo tables (the remaining
// blocks) in one launch: both need only the final cell starts, and neither
// reads what the other writes. A rebuild stage fewer also means one launch
// fewer on every step without a rebuild.
struct PermHaloArgs {
    PermArgs p;
    BrickBuildArgs h;
    int gp; // permutation blocks
};

struct OpPermuteHalo {
    using Args = PermHaloArgs;
    static __device__ void run(const PermHaloArgs& a, int bid, int nbk) {
        if (bid < a.gp)
            OpPermute::run(a.p, bid, a.gp);
        else
            OpHalo::run(a.h, bid - a.gp, nbk - a.gp);
    }
};

// Rebuild step 2 (one workgroup per brick, one wavefront per home cell). The
// wave lays the cell's candidates out flat — every stencil column's z-run of
// halo rows, in column order — and takes them 64 at a time, one per lane,
// testing each against the cell's atoms two at a time (their coordinates
// and running counts are wave-uniform scalars), so the candidate loop does
// not diverge. Candidates are screened in FP32 with image coordinates X
// relative to the brick centre, in the expanded form
//   r^2 - |X_i|^2 = |X_j|^2 - 2 X_i . X_j   (three fma per test)
// accepted below (rc+skin)^2 - band, rejected above (rc+skin)^2 + band, and
// only inside the band decided by the exact reference predicate — raw FP64 positions,
// the shift form of the half-stencil offset in its canonical orientation
// (refresh_half, cell_list.cpp:680-727). The pair set is therefore exactly
// the reference's. Accepted rows are compacted with a ballot in candidate
// order (column order, then halo row order: deterministic) into 16-bit rows
// in 8-wide groups; the unused tail of the last group holds the sentinel
// row. While the list capacity fits kOutCap the rows of an atom collect in
// LDS and leave as whole 16-byte groups (one store per lane); otherwise
// each accepted row is stored directly. A cell with more than kCandCap
// candidates walks its columns one at a time instead.
// Half-width of the undecided band for a brick whose halo reaches E Å from
// its centre on any axis (|X| components <= E). With unit roundoff u = 2^-24
// the screen's error in r^2 is at most 72 u E^2 + u rc^2: the FP32 rounding
// of the coordinates (24 u E^2 for the pair distances in play), of |X|^2
// (9 u E^2 each for i and j), of the three fma whose partial sums stay
// below 9 E^2 (27 u E^2) and of the threshold (3 u E^2 + u rc^2). The band
// is twice that, plus 1e-6 Å^2.
__device__ __forceinline__ float screen_band(float E, float rc2) {
    return 8.6e-6f * E * E + 1.2e-7f * rc2 + 1e-6f;
}
constexpr int kCandCap = 768; // flat candidates per wavefront (LDS); a multiple of 64
constexpr int kOutCap = 256;  // list entries per atom staged in LDS
constexpr int kOutGroups = kOutCap / kGroup + 1; // + one group that absorbs rows beyond the capacity

__device__ __forceinline__ bool exact_pair(const double* __restrict__ x, const double* __restrict__ y,
                                        const double* __restrict__ z, double rc_sq, int gi, int gj,
                                        int kind, double shx, double shy, double shz) {
    const double xi = x[gi], yi = y[gi], zi = z[gi];
    const double xj = x[gj], yj = y[gj], zj = z[gj];
    double dx, dy, dz;
    if (kind == 2) {
        dx = xj - (xi - shx);
        dy = yj - (yi - shy);
        dz = zj - (zi - shz);
    } else {
        dx = xi - (xj + shx);
        dy = yi - (yj + shy);
        dz = zi - (zj + shz);
    }
    const double r2 = dx * dx + dy * dy + dz * dz;
    return r2 < rc_sq;
}

// The reference predicate for a candidate inside the FP32 band: the halo
// cell of row q (the last cell whose first row is <= q), hence the stencil
// offset (di, dj, dk), the periodic shift of the neighbour cell, and the
// exact FP64 test in the canonical orientation. (Arguments by value: a
// reference to the kernel's argument block would be copied to scratch to
// take its address.)
__device__ __forceinline__ bool band_pair(const double* __restrict__ x, const double* __restrict__ y,
                                       const double* __restrict__ z, double rc_sq, double Lx, double Ly, double Lz,
                                       int ncx, int ncy, int ncz, const int* h_off, int ncell, int HY, int HZ, int q,
                                       int hx0, int hy0, int hz0, int cx, int cy, int cz, int gi, int gj) {
    int lo = 0, hi = ncell; // h_off[lo] <= q < h_off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (h_off[mid] <= q)
            lo = mid;
        else
            hi = mid;
    }
    const int di = lo / (HY * HZ) - hx0, dj = (lo / HZ) % HY - hy0, dk = lo % HZ - hz0;
    const int ux = cx + di, uy = cy + dj, uz = cz + dk;
    const double shx = ux < 0 ? -Lx : (ux >= ncx ? Lx : 0.0);
    const double shy = uy < 0 ? -Ly : (uy >= ncy ? Ly : 0.0);
    const double shz = uz < 0 ? -Lz : (uz >= ncz ? Lz : 0.0);
    const int colkind = (di > 0 || (di == 0 && dj > 0)) ? 1 : ((di == 0 && dj == 0) ? 0 : 2);
    const int kind = colkind != 0 ? colkind : (dk > 0 ? 1 : (dk < 0 ? 2 : 0));
    return exact_pair(x, y, z, rc_sq, gi, gj, kind, shx, shy, shz);
}

// band_pair with everything but the brick b, the home cell's packed
// brick-local coordinates, the candidate row and the home slot taken from the
// argument block and recomputed here (the brick geometry as at the brick's
// start). The caller passes the block through an opaque pointer, so none of
// these values is hoisted out of the candidate loop for this rare branch.
__device__ __forceinline__ bool band_pair_args(const BrickBuildArgs& ab, const int* h_off, int b, int lpack, int q,
                                            int gi) {
    const BrickGeom& g = ab.g;
    int bx, by, bz;
    brick_coords(g, b, bx, by, bz);
    const int o0 = bx * g.bdim[0], o1 = by * g.bdim[1], o2 = bz * g.bdim[2];
    const int d0 = min(g.bdim[0], g.nc[0] - o0);
    const int d1 = min(g.bdim[1], g.nc[1] - o1), d2 = min(g.bdim[2], g.nc[2] - o2);
    const int HY = d1 + 2 * g.R[1], HZ = d2 + 2 * g.R[2];
    const int ncell = (d0 + 2 * g.R[0]) * HY * HZ;
    const int lx = lpack & 0xff, ly = (lpack >> 8) & 0xff, lz = lpack >> 16;
    const int* rows = ab.halo_rows + (int64_t)b * kHaloCap;
    return band_pair(ab.x, ab.y, ab.z, ab.rc_sq, g.L[0], g.L[1], g.L[2], g.nc[0], g.nc[1], g.nc[2], h_off, ncell, HY,
                     HZ, q, lx + g.R[0], ly + g.R[1], lz + g.R[2], o0 + lx, o1 + ly, o2 + lz, gi,
                     rows[q] & kRowSlotMask);
}

// Inclusive prefix sum over the 64 lanes (DPP row shifts, then the row
// broadcasts of lanes 15 and 31).
__device__ __forceinline__ int wave_inclusive_sum(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true); // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true); // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true); // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true); // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false); // row_bcast:15 -> rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false); // row_bcast:31 -> rows 2, 3
    return x;
}

struct OpPairs {
    using Args = BrickBuildArgs;
    static constexpr int kThreads = 1024;
    static constexpr int kMinWaves = 8;
    static constexpr int kWaves = kThreads / 64;
    static __device__ void run(const BrickBuildArgs& a, int bid, int nbk);
};

// The same build for the batched rebuild stage. Its kernel keeps more values
// live (the active-member search), and there the opaque values of the band
// branch (see pairs_run) cost vector spills instead of saving scalar ones
// (ladder 3.06e9 -> 2.68e9 atom-steps/s), so it keeps the plain form.
struct OpPairsBatched : OpPairs {
    static __device__ void run(const BrickBuildArgs& a, int bid, int nbk);
};

constexpr int kMaxBrickCells = 100; // home cells of the largest brick shape (4 x 5 x 5)

// LDS of one OpPairs workgroup (79.2 KB: two workgroups per CU).
struct PairsLds {
    float4 sp[kHaloCap]; // image coordinates relative to the brick centre; w = their squared norm
    int h_off[kHaloCellCap + 1];
    // per home cell: first slot, halo cell, halo rows [lo, hi) packed lo | hi << 16,
    // brick-local cell coordinates packed x | y << 8 | z << 16
    int4 cells[kMaxBrickCells];
    unsigned short cand[OpPairs::kWaves][kCandCap];
    uint4 out[OpPairs::kWaves][2][kOutGroups]; // staged rows of the wave's two atoms
    long long tot[OpPairs::kWaves];
    int max[OpPairs::kWaves], over[OpPairs::kWaves];
    int next_cell; // next unclaimed cell of the brick
};

// One workgroup builds the rows of one brick (t-th in XCD-aware order). The
// grid is the brick count, so no kernel-wide value has to survive from one
// brick to the next through the candidate loop.
// BAND: how the rare band test reaches its values. kBandKernarg (k_single):
// the home slot and the argument block are opaque to the compiler, the block
// read from the kernel-argument segment; kBandPlain: the values as computed
// for the brick (the batched stage: the same barriers through the block's
// own address there cost vector spills on the per-pass path, ladder 3.06e9 ->
// 2.72e9 atom-steps/s).
constexpr int kBandPlain = 0, kBandKernarg = 1;

template <bool STAGED, int BAND>
__device__ __forceinline__ void pairs_run(const BrickBuildArgs& a, int t, PairsLds& L) {
    constexpr int kWaves = OpPairs::kWaves;
    // wave-uniform values go through readfirstlane so their loops run on
    // scalar registers
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    long long my_cnt_sum = 0;
    int my_max = 0, my_over = 0;
    const double Lx = a.g.L[0], Ly = a.g.L[1], Lz = a.g.L[2];
    const int R0 = a.g.R[0], R1 = a.g.R[1], R2 = a.g.R[2];
    const float rc2f = static_cast<float>(a.rc_sq);
    const int ncols = a.nrows;
    const int cap = a.cap;
    // stencil column of this lane (lanes >= ncols have none)
    const int4 rw = lane < ncols ? a.rows[lane] : make_int4(0, 0, 0, 0);
    unsigned short* const obuf = reinterpret_cast<unsigned short*>(&L.out[wid][0][0]);
    {
        const int b = xcd_brick(t, a.nbricks);
        const int nh = a.halo_n[b];
        const int2 br = a.brick_range[b];
        const int a0 = br.x, nat = br.y;
        if (brick_skipped(nh, nat, a.g.max_atoms))
            return; // reported by OpHalo (block-uniform)
        const int cbase = a.g.brick_cell_base[b];
        int bx, by, bz;
        brick_coords(a.g, b, bx, by, bz);
        const int o0 = bx * a.g.bdim[0], o1 = by * a.g.bdim[1], o2 = bz * a.g.bdim[2];
        const int d0 = min(a.g.bdim[0], a.g.nc[0] - o0);
        const int d1 = min(a.g.bdim[1], a.g.nc[1] - o1), d2 = min(a.g.bdim[2], a.g.nc[2] - o2);
        const int HY = d1 + 2 * R1, HZ = d2 + 2 * R2;
        const int ncell = (d0 + 2 * R0) * HY * HZ;
        const double bc0 = (o0 + 0.5 * d0) * a.g.cs[0];
        const double bc1 = (o1 + 0.5 * d1) * a.g.cs[1];
        const double bc2 = (o2 + 0.5 * d2) * a.g.cs[2];
        const int* __restrict__ hoff = a.halo_off + (int64_t)b * (kHaloCellCap + 1);
        for (int h = threadIdx.x; h <= ncell; h += blockDim.x)
            L.h_off[h] = hoff[h];
        const int* __restrict__ rows = a.halo_rows + (int64_t)b * kHaloCap;
        // coordinates per row (independent loads, so many are in flight at once)
        for (int q = threadIdx.x; q < nh; q += blockDim.x) {
            const int row = rows[q];
            const int gs = row & kRowSlotMask;
            const int code = static_cast<unsigned>(row) >> kRowSlotBits;
            const float X = static_cast<float>((a.x[gs] + shift_of(code, 0, Lx)) - bc0);
            const float Y = static_cast<float>((a.y[gs] + shift_of(code, 1, Ly)) - bc1);
            const float Z = static_cast<float>((a.z[gs] + shift_of(code, 2, Lz)) - bc2);
            L.sp[q] = make_float4(X, Y, Z, __builtin_fmaf(Z, Z, __builtin_fmaf(Y, Y, X * X)));
        }
        // the home cells' table (the divisions run once per cell and brick)
        const int nbc = d0 * d1 * d2;
        for (int c = threadIdx.x; c < nbc; c += blockDim.x) {
            const int lz = c % d2, ly = (c / d2) % d1, lx = c / (d1 * d2);
            const int hs = ((lx + R0) * HY + (ly + R1)) * HZ + lz + R2;
            L.cells[c] = make_int4(a.start[cbase + c], hs, hoff[hs] | (hoff[hs + 1] << 16), lx | (ly << 8) | (lz << 16));
        }
        // this lane's stencil column as an offset of halo cells from the home cell
        const int colofs = (rw.x * HY + rw.y) * HZ;
        // the halo reaches (d/2 + R) cells from the centre on each axis
        const float ext = static_cast<float>(fmax(fmax((0.5 * d0 + R0) * a.g.cs[0], (0.5 * d1 + R1) * a.g.cs[1]),
                                                  (0.5 * d2 + R2) * a.g.cs[2]));
        const float band_w = screen_band(ext, rc2f);
        const float lo_acc = rc2f - band_w, hi_rej = rc2f + band_w;
        if (threadIdx.x == 0) {
            L.next_cell = kWaves;
            // the sentinel row (no halo atom: nh <= kHaloRows) at |X|^2 = 1e30,
            // which every screen rejects
            L.sp[kSentinelRow] = make_float4(0.0f, 0.0f, 0.0f, 1e30f);
        }
        __syncthreads();
        uint16_t* __restrict__ lst = a.nbr16 + (int64_t)a0 * cap;
        // cells are handed out one at a time (wave w starts with cell w and
        // then takes the next unclaimed one), which evens out the waves' loads
        // when cells hold different numbers of atoms
        int c = wid;
        while (c < nbc) {
            int grab = 0;
            if (lane == 0)
                grab = atomicAdd(&L.next_cell, 1);
            [&] {
            // cells of a brick are ranked in this order, so the cell's atoms are
            // the slots from cell_start of its rank on
            const int4 cell = L.cells[c];
            const int slot0 = __builtin_amdgcn_readfirstlane(cell.x);
            const int hs = __builtin_amdgcn_readfirstlane(cell.y);
            const int hq = __builtin_amdgcn_readfirstlane(cell.z);
            const int hq0 = hq & 0xffff, hq1 = hq >> 16;
            if (hq0 == hq1)
                return;
            const int lpack = __builtin_amdgcn_readfirstlane(cell.w);
            const int lx = lpack & 0xff, ly = (lpack >> 8) & 0xff, lz = lpack >> 16;
            const int hz0 = lz + R2;
            const int cx = o0 + lx, cy = o1 + ly, cz = o2 + lz;
            // this cell's z-run of halo rows in every stencil column (lane = column)
            int qb = 0, len = 0;
            if (lane < ncols) {
                const int hcol = hs + colofs;
                qb = L.h_off[hcol - rw.z];
                len = L.h_off[hcol + rw.z + 1] - qb;
            }
            const int inc = wave_inclusive_sum(len);
            const int total = __builtin_amdgcn_readlane(inc, 63);
            const bool flat = total <= kCandCap;
            if (flat) {
                // the column's rows qb, qb+1, ...: single slots up to a multiple
                // of four, then four rows per 8-byte store (two packed pairs,
                // each advanced by 4 per store; rows stay below 2^11, so the
                // halves never carry), then the last single slots
                int p = inc - len, v = qb;
                const int end = inc;
                unsigned short* const cw = &L.cand[wid][0];
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    if ((p & 3) != 0 && p < end) {
                        cw[p] = static_cast<unsigned short>(v);
                        ++p;
                        ++v;
                    }
                const int n4 = max(end - p, 0) >> 2;
                uint2* const c64 = reinterpret_cast<uint2*>(cw) + (p >> 2);
                unsigned lo = static_cast<unsigned>(v) | (static_cast<unsigned>(v + 1) << 16);
                unsigned hi = static_cast<unsigned>(v + 2) | (static_cast<unsigned>(v + 3) << 16);
                for (int k = 0; k < n4; ++k) {
                    c64[k] = make_uint2(lo, hi);
                    lo += 0x00040004u;
                    hi += 0x00040004u;
                }
                p += n4 << 2;
                v += n4 << 2;
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    if (p < end) {
                        cw[p] = static_cast<unsigned short>(v);
                        ++p;
                        ++v;
                    }
                if (total + lane < ((total + 63) & ~63)) // kCandCap is a multiple of 64
                    L.cand[wid][total + lane] = static_cast<unsigned short>(kSentinelRow);
                __builtin_amdgcn_wave_barrier();
            }
            const int li0 = slot0 - a0; // brick-local index of the cell's first atom
            const unsigned gstride = static_cast<unsigned>(nat) * kGroup; // list entries between groups
            // the cell's atoms two at a time (coordinates and counts are
            // wave-uniform scalars; two independent test chains per candidate
            // load), their candidates 64 per step, one per lane
            for (int qa = hq0; qa < hq1; qa += 2) {
                const bool two = qa + 1 < hq1; // wave-uniform
                const int qb2 = two ? qa + 1 : qa;
                // -2 X of the two atoms and their thresholds (rc+skin)^2 -/+ band - |X|^2
                auto uni = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };
                const float4 pa = L.sp[qa], pb = L.sp[qb2];
                const float ax = uni(-2.0f * pa.x), ay = uni(-2.0f * pa.y), az = uni(-2.0f * pa.z);
                const float bx = uni(-2.0f * pb.x), by = uni(-2.0f * pb.y), bz = uni(-2.0f * pb.z);
                const float sa = pa.w, sb = pb.w;
                const float tlo_a = uni(lo_acc - sa), thi_a = uni(hi_rej - sa);
                const float tlo_b = uni(lo_acc - sb), thi_b = uni(hi_rej - sb);
                const int lia = li0 + (qa - hq0), lib = li0 + (qb2 - hq0);
                uint16_t* __restrict__ dsta = lst + static_cast<unsigned>(lia) * kGroup;
                uint16_t* __restrict__ dstb = lst + static_cast<unsigned>(lib) * kGroup;
                unsigned short* const oa = obuf;
                unsigned short* const ob = obuf + kOutGroups * kGroup;
                int ca = 0, cb = 0;
                // candidate q (one per lane) against one atom: test, then append
                // the accepted rows in lane order. The lane masks are ballots of
                // the compares themselves (one v_cmp each) combined on scalar
                // registers; the band, where the FP32 screen cannot decide, is
                // rare and takes the branch.
                auto test = [&](float qx, float qy, float qz, float qss, int q, float px, float py, float pz,
                                float tlo, float thi, int qs, int li, uint16_t* __restrict__ dst, unsigned short* ost,
                                int& count) {
                    const float v = __builtin_fmaf(px, qx, __builtin_fmaf(py, qy, __builtin_fmaf(pz, qz, qss)));
                    const unsigned long long other = __builtin_amdgcn_ballot_w64(q != qs);
                    unsigned long long m = __builtin_amdgcn_ballot_w64(v < tlo) & other;
                    const unsigned long long band = __builtin_amdgcn_ballot_w64(v <= thi) & other & ~m;
                    if (band) { // the reference predicate for the band lanes
                        bool e = false;
                        if (__builtin_amdgcn_inverse_ballot_w64(band)) {
                            // with kBandKernarg the home atom's slot is made opaque
                            // here, so the compiler cannot hoist its coordinate
                            // addresses out of the candidate loop (six 64-bit
                            // pointers held in scalar registers through it for this
                            // rare branch: C1 rebuild 689 -> 636 us)
                            int gi = a0 + li;
                            if constexpr (BAND == kBandKernarg) {
                                // and the argument block too: the band test reads it
                                // from the kernel-argument segment (the block is the
                                // only argument of k_single<OpPairs>, at offset 0;
                                // taking the address of the argument itself would copy
                                // it to scratch) and recomputes its geometry
                                // (band_pair_args)
                                asm volatile("" : "+v"(gi));
                                auto kp = __builtin_amdgcn_kernarg_segment_ptr();
                                asm volatile("" : "+s"(kp));
                                e = band_pair_args(*(const BrickBuildArgs*)kp, L.h_off, b, L.cells[c].w, q, gi);
                            } else {
                                e = band_pair(a.x, a.y, a.z, a.rc_sq, Lx, Ly, Lz, a.g.nc[0], a.g.nc[1], a.g.nc[2],
                                              L.h_off, ncell, HY, HZ, q, lx + R0, ly + R1, hz0, cx, cy, cz, gi,
                                              rows[q] & kRowSlotMask);
                            }
                        }
                        m |= __builtin_amdgcn_ballot_w64(e);
                    }
                    if (__builtin_amdgcn_inverse_ballot_w64(m)) {
                        // the running count enters as mbcnt's addend
                        const unsigned pos = __builtin_amdgcn_mbcnt_hi(
                            static_cast<unsigned>(m >> 32),
                            __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(m), static_cast<unsigned>(count)));
                        if (STAGED) // rows beyond the capacity land in the last group, which is never stored
                            ost[min(pos, static_cast<unsigned>(kOutCap))] = static_cast<unsigned short>(q);
                        else if (pos < static_cast<unsigned>(cap))
                            dst[__umul24(pos >> 3, gstride) + (pos & 7)] = static_cast<uint16_t>(q);
                    }
                    count += static_cast<int>(__popcll(m));
                };
                // lanes past the candidates take the sentinel row, staged far
                // beyond every threshold, so no lane mask is needed
                auto step = [&](int q) {
                    const float4 qv = L.sp[q];
                    test(qv.x, qv.y, qv.z, qv.w, q, ax, ay, az, tlo_a, thi_a, qa, lia, dsta, oa, ca);
                    if (two)
                        test(qv.x, qv.y, qv.z, qv.w, q, bx, by, bz, tlo_b, thi_b, qb2, lib, dstb, ob, cb);
                };
                if (flat) {
                    // the candidate slots past `total` up to the next multiple of
                    // 64 hold the sentinel row
                    const unsigned short* cp = &L.cand[wid][lane];
                    for (int k0 = 0; k0 < total; k0 += 64, cp += 64)
                        step(static_cast<int>(*cp));
                } else {
                    for (int col = 0; col < ncols; ++col) {
                        const int cqb = __builtin_amdgcn_readlane(qb, col);
                        const int clen = __builtin_amdgcn_readlane(len, col);
                        for (int k0 = 0; k0 < clen; k0 += 64)
                            step(k0 + lane < clen ? cqb + k0 + lane : kSentinelRow);
                    }
                }
                // pad the last group with the sentinel row (staged out of reach by
                // the pair kernels) so they can read whole groups; counts, stats
                auto finish = [&](int qs, int li, uint16_t* __restrict__ dst, int count) {
                    if (!STAGED) {
                        const int pos = count + lane;
                        if (pos < ((count + kGroup - 1) & ~(kGroup - 1)) && pos < cap)
                            dst[__umul24(static_cast<unsigned>(pos) >> 3, gstride) + (pos & 7)] =
                                static_cast<uint16_t>(kSentinelRow);
                    }
                    if (lane == 0) {
                        a.numneigh[a0 + li] = count < cap ? count : cap;
                        a.self16[a0 + li] = static_cast<uint16_t>(qs);
                        my_cnt_sum += count;
                        my_max = max(my_max, count);
                        if (count > cap)
                            ++my_over;
                    }
                };
                finish(qa, lia, dsta, ca);
                if (two)
                    finish(qb2, lib, dstb, cb);
                if (STAGED) {
                    // lanes 0-31 hold atom a's groups, lanes 32-63 atom b's (at
                    // most kOutCap / kGroup = 32 each); the kept rows are the
                    // first min(count, cap) (cap is a multiple of kGroup), the
                    // tail of the last group is the sentinel
                    const int hb = lane >> 5, g = lane & 31;
                    const int cnt = hb ? cb : ca;
                    const int kept = min(cnt, cap);
                    const int pad = (kept + kGroup - 1) & ~(kGroup - 1);
                    const bool mine = hb == 0 || two;
                    if (mine && g < pad - kept)
                        (hb ? ob : oa)[kept + g] = static_cast<unsigned short>(kSentinelRow);
                    __builtin_amdgcn_wave_barrier();
                    if (mine && g < (pad >> 3)) {
                        const uint4 v = L.out[wid][hb][g];
                        *reinterpret_cast<uint4*>((hb ? dstb : dsta) + __umul24(static_cast<unsigned>(g), gstride)) = v;
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            }();
            c = __builtin_amdgcn_readfirstlane(grab);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        my_cnt_sum += __shfl_down(my_cnt_sum, off, 64);
        my_max = max(my_max, __shfl_down(my_max, off, 64));
        my_over += __shfl_down(my_over, off, 64);
    }
    if (lane == 0) {
        L.tot[wid] = my_cnt_sum;
        L.max[wid] = my_max;
        L.over[wid] = my_over;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long tt = 0;
        int m = 0, ov = 0;
        for (int w = 0; w < kWaves; ++w) {
            tt += L.tot[w];
            m = max(m, L.max[w]);
            ov += L.over[w];
        }
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.ctl->nbr_total), static_cast<unsigned long long>(tt));
        atomicMax(&a.ctl->max_nbr, m);
        if (ov)
            atomicAdd(&a.ctl->overflow, ov);
    }
}

// The grid is the brick count (the pair stage of nl_stage_grids): one brick
// per workgroup.
__device__ void OpPairs::run(const BrickBuildArgs& a, int bid, int) {
    if (!*a.pred || bid >= a.nbricks)
        return;
    __shared__ PairsLds lds;
    if (a.cap <= kOutCap)
        pairs_run<true, kBandKernarg>(a, bid, lds);
    else
        pairs_run<false, kBandKernarg>(a, bid, lds);
}

__device__ void OpPairsBatched::run(const BrickBuildArgs& a, int bid, int) {
    if (!*a.pred || bid >= a.nbricks)
        return;
    __shared__ PairsLds lds;
    if (a.cap <= kOutCap)
        pairs_run<true, kBandPlain>(a, bid, lds);
    else
        pairs_run<false, kBandPlain>(a, bid, lds);
}

// Brick-mode export: decode halo rows back to global slots, emit tag pairs a<b.
struct BrickExportArgs {
    int nbricks;
    const int* halo_rows;
    const int2* brick_range;
    const int* tag;
    const uint16_t* nbr16;
    const int* numneigh;
    int cap;
    const long long* offs; // per slot output offset (null: count only)
    long long* cnt;        // per slot count output (count pass)
    long long* out;
};

__global__ __launch_bounds__(kBlock) void k_nl_export_brick(BrickExportArgs a) {
    const int b = blockIdx.x;
    const int2 br = a.brick_range[b];
    const int a0 = br.x, nat = br.y;
    const int* rows = a.halo_rows + (int64_t)b * kHaloCap;
    const uint16_t* lst = a.nbr16 + (int64_t)a0 * a.cap;
    for (int li = threadIdx.x; li < nat; li += blockDim.x) {
        const int i = a0 + li;
        const int ti = a.tag[i];
        const int nn = a.numneigh[i];
        long long w = a.offs ? a.offs[i] : 0;
        long long c = 0;
        for (int k = 0; k < nn; ++k) {
            const int q = lst[((int64_t)(k >> 3) * nat + li) * kGroup + (k & 7)];
            const int j = rows[q] & kRowSlotMask;
            const int tj = a.tag[j];
            if (ti < tj) {
                if (a.offs) {
                    a.out[2 * w] = ti;
                    a.out[2 * w + 1] = tj;
                    ++w;
                }
                ++c;
            }
        }
        if (!a.offs)
            a.cnt[i] = c;
    }
}

// LAMMPS Neighbor::decide for one step (neighbor_list.h:128-141) — the
// displacement half (needs_rebuild, cell_list.cpp:372-391) is folded into the
// initial-integrate kernel, which raises disp_flag.
__global__ void k_nl_decide(DevCtl* ctl, int every, int delay, int check) {
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    ctl->ago += 1;
    const bool delay_ok = ctl->ago >= delay;
    const bool every_ok = (ctl->ago % every) == 0;
    const bool rb = delay_ok && every_ok && (!check || ctl->disp_flag != 0);
    ctl->disp_flag = 0;
    ctl->rebuild = rb ? 1 : 0;
    if (rb) {
        ctl->ago = 0;
        ctl->nrebuilds += 1;
        ctl->nbr_total = 0;
        ctl->max_nbr = 0;
    }
}

__global__ void k_nl_stats_reset(DevCtl* ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->nbr_total = 0;
        ctl->max_nbr = 0;
        ctl->overflow = 0;
        ctl->halo_overflow = 0;
    }
}

__global__ void k_nl_export(int64_t n, const int* __restrict__ tag, const int* __restrict__ numneigh,
                            const int* __restrict__ nbr, int64_t npad,
                            const long long* __restrict__ offs, long long* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ti = tag[i];
        long long w = offs[i];
        const int k_n = numneigh[i];
        for (int k = 0; k < k_n; ++k) {
            const int j = nbr[(int64_t)k * npad + i];
            const int tj = tag[j];
            if (ti < tj) {
                out[2 * w] = ti;
                out[2 * w + 1] = tj;
                ++w;
            }
        }
    }
}

__global__ void k_nl_export_count(int64_t n, const int* __restrict__ tag,
                                  const int* __restrict__ numneigh, const int* __restrict__ nbr,
                                  int64_t npad, long long* __restrict__ cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ti = tag[i];
        long long c = 0;
        const int k_n = numneigh[i];
        for (int k = 0; k < k_n; ++k)
            if (ti < tag[nbr[(int64_t)k * npad + i]])
                ++c;
        cnt[i] = c;
    }
}

} // namespace

namespace atvh {

// Brick size: the largest candidate whose expected halo fills at most 70% of
// the LDS rows (uniform density) and whose expected atoms fit a 256-thread
// block, never larger than the grid. The largest shape, 4x5x5, may fill 78%:
// at about 2.4 atoms per cell (dense LJ liquids at the usual cutoff) it keeps
// two threads per atom busy in the 512-thread mixed pair kernel (4x5x5 vs
// 4x4x5 at 2.048M atoms: 141 vs 158 us); the list build then measures the
// real fill and steps down when the input is denser than uniform.
constexpr int kBrickCand[kNumBrickCands][3] = {{4, 5, 5}, {4, 4, 5}, {4, 4, 4}, {3, 4, 4},
                                               {3, 3, 4}, {3, 3, 3}, {2, 3, 3}, {2, 2, 3},
                                               {2, 2, 2}, {1, 2, 2}, {1, 1, 2}, {1, 1, 1}};
constexpr bool brick_cells_fit() {
    for (const auto& c : kBrickCand)
        if (c[0] * c[1] * c[2] > kMaxBrickCells)
            return false;
    return true;
}
static_assert(brick_cells_fit(), "the list build's LDS cell table holds every brick shape's home cells");

static void choose_bricks(atvh_system* s, int first) {
    const double vol = s->box[0] * s->box[1] * s->box[2];
    const double per_cell = double(s->n) / vol * s->cs[0] * s->cs[1] * s->cs[2];
    const auto& cand = kBrickCand;
    int pick = kLastBrickCand;
    for (int k = first < kLastBrickCand ? first : kLastBrickCand; k < kNumBrickCands; ++k) {
        int d[3], hc = 1, bc = 1;
        for (int a = 0; a < 3; ++a) {
            d[a] = std::min(cand[k][a], s->nc[a]);
            hc *= d[a] + 2 * s->reach[a];
            bc *= d[a];
        }
        const double fill = k == 0 ? 0.78 : 0.7;
        if (hc <= kHaloCellCap && per_cell * hc <= fill * kHaloCap && per_cell * bc <= kBlock) {
            pick = k;
            break;
        }
    }
    for (int a = 0; a < 3; ++a) {
        s->bdim[a] = std::min(cand[pick][a], s->nc[a]);
        s->nb[a] = (s->nc[a] + s->bdim[a] - 1) / s->bdim[a];
    }
    s->nbricks = s->nb[0] * s->nb[1] * s->nb[2];
    s->brick_pick = pick;
}

// Rank of every reference cell in the brick-major layout.
static void build_rank_tables(atvh_system* s, std::vector<int>& rank, std::vector<int>& base) {
    rank.assign(s->ncells, 0);
    base.assign(s->nbricks + 1, 0);
    int next = 0;
    for (int bx = 0; bx < s->nb[0]; ++bx)
        for (int by = 0; by < s->nb[1]; ++by)
            for (int bz = 0; bz < s->nb[2]; ++bz) {
                const int b = (bx * s->nb[1] + by) * s->nb[2] + bz;
                base[b] = next;
                const int o[3] = {bx * s->bdim[0], by * s->bdim[1], bz * s->bdim[2]};
                int d[3];
                for (int a = 0; a < 3; ++a)
                    d[a] = std::min(s->bdim[a], s->nc[a] - o[a]);
                for (int lx = 0; lx < d[0]; ++lx)
                    for (int ly = 0; ly < d[1]; ++ly)
                        for (int lz = 0; lz < d[2]; ++lz) {
                            const int c = ((o[0] + lx) * s->nc[1] + (o[1] + ly)) * s->nc[2] + (o[2] + lz);
                            rank[c] = next++;
                        }
            }
    base[s->nbricks] = next;
}

// Static halo table of every brick: for each halo cell (z-fastest over the
// brick widened by the reach) the layout rank of the periodic cell it images
// and the shift code of that image (per axis 0: -L, 1: none, 2: +L).
static void build_halo_cell_tables(atvh_system* s, const std::vector<int>& rank, std::vector<int>& out) {
    if (s->ncells > kRowSlotMask)
        throw Error(ATVH_E_ARG, "brick layout supports at most 2^26 cells");
    out.assign(size_t(s->nbricks) * kHaloCellCap, 0);
    for (int bx = 0; bx < s->nb[0]; ++bx)
        for (int by = 0; by < s->nb[1]; ++by)
            for (int bz = 0; bz < s->nb[2]; ++bz) {
                const int b = (bx * s->nb[1] + by) * s->nb[2] + bz;
                const int o[3] = {bx * s->bdim[0], by * s->bdim[1], bz * s->bdim[2]};
                int dims[3];
                for (int a = 0; a < 3; ++a)
                    dims[a] = std::min(s->bdim[a], s->nc[a] - o[a]) + 2 * s->reach[a];
                if (dims[0] * dims[1] * dims[2] > kHaloCellCap)
                    throw Error(ATVH_E_ARG, "brick halo has too many cells");
                int h = 0;
                for (int hx = 0; hx < dims[0]; ++hx)
                    for (int hy = 0; hy < dims[1]; ++hy)
                        for (int hz = 0; hz < dims[2]; ++hz, ++h) {
                            int u[3] = {o[0] + hx - s->reach[0], o[1] + hy - s->reach[1], o[2] + hz - s->reach[2]};
                            int code = 0;
                            for (int a = 0; a < 3; ++a) {
                                int sc = 1;
                                if (u[a] < 0) {
                                    u[a] += s->nc[a];
                                    sc = 0;
                                } else if (u[a] >= s->nc[a]) {
                                    u[a] -= s->nc[a];
                                    sc = 2;
                                }
                                code |= sc << (2 * a);
                            }
                            const int c = (u[0] * s->nc[1] + u[1]) * s->nc[2] + u[2];
                            out[size_t(b) * kHaloCellCap + h] = rank[c] | (code << kRowSlotBits);
                        }
            }
}

BrickGeom brick_geom(const atvh_system* s) {
    BrickGeom g{};
    for (int a = 0; a < 3; ++a) {
        g.nc[a] = s->nc[a];
        g.R[a] = s->reach[a];
        g.bdim[a] = s->bdim[a];
        g.nb[a] = s->nb[a];
        g.L[a] = s->box[a];
        g.cs[a] = s->cs[a];
    }
    g.cell_rank = s->cell_rank;
    g.brick_cell_base = s->brick_cell_base;
    g.max_atoms = brick_atom_limit(s);
    return g;
}

int64_t nl_device_bytes(const atvh_system* s) {
    if (!s->nl_configured || s->ncells == 0)
        return 0;
    const int64_t n = s->n, nc = s->ncells, nbk = s->nbricks;
    int64_t b = s->brick_mode ? int64_t(s->cap) * n * 2 : int64_t(s->cap) * s->npad * 4;
    b += n * (4 * 4 + 2);                                   // cell_of, acell_of, perm, numneigh, self16
    b += nc * 4 * 4 + 4;                                    // cell_rank, cell_count, cell_fill, cell_start
    b += int64_t(s->nscan_blocks) * 4 + (nbk + 1) * 4;      // scan_sums, brick_cell_base
    b += int64_t(s->nstencil) * 16 + int64_t(s->ncols) * 16; // stencil, columns
    if (s->brick_mode)
        b += nbk * (int64_t(kHaloCellCap) * 4 + int64_t(kHaloCap) * 4 + (kHaloCellCap + 1) * 4 + 4 + 8);
    return b;
}

void nl_realloc_capacity(atvh_system* s, int cap) {
    dev_free(s->nbr);
    dev_free(s->nbr16);
    s->cap = (cap + 7) / 8 * 8;
    if (s->brick_mode)
        dev_alloc(&s->nbr16, size_t(s->cap) * s->n);
    else
        dev_alloc(&s->nbr, size_t(s->cap) * s->npad);
}

// CellList::compute_grid + build_stencil on the host (box-change only).
void nl_setup_grid(atvh_system* s) {
    const double cutoff_skin = s->cutoff_skin;
    for (int d = 0; d < 3; ++d) {
        const double L = s->box[d];
        const int nc = std::max(int(1), static_cast<int>(std::floor(L / (0.5 * cutoff_skin))));
        s->nc[d] = nc;
        s->cs[d] = L / nc;
    }
    for (int d = 0; d < 3; ++d)
        s->reach[d] = std::max(1, static_cast<int>(std::ceil(cutoff_skin / s->cs[d] - 1e-12)));
    const int64_t ncells64 = int64_t(s->nc[0]) * s->nc[1] * s->nc[2];
    if (ncells64 > (int64_t(1) << 28))
        throw Error(ATVH_E_ARG, "neighbor grid too large");
    s->ncells = static_cast<int>(ncells64);
    s->use_half = true;
    for (int d = 0; d < 3; ++d)
        if (s->nc[d] < 2 * s->reach[d] + 1)
            s->use_half = false;
    const double rc_sq = cutoff_skin * cutoff_skin;
    auto min_dist_sq = [&](int di, int dj, int dk) {
        const double ddx = di > 0 ? (di - 1) * s->cs[0] : (di < 0 ? (di + 1) * s->cs[0] : 0.0);
        const double ddy = dj > 0 ? (dj - 1) * s->cs[1] : (dj < 0 ? (dj + 1) * s->cs[1] : 0.0);
        const double ddz = dk > 0 ? (dk - 1) * s->cs[2] : (dk < 0 ? (dk + 1) * s->cs[2] : 0.0);
        return ddx * ddx + ddy * ddy + ddz * ddz;
    };
    std::vector<int> st;
    for (int di = -s->reach[0]; di <= s->reach[0]; ++di)
        for (int dj = -s->reach[1]; dj <= s->reach[1]; ++dj)
            for (int dk = -s->reach[2]; dk <= s->reach[2]; ++dk) {
                if (min_dist_sq(di, dj, dk) > rc_sq)
                    continue;
                const bool home = (di == 0 && dj == 0 && dk == 0);
                const bool forward = (di > 0) || (di == 0 && dj > 0) || (di == 0 && dj == 0 && dk > 0);
                st.push_back(di);
                st.push_back(dj);
                st.push_back(dk);
                st.push_back(home ? 0 : (forward ? 1 : 2));
            }
    s->nstencil = static_cast<int>(st.size() / 4);
    // the stencil as columns (di, dj, kmax): for fixed (di, dj) the kept dk
    // form the symmetric range -kmax..kmax (min distance grows with |dk|)
    std::vector<int> cols;
    for (int di = -s->reach[0]; di <= s->reach[0]; ++di)
        for (int dj = -s->reach[1]; dj <= s->reach[1]; ++dj) {
            int kmax = -1;
            for (int dk = 0; dk <= s->reach[2]; ++dk)
                if (min_dist_sq(di, dj, dk) <= rc_sq)
                    kmax = dk;
            if (kmax < 0)
                continue;
            cols.push_back(di);
            cols.push_back(dj);
            cols.push_back(kmax);
            cols.push_back(0);
        }
    s->ncols = static_cast<int>(cols.size() / 4);
    if (s->ncols > kMaxCols)
        throw Error(ATVH_E_ARG, "stencil has more columns than the list build supports");
    nl_free(s);
    dev_alloc(&s->d_cols, cols.size());
    ATVH_HIP(hipMemcpy(s->d_cols, cols.data(), cols.size() * sizeof(int), hipMemcpyHostToDevice));
    // brick layout when the half-stencil arithmetic applies, else one brick
    // per cell and the global-slot fallback list
    s->brick_mode = s->use_half;
    if (s->brick_mode) {
        choose_bricks(s, s->brick_pick_min);
    } else {
        for (int a = 0; a < 3; ++a) {
            s->bdim[a] = 1;
            s->nb[a] = s->nc[a];
        }
        s->nbricks = s->ncells;
    }
    std::vector<int> rank, base;
    build_rank_tables(s, rank, base);
    std::vector<int> hcells;
    if (s->brick_mode)
        build_halo_cell_tables(s, rank, hcells);
    // per-block reduction partials: one row per pair workgroup (bricks) or grid block
    const int need = std::max(kMaxGrid, s->nbricks);
    if (need > s->npartials) {
        dev_free(s->partials);
        s->npartials = need;
        dev_alloc(&s->partials, size_t(2 * s->npartials) * kPartialStride); // pair | kinetic
    }
    dev_alloc(&s->d_stencil, st.size());
    ATVH_HIP(hipMemcpy(s->d_stencil, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice));
    if (s->brick_mode) {
        dev_alloc(&s->halo_cells, hcells.size());
        ATVH_HIP(hipMemcpy(s->halo_cells, hcells.data(), hcells.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    dev_alloc(&s->cell_rank, s->ncells);
    ATVH_HIP(hipMemcpy(s->cell_rank, rank.data(), rank.size() * sizeof(int), hipMemcpyHostToDevice));
    dev_alloc(&s->brick_cell_base, s->nbricks + 1);
    ATVH_HIP(hipMemcpy(s->brick_cell_base, base.data(), base.size() * sizeof(int), hipMemcpyHostToDevice));
    dev_alloc(&s->cell_of, s->n);
    dev_alloc(&s->acell_of, s->n);
    dev_alloc(&s->perm, s->n);
    dev_alloc(&s->cell_count, s->ncells);
    dev_alloc(&s->cell_fill, s->ncells);
    dev_alloc(&s->cell_start, s->ncells + 1);
    s->nscan_blocks = (s->ncells + kScanTile - 1) / kScanTile;
    dev_alloc(&s->scan_sums, s->nscan_blocks);
    ATVH_HIP(hipMemset(s->cell_count, 0, sizeof(int) * s->ncells));
    ATVH_HIP(hipMemset(s->cell_fill, 0, sizeof(int) * s->ncells));
    dev_alloc(&s->numneigh, s->n);
    dev_alloc(&s->self16, s->n);
    if (s->brick_mode) {
        if (s->n > kRowSlotMask)
            throw Error(ATVH_E_ARG, "brick layout supports at most 2^26 atoms per system");
        dev_alloc(&s->halo_rows, size_t(s->nbricks) * kHaloCap);
        dev_alloc(&s->halo_off, size_t(s->nbricks) * (kHaloCellCap + 1));
        dev_alloc(&s->halo_n, s->nbricks);
        dev_alloc(&s->brick_range, s->nbricks);
    }
    // capacity: 1.5x the mean full-list count at this density, plus margin
    const double vol = s->box[0] * s->box[1] * s->box[2];
    const double mean = double(s->n) / vol * 4.18879020478639 * cutoff_skin * cutoff_skin * cutoff_skin;
    int cap = static_cast<int>(std::ceil(1.5 * mean + 32));
    cap = (cap + 7) / 8 * 8;
    if (cap > s->n)
        cap = static_cast<int>(std::max<int64_t>(8, (s->n + 7) / 8 * 8));
    nl_realloc_capacity(s, cap);
}

void nl_free(atvh_system* s) {
    dev_free(s->d_stencil);
    dev_free(s->halo_cells);
    dev_free(s->d_cols);
    dev_free(s->cell_rank);
    dev_free(s->brick_cell_base);
    dev_free(s->cell_of);
    dev_free(s->acell_of);
    dev_free(s->perm);
    dev_free(s->cell_count);
    dev_free(s->cell_fill);
    dev_free(s->cell_start);
    dev_free(s->scan_sums);
    dev_free(s->numneigh);
    dev_free(s->self16);
    dev_free(s->nbr);
    dev_free(s->nbr16);
    dev_free(s->halo_rows);
    dev_free(s->halo_off);
    dev_free(s->halo_n);
    dev_free(s->brick_range);
}

SortArgs nl_sort_args(atvh_system* s, const int* pred, bool permute_forces) {
    SortArgs a{};
    a.pred = pred;
    a.n = s->n;
    a.ncells = s->ncells;
    a.x = s->x;
    a.y = s->y;
    a.z = s->z;
    a.lx = s->box[0];
    a.ly = s->box[1];
    a.lz = s->box[2];
    a.csx = s->cs[0];
    a.csy = s->cs[1];
    a.csz = s->cs[2];
    a.ncx = s->nc[0];
    a.ncy = s->nc[1];
    a.ncz = s->nc[2];
    a.cell_rank = s->cell_rank;
    a.cell_of = s->acell_of; // pre-sort order; the permutation writes s->cell_of
    a.count = s->cell_count;
    a.start = s->cell_start;
    a.tile_sums = s->scan_sums;
    a.ntiles = s->nscan_blocks;
    a.fill = s->cell_fill;
    a.perm = s->perm;
    a.tag = s->tag;
    a.vx = s->vx;
    a.vy = s->vy;
    a.vz = s->vz;
    a.fx = s->fx;
    a.fy = s->fy;
    a.fz = s->fz;
    a.q = s->q;
    a.type = s->type;
    a.ox = s->ax;
    a.oy = s->ay;
    a.oz = s->az;
    a.ovx = s->avx;
    a.ovy = s->avy;
    a.ovz = s->avz;
    a.ofx = s->afx;
    a.ofy = s->afy;
    a.ofz = s->afz;
    a.oq = s->aq;
    a.otype = s->atype;
    a.otag = s->atag;
    a.with_forces = permute_forces ? 1 : 0;
    a.with_q = s->has_charge ? 1 : 0;
    a.with_type = s->ntypes > 1 ? 1 : 0;
    return a;
}

PermArgs nl_perm_args(atvh_system* s, const int* pred, bool permute_forces) {
    PermArgs a{};
    a.pred = pred;
    a.n = s->n;
    a.perm = s->perm;
    a.ox = s->ax;
    a.oy = s->ay;
    a.oz = s->az;
    a.ovx = s->avx;
    a.ovy = s->avy;
    a.ovz = s->avz;
    a.ofx = s->afx;
    a.ofy = s->afy;
    a.ofz = s->afz;
    a.oq = s->aq;
    a.otype = s->atype;
    a.otag = s->atag;
    a.ocell_of = s->acell_of;
    a.x = s->x;
    a.y = s->y;
    a.z = s->z;
    a.vx = s->vx;
    a.vy = s->vy;
    a.vz = s->vz;
    a.fx = s->fx;
    a.fy = s->fy;
    a.fz = s->fz;
    a.q = s->q;
    a.xs = s->xs;
    a.ys = s->ys;
    a.zs = s->zs;
    a.type = s->type;
    a.tag = s->tag;
    a.cell_of = s->cell_of;
    a.with_forces = permute_forces ? 1 : 0;
    a.with_q = s->has_charge ? 1 : 0;
    a.with_type = s->ntypes > 1 ? 1 : 0;
    a.with_cell_of = s->brick_mode ? 0 : 1;
    return a;
}

BrickBuildArgs nl_brick_args(atvh_system* s, const int* pred) {
    BrickBuildArgs ba{};
    ba.pred = pred;
    ba.g = brick_geom(s);
    ba.halo_cells = s->halo_cells;
    ba.nbricks = s->nbricks;
    ba.halo_rows = s->halo_rows;
    ba.halo_n = s->halo_n;
    ba.brick_range = s->brick_range;
    ba.x = s->x;
    ba.y = s->y;
    ba.z = s->z;
    ba.cell_of = s->cell_of;
    ba.start = s->cell_start;
    ba.rows = reinterpret_cast<const int4*>(s->d_cols);
    ba.nrows = s->ncols;
    ba.rc_sq = s->cutoff_skin * s->cutoff_skin;
    ba.rcs = s->cutoff_skin;
    ba.nbr16 = s->nbr16;
    ba.self16 = s->self16;
    ba.numneigh = s->numneigh;
    ba.cap = s->cap;
    ba.ctl = s->ctl;
    ba.halo_off = s->halo_off;
    return ba;
}

// Launch grids of the rebuild stages for one system (the batched rebuild
// gives every system the same block counts).
// Every stage but the pair build strides over its units, so their grids are
// capped at kStageGridCap blocks (a multiple of the 8 XCDs, which keeps the
// XCD-aware brick order): on the steps without a rebuild each stage is
// launched and leaves at its first instruction, and a smaller grid costs less
// to drain. The pair build takes one workgroup per brick: with no loop over
// bricks its candidate loop keeps more scalar registers (C1 rebuild 751 ->
// 711 us, 512k-ion melt 318 -> 282 us), worth the ~3 us more that its idle
// launch takes to drain.
constexpr int kStageGridCap = 2048;

// Block counts of the rebuild kernels, one more than the stages: in brick
// mode the permutation (g[5]) and the halo tables (g[6]) share a launch.
void nl_stage_grids(const atvh_system* s, int g[kNlStages + 1]) {
    const int ga = std::min(grid_capped(s->n), kStageGridCap);
    g[0] = ga;                                                   // bin
    g[1] = s->nscan_blocks;                                      // scan tiles
    g[2] = 1;                                                    // scan sums
    g[3] = ga;                                                   // scatter
    g[4] = std::min(grid_capped(s->ncells), kStageGridCap);      // in-cell order (final cell starts)
    g[5] = ga;                                                   // permute (+ x_save); the halo
    g[6] = std::min(s->nbricks, kStageGridCap);                  //   tables' blocks follow in its launch
    g[7] = s->nbricks;                                           // pair build: one workgroup per brick
}

void nl_enqueue_rebuild(atvh_system* s, const int* pred, bool permute_forces) {
    ScopedTimer tm(s->ctx, "nlist");
    hipStream_t st = s->ctx->stream;
    int g[kNlStages + 1];
    nl_stage_grids(s, g);
    const SortArgs sa = nl_sort_args(s, pred, permute_forces);
    k_single<OpBin><<<g[0], kBlock, 0, st>>>(sa);
    k_single<OpScanTiles><<<g[1], kBlock, 0, st>>>(sa);
    k_single<OpScanSums><<<g[2], kBlock, 0, st>>>(sa);
    k_single<OpScatter><<<g[3], kBlock, 0, st>>>(sa);
    k_single<OpCellOrder><<<g[4], kBlock, 0, st>>>(sa);
    if (s->brick_mode) {
        const BrickBuildArgs ba = nl_brick_args(s, pred);
        const PermHaloArgs ph{nl_perm_args(s, pred, permute_forces), ba, g[5]};
        k_single<OpPermuteHalo><<<g[5] + g[6], kBlock, 0, st>>>(ph);
        k_single<OpPairs><<<g[7], OpPairs::kThreads, 0, st>>>(ba);
    } else {
        k_single<OpPermute><<<g[5], kBlock, 0, st>>>(nl_perm_args(s, pred, permute_forces));
        BuildArgs ba{s->x, s->y, s->z, s->cell_of, s->cell_rank, s->cell_start,
                     reinterpret_cast<const int4*>(s->d_stencil), s->nstencil, s->nc[0], s->nc[1],
                     s->nc[2], s->box[0], s->box[1], s->box[2], s->cutoff_skin * s->cutoff_skin,
                     0, s->nbr, s->numneigh, s->cap, s->npad, s->ctl};
        k_nl_build<false><<<grid_for(s->n), kBlock, 0, st>>>(pred, s->n, ba);
    }
    ATVH_HIP(hipGetLastError());
}

// Batched rebuild: the seven stages of a brick-mode rebuild, one launch each
// over every member; a member whose rebuild flag is clear leaves each stage
// at its first instruction.
void nl_batch_build(atvh_batch* b) {
    std::vector<SortArgs> sa;
    std::vector<PermHaloArgs> pa;
    std::vector<BrickBuildArgs> ba;
    std::vector<int> g[kNlStages];
    for (atvh_system* s : b->sys) {
        if (!s->brick_mode)
            throw Error(ATVH_E_ARG, "batch: every member needs the brick layout (>= 2*reach+1 cells per axis)");
        const int* pred = &s->ctl->rebuild;
        sa.push_back(nl_sort_args(s, pred, false));
        ba.push_back(nl_brick_args(s, pred));
        int gs[kNlStages + 1];
        nl_stage_grids(s, gs);
        pa.push_back(PermHaloArgs{nl_perm_args(s, pred, false), ba.back(), gs[5]});
        gs[5] += gs[6]; // the halo tables' blocks ride in the permutation's launch
        gs[6] = gs[7];
        for (int k = 0; k < kNlStages; ++k)
            g[k].push_back(gs[k]);
    }
    batch_stage(b, "nl_bin", sa, g[0]);
    batch_stage(b, "nl_scan_tiles", sa, g[1]);
    batch_stage(b, "nl_scan_sums", sa, g[2]);
    batch_stage(b, "nl_scatter", sa, g[3]);
    batch_stage(b, "nl_cell_order", sa, g[4]);
    batch_stage(b, "nl_permute_halo", pa, g[5]);
    batch_stage(b, "nl_pairs", ba, g[6]);
    const size_t nsys = b->sys.size();
    std::vector<int> nb;
    for (int k = 0; k < kNlStages; ++k)
        nb.insert(nb.end(), g[k].begin(), g[k].end());
    dev_free(b->nl_nb);
    dev_free(b->nl_act);
    dev_free(b->nl_flags);
    dev_free(b->nl_next);
    dev_alloc(&b->nl_nb, nb.size());
    dev_alloc(&b->nl_act, size_t(kNlStages) * (nsys + 1));
    dev_alloc(&b->nl_flags, nsys);
    dev_alloc(&b->nl_next, 1);
    ATVH_HIP(hipMemcpy(b->nl_nb, nb.data(), nb.size() * sizeof(int), hipMemcpyHostToDevice));
}

// This step's prefixes, per rebuild stage, of the block counts of the members
// whose rebuild flag is set: the flags are read once, then wavefront k scans
// stage k in member order, 64 members at a time (DPP prefix sums).
constexpr int kActiveThreads = 64 * kNlStages;

__global__ __launch_bounds__(kActiveThreads) void k_nl_active_prefix(DevCtl* const* __restrict__ ctls, int nsys,
                                                                      const int* __restrict__ nb,
                                                                      int* __restrict__ act,
                                                                      unsigned char* __restrict__ flags,
                                                                      int* __restrict__ next) {
    if (threadIdx.x == 0)
        *next = 0; // the pair build's work counter (a kernel write: graph replays need no memset)
    for (int m = threadIdx.x; m < nsys; m += blockDim.x)
        flags[m] = ctls[m]->rebuild ? 1 : 0;
    __syncthreads();
    const int k = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int* nbk = nb + (int64_t)k * nsys;
    int* ak = act + (int64_t)k * (nsys + 1);
    if (lane == 0)
        ak[0] = 0;
    int carry = 0;
    for (int c0 = 0; c0 < nsys; c0 += 64) {
        const int m = c0 + lane;
        const int v = (m < nsys && flags[m]) ? nbk[m] : 0;
        const int inc = wave_inclusive_sum(v);
        if (m < nsys)
            ak[m + 1] = carry + inc;
        carry += __builtin_amdgcn_readlane(inc, 63);
    }
}

// One rebuild stage over the rebuilding members only: a grid of about eight
// 256-thread blocks' worth per CU strides over their blocks.
template <class Op>
static void nl_batch_one(atvh_batch* b, const char* name, int stage) {
    const BatchStage& st = b->stages.at(name);
    const int nsys = static_cast<int>(b->sys.size());
    const int grid = std::max(1, std::min(st.nblocks, 2048 * kBlock / op_threads<Op>::value));
    k_batched_active<Op><<<grid, op_threads<Op>::value, 0, b->ctx->stream>>>(
        static_cast<const typename Op::Args*>(st.args), b->nl_act + (int64_t)stage * (nsys + 1), nsys);
}

void nl_batch_launch_rebuild(atvh_batch* b) {
    k_nl_active_prefix<<<1, kActiveThreads, 0, b->ctx->stream>>>(b->d_ctls, static_cast<int>(b->sys.size()),
                                                                 b->nl_nb, b->nl_act, b->nl_flags, b->nl_next);
    nl_batch_one<OpBin>(b, "nl_bin", 0);
    nl_batch_one<OpScanTiles>(b, "nl_scan_tiles", 1);
    nl_batch_one<OpScanSums>(b, "nl_scan_sums", 2);
    nl_batch_one<OpScatter>(b, "nl_scatter", 3);
    nl_batch_one<OpCellOrder>(b, "nl_cell_order", 4);
    nl_batch_one<OpPermuteHalo>(b, "nl_permute_halo", 5);
    {
        // the pair build: bricks differ widely in cost (member size, density),
        // so its resident blocks take them from a shared counter
        const BatchStage& st = b->stages.at("nl_pairs");
        const int nsys = static_cast<int>(b->sys.size());
        const int grid = std::max(1, std::min(st.nblocks, 2048 * kBlock / OpPairsBatched::kThreads));
        k_batched_queue<OpPairsBatched><<<grid, OpPairsBatched::kThreads, 0, b->ctx->stream>>>(
            static_cast<const BrickBuildArgs*>(st.args), b->nl_act + (int64_t)6 * (nsys + 1), nsys, b->nl_next);
    }
    ATVH_HIP(hipGetLastError());
}

void nl_enqueue_decide(atvh_system* s) {
    k_nl_decide<<<1, 64, 0, s->ctx->stream>>>(s->ctl, s->every, s->delay, s->check);
    ATVH_HIP(hipGetLastError());
}

void nl_check_ctl(const DevCtl& h) {
    if (h.halo_overflow)
        throw Error(ATVH_E_CAPACITY, "a brick halo exceeded the LDS capacity (density far above uniform)");
    if (h.overflow)
        throw Error(ATVH_E_CAPACITY, "neighbor capacity exceeded during a rebuild");
}

// Fullest halo and brick of the last build (per-brick counts read back; a
// list build only).
static void brick_fill(atvh_system* s, int& max_halo, int& max_atoms) {
    std::vector<int> hn(s->nbricks);
    std::vector<int2> br(s->nbricks);
    ATVH_HIP(hipMemcpy(hn.data(), s->halo_n, hn.size() * sizeof(int), hipMemcpyDeviceToHost));
    ATVH_HIP(hipMemcpy(br.data(), s->brick_range, br.size() * sizeof(int2), hipMemcpyDeviceToHost));
    max_halo = 0;
    max_atoms = 0;
    for (int b = 0; b < s->nbricks; ++b) {
        max_halo = std::max(max_halo, hn[b]);
        max_atoms = std::max(max_atoms, br[b].y);
    }
}

void nl_build(atvh_system* s, bool keep_counts) {
    atvh_ctx* c = s->ctx;
    for (int attempt = 0; attempt < 12; ++attempt) {
        k_nl_stats_reset<<<1, 64, 0, c->stream>>>(s->ctl);
        // the initial build also orders the frame's forces (first kick uses them)
        nl_enqueue_rebuild(s, &s->ctl->always, true);
        ATVH_HIP(hipStreamSynchronize(c->stream));
        DevCtl h{};
        ATVH_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
        // a brick shape whose fullest halo or brick (measured on these
        // positions) leaves less than a sixteenth of the LDS rows or of the
        // brick's atom slots free is given up for the next smaller one:
        // denser-than-uniform inputs keep some headroom for the rebuilds of
        // the run
        bool tight = false;
        if (s->brick_mode && s->brick_pick < kLastBrickCand && h.halo_overflow == 0) {
            int max_halo = 0, max_atoms = 0;
            brick_fill(s, max_halo, max_atoms);
            tight = max_halo > kHaloFillMax || max_atoms > brick_atom_limit(s) * 15 / 16;
        }
        if (h.overflow == 0 && h.halo_overflow == 0 && !tight) {
            h.ago = 0;
            if (!keep_counts)
                h.nrebuilds = 1;
            h.disp_flag = 0;
            h.rebuild = 0;
            ATVH_HIP(hipMemcpy(s->ctl, &h, sizeof(h), hipMemcpyHostToDevice));
            s->nl_built = true;
            for (int d = 0; d < 3; ++d)
                s->nl_box[d] = s->box[d];
            return;
        }
        if (h.halo_overflow || tight) {
            // a denser-than-uniform region: fall back to the next smaller brick
            if (s->brick_pick >= kLastBrickCand)
                break;
            s->brick_pick_min = s->brick_pick + 1;
            nl_setup_grid(s);
            continue;
        }
        // grow and rebuild from the (now sorted) positions
        nl_realloc_capacity(s, static_cast<int>(h.max_nbr * 1.25) + 16);
    }
    throw Error(ATVH_E_CAPACITY, "neighbor list build: the capacity could not be satisfied");
}

} // namespace atvh

#define API_TRY try {
#define API_CATCH(C)                                                                    \
    }                                                                                   \
    catch (const atvh::Error& e_) {                                                     \
        if (C)                                                                          \
            (C)->err = e_.what();                                                       \
        return e_.code;                                                                 \
    }                                                                                   \
    catch (const std::exception& e_) {                                                  \
        if (C)                                                                          \
            (C)->err = e_.what();                                                       \
        return ATVH_E_ARG;                                                              \
    }

extern "C" {

int atvh_nlist_configure(atvh_system* s, double cutoff, double skin, int every, int delay,
                         int check) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (!(cutoff > 0.0))
        throw Error(ATVH_E_ARG, "atvh_nlist_configure: cutoff must be positive");
    ATVH_HIP(hipSetDevice(c->device));
    ATVH_HIP(hipStreamSynchronize(c->stream));
    // NeighborList(cutoff, skin, every, delay, check) (neighbor_list.h:79-83),
    // CellList(cutoff, skin): skin < 0 -> 0.3*cutoff (cell_list.cpp:24-26)
    s->cutoff = cutoff;
    s->skin = skin < 0 ? 0.3 * cutoff : skin;
    s->cutoff_skin = cutoff + s->skin;
    s->every = every < 1 ? 1 : every;
    s->delay = delay < 0 ? 0 : delay;
    s->check = check ? 1 : 0;
    nl_setup_grid(s);
    s->nl_configured = true;
    s->nl_built = false;
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_nlist_build(atvh_system* s) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (!s->nl_configured)
        throw Error(ATVH_E_STATE, "atvh_nlist_build: neighbor list not configured");
    ATVH_HIP(hipSetDevice(c->device));
    nl_build(s, false);
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_nlist_set_capacity(atvh_system* s, int capacity) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (!s->nl_configured)
        throw Error(ATVH_E_STATE, "atvh_nlist_set_capacity: neighbor list not configured");
    if (capacity < 1)
        throw Error(ATVH_E_ARG, "atvh_nlist_set_capacity: capacity must be positive");
    ATVH_HIP(hipSetDevice(c->device));
    ATVH_HIP(hipStreamSynchronize(c->stream));
    nl_realloc_capacity(s, capacity);
    s->nl_built = false;
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_nlist_brick_shape(atvh_system* s, int* candidate, int* bdim) {
    if (!s)
        return ATVH_E_ARG;
    if (candidate)
        *candidate = s->brick_mode ? s->brick_pick : -1;
    if (bdim)
        for (int a = 0; a < 3; ++a)
            bdim[a] = s->bdim[a];
    return ATVH_OK;
}

int atvh_nlist_set_brick_shape(atvh_system* s, int candidate) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (!s->nl_configured)
        throw Error(ATVH_E_STATE, "atvh_nlist_set_brick_shape: neighbor list not configured");
    if (candidate < 0 || candidate > kLastBrickCand)
        throw Error(ATVH_E_ARG, "atvh_nlist_set_brick_shape: candidate out of range [0, " +
                                    std::to_string(kLastBrickCand) + "]");
    ATVH_HIP(hipSetDevice(c->device));
    ATVH_HIP(hipStreamSynchronize(c->stream));
    if (s->brick_mode) {
        s->brick_pick_min = candidate;
        nl_setup_grid(s);
        s->nl_built = false;
    }
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_nlist_brick_fill(atvh_system* s, int* max_halo, int* max_atoms) {
    if (!s || !max_halo || !max_atoms)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    *max_halo = *max_atoms = 0;
    if (!s->brick_mode)
        return ATVH_OK;
    if (!s->nl_built)
        throw Error(ATVH_E_STATE, "atvh_nlist_brick_fill: no neighbor list");
    ATVH_HIP(hipSetDevice(c->device));
    ATVH_HIP(hipStreamSynchronize(c->stream));
    brick_fill(s, *max_halo, *max_atoms);
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_nlist_set_brick_limit(atvh_system* s, int max_atoms) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (max_atoms < 0 || max_atoms > kMaxBrickAtoms)
        throw Error(ATVH_E_ARG, "atvh_nlist_set_brick_limit: max_atoms must be in [0, " +
                                    std::to_string(kMaxBrickAtoms) + "] (0: " + std::to_string(kMaxBrickAtoms) +
                                    ")");
    ATVH_HIP(hipSetDevice(c->device));
    ATVH_HIP(hipStreamSynchronize(c->stream));
    s->brick_atom_limit = max_atoms;
    if (s->brick_mode && s->nl_built) {
        // the list stays valid while no brick it holds is over the new limit
        int mh = 0, ma = 0;
        brick_fill(s, mh, ma);
        if (ma > brick_atom_limit(s))
            s->nl_built = false;
    }
    return ATVH_OK;
    API_CATCH(c)
}

// Brick-shape autotuner: for the current shape and the next two smaller
// valid ones, build the list, time one rebuild and `trials` pair launches
// with HIP events, and keep the shape of least (pair + rebuild / 8) time
// (a rebuild every ~8 steps is typical at skin 0.3). The forces of the frame
// are restored after every timing; the list is rebuilt with the winner and
// the rebuild counter restarts at 1.
int atvh_nlist_autotune(atvh_system* s, int precision, int trials, int* chosen, double* best_us) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (!s->nl_configured || !s->pair_configured)
        throw Error(ATVH_E_STATE, "atvh_nlist_autotune: configure the neighbor list and the pair style first");
    if (trials < 1)
        throw Error(ATVH_E_ARG, "atvh_nlist_autotune: trials must be >= 1");
    ATVH_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    if (!s->brick_mode) {
        if (chosen)
            *chosen = -1;
        return ATVH_OK;
    }
    const int k0 = s->brick_pick;
    double* save = nullptr;
    dev_alloc(&save, size_t(3) * s->n);
    hipEvent_t ev[4];
    for (auto& e : ev)
        ATVH_HIP(hipEventCreate(&e));
    int best = k0;
    double best_t = 1e300;
    std::vector<int> tried;
    try {
        for (int k = k0; k <= std::min(k0 + 2, kLastBrickCand); ++k) {
            s->brick_pick_min = k;
            nl_setup_grid(s);
            if (std::find(tried.begin(), tried.end(), s->brick_pick) != tried.end())
                continue;
            tried.push_back(s->brick_pick);
            s->nl_built = false;
            if (atvh_nlist_build(s) != ATVH_OK)
                continue;
            ATVH_HIP(hipEventRecord(ev[0], st));
            nl_enqueue_rebuild(s, &s->ctl->always, true);
            ATVH_HIP(hipEventRecord(ev[1], st));
            const size_t nb = size_t(s->n) * sizeof(double);
            ATVH_HIP(hipMemcpyAsync(save, s->fx, nb, hipMemcpyDeviceToDevice, st));
            ATVH_HIP(hipMemcpyAsync(save + s->n, s->fy, nb, hipMemcpyDeviceToDevice, st));
            ATVH_HIP(hipMemcpyAsync(save + 2 * s->n, s->fz, nb, hipMemcpyDeviceToDevice, st));
            ATVH_HIP(hipEventRecord(ev[2], st));
            for (int t = 0; t < trials; ++t)
                pair_enqueue(s, precision, false, false);
            ATVH_HIP(hipEventRecord(ev[3], st));
            ATVH_HIP(hipMemcpyAsync(s->fx, save, nb, hipMemcpyDeviceToDevice, st));
            ATVH_HIP(hipMemcpyAsync(s->fy, save + s->n, nb, hipMemcpyDeviceToDevice, st));
            ATVH_HIP(hipMemcpyAsync(s->fz, save + 2 * s->n, nb, hipMemcpyDeviceToDevice, st));
            ATVH_HIP(hipStreamSynchronize(st));
            float t_rb = 0.f, t_pair = 0.f;
            ATVH_HIP(hipEventElapsedTime(&t_rb, ev[0], ev[1]));
            ATVH_HIP(hipEventElapsedTime(&t_pair, ev[2], ev[3]));
            const double cost = 1e3 * (double(t_pair) / trials + double(t_rb) / 8.0);
            if (cost < best_t) {
                best_t = cost;
                best = s->brick_pick;
            }
        }
    } catch (...) {
        for (auto& e : ev)
            (void)hipEventDestroy(e);
        dev_free(save);
        throw;
    }
    for (auto& e : ev)
        (void)hipEventDestroy(e);
    dev_free(save);
    s->brick_pick_min = best;
    nl_setup_grid(s);
    s->nl_built = false;
    const int rc = atvh_nlist_build(s);
    if (rc != ATVH_OK)
        return rc;
    if (chosen)
        *chosen = best;
    if (best_us)
        *best_us = best_t;
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_nlist_decide_and_rebuild(atvh_system* s) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (!s->nl_built)
        throw Error(ATVH_E_STATE, "atvh_nlist_decide_and_rebuild: call atvh_nlist_build first");
    nl_enqueue_decide(s);
    nl_enqueue_rebuild(s, &s->ctl->rebuild, false);
    // the overflow probe is read back here, as the reference's list rebuild
    // reads its pair count (src/cuda/neighbor/neighbor_list.cu:544-553): an
    // overflowing rebuild is redone at once with a grown capacity
    ATVH_HIP(hipStreamSynchronize(c->stream));
    DevCtl h{};
    ATVH_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
    if (h.overflow || h.halo_overflow)
        nl_build(s, true);
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_nlist_stats(atvh_system* s, int64_t* nrebuilds, int64_t* npairs_half, int* max_neighbors,
                     int* capacity) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    ATVH_HIP(hipStreamSynchronize(c->stream));
    DevCtl h{};
    ATVH_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
    if (nrebuilds)
        *nrebuilds = h.nrebuilds;
    if (npairs_half)
        *npairs_half = h.nbr_total / 2;
    if (max_neighbors)
        *max_neighbors = h.max_nbr;
    if (capacity)
        *capacity = s->cap;
    nl_check_ctl(h);
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_nlist_export_pairs(atvh_system* s, int64_t* pairs, int64_t cap, int64_t* npairs) {
    if (!s || !npairs)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (!s->nl_built)
        throw Error(ATVH_E_STATE, "atvh_nlist_export_pairs: no neighbor list");
    ATVH_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int64_t n = s->n;
    {
        DevCtl hc{};
        ATVH_HIP(hipStreamSynchronize(st));
        ATVH_HIP(hipMemcpy(&hc, s->ctl, sizeof(hc), hipMemcpyDeviceToHost));
        nl_check_ctl(hc);
    }
    long long* d_cnt = nullptr;
    dev_alloc(&d_cnt, n);
    BrickExportArgs ea;
    if (s->brick_mode) {
        ea.nbricks = s->nbricks;
        ea.halo_rows = s->halo_rows;
        ea.brick_range = s->brick_range;
        ea.tag = s->tag;
        ea.nbr16 = s->nbr16;
        ea.numneigh = s->numneigh;
        ea.cap = s->cap;
        ea.offs = nullptr;
        ea.cnt = d_cnt;
        ea.out = nullptr;
        k_nl_export_brick<<<s->nbricks, kBlock, 0, st>>>(ea);
    } else {
        k_nl_export_count<<<grid_capped(n), kBlock, 0, st>>>(n, s->tag, s->numneigh, s->nbr, s->npad,
                                                             d_cnt);
    }
    ATVH_HIP(hipGetLastError());
    std::vector<long long> h(n);
    ATVH_HIP(hipMemcpyAsync(h.data(), d_cnt, n * sizeof(long long), hipMemcpyDeviceToHost, st));
    ATVH_HIP(hipStreamSynchronize(st));
    long long tot = 0;
    for (int64_t i = 0; i < n; ++i) {
        const long long v = h[i];
        h[i] = tot;
        tot += v;
    }
    *npairs = tot;
    if (!pairs || cap < tot) {
        dev_free(d_cnt);
        if (!pairs && cap == 0)
            return ATVH_OK;
        throw Error(ATVH_E_CAPACITY, "atvh_nlist_export_pairs: output capacity too small");
    }
    long long* d_out = nullptr;
    dev_alloc(&d_out, size_t(2 * std::max<long long>(tot, 1)));
    ATVH_HIP(hipMemcpyAsync(d_cnt, h.data(), n * sizeof(long long), hipMemcpyHostToDevice, st));
    if (s->brick_mode) {
        ea.offs = d_cnt;
        ea.cnt = nullptr;
        ea.out = d_out;
        k_nl_export_brick<<<s->nbricks, kBlock, 0, st>>>(ea);
    } else {
        k_nl_export<<<grid_capped(n), kBlock, 0, st>>>(n, s->tag, s->numneigh, s->nbr, s->npad, d_cnt,
                                                       d_out);
    }
    ATVH_HIP(hipGetLastError());
    std::vector<long long> out(2 * tot);
    if (tot)
        ATVH_HIP(hipMemcpyAsync(out.data(), d_out, 2 * tot * sizeof(long long),
                                hipMemcpyDeviceToHost, st));
    ATVH_HIP(hipStreamSynchronize(st));
    dev_free(d_out);
    dev_free(d_cnt);
    // sort ascending by (a, b)
    std::vector<int64_t> key(tot);
    for (long long k = 0; k < tot; ++k)
        key[k] = out[2 * k] * int64_t(n) + out[2 * k + 1];
    std::sort(key.begin(), key.end());
    for (long long k = 0; k < tot; ++k) {
        pairs[2 * k] = key[k] / n;
        pairs[2 * k + 1] = key[k] % n;
    }
    return ATVH_OK;
    API_CATCH(c)
}

} // extern "C"
