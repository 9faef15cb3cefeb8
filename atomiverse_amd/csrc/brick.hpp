// brick.hpp — brick decomposition of the cell grid and the per-workgroup halo
// staged in LDS (shared by the list build and the pair kernels).
//
// The reference's cell grid (CellList::compute_grid, cell_list.cpp:83-154) is
// kept exactly; only the ORDER in which cells (and their atoms) are laid out
// in memory changes: cells are grouped into bricks of bdim cells per axis,
// bricks are numbered z-fastest, and cells inside a brick z-fastest. A brick's
// atoms are therefore one contiguous slot range, and its halo — the brick
// widened by the stencil reach R on every side — is a small cube of cells
// whose atoms one workgroup copies into LDS. Neighbour indices are 16-bit
// positions inside that halo, and halo positions are stored as the periodic
// images the pair actually uses, so the force loop needs no minimum image.
#pragma once

#include "atvh_internal.hpp"

namespace atvh {

constexpr int kHaloCap = 2048;      // LDS rows of a staged halo
constexpr int kSentinelRow = kHaloCap - 1; // staged far away; pads the last list group
constexpr int kHaloRows = kHaloCap - 1;    // halo atoms a brick may hold
constexpr int kHaloFillMax = kHaloRows * 15 / 16; // fullest halo a list build accepts for its shape
constexpr int kMaxBrickAtoms = 320; // home atoms a brick's list build can take
// Home atoms a list build takes per brick on this system: kMaxBrickAtoms
// unless lowered (atvh_nlist_set_brick_limit).
inline int brick_atom_limit(const atvh_system* s) {
    return s->brick_atom_limit > 0 ? s->brick_atom_limit : kMaxBrickAtoms;
}
// A brick the list build left out (halo or home atoms over capacity; the
// rebuild is reported as overflowed and the system halts at that step):
// its list rows are not this halo's, so a pair kernel reads none of them.
// Every reader passes the limit the build used.
__host__ __device__ inline bool brick_skipped(int nh, int nat, int max_atoms) {
    return nh > kHaloRows || nat > max_atoms;
}
// Sentinel coordinates: every pair with the sentinel is far beyond any cutoff
// (r^2 ~ 3e30 Å^2 in FP32 relative coordinates, ~3e36 in FP64 absolute ones)
// and its arithmetic stays finite.
constexpr double kSentinelFar = 1.0e18;
constexpr float kSentinelFarF = 1.0e15f;
constexpr int kHaloCellCap = 768;   // halo cells per brick ((4+4)*(5+4)*(5+4)=648)
constexpr int kRowSlotBits = 26;    // halo row = global slot | image code << 26
constexpr int kRowSlotMask = (1 << kRowSlotBits) - 1;
constexpr int kNumBrickCands = 12;                 // brick shapes, largest first
constexpr int kLastBrickCand = kNumBrickCands - 1; // the one-cell brick
constexpr int kGroup = 8;           // neighbour indices per 16-byte load

struct BrickGeom {
    int nc[3];    // cells per axis (reference grid)
    int R[3];     // stencil reach per axis
    int bdim[3];  // nominal brick size in cells
    int nb[3];    // bricks per axis
    double L[3];
    double cs[3];
    const int* cell_rank;       // reference cell index -> layout rank
    const int* brick_cell_base; // [nbricks+1] first rank of each brick
    int max_atoms;              // home atoms the build takes per brick
};

struct HaloTable {
    int dims[3];   // halo cells per axis
    int o[3];      // brick origin (cells)
    int d[3];      // brick size (cells)
    int ncell;     // halo cells
    int natoms;    // halo atoms
};

// Brick taken by work index t of n. The hardware deals consecutive
// workgroups round-robin to the 8 XCDs, each with its own L2; this mapping
// gives XCD k (t % 8 == k) one contiguous run of bricks, so bricks that
// share halo atoms (numbered z-fastest, neighbours are close in index) are
// read through the same L2. A bijection of [0, n) for any n.
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_brick(int t, int n) {
    const int per = n / kXcds, rem = n % kXcds;
    const int k = t % kXcds, j = t / kXcds;
    return k * per + min(k, rem) + j;
}

__device__ __forceinline__ void brick_coords(const BrickGeom& g, int b, int& bx, int& by, int& bz) {
    bz = b % g.nb[2];
    by = (b / g.nb[2]) % g.nb[1];
    bx = b / (g.nb[1] * g.nb[2]);
}

// Fill the halo table of brick b into LDS: per halo cell its first global
// slot (h_start), its LDS offset (h_off, exclusive scan of the counts) and the
// periodic shift code of its image (h_shift: 3 x 2 bits, 0 = -L, 1 = 0,
// 2 = +L). `cells` is the static per-brick table (rank | code << 26) built on
// the host, so every halo cell costs two independent loads of cell_start.
// All threads of the block participate; ends with a barrier.
__device__ __forceinline__ void build_halo_table(const BrickGeom& g, const int* __restrict__ cells,
                                                 const int* __restrict__ cell_start, int b,
                                                 HaloTable& ht, int* h_start, int* h_off,
                                                 int* h_shift, int* scan_tmp) {
    int bx, by, bz;
    brick_coords(g, b, bx, by, bz);
    ht.o[0] = bx * g.bdim[0];
    ht.o[1] = by * g.bdim[1];
    ht.o[2] = bz * g.bdim[2];
    for (int a = 0; a < 3; ++a) {
        ht.d[a] = min(g.bdim[a], g.nc[a] - ht.o[a]);
        ht.dims[a] = ht.d[a] + 2 * g.R[a];
    }
    ht.ncell = ht.dims[0] * ht.dims[1] * ht.dims[2];
    const int nthr = blockDim.x;
    const int* __restrict__ tab = cells + (int64_t)b * kHaloCellCap;
    for (int h = threadIdx.x; h < ht.ncell; h += nthr) {
        const int e = tab[h];
        const int r = e & kRowSlotMask;
        const int s0 = cell_start[r], s1 = cell_start[r + 1];
        h_start[h] = s0;
        h_shift[h] = static_cast<unsigned>(e) >> kRowSlotBits;
        h_off[h] = s1 - s0; // count, scanned below
    }
    __syncthreads();
    // block exclusive scan: per-thread chunk of consecutive cells, then chunk sums
    const int per = (ht.ncell + nthr - 1) / nthr;
    const int h0 = threadIdx.x * per;
    const int h1 = min(ht.ncell, h0 + per);
    int local = 0;
    for (int h = h0; h < h1; ++h)
        local += h_off[h];
    scan_tmp[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < nthr; off <<= 1) {
        const int add = threadIdx.x >= off ? scan_tmp[threadIdx.x - off] : 0;
        __syncthreads();
        scan_tmp[threadIdx.x] += add;
        __syncthreads();
    }
    int run = scan_tmp[threadIdx.x] - local;
    ht.natoms = scan_tmp[nthr - 1];
    for (int h = h0; h < h1; ++h) {
        const int cnt = h_off[h];
        h_off[h] = run;
        run += cnt;
    }
    if (threadIdx.x == 0)
        h_off[ht.ncell] = ht.natoms;
    __syncthreads();
}

__device__ __forceinline__ double shift_of(int code, int axis, double L) {
    const int s = (code >> (2 * axis)) & 3;
    return s == 0 ? -L : (s == 2 ? L : 0.0);
}

} // namespace atvh
