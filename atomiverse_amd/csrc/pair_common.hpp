// pair_common.hpp — pieces shared by the brick pair kernels: wavefront and
// fixed-order block reductions of (energy, virial) partials, and the periodic
// image of a halo row.
#pragma once

#include "atvh_internal.hpp"
#include "brick.hpp"

namespace atvh {

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_down(v, off, 64);
    return v;
}

// Fixed-order block reduction of 7 doubles into partials[slot*8 + k]
// (kThreads-thread blocks: the waves summed in order by thread k); slot is
// the block's index within its system's launch.
template <int kThreads = kBlock>
__device__ __forceinline__ void block_partials(double (&acc)[7], double* partials, int slot) {
    __shared__ double sh[kThreads / 64][7];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 7; ++k)
        acc[k] = wave_sum(acc[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 7; ++k)
            sh[wid][k] = acc[k];
    __syncthreads();
    if (threadIdx.x < 7) {
        double t = 0.0;
        for (int w = 0; w < kThreads / 64; ++w)
            t += sh[w][threadIdx.x];
        partials[slot * kPartialStride + threadIdx.x] = t;
    }
}

// Image position of a halo row: the global slot's position plus the periodic
// shift recorded at the build (positions are not re-wrapped between builds).
__device__ __forceinline__ void row_image(const double* __restrict__ x, const double* __restrict__ y,
                                          const double* __restrict__ z, double lx, double ly, double lz,
                                          int row, int& gs, double& xx, double& yy, double& zz) {
    gs = row & kRowSlotMask;
    const int code = static_cast<unsigned>(row) >> kRowSlotBits;
    xx = x[gs] + shift_of(code, 0, lx);
    yy = y[gs] + shift_of(code, 1, ly);
    zz = z[gs] + shift_of(code, 2, lz);
}

// Centre of brick b (FP64), the reference point of the FP32 relative
// coordinates of the mixed kernels.
__device__ __forceinline__ void brick_centre(int b, const int nb[3], const int bdim[3], const int nc[3],
                                             const double cs[3], double& cx, double& cy, double& cz) {
    const int bz = b % nb[2], by = (b / nb[2]) % nb[1], bx = b / (nb[1] * nb[2]);
    const int ox = bx * bdim[0], oy = by * bdim[1], oz = bz * bdim[2];
    cx = (ox + 0.5 * min(bdim[0], nc[0] - ox)) * cs[0];
    cy = (oy + 0.5 * min(bdim[1], nc[1] - oy)) * cs[1];
    cz = (oz + 0.5 * min(bdim[2], nc[2] - oz)) * cs[2];
}

// Unpack entry u (0..7) of an 8-wide group of 16-bit halo rows.
__device__ __forceinline__ int group_entry(const uint4& g, int u) {
    const unsigned long long lo = g.x | (static_cast<unsigned long long>(g.y) << 32);
    const unsigned long long hi = g.z | (static_cast<unsigned long long>(g.w) << 32);
    return static_cast<int>(((u < 4 ? lo : hi) >> (16 * (u & 3))) & 0xffffu);
}

} // namespace atvh
