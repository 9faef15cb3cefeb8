// batch.hpp — one launch over many independent systems.
//
// A kernel body is written once as `Op::run(args, bid, nbk)`, where bid/nbk
// stand for blockIdx.x/gridDim.x of a launch over one system. The
// single-system kernel calls it with the hardware indices; the batched
// kernel gives every system a contiguous range of blocks (prefix table
// `base`, nsys+1 entries), reads its system from a per-block table (`owner`,
// built with the stage) and calls the same body with the system's own
// argument block. Every system therefore
// runs exactly the code (and the reduction order) it runs alone.
#pragma once

#include "atvh_internal.hpp"

#include <type_traits>

namespace atvh {

// Workgroup size of an op: Op::kThreads where the op declares one, else kBlock.
template <class Op, class = void>
struct op_threads {
    static constexpr int value = kBlock;
};
template <class Op>
struct op_threads<Op, std::void_t<decltype(Op::kThreads)>> {
    static constexpr int value = Op::kThreads;
};

// Minimum waves per SIMD an op asks the register allocator to allow:
// Op::kMinWaves where declared, else 1 (no constraint).
template <class Op, class = void>
struct op_waves {
    static constexpr int value = 1;
};
template <class Op>
struct op_waves<Op, std::void_t<decltype(Op::kMinWaves)>> {
    static constexpr int value = Op::kMinWaves;
};

template <class Op>
__global__ __launch_bounds__(op_threads<Op>::value) __attribute__((amdgpu_waves_per_eu(op_waves<Op>::value)))
void k_single(typename Op::Args a) {
    Op::run(a, blockIdx.x, gridDim.x);
}

template <class Op>
__global__ __launch_bounds__(op_threads<Op>::value) __attribute__((amdgpu_waves_per_eu(op_waves<Op>::value)))
void k_batched(const typename Op::Args* __restrict__ args, const int* __restrict__ base,
               const int* __restrict__ owner) {
    const int blk = blockIdx.x;
    const int lo = owner[blk]; // the member whose range [base[lo], base[lo + 1]) holds blk
    // by reference: the argument block stays in global memory (a private copy
    // would be indexed dynamically for per-type tables and land in scratch)
    Op::run(args[lo], blk - base[lo], base[lo + 1] - base[lo]);
}

// A predicated stage over the members whose flag is set: `act` (nsys+1
// entries, computed on the device for this step) is the prefix of the active
// members' block counts (0 for an inactive member). A fixed grid strides over
// the active blocks only, so members that skip the stage cost nothing beyond
// their zero-length range. Every member still runs its own single-system
// code, block for block.
template <class Op>
__global__ __launch_bounds__(op_threads<Op>::value) __attribute__((amdgpu_waves_per_eu(op_waves<Op>::value)))
void k_batched_active(const typename Op::Args* __restrict__ args, const int* __restrict__ act, int nsys) {
    const int total = act[nsys];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        if (t != static_cast<int>(blockIdx.x))
            __syncthreads(); // the previous block's shared memory is reused
        int lo = 0, hi = nsys - 1; // largest s with act[s] <= t: an active member
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (act[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        Op::run(args[lo], t - act[lo], act[lo + 1] - act[lo]);
    }
}

// k_batched_active with the active blocks handed out from a counter instead
// of a fixed stride: block b first takes active block b, then each next one
// from `next` (zeroed on the device before the launch), offset by the grid.
// A block that draws a value past the active total leaves, so every block
// ends. For stages whose blocks differ widely in cost.
template <class Op>
__global__ __launch_bounds__(op_threads<Op>::value) __attribute__((amdgpu_waves_per_eu(op_waves<Op>::value)))
void k_batched_queue(const typename Op::Args* __restrict__ args, const int* __restrict__ act, int nsys,
                     int* __restrict__ next) {
    __shared__ int s_next;
    const int total = act[nsys];
    int t = blockIdx.x;
    while (t < total) {
        int lo = 0, hi = nsys - 1; // largest s with act[s] <= t: an active member
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (act[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        Op::run(args[lo], t - act[lo], act[lo + 1] - act[lo]);
        __syncthreads(); // shared memory (and s_next) is reused
        if (threadIdx.x == 0)
            s_next = static_cast<int>(gridDim.x) + atomicAdd(next, 1);
        __syncthreads();
        t = s_next;
    }
}

} // namespace atvh
