// pair_lj.hip — shifted Lennard-Jones force, energy and virial on the full
// ELL Verlet list.
//
// Physics: PairLJCut (src/cpu/potentials/pair_lj_cut.cpp:31-42 constructor
// shift, :257-298 per-type table, :300-830 compute):
//   r6inv = r2inv^3 * sigma^6,  fpair = 24 eps r2inv (2 r6inv^2 - r6inv),
//   E = 4 eps (r6inv^2 - r6inv) - V(rc)   (energy shifted, force not),
//   virial Voigt (dx fx, dy fy, dz fz, dx fy, dx fz, dy fz).
// CUDA counterpart replaced: lj_force_tile (src/cuda/potentials/pair_lj_cut.cu:89-246).
//
// MI355X design: one workgroup per brick (brick.hpp) stages the brick's halo
// in LDS as periodic images, then one thread per brick atom walks its FULL
// neighbour list of 16-bit halo rows: every force row has a single writer (no
// atomics, deterministic, batch-invariant), no per-pair minimum image and no
// global gathers; energy and virial are halved. Per-block partials of
// (E, virial) are summed in a fixed order by the finalize kernel. Precision:
//   FP64  — everything in double, IEEE division (the plain drop-in path);
//   MIXED — LDS holds FP32 coordinates relative to the brick centre (formed in
//           FP64, |x| < ~7 Å so the rounding is <= 2.4e-7 Å); r2, 1/r2
//           (v_rcp_f32), the LJ polynomial and the per-atom force sum are FP32,
//           energy and virial are accumulated in FP64.
// Grids with fewer than 2*reach+1 cells on an axis use the global-slot list
// with per-pair minimum image (k_lj_fp64 / k_lj_mixed below).

#include "atvh_internal.hpp"
#include "brick.hpp"
#include "pair_common.hpp"
#include "batch.hpp"

#include <cmath>

using namespace atvh;

namespace {

struct LJArgs {
    const double *x, *y, *z;
    const int* type;
    const int* nbr;
    const int* numneigh;
    int64_t npad;
    double lx, ly, lz, ilx, ily, ilz;
    double rc_sq;
    // single type
    double eps24, eps4, sig6, vshift;
    // per type: double4 {eps24, eps4, sig6, shift} per (ti*ntypes + tj)
    const double4* tab;
    int ntypes;
    double *fx, *fy, *fz;
    double* partials;
};

template <bool EF, bool VF, bool TYPED>
__global__ __launch_bounds__(kBlock) void k_lj_fp64(int64_t n, LJArgs a) {
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
        const int ti = TYPED ? a.type[i] : 0;
        const int nn = a.numneigh[i];
        double fxi = 0.0, fyi = 0.0, fzi = 0.0;
        const int* __restrict__ col = a.nbr + i;
        for (int k = 0; k < nn; ++k) {
            const int j = col[(int64_t)k * a.npad];
            double dx = xi - a.x[j];
            double dy = yi - a.y[j];
            double dz = zi - a.z[j];
            dx -= a.lx * rint(dx * a.ilx);
            dy -= a.ly * rint(dy * a.ily);
            dz -= a.lz * rint(dz * a.ilz);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < a.rc_sq) {
                double eps24 = a.eps24, eps4 = a.eps4, sig6 = a.sig6, shift = a.vshift;
                if (TYPED) {
                    const double4 t = a.tab[ti * a.ntypes + a.type[j]];
                    eps24 = t.x;
                    eps4 = t.y;
                    sig6 = t.z;
                    shift = t.w;
                }
                const double r2inv = 1.0 / r2;
                const double r6inv = r2inv * r2inv * r2inv * sig6;
                const double fpair = eps24 * r2inv * fma(2.0 * r6inv, r6inv, -r6inv);
                const double fxp = fpair * dx, fyp = fpair * dy, fzp = fpair * dz;
                fxi += fxp;
                fyi += fyp;
                fzi += fzp;
                if (EF)
                    acc[0] += eps4 * (r6inv * r6inv - r6inv) - shift;
                if (VF) {
                    acc[1] += dx * fxp;
                    acc[2] += dy * fyp;
                    acc[3] += dz * fzp;
                    acc[4] += dx * fyp;
                    acc[5] += dx * fzp;
                    acc[6] += dy * fzp;
                }
            }
        }
        a.fx[i] = fxi;
        a.fy[i] = fyi;
        a.fz[i] = fzi;
    }
    if (EF || VF) {
#pragma unroll
        for (int k = 0; k < 7; ++k)
            acc[k] *= 0.5; // full list: every pair seen from both ends
        block_partials(acc, a.partials, blockIdx.x);
    }
}

template <bool EF, bool VF, bool TYPED>
__global__ __launch_bounds__(kBlock) void k_lj_mixed(int64_t n, LJArgs a) {
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const float eps24s = static_cast<float>(a.eps24), eps4s = static_cast<float>(a.eps4);
    const float sig6s = static_cast<float>(a.sig6), shifts = static_cast<float>(a.vshift);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
        const int ti = TYPED ? a.type[i] : 0;
        const int nn = a.numneigh[i];
        float fxi = 0.f, fyi = 0.f, fzi = 0.f;
        const int* __restrict__ col = a.nbr + i;
        for (int k = 0; k < nn; ++k) {
            const int j = col[(int64_t)k * a.npad];
            double dxd = xi - a.x[j];
            double dyd = yi - a.y[j];
            double dzd = zi - a.z[j];
            dxd -= a.lx * rint(dxd * a.ilx);
            dyd -= a.ly * rint(dyd * a.ily);
            dzd -= a.lz * rint(dzd * a.ilz);
            // the cutoff test stays in FP64 so in/out decisions match the plain path
            const double r2d = dxd * dxd + dyd * dyd + dzd * dzd;
            if (r2d < a.rc_sq) {
                const float dx = static_cast<float>(dxd), dy = static_cast<float>(dyd),
                            dz = static_cast<float>(dzd);
                const float r2 = static_cast<float>(r2d);
                float eps24 = eps24s, eps4 = eps4s, sig6 = sig6s, shift = shifts;
                if (TYPED) {
                    const double4 t = a.tab[ti * a.ntypes + a.type[j]];
                    eps24 = static_cast<float>(t.x);
                    eps4 = static_cast<float>(t.y);
                    sig6 = static_cast<float>(t.z);
                    shift = static_cast<float>(t.w);
                }
                const float r2inv = __builtin_amdgcn_rcpf(r2);
                const float r6inv = r2inv * r2inv * r2inv * sig6;
                const float fpair = eps24 * r2inv * __builtin_fmaf(2.0f * r6inv, r6inv, -r6inv);
                const float fxp = fpair * dx, fyp = fpair * dy, fzp = fpair * dz;
                fxi += fxp;
                fyi += fyp;
                fzi += fzp;
                if (EF)
                    acc[0] += static_cast<double>(eps4 * (r6inv * r6inv - r6inv) - shift);
                if (VF) {
                    acc[1] += static_cast<double>(dx * fxp);
                    acc[2] += static_cast<double>(dy * fyp);
                    acc[3] += static_cast<double>(dz * fzp);
                    acc[4] += static_cast<double>(dx * fyp);
                    acc[5] += static_cast<double>(dx * fzp);
                    acc[6] += static_cast<double>(dy * fzp);
                }
            }
        }
        a.fx[i] = static_cast<double>(fxi);
        a.fy[i] = static_cast<double>(fyi);
        a.fz[i] = static_cast<double>(fzi);
    }
    if (EF || VF) {
#pragma unroll
        for (int k = 0; k < 7; ++k)
            acc[k] *= 0.5;
        block_partials(acc, a.partials, blockIdx.x);
    }
}

struct BrickLJArgs {
    const double *x, *y, *z;
    const int* type;
    const int* halo_rows;
    const int* halo_n;
    const int2* brick_range;
    const uint16_t* nbr16;
    const uint16_t* self16;
    const int* numneigh;
    int cap;
    double L[3];
    double cs[3];
    int nc[3], nb[3], bdim[3];
    double rc_sq;
    double eps24, eps4, sig6, vshift;
    double inv_sigma; // 1/sigma of the single-type form (coordinates in sigma units)
    float band;       // mixed: |r^2 - rc^2| below which the FP64 predicate decides (kernel units)
    int max_atoms;    // home atoms the list build took per brick
    const double4* tab;
    int ntypes;
    double *fx, *fy, *fz;
    double* partials;
};

template <bool EF, bool VF, bool TYPED>
struct OpLjBrickFp64 {
    using Args = BrickLJArgs;
    static constexpr int kThreads = kBlock; // (512 measured slower: 311 vs 283 us at 2.048M atoms)
    static __device__ void run(const BrickLJArgs& a, int bid, int nbk);
};

template <bool EF, bool VF, bool TYPED>
__device__ void OpLjBrickFp64<EF, VF, TYPED>::run(const BrickLJArgs& a, int bid, int nbk) {
    __shared__ double sx[kHaloCap], sy[kHaloCap], sz[kHaloCap];
    __shared__ int stype[TYPED ? kHaloCap : 1];
    const int b = xcd_brick(bid, nbk); // one launch block per brick
    const int nh = a.halo_n[b];
    const int2 br = a.brick_range[b];
    const int a0 = br.x, nat = br.y;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (brick_skipped(nh, nat, a.max_atoms)) { // reported by the build; nothing valid to compute
        if (EF || VF)
            block_partials<kThreads>(acc, a.partials, bid);
        return;
    }
    const int* __restrict__ rows = a.halo_rows + (int64_t)b * kHaloCap;
    for (int q = threadIdx.x; q < nh; q += blockDim.x) {
        int gs;
        double xx, yy, zz;
        row_image(a.x, a.y, a.z, a.L[0], a.L[1], a.L[2], rows[q], gs, xx, yy, zz);
        sx[q] = xx;
        sy[q] = yy;
        sz[q] = zz;
        if (TYPED)
            stype[q] = a.type[gs];
    }
    if (threadIdx.x == 0) { // the list's padding row: out of every cutoff
        sx[kSentinelRow] = sy[kSentinelRow] = sz[kSentinelRow] = kSentinelFar;
        if (TYPED)
            stype[kSentinelRow] = 0;
    }
    __syncthreads();
    // kernel arguments copied to registers: the per-pair lambda below must not
    // take the address of the argument block
    const double rc_sq = a.rc_sq;
    const double p_eps24 = a.eps24, p_eps4 = a.eps4, p_sig6 = a.sig6, p_shift = a.vshift;
    const double4* __restrict__ tab = a.tab;
    const int ntypes = a.ntypes;
    const uint4* __restrict__ base = reinterpret_cast<const uint4*>(a.nbr16 + (int64_t)a0 * a.cap);
    // two threads per atom when the brick has at most half a block of atoms:
    // lanes 2l and 2l+1 take alternate neighbour groups of atom l
    const int tpa = (2 * nat <= static_cast<int>(blockDim.x)) ? 2 : 1;
    for (int t = threadIdx.x; t < nat * tpa; t += blockDim.x) {
        const int li = tpa == 2 ? (t >> 1) : t;
        const int h = tpa == 2 ? (t & 1) : 0;
        const int i = a0 + li;
        const int qs = a.self16[i];
        const double xi = sx[qs], yi = sy[qs], zi = sz[qs];
        const int ti = TYPED ? stype[qs] : 0;
        const int nn = a.numneigh[i];
        const uint4* __restrict__ grp = base + li;
        const int ng = (nn + kGroup - 1) / kGroup;
        double fxi = 0.0, fyi = 0.0, fzi = 0.0;
        // group indices clamp to the last group (to group 0 when nn == 0: the
        // brick's rows always exist); loads run two groups ahead
        const int gl = max(ng - 1, 0);
        uint4 cur = grp[__umul24(min(h, gl), nat)];
        uint4 nx1 = grp[__umul24(min(h + tpa, gl), nat)];
        for (int g = h; g < ng; g += tpa) {
            const uint4 nx2 = grp[__umul24(min(g + 2 * tpa, gl), nat)];
            // the whole group's LDS reads issue before any arithmetic; padding
            // rows (the far sentinel) and pairs beyond r_c add zero
            const unsigned w[4] = {cur.x, cur.y, cur.z, cur.w};
            double jx[kGroup], jy[kGroup], jz[kGroup];
            int tj[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const int q = (w[u >> 1] >> (16 * (u & 1))) & 0xffff;
                jx[u] = sx[q];
                jy[u] = sy[q];
                jz[u] = sz[q];
                if (TYPED)
                    tj[u] = stype[q];
            }
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const double dx = xi - jx[u];
                const double dy = yi - jy[u];
                const double dz = zi - jz[u];
                const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                const bool in = r2 < rc_sq && r2 > 0.0;
                double eps24 = p_eps24, eps4 = p_eps4, sig6 = p_sig6, shift = p_shift;
                if (TYPED) {
                    const double4 t = tab[ti * ntypes + tj[u]];
                    eps24 = t.x;
                    eps4 = t.y;
                    sig6 = t.z;
                    shift = t.w;
                }
                const double r2inv = 1.0 / r2;
                const double r6inv = r2inv * r2inv * r2inv * sig6;
                const double fpair = in ? eps24 * r2inv * fma(2.0 * r6inv, r6inv, -r6inv) : 0.0;
                fxi = fma(fpair, dx, fxi);
                fyi = fma(fpair, dy, fyi);
                fzi = fma(fpair, dz, fzi);
                if (EF && in)
                    acc[0] += fma(eps4, r6inv * r6inv - r6inv, -shift);
                if (VF && in) {
                    const double fxp = fpair * dx, fyp = fpair * dy, fzp = fpair * dz;
                    acc[1] += dx * fxp;
                    acc[2] += dy * fyp;
                    acc[3] += dz * fzp;
                    acc[4] += dx * fyp;
                    acc[5] += dx * fzp;
                    acc[6] += dy * fzp;
                }
            }
            cur = nx1;
            nx1 = nx2;
        }
        if (tpa == 2) {
            fxi += __shfl_xor(fxi, 1, 64);
            fyi += __shfl_xor(fyi, 1, 64);
            fzi += __shfl_xor(fzi, 1, 64);
        }
        if (h != 0)
            continue;
        a.fx[i] = fxi;
        a.fy[i] = fyi;
        a.fz[i] = fzi;
    }
    if (EF || VF) {
#pragma unroll
        for (int k = 0; k < 7; ++k)
            acc[k] *= 0.5;
        block_partials<kThreads>(acc, a.partials, bid);
    }
}

template <bool EF, bool VF, bool TYPED>
struct OpLjBrickMixed {
    using Args = BrickLJArgs;
    static constexpr int kThreads = 512; // two threads per atom up to 256-atom bricks (157 vs 181 us at 2.048M atoms)
    static __device__ void run(const BrickLJArgs& a, int bid, int nbk);
};

template <bool EF, bool VF, bool TYPED>
__device__ void OpLjBrickMixed<EF, VF, TYPED>::run(const BrickLJArgs& a, int bid, int nbk) {
    __shared__ float4 sp[kHaloCap]; // x, y, z relative to the brick centre; w = type bits
    const int b = xcd_brick(bid, nbk); // one launch block per brick
    const int nh = a.halo_n[b];
    const int2 br = a.brick_range[b];
    const int a0 = br.x, nat = br.y;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (brick_skipped(nh, nat, a.max_atoms)) { // reported by the build; nothing valid to compute
        if (EF || VF)
            block_partials<kThreads>(acc, a.partials, bid);
        return;
    }
    const int* __restrict__ rows = a.halo_rows + (int64_t)b * kHaloCap;
    // reference point: the brick centre (every halo row lies within the brick
    // widened by the stencil reach, so |x - centre| < ~7 Å and the FP32
    // rounding of the relative coordinate is <= 2.4e-7 Å)
    // Single type: coordinates in units of sigma, so the pair term is
    // r6 (2 r6 - 1) / r^2 with r6 = r^-6, and 24 eps / sigma is applied once
    // per atom (the virial takes 24 eps); the typed form keeps Å and the
    // per-pair coefficients.
    double r0x, r0y, r0z;
    brick_centre(b, a.nb, a.bdim, a.nc, a.cs, r0x, r0y, r0z);
    const double us = TYPED ? 1.0 : a.inv_sigma;
    for (int q = threadIdx.x; q < nh; q += blockDim.x) {
        int gs;
        double xx, yy, zz;
        row_image(a.x, a.y, a.z, a.L[0], a.L[1], a.L[2], rows[q], gs, xx, yy, zz);
        float4 v;
        v.x = static_cast<float>((xx - r0x) * us);
        v.y = static_cast<float>((yy - r0y) * us);
        v.z = static_cast<float>((zz - r0z) * us);
        v.w = TYPED ? __int_as_float(a.type[gs]) : 0.0f;
        sp[q] = v;
    }
    if (threadIdx.x == 0) // the list's padding row: out of every cutoff
        sp[kSentinelRow] = make_float4(kSentinelFarF, kSentinelFarF, kSentinelFarF, __int_as_float(0));
    __syncthreads();
    const float rc_sq = static_cast<float>(a.rc_sq * us * us);
    const float eps4s = static_cast<float>(a.eps4), shifts = static_cast<float>(a.vshift);
    const float fscale = TYPED ? 1.0f : static_cast<float>(a.eps24 * a.inv_sigma);
    const double4* __restrict__ tab = a.tab;
    const int ntypes = a.ntypes;
    const uint4* __restrict__ base = reinterpret_cast<const uint4*>(a.nbr16 + (int64_t)a0 * a.cap);
    // the cutoff decision inside the FP32 rounding band [rc_lo, rc_hi) around
    // rc^2 is the plain kernel's FP64 test on the same images: the main loop
    // counts the pairs below rc_lo and the fix-up below adds the band pairs the
    // FP64 test accepts, so a rejected pair is never touched. Both sides use
    // the same two compares, so every pair below rc_hi is taken exactly once.
    // (Arguments copied to registers: the lambda must not take the address of
    // the argument block.)
    const float rc_lo = rc_sq - a.band, rc_hi = rc_sq + a.band;
    const double rc_sq_fp64 = a.rc_sq;
    const double* __restrict__ gx = a.x;
    const double* __restrict__ gy = a.y;
    const double* __restrict__ gz = a.z;
    const double Lx = a.L[0], Ly = a.L[1], Lz = a.L[2];
    // two threads per atom when the brick has at most half a block of atoms:
    // lanes 2l and 2l+1 take alternate neighbour groups of atom l
    const int tpa = (2 * nat <= static_cast<int>(blockDim.x)) ? 2 : 1;
    for (int t = threadIdx.x; t < nat * tpa; t += blockDim.x) {
        const int li = tpa == 2 ? (t >> 1) : t;
        const int h = tpa == 2 ? (t & 1) : 0;
        const int i = a0 + li;
        const int qs = a.self16[i];
        const float4 pi = sp[qs];
        const int ti = TYPED ? __float_as_int(pi.w) : 0;
        const int nn = a.numneigh[i];
        const uint4* __restrict__ grp = base + li;
        const int ng = (nn + kGroup - 1) / kGroup;
        float fxi = 0.f, fyi = 0.f, fzi = 0.f;
        bool near = false; // some pair's FP32 r^2 lies in [rc_lo, rc_hi)
        // group indices clamp to the last group (to group 0 when nn == 0: the
        // brick's rows always exist); loads run two groups ahead in three
        // rotating slots, so no group is moved between registers
        const int gl = max(ng - 1, 0);
        auto load = [&](int g) { return grp[__umul24(min(g, gl), nat)]; };
        auto group = [&](const uint4 cur) {
            // the whole group's LDS reads issue before any arithmetic; padding
            // rows (the far sentinel) and pairs beyond r_c add zero
            const unsigned w[4] = {cur.x, cur.y, cur.z, cur.w};
            float4 pj[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u)
                pj[u] = sp[(w[u >> 1] >> (16 * (u & 1))) & 0xffff];
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const float dx = pi.x - pj[u].x;
                const float dy = pi.y - pj[u].y;
                const float dz = pi.z - pj[u].z;
                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                const bool in = r2 < rc_lo; // padding rows are the far sentinel
                // rc_lo < rc_hi, so the band is (r2 < rc_hi) xor in: one more
                // compare, the lane masks combine on the scalar unit
                near = near || ((r2 < rc_hi) != in);
                float eps4 = eps4s, shift = shifts;
                const float r2inv = __builtin_amdgcn_rcpf(r2);
                float r6inv, fpair;
                if (TYPED) {
                    const double4 t = tab[ti * ntypes + __float_as_int(pj[u].w)];
                    const float eps24 = static_cast<float>(t.x);
                    eps4 = static_cast<float>(t.y);
                    shift = static_cast<float>(t.w);
                    r6inv = r2inv * r2inv * r2inv * static_cast<float>(t.z);
                    fpair = in ? eps24 * r2inv * __builtin_fmaf(2.0f * r6inv, r6inv, -r6inv) : 0.0f;
                } else {
                    r6inv = r2inv * r2inv * r2inv;
                    fpair = r2inv * (r6inv * __builtin_fmaf(2.0f, r6inv, -1.0f));
                }
                if (in) {
                    fxi = __builtin_fmaf(fpair, dx, fxi);
                    fyi = __builtin_fmaf(fpair, dy, fyi);
                    fzi = __builtin_fmaf(fpair, dz, fzi);
                }
                if (EF && in)
                    acc[0] += static_cast<double>(__builtin_fmaf(eps4, r6inv * r6inv - r6inv, -shift));
                if (VF && in) {
                    const float fxp = fpair * dx, fyp = fpair * dy, fzp = fpair * dz;
                    acc[1] += static_cast<double>(dx * fxp);
                    acc[2] += static_cast<double>(dy * fyp);
                    acc[3] += static_cast<double>(dz * fzp);
                    acc[4] += static_cast<double>(dx * fyp);
                    acc[5] += static_cast<double>(dx * fzp);
                    acc[6] += static_cast<double>(dy * fzp);
                }
            }
        };
        uint4 s0 = load(h), s1 = load(h + tpa), s2;
        for (int g = h;;) {
            if (g >= ng)
                break;
            s2 = load(g + 2 * tpa);
            group(s0);
            g += tpa;
            if (g >= ng)
                break;
            s0 = load(g + 2 * tpa);
            group(s1);
            g += tpa;
            if (g >= ng)
                break;
            s1 = load(g + 2 * tpa);
            group(s2);
            g += tpa;
        }
        if (near) {
            // Rare: walk this thread's groups again and, for each band pair,
            // take the plain kernel's FP64 test on the same images; where it
            // accepts the pair, add the pair's FP32 term. Run after the main
            // loop, its FP64 registers are never live beside the main loop's.
            int gsi;
            double xi, yi, zi;
            row_image(gx, gy, gz, Lx, Ly, Lz, rows[qs], gsi, xi, yi, zi);
            for (int g = h; g < ng; g += tpa) {
                const uint4 cur = load(g);
                for (int u = 0; u < kGroup; ++u) {
                    const int q = group_entry(cur, u);
                    if (q >= nh) // the padding row has no halo row
                        continue;
                    const float4 pj = sp[q];
                    const float dx = pi.x - pj.x;
                    const float dy = pi.y - pj.y;
                    const float dz = pi.z - pj.z;
                    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    if (r2 < rc_lo || !(r2 < rc_hi))
                        continue;
                    int gsj;
                    double xj, yj, zj;
                    row_image(gx, gy, gz, Lx, Ly, Lz, rows[q], gsj, xj, yj, zj);
                    const double ddx = xi - xj, ddy = yi - yj, ddz = zi - zj;
                    const bool exact = fma(ddz, ddz, fma(ddy, ddy, ddx * ddx)) < rc_sq_fp64;
                    if (!exact)
                        continue;
                    float eps24 = 1.0f, eps4 = eps4s, shift = shifts, sig6 = 1.0f;
                    if (TYPED) {
                        const double4 t = tab[ti * ntypes + __float_as_int(pj.w)];
                        eps24 = static_cast<float>(t.x);
                        eps4 = static_cast<float>(t.y);
                        sig6 = static_cast<float>(t.z);
                        shift = static_cast<float>(t.w);
                    }
                    const float r2inv = __builtin_amdgcn_rcpf(r2);
                    const float r6inv = r2inv * r2inv * r2inv * sig6;
                    const float fpair = TYPED ? eps24 * r2inv * __builtin_fmaf(2.0f * r6inv, r6inv, -r6inv)
                                              : r2inv * (r6inv * __builtin_fmaf(2.0f, r6inv, -1.0f));
                    fxi = __builtin_fmaf(fpair, dx, fxi);
                    fyi = __builtin_fmaf(fpair, dy, fyi);
                    fzi = __builtin_fmaf(fpair, dz, fzi);
                    if (EF)
                        acc[0] += static_cast<double>(__builtin_fmaf(eps4, r6inv * r6inv - r6inv, -shift));
                    if (VF) {
                        const float fxp = fpair * dx, fyp = fpair * dy, fzp = fpair * dz;
                        acc[1] += static_cast<double>(dx * fxp);
                        acc[2] += static_cast<double>(dy * fyp);
                        acc[3] += static_cast<double>(dz * fzp);
                        acc[4] += static_cast<double>(dx * fyp);
                        acc[5] += static_cast<double>(dx * fzp);
                        acc[6] += static_cast<double>(dy * fzp);
                    }
                }
            }
        }
        if (tpa == 2) {
            fxi += __shfl_xor(fxi, 1, 64);
            fyi += __shfl_xor(fyi, 1, 64);
            fzi += __shfl_xor(fzi, 1, 64);
        }
        if (h != 0)
            continue;
        a.fx[i] = static_cast<double>(fxi * fscale);
        a.fy[i] = static_cast<double>(fyi * fscale);
        a.fz[i] = static_cast<double>(fzi * fscale);
    }
    if (EF || VF) {
        acc[0] *= 0.5;
        const double vscale = TYPED ? 0.5 : 0.5 * a.eps24;
#pragma unroll
        for (int k = 1; k < 7; ++k)
            acc[k] *= vscale;
        block_partials<kThreads>(acc, a.partials, bid);
    }
}

// The brick kernels, launched through launch_timed (family "pair").
template <class OpT>
static void lj_brick_launch(atvh_ctx* c, int grid, hipStream_t st, const BrickLJArgs& b) {
    launch_timed(c, "pair", k_single<OpT>, dim3(grid), dim3(op_threads<OpT>::value), st, b);
}
template <template <bool, bool, bool> class OP>
static void lj_brick_dispatch(atvh_ctx* c, bool ef, bool vf, bool ty, int grid, hipStream_t st,
                              const BrickLJArgs& b) {
    if (ef) {
        if (vf)
            ty ? lj_brick_launch<OP<true, true, true>>(c, grid, st, b) : lj_brick_launch<OP<true, true, false>>(c, grid, st, b);
        else
            ty ? lj_brick_launch<OP<true, false, true>>(c, grid, st, b) : lj_brick_launch<OP<true, false, false>>(c, grid, st, b);
    } else {
        if (vf)
            ty ? lj_brick_launch<OP<false, true, true>>(c, grid, st, b) : lj_brick_launch<OP<false, true, false>>(c, grid, st, b);
        else
            ty ? lj_brick_launch<OP<false, false, true>>(c, grid, st, b) : lj_brick_launch<OP<false, false, false>>(c, grid, st, b);
    }
}

#define LJ_DISPATCH(KERNEL, EF, VF, TY, GRID, ST, N, ARGS)                                   \
    do {                                                                                      \
        if (EF) {                                                                             \
            if (VF) {                                                                         \
                if (TY) KERNEL<true, true, true><<<GRID, kBlock, 0, ST>>>(N, ARGS);           \
                else KERNEL<true, true, false><<<GRID, kBlock, 0, ST>>>(N, ARGS);             \
            } else {                                                                          \
                if (TY) KERNEL<true, false, true><<<GRID, kBlock, 0, ST>>>(N, ARGS);          \
                else KERNEL<true, false, false><<<GRID, kBlock, 0, ST>>>(N, ARGS);            \
            }                                                                                 \
        } else {                                                                              \
            if (VF) {                                                                         \
                if (TY) KERNEL<false, true, true><<<GRID, kBlock, 0, ST>>>(N, ARGS);          \
                else KERNEL<false, true, false><<<GRID, kBlock, 0, ST>>>(N, ARGS);            \
            } else {                                                                          \
                if (TY) KERNEL<false, false, true><<<GRID, kBlock, 0, ST>>>(N, ARGS);         \
                else KERNEL<false, false, false><<<GRID, kBlock, 0, ST>>>(N, ARGS);           \
            }                                                                                 \
        }                                                                                     \
    } while (0)

} // namespace

namespace atvh {

int pair_grid(const atvh_system* s) { return s->brick_mode ? s->nbricks : grid_capped(s->n); }

static void lj_enqueue(atvh_system* s, int precision, bool eflag, bool vflag);

// The force phase: the pair style (PairLJCut or PairLJCutCoulLong) writes the
// forces, then KSpacePME adds its part when configured (the reference's
// evaluator order at Phase::force).
void pair_enqueue(atvh_system* s, int precision, bool eflag, bool vflag) {
    if (!s->pair_configured)
        throw Error(ATVH_E_STATE, "pair style not configured");
    if (!s->nl_built)
        throw Error(ATVH_E_STATE, "pair style: no neighbor list (add a NeighborList and build it)");
    if (s->lj_cutoff > s->cutoff + 1e-12)
        throw Error(ATVH_E_ARG, "pair style: pair cutoff exceeds the neighbor-list cutoff");
    if (s->pair_style == 1) {
        ScopedTimer tm(s->ctx, "pair");
        ljc_enqueue(s, precision, eflag, vflag);
    } else {
        lj_enqueue(s, precision, eflag, vflag); // times its own launch
    }
    if (s->pme)
        pme_enqueue(s, precision, eflag || vflag);
}

static LJArgs lj_args(atvh_system* s) {
    LJArgs a{};
    a.x = s->x;
    a.y = s->y;
    a.z = s->z;
    a.type = s->type;
    a.nbr = s->nbr;
    a.numneigh = s->numneigh;
    a.npad = s->npad;
    a.lx = s->box[0];
    a.ly = s->box[1];
    a.lz = s->box[2];
    a.ilx = 1.0 / s->box[0];
    a.ily = 1.0 / s->box[1];
    a.ilz = 1.0 / s->box[2];
    const double rc = s->lj_cutoff;
    a.rc_sq = rc * rc;
    const double sig6 = s->lj_sigma * s->lj_sigma * s->lj_sigma * s->lj_sigma * s->lj_sigma * s->lj_sigma;
    const double r2inv = 1.0 / a.rc_sq;
    const double r6inv = r2inv * r2inv * r2inv * sig6;
    a.eps24 = 24.0 * s->lj_eps;
    a.eps4 = 4.0 * s->lj_eps;
    a.sig6 = sig6;
    a.vshift = 4.0 * s->lj_eps * (r6inv * r6inv - r6inv);
    a.tab = reinterpret_cast<const double4*>(s->d_ljtab);
    a.ntypes = s->lj_ntypes;
    a.fx = s->fx;
    a.fy = s->fy;
    a.fz = s->fz;
    a.partials = s->partials;
    return a;
}

static BrickLJArgs brick_lj_args(atvh_system* s, const LJArgs& a) {
    BrickLJArgs b{};
    b.x = s->x;
    b.y = s->y;
    b.z = s->z;
    b.type = s->type;
    b.halo_rows = s->halo_rows;
    b.halo_n = s->halo_n;
    b.brick_range = s->brick_range;
    b.nbr16 = s->nbr16;
    b.self16 = s->self16;
    b.numneigh = s->numneigh;
    b.cap = s->cap;
    b.max_atoms = brick_atom_limit(s);
    for (int d = 0; d < 3; ++d) {
        b.L[d] = s->box[d];
        b.cs[d] = s->cs[d];
        b.nc[d] = s->nc[d];
        b.nb[d] = s->nb[d];
        b.bdim[d] = s->bdim[d];
    }
    b.rc_sq = a.rc_sq;
    b.eps24 = a.eps24;
    b.eps4 = a.eps4;
    b.sig6 = a.sig6;
    b.vshift = a.vshift;
    b.inv_sigma = a.sig6 > 0.0 ? std::pow(a.sig6, -1.0 / 6.0) : 1.0;
    {
        // The mixed kernel's r^2 comes from FP32 coordinates relative to the
        // brick centre, |X| <= E (the brick's half extent plus the stencil
        // reach and the skin an atom may drift before a rebuild). With unit
        // roundoff u = 2^-24: each difference carries <= 2Eu + |d|u, and the
        // fma chain adds <= 3u r^2, so near the cutoff
        //   |r2_fp32 - r2| <= u (4 sqrt(3) rc E + 5 rc^2).
        // Pairs within twice that of rc^2 take the plain kernel's FP64 test;
        // the roundings of rc^2 and of rc^2 -/+ band (<= 2u rc^2 together)
        // stay inside the second half.
        double E = 0.0;
        for (int d = 0; d < 3; ++d)
            E = std::max(E, (0.5 * s->bdim[d] + s->reach[d]) * s->cs[d] + s->skin);
        const bool typed = a.ntypes > 0 && a.tab;
        const double us = typed ? 1.0 : b.inv_sigma;
        const double rc = std::sqrt(a.rc_sq) * us;
        E *= us;
        const double u = std::ldexp(1.0, -24);
        b.band = static_cast<float>(2.0 * u * (4.0 * std::sqrt(3.0) * rc * E + 5.0 * rc * rc));
    }
    b.tab = a.tab;
    b.ntypes = a.ntypes;
    b.fx = s->fx;
    b.fy = s->fy;
    b.fz = s->fz;
    b.partials = s->partials;
    return b;
}

static void lj_enqueue(atvh_system* s, int precision, bool eflag, bool vflag) {
    if (s->lj_ntypes > 0 && s->ntypes > s->lj_ntypes)
        throw Error(ATVH_E_ARG, "PairLJCut: the frame has atom types beyond the coefficient table");
    const LJArgs a = lj_args(s);
    const bool typed = s->lj_ntypes > 0;
    const int g = pair_grid(s);
    hipStream_t st = s->ctx->stream;
    if (s->brick_mode) {
        const BrickLJArgs b = brick_lj_args(s, a);
        atvh_ctx* c = s->ctx;
        if (precision == ATVH_PREC_MIXED)
            lj_brick_dispatch<OpLjBrickMixed>(c, eflag, vflag, typed, g, st, b);
        else
            lj_brick_dispatch<OpLjBrickFp64>(c, eflag, vflag, typed, g, st, b);
        ATVH_HIP(hipGetLastError());
        return;
    }
    ScopedTimer tm(s->ctx, "pair");
    if (precision == ATVH_PREC_MIXED)
        LJ_DISPATCH(k_lj_mixed, eflag, vflag, typed, g, st, s->n, a);
    else
        LJ_DISPATCH(k_lj_fp64, eflag, vflag, typed, g, st, s->n, a);
    ATVH_HIP(hipGetLastError());
}

// Batched pair stage: one workgroup per brick of every member (PairLJCut).
// The force stage of a batched step: PairLJCut members (brick layout, no
// k-space) share one launch per kernel; every other member (LJ + Coulomb,
// with or without KSpacePME, or a small-grid layout) runs its own force call
// inside the same step, in member order.
void pair_batch_build(atvh_batch* b) {
    std::vector<BrickLJArgs> args;
    std::vector<int> blocks;
    bool typed_any = false, typed_all = true;
    b->solo.clear();
    for (atvh_system* s : b->sys) {
        if (!s->pair_configured)
            throw Error(ATVH_E_ARG, "batch: every member needs a pair style");
        if (s->pair_style != 0 || s->pme || !s->brick_mode) {
            b->solo.push_back(s);
            continue;
        }
        const bool t = s->lj_ntypes > 0;
        typed_any |= t;
        typed_all &= t;
        args.push_back(brick_lj_args(s, lj_args(s)));
        blocks.push_back(s->nbricks);
    }
    if (typed_any && !typed_all)
        throw Error(ATVH_E_ARG, "batch: PairLJCut members must all be single-type or all per-type");
    b->typed = typed_any;
    if (!args.empty())
        batch_stage(b, "pair", args, blocks);
}

template <class Op>
static void pair_batch_one(atvh_batch* b) {
    // the stage's table holds only the PairLJCut members (the others queue
    // their own force calls); every block's member comes from its owner table
    const BatchStage& st = b->stages.at("pair");
    k_batched<Op><<<st.nblocks, op_threads<Op>::value, 0, b->ctx->stream>>>(static_cast<const BrickLJArgs*>(st.args), st.base,
                                                                         st.owner);
}

void pair_batch_launch(atvh_batch* b, int precision, bool ef) {
    for (atvh_system* s : b->solo)
        pair_enqueue(s, precision, ef, ef);
    if (b->stages.find("pair") == b->stages.end())
        return;
    const bool mixed = precision == ATVH_PREC_MIXED;
    if (mixed) {
        if (ef)
            b->typed ? pair_batch_one<OpLjBrickMixed<true, true, true>>(b)
                     : pair_batch_one<OpLjBrickMixed<true, true, false>>(b);
        else
            b->typed ? pair_batch_one<OpLjBrickMixed<false, false, true>>(b)
                     : pair_batch_one<OpLjBrickMixed<false, false, false>>(b);
    } else {
        if (ef)
            b->typed ? pair_batch_one<OpLjBrickFp64<true, true, true>>(b)
                     : pair_batch_one<OpLjBrickFp64<true, true, false>>(b);
        else
            b->typed ? pair_batch_one<OpLjBrickFp64<false, false, true>>(b)
                     : pair_batch_one<OpLjBrickFp64<false, false, false>>(b);
    }
    ATVH_HIP(hipGetLastError());
}

void pair_free(atvh_system* s) { dev_free(s->d_ljtab); }

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

namespace atvh {
int64_t pair_device_bytes(const atvh_system* s) {
    if (!s->pair_configured)
        return 0;
    if (s->pair_style == 1)
        return coul_device_bytes(s);
    const int nt = std::max(1, s->lj_ntypes);
    return s->d_ljtab ? int64_t(nt) * nt * 4 * 8 : 0;
}
} // namespace atvh

extern "C" {

int atvh_pair_lj_configure(atvh_system* s, double eps, double sigma, double cutoff, int ntypes,
                           const double* eps_tab, const double* sigma_tab) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (!(cutoff > 0.0))
        throw Error(ATVH_E_ARG, "PairLJCut: cutoff must be positive");
    if (ntypes < 0 || ntypes > kMaxTypes)
        throw Error(ATVH_E_ARG, "PairLJCut: ntypes out of range");
    if (ntypes > 0 && (!eps_tab || !sigma_tab))
        throw Error(ATVH_E_ARG, "PairLJCut: the per-type form needs eps and sigma tables");
    ATVH_HIP(hipSetDevice(c->device));
    s->lj_eps = eps;
    s->lj_sigma = sigma;
    s->lj_cutoff = cutoff;
    s->lj_ntypes = ntypes;
    s->pair_style = 0;
    dev_free(s->d_ljtab);
    if (ntypes > 0) {
        // finalize_type_table_ (pair_lj_cut.cpp:257-298): sigma^6 and the
        // energy shift per pair; 24eps and 4eps prefactors.
        std::vector<double> tab(4 * ntypes * ntypes);
        for (int ab = 0; ab < ntypes * ntypes; ++ab) {
            const double e = eps_tab[ab], sg = sigma_tab[ab];
            const double s2 = sg * sg;
            const double s6 = s2 * s2 * s2;
            const double r2inv = 1.0 / (cutoff * cutoff);
            const double r6inv = r2inv * r2inv * r2inv * s6;
            tab[4 * ab + 0] = 24.0 * e;
            tab[4 * ab + 1] = 4.0 * e;
            tab[4 * ab + 2] = s6;
            tab[4 * ab + 3] = 4.0 * e * (r6inv * r6inv - r6inv);
        }
        dev_alloc(&s->d_ljtab, tab.size());
        ATVH_HIP(hipMemcpy(s->d_ljtab, tab.data(), tab.size() * sizeof(double),
                           hipMemcpyHostToDevice));
    }
    s->pair_configured = true;
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_pair_compute(atvh_system* s, int precision, int eflag, int vflag) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (precision != ATVH_PREC_FP64 && precision != ATVH_PREC_MIXED)
        throw Error(ATVH_E_ARG, "atvh_pair_compute: unknown precision");
    ATVH_HIP(hipSetDevice(c->device));
    pair_enqueue(s, precision, eflag != 0, vflag != 0);
    if (eflag || vflag)
        energy_enqueue_finalize(s, nullptr, 0);
    return ATVH_OK;
    API_CATCH(c)
}

} // extern "C"
