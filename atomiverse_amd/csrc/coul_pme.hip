// coul_pme.hip — charged systems: the fused LJ + real-space Coulomb pair
// style and smooth particle-mesh Ewald.
//
// PairLJCutCoulLong (src/cpu/potentials/pair_lj_cut_coul_long.cpp):
//   per type pair lj1..lj4 = 48 eps s^12, 24 eps s^6, 4 eps s^12, 4 eps s^6 and
//   the energy shift at the cutoff (:132-186, Lorentz-Berthelot for unset
//   off-diagonals); per pair inside r_c (r2 != 0):
//     r6inv = r2inv^3, fpair = r2inv (lj1 r6inv^2 - lj2 r6inv) + qqr2e qi qj f(r2),
//     E = lj3 r6inv^2 - lj4 r6inv - shift + qqr2e qi qj e(r2),
//   e, f linearly interpolated in r^2 over the ErfcTable (erfc_table.h:44-90,
//   10000 bins, built on the host with libm erfc/exp exactly as the reference).
//   Kernels: one workgroup per brick (same list and halo as the LJ kernels),
//   full list, energy/virial halved, per-type coefficients in LDS for up to 4
//   types. The FP64 kernel reads the table, stored per bin as
//   {f_i, f_{i+1} - f_i, e_i, e_{i+1} - e_i} (one 32-byte load per pair, the
//   differences formed on the host exactly as the reference forms them per
//   pair); the mixed kernel evaluates erfc in closed form (Abramowitz &
//   Stegun 7.1.26, |error| <= 1.5e-7) instead of loading a table row.
//
// KSpacePME (src/cpu/longrange/pme.cpp:147-253, pme_spline.h, pme_influence.h):
//   1. spread   one workgroup per brick of the neighbour list (or per chunk
//               of consecutive slots): fractional coordinate and base index in
//               FP64 (the reference's floor/wrap), B-spline weights (Essmann
//               recursion) in the compute precision, n^3 contributions added to
//               a 64-bit fixed-point copy of the brick's box of grid points in
//               LDS, then each touched point added once to the global
//               fixed-point grid. Integer atomics make the grid exact and
//               order-independent, so it is bitwise reproducible. The scale is
//               2^(61 - ceil(log2 sum|q|)), which no grid point can overflow
//               (a point holds at most sum|q_i|).
//   2. unfix    fixed point -> real grid (FP64 or FP32), clearing the
//               fixed-point grid for the next call.
//   3. hipFFT   R2C (D2Z / R2C), unnormalised.
//   4. influence  H(m) (host-built table, FP64, half spectrum) times Q(m);
//               energy and virial summed in FP64 over the half spectrum with
//               the Hermitian partner of each interior mode folded in.
//   5. hipFFT   C2R (Z2D / C2R).
//   6. gather   one thread per atom: weights and derivatives again, n^3 reads
//               of the (L2-resident) grid,
//               F = -(k_e / (pi V)) q sum_d g_d K_d / L_d, added to the forces.
//   Self and net-charge energies are host constants of the uploaded charges.

#include "atvh_internal.hpp"
#include "brick.hpp"
#include "pair_common.hpp"
#include "pme_fft.hpp"

#include <hipfft/hipfft.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <type_traits>
#include <complex>
#include <vector>

namespace atvh {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEwaldF = 1.1283791670955126; // 2/sqrt(pi), erfc_table.h:24
constexpr int kPmeMaxOrder = 8;
constexpr int kPmeGridCap = 1024; // influence kernel blocks (energy partials)

struct CoulState {
    int ntypes = 0;
    double cutoff = 0, alpha = 0;
    int nbin = 0;
    double inv_dr2 = 0;
    double* tab = nullptr;   // [ntypes^2][8] lj1 lj2 lj3 lj4 shift 0 0 0
    float* tab32 = nullptr;  // same, FP32
    double4* erfc = nullptr; // [nbin] {f, df, e, de}
    float4* erfc32 = nullptr;
    std::vector<double> eps_mixed, sigma_mixed; // finalize_type_table_ result (host copy)
};

struct PmeState {
    double alpha = 0, spacing = 0;
    int order = 0;
    int K[3] = {0, 0, 0};
    int Kh = 0; // K2/2 + 1
    double vol = 0, ic[3] = {0, 0, 0};
    double fix_scale = 0;
    unsigned long long* fix = nullptr;
    double* real64 = nullptr;
    double2* cplx64 = nullptr;
    float* real32 = nullptr;
    float2* cplx32 = nullptr;
    double* H = nullptr; // [K0][K1][Kh]
    double* partials = nullptr;
    int nb = 0;
    hipfftHandle f64 = nullptr, b64 = nullptr, f32 = nullptr, b32 = nullptr;
    // power-of-two meshes transform in FP32 with this library's own kernels
    // (pme_fft.hpp; FP64 keeps the FFT library, faster there: 197 vs 227 us
    // of PME at 128^3): twiddle tables exp(-2 pi i k / K_a), k <= K_a / 2
    bool native = false;
    float2* twf[3] = {nullptr, nullptr, nullptr};
};

} // namespace atvh

using namespace atvh;

namespace {

// ---------------------------------------------------------------- pair style

struct LjcArgs {
    const double *x, *y, *z, *q;
    const int* type;
    const int* halo_rows;
    const int* halo_n;
    const int2* brick_range;
    const uint16_t* nbr16;
    const uint16_t* self16;
    const int* numneigh;
    int cap;
    int max_atoms; // home atoms the list build took per brick
    double L[3], cs[3];
    int nc[3], nb[3], bdim[3];
    double rc_sq;
    int ntypes;
    const double* tab;
    const float* tab32;
    const double4* erfc;
    const float4* erfc32;
    int nbin;
    double inv_dr2;
    double alpha; // Ewald splitting parameter (the mixed kernels' closed-form erfc)
    double *fx, *fy, *fz;
    double* partials;
    // small-grid form (global slots, per-pair minimum image)
    const int* nbr;
    int64_t npad;
    double il[3];
    int64_t n;
};

// Force arithmetic is written with explicit fma and no automatic
// contraction, so it rounds identically in every EF/VF instantiation: a
// trajectory does not depend on which steps asked for energies.
constexpr int kLjcLdsTypes = 4; // type tables up to this size are staged in LDS

template <bool EF, bool VF, bool SMALLTAB>
__global__ __launch_bounds__(kBlock) void k_ljc_brick_fp64(LjcArgs a) {
    __shared__ double sx[kHaloCap], sy[kHaloCap], sz[kHaloCap], sq[kHaloCap];
    __shared__ int stype[kHaloCap];
    __shared__ double ljt[SMALLTAB ? kLjcLdsTypes * kLjcLdsTypes * 8 : 1]; // [ti][tj] lj1..lj4, shift
    const int b = xcd_brick(blockIdx.x, gridDim.x); // one launch block per brick
    const int nh = a.halo_n[b];
    const int2 br = a.brick_range[b];
    const int a0 = br.x, nat = br.y;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (brick_skipped(nh, nat, a.max_atoms)) { // reported by the build; nothing valid to compute
        if (EF || VF)
            block_partials(acc, a.partials, blockIdx.x);
        return;
    }
    const int* __restrict__ rows = a.halo_rows + (int64_t)b * kHaloCap;
    for (int r = threadIdx.x; r < nh; r += blockDim.x) {
        int gs;
        double xx, yy, zz;
        row_image(a.x, a.y, a.z, a.L[0], a.L[1], a.L[2], rows[r], gs, xx, yy, zz);
        sx[r] = xx;
        sy[r] = yy;
        sz[r] = zz;
        sq[r] = a.q[gs];
        stype[r] = a.type[gs];
    }
    if (threadIdx.x == 0) { // the list's padding row: out of every cutoff
        sx[kSentinelRow] = sy[kSentinelRow] = sz[kSentinelRow] = kSentinelFar;
        sq[kSentinelRow] = 0.0;
        stype[kSentinelRow] = 0;
    }
    if (SMALLTAB)
        for (int k = threadIdx.x; k < a.ntypes * a.ntypes * 8; k += blockDim.x)
            ljt[(k / (a.ntypes * 8)) * kLjcLdsTypes * 8 + k % (a.ntypes * 8)] = a.tab[k];
    __syncthreads();
    const double rc_sq = a.rc_sq, inv_dr2 = a.inv_dr2;
    const int ntypes = a.ntypes, nbin1 = a.nbin - 1;
    const double* __restrict__ tab = a.tab;
    const double4* __restrict__ etab = a.erfc;
    const uint4* __restrict__ base = reinterpret_cast<const uint4*>(a.nbr16 + (int64_t)a0 * a.cap);
    // two threads per atom when the brick has at most half a block of atoms
    const int tpa = (2 * nat <= static_cast<int>(blockDim.x)) ? 2 : 1;
    for (int t = threadIdx.x; t < nat * tpa; t += blockDim.x) {
        const int li = tpa == 2 ? (t >> 1) : t;
        const int h = tpa == 2 ? (t & 1) : 0;
        const int i = a0 + li;
        const int qs = a.self16[i];
        const double xi = sx[qs], yi = sy[qs], zi = sz[qs];
        const double qi_scale = coulomb * sq[qs];
        const double* __restrict__ trow =
            SMALLTAB ? ljt + stype[qs] * kLjcLdsTypes * 8 : tab + stype[qs] * ntypes * 8;
        const int nn = a.numneigh[i];
        const uint4* __restrict__ grp = base + li;
        const int ng = (nn + kGroup - 1) / kGroup;
        double fxi = 0.0, fyi = 0.0, fzi = 0.0;
        const int gl = max(ng - 1, 0); // group 0 exists even when nn == 0
        uint4 cur = grp[__umul24(min(h, gl), nat)];
        uint4 nx1 = grp[__umul24(min(h + tpa, gl), nat)];
        for (int g = h; g < ng; g += tpa) {
            const uint4 nx2 = grp[__umul24(min(g + 2 * tpa, gl), nat)];
            // the whole group's LDS reads issue before any arithmetic; padding
            // rows are the far sentinel (skipped like r >= r_c)
            double jx[kGroup], jy[kGroup], jz[kGroup], jq[kGroup];
            int jt[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const int q = group_entry(cur, u);
                jx[u] = sx[q];
                jy[u] = sy[q];
                jz[u] = sz[q];
                jq[u] = sq[q];
                jt[u] = stype[q];
            }
            // every pair's table row is read before any pair's arithmetic, so
            // the group's eight table loads overlap (a pair beyond the cutoff
            // reads the row of r_c and discards it; within it fmin is r^2)
            double4 evs[kGroup];
            double tts[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const double dx = xi - jx[u];
                const double dy = yi - jy[u];
                const double dz = zi - jz[u];
                const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                const double xb = fmin(r2, rc_sq) * inv_dr2;
                const int bi = min(static_cast<int>(xb), nbin1);
                tts[u] = xb - static_cast<double>(bi);
                evs[u] = etab[bi];
            }
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const double dx = xi - jx[u];
                const double dy = yi - jy[u];
                const double dz = zi - jz[u];
                const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                if (r2 < rc_sq && r2 != 0.0) {
                    const double* __restrict__ t = trow + jt[u] * 8;
                    const double r2inv = 1.0 / r2;
                    const double r6inv = r2inv * r2inv * r2inv;
                    const double tt = tts[u];
                    const double4 ev = evs[u];
                    const double qq = qi_scale * jq[u];
                    const double flj = r2inv * fma(t[0] * r6inv, r6inv, -(t[1] * r6inv));
                    const double fpair = fma(qq, fma(tt, ev.y, ev.x), flj);
                    fxi = fma(fpair, dx, fxi);
                    fyi = fma(fpair, dy, fyi);
                    fzi = fma(fpair, dz, fzi);
                    if (EF)
                        acc[0] += (t[2] * r6inv * r6inv - t[3] * r6inv - t[4]) + qq * (ev.z + tt * ev.w);
                    if (VF) {
                        const double fxp = fpair * dx, fyp = fpair * dy, fzp = fpair * dz;
                        acc[1] += dx * fxp;
                        acc[2] += dy * fyp;
                        acc[3] += dz * fzp;
                        acc[4] += dx * fyp;
                        acc[5] += dx * fzp;
                        acc[6] += dy * fzp;
                    }
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
            acc[k] *= 0.5; // full list: every pair seen from both ends
        block_partials(acc, a.partials, blockIdx.x);
    }
}

// erfc(x) / r and exp(-x^2) in FP32 for the mixed path: Abramowitz & Stegun,
// Handbook of Mathematical Functions (1964), eq. 7.1.26,
// erfc(x) = t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-x^2), t = 1/(1 + p x),
// |error| <= 1.5e-7 for x >= 0 (below the FP32 table's own interpolation
// error at the mixed path's force bar). exp via v_exp_f32 (2^y).
__device__ __forceinline__ float erfc_over_r(float x, float rinv, float& ex) {
    constexpr float p = 0.3275911f;
    constexpr float a1 = 0.254829592f, a2 = -0.284496736f, a3 = 1.421413741f, a4 = -1.453152027f,
                    a5 = 1.061405429f;
    constexpr float log2e = 1.4426950408889634f;
    const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(p, x, 1.0f));
    const float poly =
        t * __builtin_fmaf(t, __builtin_fmaf(t, __builtin_fmaf(t, __builtin_fmaf(t, a5, a4), a3), a2), a1);
    ex = __builtin_amdgcn_exp2f(-(x * x) * log2e);
    return poly * ex * rinv;
}

// 512 threads and at most 64 VGPRs (half groups): 8 waves per SIMD for the
// force-only kernel (112 -> 107 us on the 512k-ion melt)
constexpr int kLjcMixedThreads = 512;

template <bool EF, bool VF, bool SMALLTAB>
__global__ __launch_bounds__(kLjcMixedThreads) __attribute__((amdgpu_waves_per_eu((EF || VF) ? 4 : 8)))
void k_ljc_brick_mixed(LjcArgs a) {
    __shared__ float4 sp[kHaloCap]; // x, y, z relative to the brick centre; w = charge
    __shared__ unsigned char stype[kHaloCap];
    __shared__ float ljt[SMALLTAB ? kLjcLdsTypes * kLjcLdsTypes * 8 : 1]; // [ti][tj] lj1..lj4, shift
    const int b = xcd_brick(blockIdx.x, gridDim.x); // one launch block per brick
    const int nh = a.halo_n[b];
    const int2 br = a.brick_range[b];
    const int a0 = br.x, nat = br.y;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (brick_skipped(nh, nat, a.max_atoms)) { // reported by the build; nothing valid to compute
        if (EF || VF)
            block_partials<kLjcMixedThreads>(acc, a.partials, blockIdx.x);
        return;
    }
    const int* __restrict__ rows = a.halo_rows + (int64_t)b * kHaloCap;
    double r0x, r0y, r0z;
    brick_centre(b, a.nb, a.bdim, a.nc, a.cs, r0x, r0y, r0z);
    for (int r = threadIdx.x; r < nh; r += blockDim.x) {
        int gs;
        double xx, yy, zz;
        row_image(a.x, a.y, a.z, a.L[0], a.L[1], a.L[2], rows[r], gs, xx, yy, zz);
        float4 v;
        v.x = static_cast<float>(xx - r0x);
        v.y = static_cast<float>(yy - r0y);
        v.z = static_cast<float>(zz - r0z);
        v.w = static_cast<float>(a.q[gs]);
        sp[r] = v;
        stype[r] = static_cast<unsigned char>(a.type[gs]);
    }
    if (threadIdx.x == 0) { // the list's padding row: out of every cutoff
        sp[kSentinelRow] = make_float4(kSentinelFarF, kSentinelFarF, kSentinelFarF, 0.0f);
        stype[kSentinelRow] = 0;
    }
    if (SMALLTAB)
        for (int k = threadIdx.x; k < a.ntypes * a.ntypes * 8; k += blockDim.x)
            ljt[(k / (a.ntypes * 8)) * kLjcLdsTypes * 8 + k % (a.ntypes * 8)] = a.tab32[k];
    __syncthreads();
    const float rc_sq = static_cast<float>(a.rc_sq);
    const float alpha = static_cast<float>(a.alpha);
    const float two_a_rtpi = static_cast<float>(kEwaldF * a.alpha);
    const int ntypes = a.ntypes;
    const float* __restrict__ tab = a.tab32;
    const uint4* __restrict__ base = reinterpret_cast<const uint4*>(a.nbr16 + (int64_t)a0 * a.cap);
    // two threads per atom when the brick has at most half a block of atoms
    const int tpa = (2 * nat <= static_cast<int>(blockDim.x)) ? 2 : 1;
    for (int t = threadIdx.x; t < nat * tpa; t += blockDim.x) {
        const int li = tpa == 2 ? (t >> 1) : t;
        const int h = tpa == 2 ? (t & 1) : 0;
        const int i = a0 + li;
        const int qs = a.self16[i];
        const float4 pi = sp[qs];
        const float qi_scale = static_cast<float>(coulomb) * pi.w;
        const float* __restrict__ trow = SMALLTAB ? ljt + stype[qs] * kLjcLdsTypes * 8 : tab + stype[qs] * ntypes * 8;
        const int nn = a.numneigh[i];
        const uint4* __restrict__ grp = base + li;
        const int ng = (nn + kGroup - 1) / kGroup;
        float fxi = 0.f, fyi = 0.f, fzi = 0.f;
        const int gl = max(ng - 1, 0); // group 0 exists even when nn == 0
        uint4 cur = grp[__umul24(min(h, gl), nat)];
        uint4 nx1 = grp[__umul24(min(h + tpa, gl), nat)];
        for (int g = h; g < ng; g += tpa) {
            const uint4 nx2 = grp[__umul24(min(g + 2 * tpa, gl), nat)];
            // each half group's LDS reads issue before its arithmetic (half
            // groups keep the kernel within 64 VGPRs: 8 waves per SIMD); padding
            // rows are the far sentinel (skipped like r >= r_c)
#pragma unroll
            for (int hg = 0; hg < kGroup; hg += kGroup / 2) {
                float4 pjs[kGroup / 2];
                int jt[kGroup / 2];
#pragma unroll
                for (int u = 0; u < kGroup / 2; ++u) {
                    const int q = group_entry(cur, hg + u);
                    pjs[u] = sp[q];
                    jt[u] = stype[q];
                }
#pragma unroll
                for (int u = 0; u < kGroup / 2; ++u) {
                    const float4 pj = pjs[u];
                    const float dx = pi.x - pj.x;
                    const float dy = pi.y - pj.y;
                    const float dz = pi.z - pj.z;
                    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    if (r2 < rc_sq && r2 != 0.0f) {
                        const float4 t = *reinterpret_cast<const float4*>(trow + jt[u] * 8);
                        // real-space Coulomb in closed form (no table load):
                        // e = erfc(a r) / r, f = (e + 2/sqrt(pi) a exp(-a^2 r^2)) / r^2
                        const float rinv = __builtin_amdgcn_rsqf(r2);
                        const float r2inv = rinv * rinv;
                        const float r6inv = r2inv * r2inv * r2inv;
                        float ex;
                        const float ecoul = erfc_over_r(r2 * rinv * alpha, rinv, ex);
                        const float fcoul = __builtin_fmaf(two_a_rtpi, ex, ecoul) * r2inv;
                        const float qq = qi_scale * pj.w;
                        const float fpair = r2inv * __builtin_fmaf(t.x * r6inv, r6inv, -t.y * r6inv) + qq * fcoul;
                        fxi = __builtin_fmaf(fpair, dx, fxi);
                        fyi = __builtin_fmaf(fpair, dy, fyi);
                        fzi = __builtin_fmaf(fpair, dz, fzi);
                        if (EF) {
                            const float sh = trow[jt[u] * 8 + 4];
                            acc[0] += static_cast<double>(__builtin_fmaf(t.z * r6inv, r6inv, -t.w * r6inv) - sh) +
                                      static_cast<double>(qq * ecoul);
                        }
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
        a.fx[i] = static_cast<double>(fxi);
        a.fy[i] = static_cast<double>(fyi);
        a.fz[i] = static_cast<double>(fzi);
    }
    if (EF || VF) {
#pragma unroll
        for (int k = 0; k < 7; ++k)
            acc[k] *= 0.5;
        block_partials<kLjcMixedThreads>(acc, a.partials, blockIdx.x);
    }
}

template <bool SMALLTAB>
void launch_ljc_fp64(bool ef, bool vf, int grid, hipStream_t st, const LjcArgs& a) {
    if (ef) {
        if (vf)
            k_ljc_brick_fp64<true, true, SMALLTAB><<<grid, kBlock, 0, st>>>(a);
        else
            k_ljc_brick_fp64<true, false, SMALLTAB><<<grid, kBlock, 0, st>>>(a);
    } else {
        if (vf)
            k_ljc_brick_fp64<false, true, SMALLTAB><<<grid, kBlock, 0, st>>>(a);
        else
            k_ljc_brick_fp64<false, false, SMALLTAB><<<grid, kBlock, 0, st>>>(a);
    }
}

template <bool SMALLTAB>
void launch_ljc_mixed(bool ef, bool vf, int grid, hipStream_t st, const LjcArgs& a) {
    if (ef) {
        if (vf)
            k_ljc_brick_mixed<true, true, SMALLTAB><<<grid, kLjcMixedThreads, 0, st>>>(a);
        else
            k_ljc_brick_mixed<true, false, SMALLTAB><<<grid, kLjcMixedThreads, 0, st>>>(a);
    } else {
        if (vf)
            k_ljc_brick_mixed<false, true, SMALLTAB><<<grid, kLjcMixedThreads, 0, st>>>(a);
        else
            k_ljc_brick_mixed<false, false, SMALLTAB><<<grid, kLjcMixedThreads, 0, st>>>(a);
    }
}

// Small grids (< 2*reach+1 cells on an axis): global-slot ELL list with the
// reference's per-pair minimum image d -= L nearbyint(d / L) (1/L hoisted).
template <bool EF, bool VF, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_ljc_global(LjcArgs a) {
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const int ntypes = a.ntypes, nbin1 = a.nbin - 1;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
        const double qi_scale = coulomb * a.q[i];
        const int ti = a.type[i];
        const int nn = a.numneigh[i];
        double fxi = 0.0, fyi = 0.0, fzi = 0.0;
        for (int k = 0; k < nn; ++k) {
            const int j = a.nbr[(int64_t)k * a.npad + i];
            double dx = xi - a.x[j];
            double dy = yi - a.y[j];
            double dz = zi - a.z[j];
            dx -= a.L[0] * rint(dx * a.il[0]);
            dy -= a.L[1] * rint(dy * a.il[1]);
            dz -= a.L[2] * rint(dz * a.il[2]);
            const double r2 = dx * dx + dy * dy + dz * dz; // FP64 in both modes: same survivors
            if (!(r2 < a.rc_sq && r2 != 0.0))
                continue;
            const int ab = (ti * ntypes + a.type[j]) * 8;
            double fpair, epair = 0.0;
            if (MIXED) {
                const float r2f = static_cast<float>(r2);
                const float r2inv = __builtin_amdgcn_rcpf(r2f);
                const float r6inv = r2inv * r2inv * r2inv;
                const float xb = r2f * static_cast<float>(a.inv_dr2);
                const int bi = min(static_cast<int>(xb), nbin1);
                const float tt = xb - static_cast<float>(bi);
                const float4 ev = a.erfc32[bi];
                const float qq = static_cast<float>(qi_scale) * static_cast<float>(a.q[j]);
                const float* t = a.tab32 + ab;
                fpair = static_cast<double>(r2inv * (t[0] * r6inv * r6inv - t[1] * r6inv) + qq * (ev.x + tt * ev.y));
                if (EF)
                    epair = static_cast<double>((t[2] * r6inv * r6inv - t[3] * r6inv - t[4]) + qq * (ev.z + tt * ev.w));
            } else {
                const double r2inv = 1.0 / r2;
                const double r6inv = r2inv * r2inv * r2inv;
                const double xb = r2 * a.inv_dr2;
                const int bi = min(static_cast<int>(xb), nbin1);
                const double tt = xb - static_cast<double>(bi);
                const double4 ev = a.erfc[bi];
                const double qq = qi_scale * a.q[j];
                const double* t = a.tab + ab;
                fpair = r2inv * (t[0] * r6inv * r6inv - t[1] * r6inv) + qq * (ev.x + tt * ev.y);
                if (EF)
                    epair = (t[2] * r6inv * r6inv - t[3] * r6inv - t[4]) + qq * (ev.z + tt * ev.w);
            }
            const double fxp = fpair * dx, fyp = fpair * dy, fzp = fpair * dz;
            fxi += fxp;
            fyi += fyp;
            fzi += fzp;
            if (EF)
                acc[0] += epair;
            if (VF) {
                acc[1] += dx * fxp;
                acc[2] += dy * fyp;
                acc[3] += dz * fzp;
                acc[4] += dx * fyp;
                acc[5] += dx * fzp;
                acc[6] += dy * fzp;
            }
        }
        a.fx[i] = fxi;
        a.fy[i] = fyi;
        a.fz[i] = fzi;
    }
    if (EF || VF) {
#pragma unroll
        for (int k = 0; k < 7; ++k)
            acc[k] *= 0.5;
        block_partials(acc, a.partials, blockIdx.x);
    }
}

// ---------------------------------------------------------------- PME

struct PmeArgs {
    int64_t n;
    const double *x, *y, *z, *q;
    double *fx, *fy, *fz;
    int K0, K1, K2, Kh;
    double ic0, ic1, ic2;
    double fix_scale, inv_fix;
    unsigned long long* fix;
    double fscale;      // k_e / (pi V)
    double pi2_a2;      // pi^2 / alpha^2
    const double* H;
    double* partials;
};

__device__ __forceinline__ int pme_wrap(int g, int K) {
    // pme_spline.h:149-160: at most one period either side
    if (g >= K) g -= K;
    if (g < 0) g += K;
    if (g >= K) g -= K;
    if (g < 0) g += K;
    return g;
}

// bspline (pme_spline.h:29-55) in precision T, order N.
template <class T, int N>
__host__ __device__ __forceinline__ void bspline_t(T frac, T (&th)[N], T (&dth)[N]) {
    th[N - 1] = T(0);
    th[1] = frac;
    th[0] = T(1) - frac;
#pragma unroll
    for (int k = 3; k < N; ++k) {
        const T div = T(1) / T(k - 1);
        th[k - 1] = div * frac * th[k - 2];
#pragma unroll
        for (int j = 1; j < k - 1; ++j)
            th[k - 1 - j] = div * ((frac + T(j)) * th[k - 2 - j] + (T(k - j) - frac) * th[k - 1 - j]);
        th[0] = div * (T(1) - frac) * th[0];
    }
    dth[0] = -th[0];
#pragma unroll
    for (int j = 1; j < N; ++j)
        dth[j] = th[j - 1] - th[j];
    const T div = T(1) / T(N - 1);
    th[N - 1] = div * frac * th[N - 2];
#pragma unroll
    for (int j = 1; j < N - 1; ++j)
        th[N - 1 - j] = div * ((frac + T(j)) * th[N - 2 - j] + (T(N - j) - frac) * th[N - 1 - j]);
    th[0] = div * (T(1) - frac) * th[0];
}

// pme_fill_atom_splines (pme_spline.h:131-146), one axis: u = (x/L) K wrapped
// into [0, K) by u - floor(u/K) K, in FP64; the weights in precision T.
__device__ __forceinline__ int axis_cell(double xv, double ic, int K, double& frac) {
    const double s = ic * xv;
    double u = s * K;
    u -= floor(u / K) * K;
    // u lies in [0, K]; the clamp only keeps a non-finite coordinate in bounds
    const int u0 = min(max(static_cast<int>(floor(u)), 0), K);
    frac = u - u0;
    return u0;
}

template <class T, int N>
__device__ __forceinline__ int axis_spline(double xv, double ic, int K, T (&th)[N], T (&dth)[N]) {
    double frac;
    const int u0 = axis_cell(xv, ic, K, frac);
    bspline_t<T, N>(static_cast<T>(frac), th, dth);
    return u0 - (N - 1);
}

// Fixed-point value of one weight: the same rounding on every path, so the
// grid does not depend on which path (or order) added it. The scale keeps
// |w| scale below 2^50, so w scale + 1.5 2^52 lies in the binade whose
// doubles are the integers: the FMA's one rounding gives
// rint(w scale) + 1.5 2^52 (ties to even, the offset being even), and
// subtracting the offset's bit pattern leaves rint(w scale) in two's
// complement (three instructions instead of rint and a double -> int64
// conversion).
template <class T>
__device__ __forceinline__ unsigned long long fix_of(T w, double scale) {
    constexpr double kRound = 0x1.8p52;
    const double d = __builtin_fma(static_cast<double>(w), scale, kRound);
    return static_cast<unsigned long long>(__double_as_longlong(d)) -
           static_cast<unsigned long long>(__double_as_longlong(kRound));
}

// All n^3 contributions of atom i added straight to the global grid.
template <class T, int N>
__device__ __forceinline__ void spread_atom_global(const PmeArgs& a, int64_t i) {
    const double qi = a.q[i];
    if (qi == 0.0)
        return;
    T tx[N], ty[N], tz[N], dx[N], dy[N], dz[N];
    const int b0 = axis_spline<T, N>(a.x[i], a.ic0, a.K0, tx, dx);
    const int b1 = axis_spline<T, N>(a.y[i], a.ic1, a.K1, ty, dy);
    const int b2 = axis_spline<T, N>(a.z[i], a.ic2, a.K2, tz, dz);
    const T qt = static_cast<T>(qi);
    int g2s[N];
#pragma unroll
    for (int c = 0; c < N; ++c)
        g2s[c] = pme_wrap(b2 + c, a.K2);
#pragma unroll
    for (int p = 0; p < N; ++p) {
        const int g0 = pme_wrap(b0 + p, a.K0);
        const T wa = qt * tx[p];
#pragma unroll
        for (int r = 0; r < N; ++r) {
            const int g1 = pme_wrap(b1 + r, a.K1);
            const T wab = wa * ty[r];
            unsigned long long* __restrict__ row = a.fix + (int64_t)a.K2 * (g1 + (int64_t)a.K1 * g0);
#pragma unroll
            for (int c = 0; c < N; ++c)
                atomicAdd(row + g2s[c], fix_of(wab * tz[c], a.fix_scale));
        }
    }
}

// Offset of grid index b from the reference index r along a periodic axis of
// K points, folded into [-K/2, K/2] (|b - r| <= K + n on entry).
__device__ __forceinline__ int grid_offset(int b, int r, int K) {
    int d = b - r;
    if (2 * d > K)
        d -= K;
    else if (2 * d < -K)
        d += K;
    return d;
}

__device__ __forceinline__ int wave_min_i(int v) {
    for (int off = 32; off > 0; off >>= 1)
        v = min(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, __shfl_xor(v, off, 64));
    return v;
}

constexpr int kSpreadChunk = kBlock; // atoms per workgroup without brick ranges
constexpr int kSpreadCap = 6144;     // most LDS grid points per workgroup (48 KiB of int64)
constexpr int kSpreadGridCap = 8192; // workgroups of one spread launch

// The box of grid points a range of atoms touches. Base indices are taken as
// periodic offsets from the range's first atom, so a range that drifted
// across the box edge stays compact; box point (i0, i1, i2) is grid point
// (o + i) mod K on each axis. All threads of the block call it (two barriers).
struct RangeBox {
    int64_t first;
    int count;
    int ref[3]; // base index of the first atom
    int lo[3];  // smallest offset from ref
    int o[3];   // ref + lo: grid index of box point 0 (unwrapped)
    int e[3];   // box extent (points)
    int64_t total;
};

__device__ __forceinline__ bool range_of(const PmeArgs& a, const int2* __restrict__ ranges, int r, RangeBox& rb) {
    if (ranges) {
        const int2 br = ranges[r];
        rb.first = br.x;
        rb.count = br.y;
    } else {
        rb.first = (int64_t)r * kSpreadChunk;
        rb.count = static_cast<int>(min<int64_t>(kSpreadChunk, a.n - rb.first));
    }
    return rb.count > 0;
}

template <int N>
__device__ __forceinline__ void range_box(const PmeArgs& a, RangeBox& rb, int (*red)[6]) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double fr;
    rb.ref[0] = axis_cell(a.x[rb.first], a.ic0, a.K0, fr) - (N - 1);
    rb.ref[1] = axis_cell(a.y[rb.first], a.ic1, a.K1, fr) - (N - 1);
    rb.ref[2] = axis_cell(a.z[rb.first], a.ic2, a.K2, fr) - (N - 1);
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (int t = threadIdx.x; t < rb.count; t += blockDim.x) {
        const int64_t i = rb.first + t;
        const int d0 = grid_offset(axis_cell(a.x[i], a.ic0, a.K0, fr) - (N - 1), rb.ref[0], a.K0);
        const int d1 = grid_offset(axis_cell(a.y[i], a.ic1, a.K1, fr) - (N - 1), rb.ref[1], a.K1);
        const int d2 = grid_offset(axis_cell(a.z[i], a.ic2, a.K2, fr) - (N - 1), rb.ref[2], a.K2);
        lo[0] = min(lo[0], d0);
        lo[1] = min(lo[1], d1);
        lo[2] = min(lo[2], d2);
        hi[0] = max(hi[0], d0);
        hi[1] = max(hi[1], d1);
        hi[2] = max(hi[2], d2);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = wave_min_i(lo[k]);
        hi[k] = wave_max_i(hi[k]);
    }
    __syncthreads(); // red[] may still be read for the previous range
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            red[wid][k] = lo[k];
            red[wid][3 + k] = hi[k];
        }
    __syncthreads();
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int l = red[0][k], h = red[0][3 + k];
        for (int w = 1; w < nw; ++w) {
            l = min(l, red[w][k]);
            h = max(h, red[w][3 + k]);
        }
        rb.lo[k] = l;
        rb.o[k] = rb.ref[k] + l;
        rb.e[k] = h - l + N;
    }
    rb.total = (int64_t)rb.e[0] * rb.e[1] * rb.e[2];
}

__device__ __forceinline__ int wrap_any(int g, int K) {
    g %= K;
    return g < 0 ? g + K : g;
}

// Spread with the grid privatised per workgroup. Atoms are stored in cell
// order, so a brick's atoms (or a chunk of consecutive slots) touch a small
// box of grid points: the workgroup accumulates all its contributions to that
// box in LDS with 64-bit integer atomics and adds every touched point to the
// global grid once. The LDS box holds `cap` points (dynamic shared memory
// sized by the host from the brick extent, so small bricks leave room for
// more resident workgroups); ranges whose box exceeds it go straight to the
// global grid. Integer addition makes both paths give the same grid
// bit for bit. Work items are (atom, first-axis offset p) pairs so a range of
// up to n * count items spreads evenly over the workgroup.
template <class T, int N>
__global__ __launch_bounds__(kBlock) void k_pme_spread_lds(PmeArgs a, const int2* __restrict__ ranges,
                                                           int nranges, int cap) {
    extern __shared__ unsigned long long acc[]; // cap points
    __shared__ int red[kBlock / 64][6];
    for (int t = blockIdx.x; t < nranges; t += gridDim.x) {
        const int r = xcd_brick(t, nranges);
        RangeBox rb;
        if (!range_of(a, ranges, r, rb))
            continue;
        range_box<N>(a, rb, red);
        if (rb.total > cap) {
            for (int t = threadIdx.x; t < rb.count; t += blockDim.x)
                spread_atom_global<T, N>(a, rb.first + t);
            continue;
        }
        const int nbox = static_cast<int>(rb.total);
        const int e1 = rb.e[1], e2 = rb.e[2];
        for (int k = threadIdx.x; k < nbox; k += blockDim.x)
            acc[k] = 0ull;
        __syncthreads();
        for (int item = threadIdx.x; item < rb.count * N; item += blockDim.x) {
            const int t = item / N, p = item - t * N;
            const int64_t i = rb.first + t;
            const double qi = a.q[i];
            if (qi == 0.0)
                continue;
            T tx[N], ty[N], tz[N], dx[N], dy[N], dz[N];
            const int b0 = axis_spline<T, N>(a.x[i], a.ic0, a.K0, tx, dx);
            const int b1 = axis_spline<T, N>(a.y[i], a.ic1, a.K1, ty, dy);
            const int b2 = axis_spline<T, N>(a.z[i], a.ic2, a.K2, tz, dz);
            T txp = tx[0];
#pragma unroll
            for (int k = 1; k < N; ++k)
                txp = (k == p) ? tx[k] : txp;
            const int l0 = grid_offset(b0, rb.ref[0], a.K0) - rb.lo[0] + p;
            const int l1 = grid_offset(b1, rb.ref[1], a.K1) - rb.lo[1];
            const int l2 = grid_offset(b2, rb.ref[2], a.K2) - rb.lo[2];
            const T wa = static_cast<T>(qi) * txp;
            unsigned long long* plane = acc + (l0 * e1 + l1) * e2 + l2;
#pragma unroll
            for (int rr = 0; rr < N; ++rr) {
                const T wab = wa * ty[rr];
#pragma unroll
                for (int c = 0; c < N; ++c)
                    atomicAdd(plane + rr * e2 + c, fix_of(wab * tz[c], a.fix_scale));
            }
        }
        __syncthreads();
        // box point k = (i0 e1 + i1) e2 + i2 is walked with the block's stride:
        // the indices advance by the stride's own digits (at most one carry
        // each), so the walk divides once per range, not per point
        const int st = static_cast<int>(blockDim.x);
        const int s2 = st % e2, s1 = (st / e2) % e1, s0 = st / (e2 * e1);
        int i2 = static_cast<int>(threadIdx.x) % e2;
        int i1 = (static_cast<int>(threadIdx.x) / e2) % e1;
        int i0 = static_cast<int>(threadIdx.x) / (e2 * e1);
        const int w0 = wrap_any(rb.o[0], a.K0), w1 = wrap_any(rb.o[1], a.K1), w2 = wrap_any(rb.o[2], a.K2);
        for (int k = threadIdx.x; k < nbox; k += st) {
            const unsigned long long v = acc[k];
            if (v != 0ull) {
                int g0 = w0 + i0, g1 = w1 + i1, g2 = w2 + i2;
                while (g0 >= a.K0)
                    g0 -= a.K0;
                while (g1 >= a.K1)
                    g1 -= a.K1;
                while (g2 >= a.K2)
                    g2 -= a.K2;
                atomicAdd(a.fix + (int64_t)a.K2 * (g1 + (int64_t)a.K1 * g0) + g2, v);
            }
            i2 += s2;
            if (i2 >= e2) {
                i2 -= e2;
                ++i1;
            }
            i1 += s1;
            if (i1 >= e1) {
                i1 -= e1;
                ++i0;
            }
            i0 += s0;
        }
        __syncthreads(); // acc[] is cleared for the next range
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pme_unfix(int64_t ntot, unsigned long long* fix, T* real,
                                                      double inv_fix) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ntot;
         i += (int64_t)gridDim.x * blockDim.x) {
        const long long v = static_cast<long long>(fix[i]);
        real[i] = static_cast<T>(static_cast<double>(v) * inv_fix);
        fix[i] = 0ull;
    }
}

// pme_apply_influence_tables (pme_influence.h:179-222) on the R2C half
// spectrum. Modes with 0 < m2 < K2/2 stand for themselves and their Hermitian
// partner (-m): energy and diagonal virial count twice; an off-diagonal term
// m_a m_b of the partner changes sign unless one of the two indices is the
// Nyquist plane of its axis (h = K/2 maps to itself), so it is folded in as
// m_a m_b (1 + s_a s_b) with s = -1, or +1 on a Nyquist plane (s_z = -1).
template <class T2, bool EV>
__global__ __launch_bounds__(kBlock) void k_pme_influence(PmeArgs a, T2* __restrict__ cplx) {
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const int64_t nmodes = (int64_t)a.K0 * a.K1 * a.Kh;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < nmodes;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const double h = a.H[idx];
        T2 c = cplx[idx];
        if (EV && h != 0.0) {
            const int m2 = static_cast<int>(idx % a.Kh);
            const int64_t r = idx / a.Kh;
            const int m1 = static_cast<int>(r % a.K1);
            const int m0 = static_cast<int>(r / a.K1);
            const int h0 = (m0 <= a.K0 / 2) ? m0 : m0 - a.K0;
            const int h1 = (m1 <= a.K1 / 2) ? m1 : m1 - a.K1;
            const int h2 = (m2 <= a.K2 / 2) ? m2 : m2 - a.K2;
            const double mx = h0 * a.ic0, my = h1 * a.ic1, mz = h2 * a.ic2;
            const double msq = mx * mx + my * my + mz * mz;
            const double re = static_cast<double>(c.x), im = static_cast<double>(c.y);
            const double q2 = re * re + im * im;
            const double vt = -2.0 * (1.0 / msq + a.pi2_a2);
            const double ug = h * q2;
            const bool pair = m2 != 0 && 2 * m2 != a.K2;
            const double w = pair ? 2.0 : 1.0;
            const double sx = (2 * m0 == a.K0) ? 1.0 : -1.0;
            const double sy = (2 * m1 == a.K1) ? 1.0 : -1.0;
            const double oxy = pair ? 1.0 + sx * sy : 1.0;
            const double oxz = pair ? 1.0 - sx : 1.0;
            const double oyz = pair ? 1.0 - sy : 1.0;
            acc[0] += w * ug;
            acc[1] += w * ug * (1.0 + vt * mx * mx);
            acc[2] += w * ug * (1.0 + vt * my * my);
            acc[3] += w * ug * (1.0 + vt * mz * mz);
            acc[4] += oxy * ug * (vt * mx * my);
            acc[5] += oxz * ug * (vt * mx * mz);
            acc[6] += oyz * ug * (vt * my * mz);
        }
        using TS = decltype(c.x);
        c.x *= static_cast<TS>(h);
        c.y *= static_cast<TS>(h);
        cplx[idx] = c;
    }
    if (EV)
        block_partials(acc, a.partials, blockIdx.x);
}

// Force on atom i from the potential grid, read one row of n points along
// the last axis at a time through row(b0, b1, b2, p, r, v) with the base
// indices b (grid index of weight 0 on each axis, before wrapping).
template <class T, int N, class Row>
__device__ __forceinline__ void gather_atom(const PmeArgs& a, int64_t i, Row row) {
    const double qi = a.q[i];
    if (qi == 0.0)
        return;
    T tx[N], ty[N], tz[N], dx[N], dy[N], dz[N];
    const int b0 = axis_spline<T, N>(a.x[i], a.ic0, a.K0, tx, dx);
    const int b1 = axis_spline<T, N>(a.y[i], a.ic1, a.K1, ty, dy);
    const int b2 = axis_spline<T, N>(a.z[i], a.ic2, a.K2, tz, dz);
    T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
    for (int p = 0; p < N; ++p) {
        T v[N][N]; // the plane's n rows, all read before its arithmetic
#pragma unroll
        for (int r = 0; r < N; ++r)
            row(b0, b1, b2, p, r, v[r]);
#pragma unroll
        for (int r = 0; r < N; ++r) {
#pragma unroll
            for (int c = 0; c < N; ++c) {
                const T cval = v[r][c];
                gx += cval * dx[p] * ty[r] * tz[c];
                gy += cval * tx[p] * dy[r] * tz[c];
                gz += cval * tx[p] * ty[r] * dz[c];
            }
        }
    }
    const double fxi = static_cast<double>(gx) * a.K0 * a.ic0;
    const double fyi = static_cast<double>(gy) * a.K1 * a.ic1;
    const double fzi = static_cast<double>(gz) * a.K2 * a.ic2;
    const double pre = -a.fscale * qi;
    a.fx[i] += pre * fxi;
    a.fy[i] += pre * fyi;
    a.fz[i] += pre * fzi;
}

// The row's first two indices are wrapped once per row, its last index per
// point (the row may cross the periodic edge).
template <class T, int N>
__device__ __forceinline__ void gather_atom_global(const PmeArgs& a, int64_t i, const T* __restrict__ real) {
    gather_atom<T, N>(a, i, [&](int b0, int b1, int b2, int p, int r, T (&v)[N]) {
        const int g0 = pme_wrap(b0 + p, a.K0), g1 = pme_wrap(b1 + r, a.K1);
        const T* __restrict__ rp = real + (int64_t)a.K2 * (g1 + (int64_t)a.K1 * g0);
#pragma unroll
        for (int c = 0; c < N; ++c)
            v[c] = rp[pme_wrap(b2 + c, a.K2)];
    });
}

// One thread per atom, grid values read from the global grid (L2-resident;
// a box staged in LDS per brick measured slower).
template <class T, int N>
__global__ __launch_bounds__(kBlock) void k_pme_gather(PmeArgs a, const T* __restrict__ real) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.n;
         i += (int64_t)gridDim.x * blockDim.x)
        gather_atom_global<T, N>(a, i, real);
}

__global__ __launch_bounds__(kBlock) void k_zero3(int64_t n, double* a, double* b, double* c) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        a[i] = 0.0;
        b[i] = 0.0;
        c[i] = 0.0;
    }
}

#define LJC_DISPATCH3(KERNEL, EF, VF, M, GRID, ST, ARGS)                                        \
    do {                                                                                       \
        if (M) {                                                                               \
            if (EF) {                                                                          \
                if (VF) KERNEL<true, true, true><<<GRID, kBlock, 0, ST>>>(ARGS);               \
                else KERNEL<true, false, true><<<GRID, kBlock, 0, ST>>>(ARGS);                 \
            } else {                                                                           \
                if (VF) KERNEL<false, true, true><<<GRID, kBlock, 0, ST>>>(ARGS);              \
                else KERNEL<false, false, true><<<GRID, kBlock, 0, ST>>>(ARGS);                \
            }                                                                                  \
        } else {                                                                               \
            if (EF) {                                                                          \
                if (VF) KERNEL<true, true, false><<<GRID, kBlock, 0, ST>>>(ARGS);              \
                else KERNEL<true, false, false><<<GRID, kBlock, 0, ST>>>(ARGS);                \
            } else {                                                                           \
                if (VF) KERNEL<false, true, false><<<GRID, kBlock, 0, ST>>>(ARGS);             \
                else KERNEL<false, false, false><<<GRID, kBlock, 0, ST>>>(ARGS);               \
            }                                                                                  \
        }                                                                                      \
    } while (0)

// Atom ranges of the privatised spread: the bricks of the current list (slot
// order is brick order until the next rebuild), else chunks of kSpreadChunk
// consecutive slots.
struct SpreadRanges {
    const int2* ranges;
    int n;
    int cap; // LDS box points per workgroup
};

// Grid points of the box a brick's atoms touch: per axis the brick's extent
// plus the skin (atoms move at most skin/2 either way between checked
// rebuilds) in grid spacings, rounded up, plus the order's n points and one
// for the floor of the lowest atom. A range whose box is larger at run time
// (unchecked rebuilds) takes the global path, so the bound only sizes the LDS.
int spread_box_cap(const atvh_system* s, const PmeState* p) {
    int64_t tot = 1;
    for (int d = 0; d < 3; ++d) {
        const double span = (s->bdim[d] * s->cs[d] + s->skin) * p->K[d] * p->ic[d];
        tot *= static_cast<int64_t>(std::ceil(span)) + p->order + 1;
    }
    return static_cast<int>(std::min<int64_t>(tot, kSpreadCap));
}

template <class T, int N>
void launch_spread_gather(bool spread, hipStream_t st, const PmeArgs& a, const T* real, const SpreadRanges& sr) {
    if (spread)
        k_pme_spread_lds<T, N><<<std::min(sr.n, kSpreadGridCap), kBlock, sr.cap * sizeof(unsigned long long), st>>>(
            a, sr.ranges, sr.n, sr.cap);
    else
        k_pme_gather<T, N><<<grid_capped(a.n), kBlock, 0, st>>>(a, real);
}

template <class T>
void dispatch_order(int order, bool spread, hipStream_t st, const PmeArgs& a, const T* real,
                    const SpreadRanges& sr) {
    switch (order) {
    case 3: launch_spread_gather<T, 3>(spread, st, a, real, sr); break;
    case 4: launch_spread_gather<T, 4>(spread, st, a, real, sr); break;
    case 5: launch_spread_gather<T, 5>(spread, st, a, real, sr); break;
    case 6: launch_spread_gather<T, 6>(spread, st, a, real, sr); break;
    case 7: launch_spread_gather<T, 7>(spread, st, a, real, sr); break;
    case 8: launch_spread_gather<T, 8>(spread, st, a, real, sr); break;
    default: throw Error(ATVH_E_ARG, "KSpacePME: order must be in [3, 8]");
    }
}

#define HIPFFT_CHECK(call)                                                                 \
    do {                                                                                   \
        hipfftResult r_ = (call);                                                          \
        if (r_ != HIPFFT_SUCCESS)                                                          \
            throw Error(ATVH_E_HIP, std::string(#call) + ": hipFFT error " + std::to_string(int(r_))); \
    } while (0)

// Host restatements of the mesh helpers (pme_spline.h:63-126).
int good_fft_size(int n) {
    auto ok = [](int m) {
        for (int p : {2, 3, 5, 7})
            while (m % p == 0)
                m /= p;
        return m == 1;
    };
    if (n < 4)
        n = 4;
    while (!ok(n))
        ++n;
    return n;
}

// Cardinal B-spline weights of order N at the grid points (frac = 0), from
// the same template the kernels evaluate, so the table and the kernels agree.
template <int N>
std::vector<double> grid_point_weights() {
    double th[N], dth[N];
    bspline_t<double, N>(0.0, th, dth);
    return std::vector<double>(th, th + N);
}

std::vector<double> grid_point_weights(int n) {
    switch (n) {
    case 3: return grid_point_weights<3>();
    case 4: return grid_point_weights<4>();
    case 5: return grid_point_weights<5>();
    case 6: return grid_point_weights<6>();
    case 7: return grid_point_weights<7>();
    case 8: return grid_point_weights<8>();
    default: throw Error(ATVH_E_ARG, "KSpacePME: spline order must be in [3, 8]");
    }
}

// Euler exponential-spline factors 1/|b(m)|^2 of one axis: b(m) is the
// discrete Fourier sum of the grid-point weights at mode m. A mode where the
// sum vanishes (|b|^2 <= 1e-12, e.g. m = K/2 for odd orders on even K) takes
// the mean of its two neighbours, the reference's repair
// (include/atomiverse/cpu/longrange/pme_spline.h:32-75).
std::vector<double> spline_moduli(int K, int n) {
    const std::vector<double> w = grid_point_weights(n);
    std::vector<double> inv_b2(K);
    for (int m = 0; m < K; ++m) {
        std::complex<double> b(0.0, 0.0);
        for (int k = 0; k + 1 < n; ++k) // the last weight is zero at frac = 0
            b += w[k] * std::polar(1.0, 2.0 * kPi * m * k / K);
        const double b2 = std::norm(b);
        inv_b2[m] = b2 > 1e-12 ? 1.0 / b2 : 0.0;
    }
    std::vector<double> out(inv_b2);
    for (int m = 0; m < K; ++m)
        if (inv_b2[m] == 0.0)
            out[m] = 0.5 * (out[(m + K - 1) % K] + out[(m + 1) % K]);
    return out;
}

void pme_release(PmeState* p) {
    if (!p)
        return;
    if (p->f64) hipfftDestroy(p->f64);
    if (p->b64) hipfftDestroy(p->b64);
    if (p->f32) hipfftDestroy(p->f32);
    if (p->b32) hipfftDestroy(p->b32);
    dev_free(p->fix);
    dev_free(p->real64);
    dev_free(p->cplx64);
    dev_free(p->real32);
    dev_free(p->cplx32);
    dev_free(p->H);
    dev_free(p->partials);
    for (int d = 0; d < 3; ++d)
        dev_free(p->twf[d]);
    delete p;
}

void coul_release(CoulState* c) {
    if (!c)
        return;
    dev_free(c->tab);
    dev_free(c->tab32);
    dev_free(c->erfc);
    dev_free(c->erfc32);
    delete c;
}

// Largest device footprint of one forward + backward plan pair seen in this
// process (admission estimates read it).
std::atomic<int64_t> g_fft_pair_bytes{0};

void note_fft_pair_bytes(int64_t v) {
    int64_t cur = g_fft_pair_bytes.load();
    while (v > cur && !g_fft_pair_bytes.compare_exchange_weak(cur, v)) {
    }
}

PmeArgs make_pme_args(atvh_system* s) {
    PmeState* p = s->pme;
    PmeArgs a{};
    a.n = s->n;
    a.x = s->x;
    a.y = s->y;
    a.z = s->z;
    a.q = s->q;
    a.fx = s->fx;
    a.fy = s->fy;
    a.fz = s->fz;
    a.K0 = p->K[0];
    a.K1 = p->K[1];
    a.K2 = p->K[2];
    a.Kh = p->Kh;
    a.ic0 = p->ic[0];
    a.ic1 = p->ic[1];
    a.ic2 = p->ic[2];
    // fixed-point scale: a grid point holds at most sum|q_i| (weights <= 1),
    // kept below 2^61; one weight at most max|q_i|, kept below 2^50 (the
    // range of the spread's single-FMA rounding)
    const double qabs = std::max(s->q_abs, 1.0);
    int fix_exp = 61 - static_cast<int>(std::ceil(std::log2(qabs)));
    if (s->q_max > 0.0)
        fix_exp = std::min(fix_exp, 50 - static_cast<int>(std::ceil(std::log2(s->q_max))));
    p->fix_scale = std::ldexp(1.0, fix_exp);
    a.fix_scale = p->fix_scale;
    a.inv_fix = 1.0 / p->fix_scale;
    a.fix = p->fix;
    a.fscale = coulomb / (kPi * p->vol);
    a.pi2_a2 = kPi * kPi / (p->alpha * p->alpha);
    a.H = p->H;
    a.partials = p->partials;
    return a;
}

} // namespace

namespace atvh {

void ljc_enqueue(atvh_system* s, int precision, bool eflag, bool vflag) {
    CoulState* c = s->coul;
    if (s->ntypes > c->ntypes)
        throw Error(ATVH_E_ARG, "PairLJCutCoulLong: the frame has atom types beyond the coefficient table");
    LjcArgs a{};
    a.x = s->x;
    a.y = s->y;
    a.z = s->z;
    a.q = s->q;
    a.type = s->type;
    a.halo_rows = s->halo_rows;
    a.halo_n = s->halo_n;
    a.brick_range = s->brick_range;
    a.nbr16 = s->nbr16;
    a.self16 = s->self16;
    a.numneigh = s->numneigh;
    a.cap = s->cap;
    a.max_atoms = brick_atom_limit(s);
    for (int d = 0; d < 3; ++d) {
        a.L[d] = s->box[d];
        a.cs[d] = s->cs[d];
        a.nc[d] = s->nc[d];
        a.nb[d] = s->nb[d];
        a.bdim[d] = s->bdim[d];
        a.il[d] = 1.0 / s->box[d];
    }
    a.rc_sq = c->cutoff * c->cutoff;
    a.ntypes = c->ntypes;
    a.tab = c->tab;
    a.tab32 = c->tab32;
    a.erfc = c->erfc;
    a.erfc32 = c->erfc32;
    a.nbin = c->nbin;
    a.inv_dr2 = c->inv_dr2;
    a.alpha = c->alpha;
    a.fx = s->fx;
    a.fy = s->fy;
    a.fz = s->fz;
    a.partials = s->partials;
    a.nbr = s->nbr;
    a.npad = s->npad;
    a.n = s->n;
    hipStream_t st = s->ctx->stream;
    const int g = pair_grid(s);
    if (s->brick_mode) {
        if (precision == ATVH_PREC_MIXED)
            if (c->ntypes <= kLjcLdsTypes)
                launch_ljc_mixed<true>(eflag, vflag, g, st, a);
            else
                launch_ljc_mixed<false>(eflag, vflag, g, st, a);
        else
            if (c->ntypes <= kLjcLdsTypes)
                launch_ljc_fp64<true>(eflag, vflag, g, st, a);
            else
                launch_ljc_fp64<false>(eflag, vflag, g, st, a);
    } else {
        LJC_DISPATCH3(k_ljc_global, eflag, vflag, precision == ATVH_PREC_MIXED, g, st, a);
    }
    ATVH_HIP(hipGetLastError());
}

// The mesh transforms run as this library's kernels when every axis is a
// power of two the LDS lines and tiles hold (x and y: 16 to 256 points, z:
// 32 to 512); otherwise through the FFT library.
bool fft_native(const int K[3]) {
    auto pow2 = [](int k) { return k > 0 && (k & (k - 1)) == 0; };
    for (int d = 0; d < 3; ++d)
        if (!pow2(K[d]))
            return false;
    // the x and y tiles hold kFftTile columns: K2/2 and the outer axes are multiples of it
    const int maxk = 1 << pmefft::kFftMaxLog2, tile = pmefft::kFftTile;
    return K[0] >= tile && K[0] <= maxk && K[1] >= tile && K[1] <= maxk && K[2] >= 2 * tile && K[2] <= 2 * maxk;
}

// Whether an evaluation of this precision runs the native transforms.
bool pme_native(const PmeState* p, int precision) { return p->native && precision == ATVH_PREC_MIXED; }

// exp(-2 pi i k / L) for k = 0 .. L/2, from long double sin and cos.
template <class T2, class T>
T2* fft_twiddles(int L) {
    std::vector<T2> h(L / 2 + 1);
    const long double two_pi = 2.0L * std::acos(-1.0L);
    for (int k = 0; k <= L / 2; ++k) {
        const long double ang = two_pi * static_cast<long double>(k) / static_cast<long double>(L);
        h[k].x = static_cast<T>(std::cos(ang));
        h[k].y = static_cast<T>(-std::sin(ang));
    }
    T2* d = nullptr;
    dev_alloc(&d, h.size());
    ATVH_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T2), hipMemcpyHostToDevice));
    return d;
}

int ilog2(int k) {
    int l = 0;
    while ((1 << l) < k)
        ++l;
    return l;
}

#define ATVH_FFT_LOG2_SWITCH(v, CALL)                                                           \
    switch (v) {                                                                                \
    case 3: CALL(3); break;                                                                     \
    case 4: CALL(4); break;                                                                     \
    case 5: CALL(5); break;                                                                     \
    case 6: CALL(6); break;                                                                     \
    case 7: CALL(7); break;                                                                     \
    case 8: CALL(8); break;                                                                     \
    default: throw Error(ATVH_E_ARG, "KSpacePME: mesh axis outside the native transform sizes"); \
    }

// Passes A-E of pme_fft.hpp on the stream: fixed-point mesh in, real
// potential mesh out, influence applied, energy/virial partials (ev) in
// gm blocks.
template <class T>
void fft_native_enqueue(atvh_system* s, const PmeArgs& pa, bool ev, int gm, hipStream_t st) {
    using namespace pmefft;
    PmeState* p = s->pme;
    const bool f32 = std::is_same<T, float>::value;
    FftArgs f{};
    f.K0 = p->K[0];
    f.K1 = p->K[1];
    f.K2 = p->K[2];
    f.Kh = p->Kh;
    f.ks = fft_row(p->Kh);
    f.inv_fix = pa.inv_fix;
    f.fix = p->fix;
    f.real = f32 ? static_cast<void*>(p->real32) : static_cast<void*>(p->real64);
    f.cplx = f32 ? static_cast<void*>(p->cplx32) : static_cast<void*>(p->cplx64);
    static_assert(std::is_same<T, float>::value, "the native transforms run in FP32");
    f.tw0 = p->twf[0];
    f.tw1 = p->twf[1];
    f.tw2 = p->twf[2];
    f.H = p->H;
    f.ic0 = pa.ic0;
    f.ic1 = pa.ic1;
    f.ic2 = pa.ic2;
    f.pi2_a2 = pa.pi2_a2;
    f.partials = p->partials;
    const int l0 = ilog2(p->K[0]), l1 = ilog2(p->K[1]), lm = ilog2(p->K[2]) - 1;
    // passes A and E: lines of M = K2/2 points, min(64, M/2) lanes each
    const int M = p->K[2] / 2, nt = std::min(64, M / 2), lines_per_wg = (kFftThreads / 64) * (64 / nt);
    const int nlines = p->K[0] * p->K[1];
    const int ga = std::min((nlines + lines_per_wg - 1) / lines_per_wg, kFftLineGrid);
    // passes B and D: in FP32, 512 workgroups (two per CU) stride over the
    // tiles, each loading its next tile while the current one transforms; in
    // FP64 one workgroup per tile (see kFftPrefetch)
    const int ntiles_y = p->K[0] * ((p->Kh - 1) / kFftTile) + p->K[0] / kFftTile;
    const int gb = std::min(ntiles_y, f32 ? kFftTileGrid : kFftLineGrid);
#define ATVH_FFT_A(L) k_fft_r2c_z<T, L><<<ga, kFftThreads, 0, st>>>(f)
#define ATVH_FFT_B(L) k_fft_y<T, L, false><<<gb, kFftTileThreads, 0, st>>>(f)
#define ATVH_FFT_C(L)                                                    \
    do {                                                                 \
        if (ev)                                                          \
            k_fft_x_influence<T, L, true><<<gm, kFftTileThreads, 0, st>>>(f); \
        else                                                             \
            k_fft_x_influence<T, L, false><<<gm, kFftTileThreads, 0, st>>>(f); \
    } while (0)
#define ATVH_FFT_D(L) k_fft_y<T, L, true><<<gb, kFftTileThreads, 0, st>>>(f)
#define ATVH_FFT_E(L) k_fft_c2r_z<T, L><<<ga, kFftThreads, 0, st>>>(f)
    ATVH_FFT_LOG2_SWITCH(lm, ATVH_FFT_A)
    ATVH_FFT_LOG2_SWITCH(l1, ATVH_FFT_B)
    ATVH_FFT_LOG2_SWITCH(l0, ATVH_FFT_C)
    ATVH_FFT_LOG2_SWITCH(l1, ATVH_FFT_D)
    ATVH_FFT_LOG2_SWITCH(lm, ATVH_FFT_E)
#undef ATVH_FFT_A
#undef ATVH_FFT_B
#undef ATVH_FFT_C
#undef ATVH_FFT_D
#undef ATVH_FFT_E
    ATVH_HIP(hipGetLastError());
}

// Blocks (and fixed-order energy partials) of the influence kernel.
int pme_grid_blocks(const atvh_system* s, int precision) {
    const PmeState* p = s->pme;
    if (pme_native(p, precision)) // pass C: one workgroup per tile, grid-stride beyond the partials' cap
        return std::min(p->K[1] * ((p->Kh - 1) / pmefft::kFftTile) + p->K[1] / pmefft::kFftTile, kPmeGridCap);
    const int64_t nmodes = (int64_t)p->K[0] * p->K[1] * p->Kh;
    return std::min(grid_for(nmodes), kPmeGridCap);
}

// The mesh buffers and transform plans of one precision.
void pme_make_buffers(atvh_system* s, int precision) {
    PmeState* p = s->pme;
    const int64_t ntot = (int64_t)p->K[0] * p->K[1] * p->K[2];
    const int64_t nmodes = (int64_t)p->K[0] * p->K[1] * p->Kh;
    if (pme_native(p, precision)) {
        if (!p->real32) {
            dev_alloc(&p->real32, ntot);
            dev_alloc(&p->cplx32, (int64_t)p->K[0] * p->K[1] * pmefft::fft_row(p->Kh));
            for (int d = 0; d < 3; ++d)
                p->twf[d] = fft_twiddles<float2, float>(p->K[d]);
        }
        return;
    }
    if (precision == ATVH_PREC_MIXED && !p->real32) {
        dev_alloc(&p->real32, ntot);
        dev_alloc(&p->cplx32, nmodes);
        HIPFFT_CHECK(hipfftPlan3d(&p->f32, p->K[0], p->K[1], p->K[2], HIPFFT_R2C));
        HIPFFT_CHECK(hipfftPlan3d(&p->b32, p->K[0], p->K[1], p->K[2], HIPFFT_C2R));
    }
    if (precision != ATVH_PREC_MIXED && !p->real64) {
        dev_alloc(&p->real64, ntot);
        dev_alloc(&p->cplx64, nmodes);
        HIPFFT_CHECK(hipfftPlan3d(&p->f64, p->K[0], p->K[1], p->K[2], HIPFFT_D2Z));
        HIPFFT_CHECK(hipfftPlan3d(&p->b64, p->K[0], p->K[1], p->K[2], HIPFFT_Z2D));
    }
}

// Everything a PME evaluation of this precision would allocate, made now:
// buffers, plans, and one execution of each transform on the scratch mesh
// (the FFT library loads its kernels and work areas at the first one), so
// that later evaluations can be captured into a graph. Leaves the forces
// untouched. Synchronises.
void pme_prepare(atvh_system* s, int precision) {
    PmeState* p = s->pme;
    if (!p)
        return;
    // the energy partials an evaluation of this precision leaves (the count
    // differs between the native and the library transforms)
    p->nb = pme_grid_blocks(s, precision);
    if (pme_ready(s, precision))
        return;
    if (pme_native(p, precision)) { // nothing to load: the transforms are this library's kernels
        pme_make_buffers(s, precision);
        p->nb = pme_grid_blocks(s, precision);
        return;
    }
    size_t free_before = 0, free_after = 0, total_mem = 0;
    ATVH_HIP(hipMemGetInfo(&free_before, &total_mem));
    pme_make_buffers(s, precision);
    hipStream_t st = s->ctx->stream;
    if (precision == ATVH_PREC_MIXED) {
        HIPFFT_CHECK(hipfftSetStream(p->f32, st));
        HIPFFT_CHECK(hipfftExecR2C(p->f32, p->real32, reinterpret_cast<hipfftComplex*>(p->cplx32)));
        HIPFFT_CHECK(hipfftSetStream(p->b32, st));
        HIPFFT_CHECK(hipfftExecC2R(p->b32, reinterpret_cast<hipfftComplex*>(p->cplx32), p->real32));
    } else {
        HIPFFT_CHECK(hipfftSetStream(p->f64, st));
        HIPFFT_CHECK(hipfftExecD2Z(p->f64, p->real64, reinterpret_cast<hipfftDoubleComplex*>(p->cplx64)));
        HIPFFT_CHECK(hipfftSetStream(p->b64, st));
        HIPFFT_CHECK(hipfftExecZ2D(p->b64, reinterpret_cast<hipfftDoubleComplex*>(p->cplx64), p->real64));
    }
    ATVH_HIP(hipStreamSynchronize(st));
    ATVH_HIP(hipMemGetInfo(&free_after, &total_mem));
    if (free_before > free_after)
        note_fft_pair_bytes(static_cast<int64_t>(free_before - free_after));
    p->nb = pme_grid_blocks(s, precision);
}

void pme_enqueue(atvh_system* s, int precision, bool ev) {
    PmeState* p = s->pme;
    if (!s->has_charge)
        throw Error(ATVH_E_STATE, "KSpacePME: the frame carries no charges");
    ScopedTimer tm(s->ctx, "pme");
    hipStream_t st = s->ctx->stream;
    const bool mixed = precision == ATVH_PREC_MIXED;
    const int64_t ntot = (int64_t)p->K[0] * p->K[1] * p->K[2];
    // buffers and plans of this precision, made on first use; the device
    // memory the FFT library takes for a plan pair (work areas, tables, the
    // kernels it loads at the first execution) is measured once it has run
    size_t free_before = 0, total_mem = 0;
    const bool fresh = mixed ? !p->real32 : !p->real64;
    if (fresh) {
        ATVH_HIP(hipMemGetInfo(&free_before, &total_mem));
        pme_make_buffers(s, precision);
    }
    PmeArgs a = make_pme_args(s);
    SpreadRanges sr;
    if (s->brick_mode && s->nl_built && s->brick_range) {
        sr.ranges = s->brick_range;
        sr.n = s->nbricks;
        sr.cap = spread_box_cap(s, p);
    } else {
        sr.ranges = nullptr;
        sr.n = static_cast<int>((s->n + kSpreadChunk - 1) / kSpreadChunk);
        sr.cap = kSpreadCap;
    }
    const int gg = grid_capped(ntot);
    const int gm = pme_grid_blocks(s, precision);
    const bool native = pme_native(p, precision);
    if (native) {
        dispatch_order<float>(p->order, true, st, a, nullptr, sr);
        fft_native_enqueue<float>(s, a, ev, gm, st);
        dispatch_order<float>(p->order, false, st, a, p->real32, sr);
    } else if (mixed) {
        dispatch_order<float>(p->order, true, st, a, nullptr, sr);
        k_pme_unfix<float><<<gg, kBlock, 0, st>>>(ntot, p->fix, p->real32, a.inv_fix);
        HIPFFT_CHECK(hipfftSetStream(p->f32, st));
        HIPFFT_CHECK(hipfftExecR2C(p->f32, p->real32, reinterpret_cast<hipfftComplex*>(p->cplx32)));
        if (ev)
            k_pme_influence<float2, true><<<gm, kBlock, 0, st>>>(a, p->cplx32);
        else
            k_pme_influence<float2, false><<<gm, kBlock, 0, st>>>(a, p->cplx32);
        HIPFFT_CHECK(hipfftSetStream(p->b32, st));
        HIPFFT_CHECK(hipfftExecC2R(p->b32, reinterpret_cast<hipfftComplex*>(p->cplx32), p->real32));
        dispatch_order<float>(p->order, false, st, a, p->real32, sr);
    } else {
        dispatch_order<double>(p->order, true, st, a, nullptr, sr);
        k_pme_unfix<double><<<gg, kBlock, 0, st>>>(ntot, p->fix, p->real64, a.inv_fix);
        HIPFFT_CHECK(hipfftSetStream(p->f64, st));
        HIPFFT_CHECK(hipfftExecD2Z(p->f64, p->real64, reinterpret_cast<hipfftDoubleComplex*>(p->cplx64)));
        if (ev)
            k_pme_influence<double2, true><<<gm, kBlock, 0, st>>>(a, p->cplx64);
        else
            k_pme_influence<double2, false><<<gm, kBlock, 0, st>>>(a, p->cplx64);
        HIPFFT_CHECK(hipfftSetStream(p->b64, st));
        HIPFFT_CHECK(hipfftExecZ2D(p->b64, reinterpret_cast<hipfftDoubleComplex*>(p->cplx64), p->real64));
        dispatch_order<double>(p->order, false, st, a, p->real64, sr);
    }
    if (fresh && !native) {
        size_t free_after = 0;
        ATVH_HIP(hipMemGetInfo(&free_after, &total_mem));
        if (free_before > free_after)
            note_fft_pair_bytes(static_cast<int64_t>(free_before - free_after));
    }
    p->nb = gm;
    ATVH_HIP(hipGetLastError());
}

int pme_partial_blocks(const atvh_system* s) { return s->pme ? s->pme->nb : 0; }
int64_t coul_device_bytes(const atvh_system* s) {
    const CoulState* c = s->coul;
    if (!c)
        return 0;
    const int64_t t2 = int64_t(c->ntypes) * c->ntypes;
    return t2 * 8 * (8 + 4) + int64_t(c->nbin) * (32 + 16);
}
int64_t pme_device_bytes(const atvh_system* s) {
    const PmeState* p = s->pme;
    if (!p)
        return 0;
    const int64_t ntot = int64_t(p->K[0]) * p->K[1] * p->K[2];
    const int64_t nmodes = int64_t(p->K[0]) * p->K[1] * p->Kh;
    int64_t b = ntot * 8 + nmodes * 8 + int64_t(kPmeGridCap) * kPartialStride * 8; // fix, H, partials
    if (p->real64)
        b += ntot * 8 + nmodes * 16 + fft_pair_bytes();
    if (p->real32) {
        if (p->native) // padded rows and three twiddle tables, no plans
            b += ntot * 4 + int64_t(p->K[0]) * p->K[1] * pmefft::fft_row(p->Kh) * 8 +
                 int64_t(p->K[0] + p->K[1] + p->K[2] + 3) * 8;
        else
            b += ntot * 4 + nmodes * 8 + fft_pair_bytes();
    }
    return b;
}
bool pme_ready(const atvh_system* s, int precision) {
    if (!s->pme)
        return true;
    return precision == ATVH_PREC_MIXED ? s->pme->real32 != nullptr : s->pme->real64 != nullptr;
}
const double* pme_partials(const atvh_system* s) { return s->pme ? s->pme->partials : nullptr; }

// escale = k_e / (2 pi V) (pme_influence.h:215); self and net-charge energies
// (pme.cpp:222-246) from the charge sums of the uploaded frame.
int64_t fft_pair_bytes() {
    const int64_t v = g_fft_pair_bytes.load();
    return v > 0 ? v : (int64_t(64) << 20);
}

void pme_energy_consts(const atvh_system* s, double& escale, double& e_self, double& e_net) {
    const PmeState* p = s->pme;
    escale = coulomb / (2.0 * kPi * p->vol);
    e_self = -coulomb * p->alpha / std::sqrt(kPi) * s->q_sq;
    e_net = -coulomb * kPi / (2.0 * p->vol * p->alpha * p->alpha) * s->q_sum * s->q_sum;
}

void coul_free(atvh_system* s) {
    coul_release(s->coul);
    s->coul = nullptr;
    pme_release(s->pme);
    s->pme = nullptr;
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

int atvh_pair_ljcoul_configure(atvh_system* s, int ntypes, const double* eps_tab, const double* sigma_tab,
                               const int* set_tab, double alpha, double cutoff, int nbin) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    // constructor checks (pair_lj_cut_coul_long.cpp:34-44)
    if (ntypes < 1 || ntypes > kMaxTypes || !eps_tab || !sigma_tab || !set_tab)
        throw Error(ATVH_E_ARG, "PairLJCutCoulLong: empty or oversized coeff table");
    if (!(alpha > 0.0))
        throw Error(ATVH_E_ARG, "PairLJCutCoulLong: alpha must be > 0");
    if (!(cutoff > 0.0))
        throw Error(ATVH_E_ARG, "PairLJCutCoulLong: cutoff must be > 0");
    if (nbin < 16)
        throw Error(ATVH_E_ARG, "ErfcTable::build: bad alpha/cut/nbin");
    for (int ab = 0; ab < ntypes * ntypes; ++ab)
        if (set_tab[ab] && (!(sigma_tab[ab] > 0.0) || !(eps_tab[ab] >= 0.0)))
            throw Error(ATVH_E_ARG, "PairLJCutCoulLong: bad epsilon/sigma");
    ATVH_HIP(hipSetDevice(c->device));
    auto* cs = new CoulState();
    try {
        cs->ntypes = ntypes;
        cs->alpha = alpha;
        cs->cutoff = cutoff;
        cs->nbin = nbin;
        // finalize_type_table_ (pair_lj_cut_coul_long.cpp:132-186)
        std::vector<double> eps(eps_tab, eps_tab + ntypes * ntypes), sig(sigma_tab, sigma_tab + ntypes * ntypes);
        std::vector<double> tab(size_t(8) * ntypes * ntypes, 0.0);
        std::vector<float> tab32(tab.size(), 0.0f);
        for (int ta = 0; ta < ntypes; ++ta)
            for (int tb = 0; tb < ntypes; ++tb) {
                const int ab = ta * ntypes + tb;
                double e = eps[ab], sg = sig[ab];
                if (!set_tab[ab]) {
                    const int aa = ta * ntypes + ta, bb = tb * ntypes + tb;
                    if (set_tab[aa] && set_tab[bb]) {
                        e = std::sqrt(eps_tab[aa] * eps_tab[bb]);
                        sg = 0.5 * (sigma_tab[aa] + sigma_tab[bb]);
                    } else {
                        e = 0.0;
                        sg = 1.0;
                    }
                    eps[ab] = e;
                    sig[ab] = sg;
                }
                const double s2 = sg * sg;
                const double s6 = s2 * s2 * s2;
                const double s12 = s6 * s6;
                double* t = &tab[size_t(8) * ab];
                t[0] = 48.0 * e * s12;
                t[1] = 24.0 * e * s6;
                t[2] = 4.0 * e * s12;
                t[3] = 4.0 * e * s6;
                if (e > 0.0) {
                    const double rc2 = cutoff * cutoff;
                    const double r6 = rc2 * rc2 * rc2;
                    const double r6inv = 1.0 / r6;
                    t[4] = t[2] * r6inv * r6inv - t[3] * r6inv;
                }
                for (int k = 0; k < 8; ++k)
                    tab32[size_t(8) * ab + k] = static_cast<float>(t[k]);
            }
        cs->eps_mixed = eps;
        cs->sigma_mixed = sig;
        // ErfcTable::build (erfc_table.h:44-72)
        const double cutoff_sq = cutoff * cutoff;
        const double dr2 = cutoff_sq / static_cast<double>(nbin);
        cs->inv_dr2 = 1.0 / dr2;
        std::vector<double> et(nbin + 1), ft(nbin + 1);
        for (int i = 1; i <= nbin; ++i) {
            const double r2 = i * dr2;
            const double r = std::sqrt(r2);
            const double grij = alpha * r;
            const double expm2 = std::exp(-grij * grij);
            const double erfcv = std::erfc(grij);
            et[i] = erfcv / r;
            ft[i] = (erfcv + kEwaldF * grij * expm2) / (r2 * r);
        }
        et[0] = et[1];
        ft[0] = ft[1];
        std::vector<double4> e4(nbin);
        std::vector<float4> e4f(nbin);
        for (int i = 0; i < nbin; ++i) {
            e4[i] = make_double4(ft[i], ft[i + 1] - ft[i], et[i], et[i + 1] - et[i]);
            e4f[i] = make_float4(static_cast<float>(e4[i].x), static_cast<float>(e4[i].y),
                                 static_cast<float>(e4[i].z), static_cast<float>(e4[i].w));
        }
        dev_alloc(&cs->tab, tab.size());
        dev_alloc(&cs->tab32, tab32.size());
        dev_alloc(&cs->erfc, e4.size());
        dev_alloc(&cs->erfc32, e4f.size());
        ATVH_HIP(hipMemcpy(cs->tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
        ATVH_HIP(hipMemcpy(cs->tab32, tab32.data(), tab32.size() * sizeof(float), hipMemcpyHostToDevice));
        ATVH_HIP(hipMemcpy(cs->erfc, e4.data(), e4.size() * sizeof(double4), hipMemcpyHostToDevice));
        ATVH_HIP(hipMemcpy(cs->erfc32, e4f.data(), e4f.size() * sizeof(float4), hipMemcpyHostToDevice));
    } catch (...) {
        coul_release(cs);
        throw;
    }
    coul_release(s->coul);
    s->coul = cs;
    s->pair_style = 1;
    s->lj_cutoff = cutoff;
    s->pair_configured = true;
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_pme_configure(atvh_system* s, double alpha, double spacing, int order) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    // constructor checks (pme.cpp:31-39), plus this backend's order limit
    if (!(alpha > 0.0))
        throw Error(ATVH_E_ARG, "KSpacePME: alpha must be > 0");
    if (!(spacing > 0.0))
        throw Error(ATVH_E_ARG, "KSpacePME: grid_spacing must be > 0");
    if (order < 3)
        throw Error(ATVH_E_ARG, "KSpacePME: order must be >= 3");
    if (order > kPmeMaxOrder)
        throw Error(ATVH_E_ARG, "KSpacePME: this backend supports spline orders up to 8");
    ATVH_HIP(hipSetDevice(c->device));
    auto* p = new PmeState();
    try {
        p->alpha = alpha;
        p->spacing = spacing;
        p->order = order;
        // ensure_grid (pme.cpp:66-74)
        for (int d = 0; d < 3; ++d) {
            p->K[d] = good_fft_size(static_cast<int>(std::ceil(s->box[d] / spacing)));
            if (p->K[d] < order + 1)
                p->K[d] = good_fft_size(order + 1);
            p->ic[d] = 1.0 / s->box[d];
        }
        p->Kh = p->K[2] / 2 + 1;
        p->native = fft_native(p->K);
        p->vol = s->box[0] * s->box[1] * s->box[2];
        const int64_t ntot = (int64_t)p->K[0] * p->K[1] * p->K[2];
        const int64_t nmodes = (int64_t)p->K[0] * p->K[1] * p->Kh;
        // pme_build_influence_tables (pme_influence.h:115-175) on the half spectrum
        std::vector<double> mod[3];
        for (int d = 0; d < 3; ++d)
            mod[d] = spline_moduli(p->K[d], order);
        const double inv_pi2_a2 = kPi * kPi / (alpha * alpha);
        std::vector<double> H(nmodes, 0.0);
        for (int m0 = 0; m0 < p->K[0]; ++m0) {
            const int h0 = (m0 <= p->K[0] / 2) ? m0 : m0 - p->K[0];
            const double bm0 = mod[0][m0];
            for (int m1 = 0; m1 < p->K[1]; ++m1) {
                const int h1 = (m1 <= p->K[1] / 2) ? m1 : m1 - p->K[1];
                const double bm01 = bm0 * mod[1][m1];
                for (int m2 = 0; m2 < p->Kh; ++m2) {
                    if (m0 == 0 && m1 == 0 && m2 == 0)
                        continue;
                    const int h2 = (m2 <= p->K[2] / 2) ? m2 : m2 - p->K[2];
                    const double mx = h0 * p->ic[0], my = h1 * p->ic[1], mz = h2 * p->ic[2];
                    const double msq = mx * mx + my * my + mz * mz;
                    const double inv_msq = 1.0 / msq;
                    const double B = bm01 * mod[2][m2];
                    H[m2 + (int64_t)p->Kh * (m1 + (int64_t)p->K[1] * m0)] =
                        std::exp(-inv_pi2_a2 * msq) * inv_msq * B;
                }
            }
        }
        dev_alloc(&p->H, H.size());
        ATVH_HIP(hipMemcpy(p->H, H.data(), H.size() * sizeof(double), hipMemcpyHostToDevice));
        dev_alloc(&p->fix, ntot);
        ATVH_HIP(hipMemset(p->fix, 0, ntot * sizeof(unsigned long long)));
        dev_alloc(&p->partials, size_t(kPmeGridCap) * kPartialStride);
        ATVH_HIP(hipMemset(p->partials, 0, size_t(kPmeGridCap) * kPartialStride * sizeof(double)));
    } catch (...) {
        pme_release(p);
        throw;
    }
    pme_release(s->pme);
    s->pme = p;
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_pme_info(atvh_system* s, int* K, double* e_parts) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (!s->pme)
        throw Error(ATVH_E_STATE, "KSpacePME not configured");
    if (K)
        for (int d = 0; d < 3; ++d)
            K[d] = s->pme->K[d];
    if (e_parts) {
        ATVH_HIP(hipStreamSynchronize(c->stream));
        DevCtl h;
        ATVH_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
        for (int k = 0; k < 3; ++k)
            e_parts[k] = h.e_kspace[k];
    }
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_pme_compute(atvh_system* s, int precision, int eflag, int vflag) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (precision != ATVH_PREC_FP64 && precision != ATVH_PREC_MIXED)
        throw Error(ATVH_E_ARG, "atvh_pme_compute: unknown precision");
    if (!s->pme)
        throw Error(ATVH_E_STATE, "KSpacePME not configured");
    ATVH_HIP(hipSetDevice(c->device));
    k_zero3<<<grid_capped(s->n), kBlock, 0, c->stream>>>(s->n, s->fx, s->fy, s->fz);
    ATVH_HIP(hipGetLastError());
    const bool ev = eflag || vflag;
    pme_enqueue(s, precision, ev);
    if (ev)
        energy_enqueue_finalize(s, nullptr, 0, false);
    return ATVH_OK;
    API_CATCH(c)
}

} // extern "C"
