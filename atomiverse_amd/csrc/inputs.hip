// inputs.hip — host-side generators of the synthetic inputs the workloads are
// defined on (lattices, Maxwell-Boltzmann velocities). Pure host C++.
//
// Velocities follow InitVelocity::build (src/cpu/dynamics/init.cpp:107-215):
// std::mt19937(seed) + std::normal_distribution<double>(0,1) drawn vx,vy,vz
// per atom and scaled by 1/sqrt(m), centre-of-mass velocity removed, then a
// rescale to the exact temperature with Ndof degrees of freedom.

#include "atvh_internal.hpp"

#include <cmath>
#include <random>

namespace {

// Plain streaming copy (16-byte vector loads and stores): the reference
// kernel for calibrating HBM byte counters against a known traffic.
__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     int64_t n16) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

} // namespace

extern "C" {

int atvh_util_stream_copy(atvh_ctx* c, int64_t bytes, int iters, double* avg_ms) {
    if (!c || bytes < 16 || iters < 1)
        return ATVH_E_ARG;
    try {
        ATVH_HIP(hipSetDevice(c->device));
        const int64_t n16 = bytes / 16;
        uint4 *src = nullptr, *dst = nullptr;
        atvh::dev_alloc(&src, n16);
        atvh::dev_alloc(&dst, n16);
        hipStream_t st = c->stream;
        ATVH_HIP(hipMemsetAsync(src, 1, n16 * 16, st));
        hipEvent_t e0, e1;
        ATVH_HIP(hipEventCreate(&e0));
        ATVH_HIP(hipEventCreate(&e1));
        const int grid = 256 * 16;
        k_stream_copy<<<grid, 256, 0, st>>>(src, dst, n16); // warm
        ATVH_HIP(hipEventRecord(e0, st));
        for (int k = 0; k < iters; ++k)
            k_stream_copy<<<grid, 256, 0, st>>>(src, dst, n16);
        ATVH_HIP(hipEventRecord(e1, st));
        ATVH_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        ATVH_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        atvh::dev_free(src);
        atvh::dev_free(dst);
        if (avg_ms)
            *avg_ms = double(ms) / iters;
        return ATVH_OK;
    } catch (const atvh::Error& e) {
        c->err = e.what();
        return e.code;
    }
}

int64_t atvh_util_fcc_lattice(int nx, int ny, int nz, double a, int64_t nmax, double* x, double* y,
                              double* z) {
    static const double basis[4][3] = {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.0}, {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(a > 0.0))
        return ATVH_E_ARG;
    int64_t k = 0;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix)
                for (int b = 0; b < 4; ++b) {
                    if (nmax >= 0 && k >= nmax)
                        return k;
                    if (x) {
                        x[k] = (ix + basis[b][0]) * a;
                        y[k] = (iy + basis[b][1]) * a;
                        z[k] = (iz + basis[b][2]) * a;
                    }
                    ++k;
                }
    return k;
}

int64_t atvh_util_rocksalt_lattice(int nx, int ny, int nz, double spacing, double* x, double* y,
                                   double* z, int* type, double* charge, double q0) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(spacing > 0.0))
        return ATVH_E_ARG;
    int64_t k = 0;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                if (x) {
                    const int odd = (ix + iy + iz) & 1;
                    x[k] = ix * spacing;
                    y[k] = iy * spacing;
                    z[k] = iz * spacing;
                    if (type)
                        type[k] = odd;
                    if (charge)
                        charge[k] = odd ? -q0 : q0;
                }
                ++k;
            }
    return k;
}

int atvh_util_init_velocity(int64_t n, const double* mass, double T, uint32_t seed, int64_t ndof,
                            double* vx, double* vy, double* vz) {
    if (n <= 0 || !mass || !vx || !vy || !vz)
        return ATVH_E_ARG;
    std::mt19937 rng(seed);
    std::normal_distribution<double> draw(0.0, 1.0);
    for (int64_t i = 0; i < n; ++i) {
        const double s = 1.0 / std::sqrt(mass[i]);
        vx[i] = draw(rng) * s;
        vy[i] = draw(rng) * s;
        vz[i] = draw(rng) * s;
    }
    double px = 0, py = 0, pz = 0, mtot = 0;
    for (int64_t i = 0; i < n; ++i) {
        px += mass[i] * vx[i];
        py += mass[i] * vy[i];
        pz += mass[i] * vz[i];
        mtot += mass[i];
    }
    if (mtot > 0) {
        px /= mtot;
        py /= mtot;
        pz /= mtot;
        for (int64_t i = 0; i < n; ++i) {
            vx[i] -= px;
            vy[i] -= py;
            vz[i] -= pz;
        }
    }
    double mv2 = 0.0;
    for (int64_t i = 0; i < n; ++i)
        mv2 += mass[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    if (ndof > 0 && mv2 > 0.0) {
        const double ke = 0.5 * atvh::mass_v2_to_energy * mv2;
        const double t_now = 2.0 * ke / (static_cast<double>(ndof) * atvh::kB_ev_per_k);
        const double f = std::sqrt(T / t_now);
        for (int64_t i = 0; i < n; ++i) {
            vx[i] *= f;
            vy[i] *= f;
            vz[i] *= f;
        }
    }
    return ATVH_OK;
}

} // extern "C"
