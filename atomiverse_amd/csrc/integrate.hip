// integrate.hip — velocity Verlet + single-chain Nose-Hoover, kinetic energy,
// the displacement half of the rebuild check, and the fused MD loop.
//
// Reference order per step (Runtime::step_verlet_, src/cpu/driver/runtime.cpp:452-502):
//   initial_integrate: VelocityVerlet (src/cpu/dynamics/velocity_verlet.cpp:41-55)
//                      v += dtf/m f; x += dt v; wrap_orthorhombic (pbc.h:37-51)
//                      (here the wrap happens at the next rebuild, see k_initial)
//                      then NoseHoover (nose_hoover.cpp:104-170):
//                      KE = 0.5 mvv2e sum m v^2; xi += dt/2 (2KE - Ndof kB T)/Q;
//                      v *= exp(-dt/2 xi)
//   decide + rebuild, zero + force,
//   final_integrate:   v += dtf/m f, then NoseHoover again; end_of_step: KE.
//
// MI355X schedule: the NH velocity scale is applied lazily — it is stored in
// the control block and multiplied into v inside the next kick, which is the
// same operation order as the reference ((v*s) + dtfm*f), so no extra pass
// over v is spent on it. Each kick kernel also emits fixed-order per-block
// kinetic partials; a one-block kernel reduces them, advances xi and (after
// the initial kick) takes the neighbor-list decision, all on the device.
// The initial kernel also performs the needs_rebuild displacement test
// (cell_list.cpp:372-391): 4*|min_image(x - x_save)|^2 > skin^2 per atom.

#include "atvh_internal.hpp"
#include "batch.hpp"
#include "pair_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

using namespace atvh;

namespace {

constexpr int kIntGrid = 1024; // integrator grid cap: few kinetic partials to reduce

struct IntArgs {
    int64_t n;
    double *x, *y, *z, *vx, *vy, *vz;
    const double *fx, *fy, *fz;
    const int* type;
    const double *xs, *ys, *zs;
    double dtfm[kMaxTypes];
    double mass[kMaxTypes];
    int ntypes;
    double dt;
    double lx, ly, lz;
    double skin_sq;
    int thermostat;
    DevCtl* ctl;
    double* ke_partials;
};

__device__ __forceinline__ double wave_sum_d(double v) {
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ void block_sum_to(double v, double* out) {
    __shared__ double sh[kBlock / 64];
    v = wave_sum_d(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        sh[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w)
            t += sh[w];
        *out = t;
    }
}

// The kicks: one atom per grid-stride step, its loads issued before its
// stores. A one-type frame takes dtfm and mass from scalars (no dependent
// per-atom table loads). (Two atoms per step, both loaded before either is
// stored, needed 142 registers for the initial kick and lost on the batched
// ladder what it gained on the 2.048M-atom fluid.)
struct KickAtom {
    double vx, vy, vz, fx, fy, fz;
    int t;
};

__device__ __forceinline__ KickAtom kick_load(const IntArgs& a, int64_t i, bool typed) {
    KickAtom k;
    k.t = typed ? a.type[i] : 0;
    k.vx = a.vx[i];
    k.vy = a.vy[i];
    k.vz = a.vz[i];
    k.fx = a.fx[i];
    k.fy = a.fy[i];
    k.fz = a.fz[i];
    return k;
}

// v <- (v s) + dtfm f (the reference's order: the pending NH scale, then the
// kick); returns m v^2 of the new velocity
__device__ __forceinline__ double kick_apply(const IntArgs& a, KickAtom& k, double s, bool typed, double dtfm0,
                                             double m0) {
    double dtfm = dtfm0, m = m0;
    if (typed) {
        dtfm = a.dtfm[k.t];
        m = a.mass[k.t];
    }
    if (a.thermostat) {
        k.vx *= s;
        k.vy *= s;
        k.vz *= s;
    }
    k.vx += dtfm * k.fx;
    k.vy += dtfm * k.fy;
    k.vz += dtfm * k.fz;
    return m * (k.vx * k.vx + k.vy * k.vy + k.vz * k.vz);
}

struct OpInitial {
    using Args = IntArgs;
    struct Pos {
        double x, y, z, xs, ys, zs;
    };
    static __device__ __forceinline__ Pos load_pos(const IntArgs& a, int64_t i) {
        return Pos{a.x[i], a.y[i], a.z[i], a.xs[i], a.ys[i], a.zs[i]};
    }
    // kick, drift, store; the displacement test of the new position
    static __device__ __forceinline__ void finish(const IntArgs& a, int64_t i, KickAtom& k, const Pos& p, double s,
                                                  bool typed, double dtfm0, double m0, double& mv2, int& moved) {
        mv2 += kick_apply(a, k, s, typed, dtfm0, m0);
        // drift; the periodic wrap of wrap_orthorhombic is deferred to the next
        // neighbor rebuild (and to downloads), which keeps the image of every
        // listed pair fixed between rebuilds
        const double x = p.x + a.dt * k.vx;
        const double y = p.y + a.dt * k.vy;
        const double z = p.z + a.dt * k.vz;
        a.x[i] = x;
        a.y[i] = y;
        a.z[i] = z;
        a.vx[i] = k.vx;
        a.vy[i] = k.vy;
        a.vz[i] = k.vz;
        double dx = x - p.xs, dy = y - p.ys, dz = z - p.zs;
        dx -= a.lx * nearbyint(dx / a.lx);
        dy -= a.ly * nearbyint(dy / a.ly);
        dz -= a.lz * nearbyint(dz / a.lz);
        const double dsq = dx * dx + dy * dy + dz * dz;
        if (dsq * 4.0 > a.skin_sq)
            moved = 1;
    }
    static __device__ void run(const IntArgs& a, int bid, int nbk) {
        if (halted(a.ctl))
            return;
        const double s = a.thermostat ? a.ctl->nh_scale : 1.0;
        const bool typed = a.ntypes > 1;
        const double dtfm0 = a.dtfm[0], m0 = a.mass[0];
        double mv2 = 0.0;
        int moved = 0;
        for (int64_t i = bid * (int64_t)blockDim.x + threadIdx.x; i < a.n; i += (int64_t)nbk * blockDim.x) {
            KickAtom k0 = kick_load(a, i, typed);
            const Pos p0 = load_pos(a, i);
            finish(a, i, k0, p0, s, typed, dtfm0, m0, mv2, moved);
        }
        if (__any(moved) && (threadIdx.x & 63) == 0)
            a.ctl->disp_flag = 1;
        block_sum_to(mv2, &a.ke_partials[bid]);
    }
};

struct OpFinal {
    using Args = IntArgs;
    static __device__ __forceinline__ void finish(const IntArgs& a, int64_t i, KickAtom& k, double s, bool typed,
                                                  double dtfm0, double m0, double& mv2) {
        mv2 += kick_apply(a, k, s, typed, dtfm0, m0);
        a.vx[i] = k.vx;
        a.vy[i] = k.vy;
        a.vz[i] = k.vz;
    }
    static __device__ void run(const IntArgs& a, int bid, int nbk) {
        if (halted(a.ctl))
            return;
        const double s = a.thermostat ? a.ctl->nh_scale : 1.0;
        const bool typed = a.ntypes > 1;
        const double dtfm0 = a.dtfm[0], m0 = a.mass[0];
        double mv2 = 0.0;
        for (int64_t i = bid * (int64_t)blockDim.x + threadIdx.x; i < a.n; i += (int64_t)nbk * blockDim.x) {
            KickAtom k0 = kick_load(a, i, typed);
            finish(a, i, k0, s, typed, dtfm0, m0, mv2);
        }
        block_sum_to(mv2, &a.ke_partials[bid]);
    }
};

// The Virial observer's kinetic part (cpu::md::Virial::launch,
// src/cpu/observe/virial_stress.cpp:14-56): sum_i m_i v_ia v_ib over the
// current velocities (the pending Nose-Hoover scale applied), in the
// reference's Voigt order xx yy zz xy xz yz, as fixed-order per-block
// partials (slots 1..6 of the block's partial row).
struct OpKinStress {
    using Args = IntArgs;
    static __device__ void run(const IntArgs& a, int bid, int nbk) {
        const double s = a.thermostat ? a.ctl->nh_scale : 1.0;
        double acc[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int64_t i = bid * (int64_t)blockDim.x + threadIdx.x; i < a.n; i += (int64_t)nbk * blockDim.x) {
            const double m = a.mass[a.ntypes > 1 ? a.type[i] : 0];
            const double vx = a.vx[i] * s, vy = a.vy[i] * s, vz = a.vz[i] * s;
            acc[1] += m * vx * vx;
            acc[2] += m * vy * vy;
            acc[3] += m * vz * vz;
            acc[4] += m * vx * vy;
            acc[5] += m * vx * vz;
            acc[6] += m * vy * vz;
        }
        block_partials<kBlock>(acc, a.ke_partials, bid);
    }
};

// Fixed-order sum of nb partials (stride `stride`) over one 256-thread block.
__device__ double block_fixed_sum(const double* p, int nb, int stride) {
    __shared__ double sh[kBlock];
    double t = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock)
        t += p[(int64_t)b * stride];
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

struct PostArgs {
    DevCtl* ctl;
    const double* ke_partials;
    int nb;
    int thermostat;
    double dt;
    // decide (initial phase only)
    int decide, every, delay, check;
    double* record_ke; // may be null
    int step;          // record index; < 0: the device counter ctl->rec_step (advanced here)
    int final_phase;
};

// NoseHoover::launch scalar part + NeighborList::decide_reneighbor.
struct OpPost {
    using Args = PostArgs;
    static __device__ void run(const PostArgs& a, int, int) {
        if (halted(a.ctl)) {
            // nothing of this step is integrated, and no rebuild is decided
            if (threadIdx.x == 0 && a.decide)
                a.ctl->rebuild = 0;
            return;
        }
        const double mv2 = block_fixed_sum(a.ke_partials, a.nb, 1);
        if (threadIdx.x != 0)
            return;
        DevCtl* c = a.ctl;
        const double ke = 0.5 * mass_v2_to_energy * mv2;
        double scale = 1.0;
        if (a.thermostat) {
            // Q and Ndof kB T live in the control block (a replica exchange moves them)
            c->xi += (a.dt * 0.5) * (2.0 * ke - c->nh_ndof_kT) / c->nh_Q;
            scale = exp(-a.dt * 0.5 * c->xi);
            c->nh_scale = scale;
        }
        if (a.final_phase) {
            c->ke = ke * scale * scale; // KineticEnergy at end_of_step (scaled v)
            const int step = a.step >= 0 ? a.step : c->rec_step;
            if (a.record_ke)
                a.record_ke[step] = c->ke;
            if (a.step < 0)
                c->rec_step = step + 1;
            c->steps_done += 1;
        } else {
            c->ke_half = ke;
        }
        if (a.decide) {
            c->ago += 1;
            const bool delay_ok = c->ago >= a.delay;
            const bool every_ok = (c->ago % a.every) == 0;
            const bool rb = delay_ok && every_ok && (!a.check || c->disp_flag != 0);
            c->disp_flag = 0;
            c->rebuild = rb ? 1 : 0;
            if (rb) {
                c->ago = 0;
                c->nrebuilds += 1;
                c->nbr_total = 0;
                c->max_nbr = 0;
            }
        }
    }
};

struct KinStressArgs {
    DevCtl* ctl;
    const double* partials;
    int nb;
};

// Totals of the kinetic-stress partials, converted to eV (mass_v2_to_energy).
struct OpKinStressSum {
    using Args = KinStressArgs;
    static __device__ void run(const KinStressArgs& a, int, int) {
        double v[6];
        for (int c = 0; c < 6; ++c)
            v[c] = block_fixed_sum(a.partials + 1 + c, a.nb, kPartialStride);
        if (threadIdx.x == 0)
            for (int c = 0; c < 6; ++c)
                a.ctl->stress_kin[c] = mass_v2_to_energy * v[c];
    }
};

struct FinalizeArgs {
    DevCtl* ctl;
    const double* partials; // pair style, per workgroup
    int nb;
    const double* kpartials; // k-space influence, per workgroup (kspace only)
    int knb;
    double escale, e_self, e_net;
    double* record_pe;
    int step; // record index; < 0: the device counter ctl->rec_step
};

// Fixed-order totals of the force phase: pair energy and virial, then the
// k-space terms in the reference's order (self first, then reciprocal + net
// charge; pme.cpp:222-246).
struct OpFinalize {
    using Args = FinalizeArgs;
    static __device__ void run(const FinalizeArgs& a, int, int) {
        double v[7], k[7];
        for (int c = 0; c < 7; ++c)
            v[c] = a.nb > 0 ? block_fixed_sum(a.partials + c, a.nb, kPartialStride) : 0.0;
        for (int c = 0; c < 7; ++c)
            k[c] = a.knb > 0 ? block_fixed_sum(a.kpartials + c, a.knb, kPartialStride) : 0.0;
        if (threadIdx.x == 0) {
            double pe = v[0];
            if (a.knb > 0) {
                const double e_recip = k[0] * a.escale;
                pe += a.e_self;
                pe += e_recip + a.e_net;
                for (int c = 0; c < 6; ++c)
                    v[c + 1] += k[c + 1] * a.escale;
                a.ctl->e_kspace[0] = e_recip;
                a.ctl->e_kspace[1] = a.e_self;
                a.ctl->e_kspace[2] = a.e_net;
            }
            a.ctl->pe = pe;
            for (int c = 0; c < 6; ++c)
                a.ctl->vir[c] = v[c + 1];
            if (a.record_pe)
                a.record_pe[a.step >= 0 ? a.step : a.ctl->rec_step] = pe;
        }
    }
};

// Start of a fused run: no step of it completed yet.
__global__ void k_run_reset(DevCtl* ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        ctl->steps_done = 0;
}

IntArgs make_int_args(atvh_system* s) {
    IntArgs a{};
    a.n = s->n;
    a.x = s->x;
    a.y = s->y;
    a.z = s->z;
    a.vx = s->vx;
    a.vy = s->vy;
    a.vz = s->vz;
    a.fx = s->fx;
    a.fy = s->fy;
    a.fz = s->fz;
    a.type = s->type;
    a.xs = s->xs;
    a.ys = s->ys;
    a.zs = s->zs;
    // VelocityVerlet: dtf = 0.5*dt*force_to_accel, dtfm = dtf / m
    const double dtf = 0.5 * s->dt * force_to_accel;
    for (int t = 0; t < kMaxTypes; ++t) {
        a.mass[t] = t < s->ntypes ? s->mass_t[t] : 1.0;
        a.dtfm[t] = dtf / a.mass[t];
    }
    a.ntypes = s->ntypes;
    a.dt = s->dt;
    a.lx = s->box[0];
    a.ly = s->box[1];
    a.lz = s->box[2];
    a.skin_sq = s->skin * s->skin;
    a.thermostat = s->thermostat;
    a.ctl = s->ctl;
    a.ke_partials = s->partials + size_t(s->npartials) * kPartialStride;
    return a;
}

PostArgs make_post_args(atvh_system* s, bool final_phase, bool decide) {
    PostArgs p{};
    p.ctl = s->ctl;
    p.ke_partials = s->partials + size_t(s->npartials) * kPartialStride;
    p.nb = std::min(grid_for(s->n), kIntGrid);
    p.thermostat = s->thermostat;
    p.dt = s->dt;
    p.decide = decide ? 1 : 0;
    p.every = s->every;
    p.delay = s->delay;
    p.check = s->check;
    p.record_ke = nullptr;
    p.step = 0;
    p.final_phase = final_phase ? 1 : 0;
    return p;
}

void check_integ(atvh_system* s) {
    if (!s->integ_configured)
        throw Error(ATVH_E_STATE, "integrator not configured");
}

} // namespace

namespace atvh {

int integ_grid(const atvh_system* s) { return std::min(grid_for(s->n), kIntGrid); }

void integ_enqueue_initial(atvh_system* s) {
    ScopedTimer tm(s->ctx, "integrate");
    IntArgs a = make_int_args(s);
    k_single<OpInitial><<<integ_grid(s), kBlock, 0, s->ctx->stream>>>(a);
    ATVH_HIP(hipGetLastError());
}

void integ_enqueue_final(atvh_system* s) {
    ScopedTimer tm(s->ctx, "integrate");
    IntArgs a = make_int_args(s);
    k_single<OpFinal><<<integ_grid(s), kBlock, 0, s->ctx->stream>>>(a);
    ATVH_HIP(hipGetLastError());
}

// Batched stages: kick/drift (+ displacement flag), NH + decide, final kick,
// NH + end-of-step KE, and the force-phase energy totals. Records go to each
// member's thermo buffer at the member's device step counter.
void integ_batch_build(atvh_batch* b) {
    std::vector<IntArgs> ia;
    std::vector<PostArgs> pi, pf;
    std::vector<FinalizeArgs> fa;
    std::vector<int> gi, g1;
    for (atvh_system* s : b->sys) {
        ia.push_back(make_int_args(s));
        gi.push_back(integ_grid(s));
        PostArgs p0 = make_post_args(s, false, true);
        PostArgs p1 = make_post_args(s, true, false);
        p1.record_ke = s->d_thermo + s->thermo_cap;
        p1.step = -1;
        pi.push_back(p0);
        pf.push_back(p1);
        FinalizeArgs f{};
        f.ctl = s->ctl;
        f.partials = s->partials;
        f.nb = pair_grid(s);
        f.kpartials = nullptr;
        f.knb = 0;
        f.escale = f.e_self = f.e_net = 0.0;
        if (s->pme) { // the k-space energy and virial partials of the member's PME call
            f.kpartials = pme_partials(s);
            f.knb = pme_partial_blocks(s); // set for this run's precision by pme_prepare
            pme_energy_consts(s, f.escale, f.e_self, f.e_net);
        }
        f.record_pe = s->d_thermo;
        f.step = -1;
        fa.push_back(f);
        g1.push_back(1);
    }
    batch_stage(b, "initial", ia, gi);
    batch_stage(b, "final", ia, gi);
    batch_stage(b, "post_initial", pi, g1);
    batch_stage(b, "post_final", pf, g1);
    batch_stage(b, "finalize", fa, g1);
}

void integ_batch_launch(atvh_batch* b, const char* stage) {
    const BatchStage& st = b->stages.at(stage);
    hipStream_t s = b->ctx->stream;
    const int* own = st.owner;
    const std::string name(stage);
    if (name == "initial")
        k_batched<OpInitial><<<st.nblocks, kBlock, 0, s>>>(static_cast<const IntArgs*>(st.args), st.base, own);
    else if (name == "final")
        k_batched<OpFinal><<<st.nblocks, kBlock, 0, s>>>(static_cast<const IntArgs*>(st.args), st.base, own);
    else if (name == "post_initial" || name == "post_final")
        k_batched<OpPost><<<st.nblocks, kBlock, 0, s>>>(static_cast<const PostArgs*>(st.args), st.base, own);
    else if (name == "finalize")
        k_batched<OpFinalize><<<st.nblocks, kBlock, 0, s>>>(static_cast<const FinalizeArgs*>(st.args), st.base,
                                                             own);
    else
        throw Error(ATVH_E_ARG, "unknown batch stage");
    ATVH_HIP(hipGetLastError());
}

ThermoRec thermo_rec(atvh_system* s) { return ThermoRec{s->d_thermo, s->d_thermo + s->thermo_cap}; }

void md_enqueue_steps(atvh_system* s, int first, int nsteps, int precision, int energy_every, bool resume,
                      ThermoRec rec) {
    atvh_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    double* rec_pe = rec.pe;
    double* rec_ke = rec.ke;
    const IntArgs ia = make_int_args(s);
    PostArgs p_init = make_post_args(s, false, true);
    PostArgs p_fin = make_post_args(s, true, false);
    p_fin.record_ke = rec_ke;
    const int g = integ_grid(s);
    for (int step = first; step < nsteps; ++step) {
        if (!(resume && step == first)) {
            {
                ScopedTimer tm(c, "integrate");
                k_single<OpInitial><<<g, kBlock, 0, st>>>(ia);
            }
            k_single<OpPost><<<1, kBlock, 0, st>>>(p_init);
            nl_enqueue_rebuild(s, &s->ctl->rebuild, false);
        }
        const bool ef = (step == nsteps - 1) || (energy_every > 0 && (step % energy_every) == 0);
        pair_enqueue(s, precision, ef, ef);
        if (ef)
            energy_enqueue_finalize(s, rec_pe, step);
        {
            ScopedTimer tm(c, "integrate");
            k_single<OpFinal><<<g, kBlock, 0, st>>>(ia);
        }
        p_fin.step = step;
        k_single<OpPost><<<1, kBlock, 0, st>>>(p_fin);
        ATVH_HIP(hipGetLastError());
    }
}

// A rebuild that overflowed halted the system at the start of that step's
// force phase (every later kernel of the run left at once). The list is
// regrown at the positions the step reached (the reference regrows inside
// the rebuild, src/cuda/neighbor/neighbor_list.cu:544-553) and the run
// resumes at that force phase, so the trajectory equals an uninterrupted
// run with a larger capacity.
void md_recover_and_finish(atvh_system* s, int nsteps, int precision, int energy_every) {
    hipStream_t st = s->ctx->stream;
    for (int attempt = 0;; ++attempt) {
        DevCtl h{};
        ATVH_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
        if (!(h.overflow || h.halo_overflow))
            return;
        if (attempt >= 8 || h.steps_done >= nsteps) {
            s->nl_built = false;
            nl_check_ctl(h);
        }
        nl_build(s, true);
        md_enqueue_steps(s, h.steps_done, nsteps, precision, energy_every, true, thermo_rec(s));
        ATVH_HIP(hipStreamSynchronize(st));
    }
}

void energy_enqueue_finalize(atvh_system* s, double* record_pe, int step, bool with_pair) {
    FinalizeArgs a{};
    a.ctl = s->ctl;
    a.partials = s->partials;
    a.nb = with_pair ? pair_grid(s) : 0;
    a.kpartials = pme_partials(s);
    a.knb = pme_partial_blocks(s);
    a.escale = a.e_self = a.e_net = 0.0;
    if (a.knb > 0)
        pme_energy_consts(s, a.escale, a.e_self, a.e_net);
    a.record_pe = record_pe;
    a.step = step;
    k_single<OpFinalize><<<1, kBlock, 0, s->ctx->stream>>>(a);
    ATVH_HIP(hipGetLastError());
}

} // namespace atvh

namespace atvh {

// Allocation-free MD run: the per-step energy record lives in this file-scope
// array instead of a per-system buffer grown with hipMalloc. Layout: potential
// energies in [0, kNoAllocSteps), kinetic energies in [kNoAllocSteps,
// 2*kNoAllocSteps). Each device holds its own copy (1 MiB) until the process
// ends. Host code never passes the array's name to a kernel or a copy (on the
// host it is only a stand-in for the symbol): it uses the device address that
// noalloc_bind looks up once per device.
constexpr int kNoAllocSteps = 1 << 16;
static __device__ __align__(256) double scratch[2 * kNoAllocSteps];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

__global__ void k_fill_records(double* rec, int nsteps, int stride) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsteps) {
        const double nan = __builtin_nan("");
        rec[i] = nan;
        rec[stride + i] = nan;
    }
}

double* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices || !g_scratch_ptr[device])
        throw Error(ATVH_E_STATE, "atvh_md_run_noalloc: the device's static record was not looked up");
    return g_scratch_ptr[device];
}
} // namespace

void noalloc_bind(int device) {
    if (device < 0 || device >= kMaxDevices)
        throw Error(ATVH_E_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device])
        return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw Error(ATVH_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double*>(ptr);
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

int atvh_integrator_configure(atvh_system* s, double dt, int thermostat, double T_target,
                              double tau_T, int64_t ndof) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (!(dt > 0.0))
        throw Error(ATVH_E_ARG, "VelocityVerlet: dt must be positive");
    if (thermostat != 0 && thermostat != 1)
        throw Error(ATVH_E_ARG, "unknown thermostat");
    if (thermostat == 1 && (!(T_target > 0.0) || !(tau_T > 0.0)))
        throw Error(ATVH_E_ARG, "NoseHoover: T and tau must be positive");
    s->dt = dt;
    s->thermostat = thermostat;
    s->T_target = T_target;
    s->tau_T = tau_T;
    // publish_global_dof (src/cpu/driver/dof.cpp:49): Ndof = 3N - 3
    s->ndof = ndof > 0 ? ndof : 3 * s->n - 3;
    npt_free(s); // VelocityVerlet (+ NoseHoover) replaces a NoseHooverNPT
    ATVH_HIP(hipSetDevice(c->device));
    ATVH_HIP(hipStreamSynchronize(c->stream));
    DevCtl h{};
    ATVH_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
    // NoseHoover::build: Q = Ndof kB T tau^2
    h.nh_Q = static_cast<double>(s->ndof) * kB_ev_per_k * T_target * tau_T * tau_T;
    h.nh_ndof_kT = static_cast<double>(s->ndof) * kB_ev_per_k * T_target;
    h.T_target = T_target;
    ATVH_HIP(hipMemcpy(s->ctl, &h, sizeof(h), hipMemcpyHostToDevice));
    s->integ_configured = true;
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_step_vvnh(atvh_system* s, int phase) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    check_integ(s);
    if (s->npt)
        throw Error(ATVH_E_STATE, "atvh_step_vvnh: a NoseHooverNPT is configured (atvh_md_run steps it)");
    ATVH_HIP(hipSetDevice(c->device));
    if (phase == ATVH_PHASE_INITIAL) {
        integ_enqueue_initial(s);
        PostArgs p = make_post_args(s, false, false);
        k_single<OpPost><<<1, kBlock, 0, c->stream>>>(p);
    } else if (phase == ATVH_PHASE_FINAL) {
        integ_enqueue_final(s);
        PostArgs p = make_post_args(s, true, false);
        k_single<OpPost><<<1, kBlock, 0, c->stream>>>(p);
    } else {
        throw Error(ATVH_E_ARG, "atvh_step_vvnh: unknown phase");
    }
    ATVH_HIP(hipGetLastError());
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_md_run(atvh_system* s, int nsteps, int precision, int energy_every) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (nsteps < 0)
        throw Error(ATVH_E_ARG, "atvh_md_run: nsteps < 0");
    if (precision != ATVH_PREC_FP64 && precision != ATVH_PREC_MIXED)
        throw Error(ATVH_E_ARG, "atvh_md_run: unknown precision");
    check_integ(s);
    if (!s->nl_configured)
        throw Error(ATVH_E_STATE, "atvh_md_run: no neighbor list configured");
    if (!s->pair_configured)
        throw Error(ATVH_E_STATE, "atvh_md_run: no force evaluator configured");
    ATVH_HIP(hipSetDevice(c->device));
    if (!s->nl_built) {
        const int rc = atvh_nlist_build(s);
        if (rc != ATVH_OK)
            return rc;
    }
    hipStream_t st = c->stream;
    if (s->thermo_cap < nsteps) {
        dev_free(s->d_thermo);
        s->thermo_cap = nsteps;
        dev_alloc(&s->d_thermo, size_t(2) * nsteps);
    }
    s->thermo_n = nsteps;
    if (s->npt) {
        // NoseHooverNPT: energies every step (its barostat reads the virial
        // and its conserved quantity the potential energy)
        std::vector<double> rec(size_t(2) * std::max(s->thermo_cap, 1), 0.0);
        npt_run(s, nsteps, precision, rec.data(), rec.data() + s->thermo_cap);
        ATVH_HIP(hipMemcpy(s->d_thermo, rec.data(), size_t(2) * s->thermo_cap * sizeof(double),
                           hipMemcpyHostToDevice));
        return ATVH_OK;
    }
    if (nsteps > 0) {
        // NaN marks steps without an energy request; layout [pe | ke], stride thermo_cap
        std::vector<double> nanv(size_t(2) * s->thermo_cap, std::nan(""));
        ATVH_HIP(hipMemcpyAsync(s->d_thermo, nanv.data(), nanv.size() * sizeof(double),
                                hipMemcpyHostToDevice, st));
        ATVH_HIP(hipStreamSynchronize(st));
    }
    k_run_reset<<<1, 64, 0, st>>>(s->ctl);
    md_enqueue_steps(s, 0, nsteps, precision, energy_every, false, thermo_rec(s));
    ATVH_HIP(hipStreamSynchronize(st));
    md_recover_and_finish(s, nsteps, precision, energy_every);
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_md_run_noalloc(atvh_system* s, int nsteps, int precision, int energy_every, double* pe, double* ke) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (nsteps < 0)
        throw Error(ATVH_E_ARG, "atvh_md_run_noalloc: nsteps < 0");
    if (nsteps > kNoAllocSteps)
        throw Error(ATVH_E_ARG, "atvh_md_run_noalloc: nsteps exceeds the static record of " +
                                    std::to_string(kNoAllocSteps) + " steps");
    if (precision != ATVH_PREC_FP64 && precision != ATVH_PREC_MIXED)
        throw Error(ATVH_E_ARG, "atvh_md_run_noalloc: unknown precision");
    check_integ(s);
    if (!s->nl_configured)
        throw Error(ATVH_E_STATE, "atvh_md_run_noalloc: no neighbor list configured");
    if (!s->pair_configured)
        throw Error(ATVH_E_STATE, "atvh_md_run_noalloc: no force evaluator configured");
    // every buffer a step touches must already exist: the first list build
    // and the first PME evaluation of a precision allocate theirs
    if (!s->nl_built)
        throw Error(ATVH_E_STATE, "atvh_md_run_noalloc: the neighbor list is not built (atvh_nlist_build first)");
    if (!pme_ready(s, precision))
        throw Error(ATVH_E_STATE, "atvh_md_run_noalloc: no PME evaluation of this precision has run yet "
                                  "(its mesh buffers and transform plans are made then)");
    if (c->timing)
        throw Error(ATVH_E_STATE, "atvh_md_run_noalloc: kernel timing is on (it takes events from a pool)");
    if (s->npt)
        throw Error(ATVH_E_STATE, "atvh_md_run_noalloc: NoseHooverNPT rebuilds its grid for each new box (allocates)");
    ATVH_HIP(hipSetDevice(c->device));
    double* rec = noalloc_scratch(c->device);
    hipStream_t st = c->stream;
    if (nsteps > 0) {
        // NaN marks steps without an energy request (a fill kernel, no memset)
        k_fill_records<<<(nsteps + kBlock - 1) / kBlock, kBlock, 0, st>>>(rec, nsteps, kNoAllocSteps);
        ATVH_HIP(hipGetLastError());
    }
    s->thermo_n = 0; // the system's own record no longer describes the last run
    k_run_reset<<<1, 64, 0, st>>>(s->ctl);
    md_enqueue_steps(s, 0, nsteps, precision, energy_every, false, ThermoRec{rec, rec + kNoAllocSteps});
    ATVH_HIP(hipStreamSynchronize(st));
    DevCtl h{};
    ATVH_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
    if (h.overflow || h.halo_overflow) {
        // regrowing the list would allocate: the system stays halted at the
        // step whose rebuild overflowed and the caller re-uploads its frame
        s->nl_built = false;
        nl_check_ctl(h);
    }
    if (nsteps > 0 && pe)
        ATVH_HIP(hipMemcpy(pe, rec, size_t(nsteps) * sizeof(double), hipMemcpyDeviceToHost));
    if (nsteps > 0 && ke)
        ATVH_HIP(hipMemcpy(ke, rec + kNoAllocSteps, size_t(nsteps) * sizeof(double), hipMemcpyDeviceToHost));
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_kinetic_stress(atvh_system* s) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    check_integ(s);
    ATVH_HIP(hipSetDevice(c->device));
    const IntArgs ia = make_int_args(s);
    const int g = integ_grid(s);
    k_single<OpKinStress><<<g, kBlock, 0, c->stream>>>(ia);
    KinStressArgs ka{s->ctl, ia.ke_partials, g};
    k_single<OpKinStressSum><<<1, kBlock, 0, c->stream>>>(ka);
    ATVH_HIP(hipGetLastError());
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_stress(atvh_system* s, double* kinetic6, double* virial6, double* pressure) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    ATVH_HIP(hipSetDevice(c->device));
    ATVH_HIP(hipStreamSynchronize(c->stream));
    DevCtl h{};
    ATVH_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
    if (kinetic6)
        for (int k = 0; k < 6; ++k)
            kinetic6[k] = h.stress_kin[k];
    if (virial6)
        for (int k = 0; k < 6; ++k)
            virial6[k] = h.vir[k];
    if (pressure) {
        // compute_pressure convention (virial_stress.h:30-45): P = (tr K + tr W) / (3V),
        // positive is compressive, in eV/Å^3 (the reference's pressure unit)
        const double vol = s->box[0] * s->box[1] * s->box[2];
        *pressure = (h.stress_kin[0] + h.stress_kin[1] + h.stress_kin[2] + h.vir[0] + h.vir[1] + h.vir[2]) /
                    (3.0 * vol);
    }
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_md_thermo(atvh_system* s, double* pe, double* ke, int nsteps) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (nsteps > s->thermo_n)
        throw Error(ATVH_E_ARG, "atvh_md_thermo: more steps requested than recorded");
    ATVH_HIP(hipStreamSynchronize(c->stream));
    if (nsteps <= 0)
        return ATVH_OK;
    if (pe)
        ATVH_HIP(hipMemcpy(pe, s->d_thermo, nsteps * sizeof(double), hipMemcpyDeviceToHost));
    if (ke)
        ATVH_HIP(hipMemcpy(ke, s->d_thermo + s->thermo_cap, nsteps * sizeof(double),
                           hipMemcpyDeviceToHost));
    return ATVH_OK;
    API_CATCH(c)
}

} // extern "C"
