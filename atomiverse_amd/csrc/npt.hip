// npt.hip — MTK Nose-Hoover NPT (isotropic or per-axis barostat, M-link
// Nose-Hoover chains on the atoms and on the barostat, optional temperature
// ramp) in place of VelocityVerlet + NoseHoover.
//
// Replaces cuda::md::NoseHooverNPT (include/atomiverse/cuda/dynamics/
// nose_hoover_npt.cuh:56-94, src/cuda/dynamics/nose_hoover_npt.cu); the
// algorithm and its Trotter order are cpu::md::NoseHooverNPT's
// (src/cpu/dynamics/nose_hoover_npt.cpp), per phase:
//   initial_integrate  P1 barostat chain, P2 atom chain (v *= s),
//                      P3 eps_dot += G_eps dt/2 (fresh per-axis KE, the last
//                      force evaluation's virial), P4 v_a *= exp(-(eps_dot_a +
//                      mtk2) dt/2), P5 v += dt/2 f/m, P6 x = x e^{eps dt} +
//                      v dt (e^{eps dt}-1)/(eps dt), box *= e^{eps dt}
//   final_integrate    P7 v += dt/2 f/m
//   end_of_step        P8 = P4, P9 = P3 with the new virial, P10 atom chain,
//                      P11 barostat chain; KE and the conserved H published.
//
// MI355X split: the chains, eps_dot and the box are a handful of FP64
// scalars advanced on the host, as the reference does; every per-atom pass
// is one fused kernel with fixed-order block partials (a sum's rounding
// does not depend on the launch), and the host reads the sums back twice per
// step (after P2 and after P8), because each chain update needs a kinetic
// energy of velocities the previous kernel produced. The P10 velocity scale
// is not a pass of its own: it is carried to the next step's first kernel
// (or applied at the end of the run), which multiplies it in before P2's, in
// the reference's order.
//
// The box moves every step, so a rebuild decided on the reference's cadence
// (NeighborList::decide_reneighbor; a box that changed since the last build
// counts as stale, src/cpu/neighbor/cell_list.cpp:156-159, 369-370) also
// recomputes the cell grid, stencil and bricks for the new box
// (cell_list.cpp:344-347). Between rebuilds the positions are not wrapped
// (the deferred wrap of the fixed-box path): an unwrapped coordinate scales
// with the box, so every listed pair keeps its periodic image, and the pair
// kernels take the current box for the image shifts.

#include "atvh_internal.hpp"
#include "batch.hpp"
#include "pair_common.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace atvh {

constexpr int kNptMaxChain = 8;
constexpr double kEpsDtMaxFraction = 0.005; // nose_hoover_npt.cpp:45

struct NptState {
    // NoseHooverNPT constructor (nose_hoover_npt.h:183-188)
    double T_target = 0, T_stop = 0, tau_T = 0, P_target = 0, tau_P = 0, drag = 0;
    int mtchain = 3, mpchain = 3;
    int couple[3] = {1, 1, 1};
    int pdim = 3;
    int64_t nfirst = 0, nlast = 0;
    // derived (build): masses, drag factors
    double kT = 0, W = 0, tdrag = 1, pdrag = 1;
    double Q_t[kNptMaxChain] = {}, Q_b[kNptMaxChain] = {};
    // extended-Lagrangian state
    bool built = false;
    double eta_t[kNptMaxChain] = {}, eta_t_dot[kNptMaxChain] = {};
    double eta_b[kNptMaxChain] = {}, eta_b_dot[kNptMaxChain] = {};
    double eps_dot[3] = {0, 0, 0};
    double ke_global = 0;
    double vir[3] = {0, 0, 0}; // virial diagonal of the last force evaluation
    double pending = 1.0;      // P10 velocity scale not yet applied to v
    int64_t step = 0;          // cumulative step index (the ramp's clock)
    // per-step records of the last run
    std::vector<double> H, V;
};

namespace {

// (e^{2x} - 1)/(2x) with a Taylor branch near 0 (nose_hoover_npt.cpp:66-73)
double expm1c(double x) {
    if (std::abs(x) < 1e-6)
        return 1.0 + x + (2.0 / 3.0) * x * x;
    return std::expm1(2.0 * x) / (2.0 * x);
}

struct NptArgs {
    int64_t n;
    double *x, *y, *z, *vx, *vy, *vz;
    const double *fx, *fy, *fz;
    const int* type;
    int ntypes;
    double mass[kMaxTypes];
    double dtfm[kMaxTypes];
    const double *xs, *ys, *zs;
    double lx, ly, lz, skin_sq;
    DevCtl* ctl;
    double* partials; // [block][kPartialStride]
    int nb;
    double s0, s1;    // uniform velocity scales applied first, in this order
    double sa[3];     // per-axis velocity scale
    double a[3];      // position scale e^{eps dt}
    double drift[3];  // (e^{eps dt} - 1)/(eps dt)
    double dt;
};

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_down(v, off, 64);
    return v;
}

// Fixed-order block sum of `cols` values per thread into partials[bid][c].
template <int kCols>
__device__ __forceinline__ void block_partials(double (&v)[kCols], double* out) {
    __shared__ double sh[kCols][kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < kCols; ++c) {
        const double t = wave_sum(v[c]);
        if (lane == 0)
            sh[c][wid] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < kCols; ++c) {
            double t = 0.0;
            for (int w = 0; w < kBlock / 64; ++w)
                t += sh[c][w];
            out[c] = t;
        }
    }
}

// P10 of the previous step (s0) and P2 (s1), then the per-axis sums
// m v_a v_a that P3 reads (compute_axis_KE, nose_hoover_npt.cpp:236-266).
__global__ __launch_bounds__(kBlock) void k_npt_chain_scale(NptArgs a) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kBlock) {
        const int t = a.ntypes > 1 ? a.type[i] : 0;
        const double m = a.mass[t];
        double vx = a.vx[i] * a.s0, vy = a.vy[i] * a.s0, vz = a.vz[i] * a.s0;
        vx *= a.s1;
        vy *= a.s1;
        vz *= a.s1;
        a.vx[i] = vx;
        a.vy[i] = vy;
        a.vz[i] = vz;
        acc[0] += m * vx * vx;
        acc[1] += m * vy * vy;
        acc[2] += m * vz * vz;
    }
    block_partials<3>(acc, a.partials + (size_t)blockIdx.x * kPartialStride);
}

// P4 + P5 + P6 per atom (scale_v_by_baro, kick_v_by_force,
// propagate_x_and_box) and the displacement half of needs_rebuild against
// the positions of the last build, in the current box.
__global__ __launch_bounds__(kBlock) void k_npt_drift(NptArgs a) {
    int moved = 0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kBlock) {
        const int t = a.ntypes > 1 ? a.type[i] : 0;
        const double dtfm = a.dtfm[t];
        double vx = a.vx[i] * a.sa[0], vy = a.vy[i] * a.sa[1], vz = a.vz[i] * a.sa[2];
        vx += dtfm * a.fx[i];
        vy += dtfm * a.fy[i];
        vz += dtfm * a.fz[i];
        const double x = a.x[i] * a.a[0] + vx * a.dt * a.drift[0];
        const double y = a.y[i] * a.a[1] + vy * a.dt * a.drift[1];
        const double z = a.z[i] * a.a[2] + vz * a.dt * a.drift[2];
        a.x[i] = x;
        a.y[i] = y;
        a.z[i] = z;
        a.vx[i] = vx;
        a.vy[i] = vy;
        a.vz[i] = vz;
        double dx = x - a.xs[i], dy = y - a.ys[i], dz = z - a.zs[i];
        dx -= a.lx * nearbyint(dx / a.lx);
        dy -= a.ly * nearbyint(dy / a.ly);
        dz -= a.lz * nearbyint(dz / a.lz);
        if ((dx * dx + dy * dy + dz * dz) * 4.0 > a.skin_sq)
            moved = 1;
    }
    if (__any(moved) && (threadIdx.x & 63) == 0)
        a.ctl->disp_flag = 1;
}

// P7 (the second half kick), the total m v^2 that end_of_step's
// compute_global_KE takes, then P8's per-axis scale and the sums P8/P9 read:
// columns {m vx^2, m vy^2, m vz^2, m v^2 before P8, m v^2 after P8}.
__global__ __launch_bounds__(kBlock) void k_npt_kick_scale(NptArgs a) {
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kBlock) {
        const int t = a.ntypes > 1 ? a.type[i] : 0;
        const double m = a.mass[t];
        const double dtfm = a.dtfm[t];
        double vx = a.vx[i] + dtfm * a.fx[i];
        double vy = a.vy[i] + dtfm * a.fy[i];
        double vz = a.vz[i] + dtfm * a.fz[i];
        acc[3] += m * (vx * vx + vy * vy + vz * vz);
        vx *= a.sa[0];
        vy *= a.sa[1];
        vz *= a.sa[2];
        a.vx[i] = vx;
        a.vy[i] = vy;
        a.vz[i] = vz;
        acc[0] += m * vx * vx;
        acc[1] += m * vy * vy;
        acc[2] += m * vz * vz;
        acc[4] += m * (vx * vx + vy * vy + vz * vz);
    }
    block_partials<5>(acc, a.partials + (size_t)blockIdx.x * kPartialStride);
}

// The pending uniform scale at the end of a run.
__global__ __launch_bounds__(kBlock) void k_npt_scale(NptArgs a) {
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kBlock) {
        a.vx[i] *= a.s0;
        a.vy[i] *= a.s0;
        a.vz[i] *= a.s0;
    }
}

// One block: column sums over the grid's partials, in block order, into
// ctl->red; also clears the displacement flag ahead of the drift kernel.
__global__ __launch_bounds__(kBlock) void k_npt_sum(const double* partials, int nb, int cols, DevCtl* ctl) {
    __shared__ double sh[kBlock];
    for (int c = 0; c < cols; ++c) {
        double t = 0.0;
        for (int b = threadIdx.x; b < nb; b += kBlock)
            t += partials[(size_t)b * kPartialStride + c];
        sh[threadIdx.x] = t;
        __syncthreads();
        for (int off = kBlock / 2; off > 0; off >>= 1) {
            if (threadIdx.x < off)
                sh[threadIdx.x] += sh[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            ctl->red[c] = sh[0];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        ctl->disp_flag = 0;
}

int npt_grid(const atvh_system* s) { return std::min(grid_for(s->n), 1024); }

NptArgs npt_args(atvh_system* s) {
    NptArgs a{};
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
    a.ntypes = s->ntypes;
    // kick_v_by_force (nose_hoover_npt.cpp:639-660): dtf = dt/2 ftm2v, dtfm = dtf/m
    const double dtf = 0.5 * s->dt * force_to_accel;
    for (int t = 0; t < kMaxTypes; ++t) {
        a.mass[t] = t < s->ntypes ? s->mass_t[t] : 1.0;
        a.dtfm[t] = dtf / a.mass[t];
    }
    a.xs = s->xs;
    a.ys = s->ys;
    a.zs = s->zs;
    a.lx = s->box[0];
    a.ly = s->box[1];
    a.lz = s->box[2];
    a.skin_sq = s->skin * s->skin;
    a.ctl = s->ctl;
    a.partials = s->partials + size_t(s->npartials) * kPartialStride; // the kinetic rows
    a.nb = npt_grid(s);
    a.s0 = a.s1 = 1.0;
    for (int d = 0; d < 3; ++d) {
        a.sa[d] = 1.0;
        a.a[d] = 1.0;
        a.drift[d] = 1.0;
    }
    a.dt = s->dt;
    return a;
}

void read_ctl(atvh_system* s, DevCtl& h) {
    ATVH_HIP(hipStreamSynchronize(s->ctx->stream));
    ATVH_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
}

void reduce(atvh_system* s, const NptArgs& a, int cols) {
    k_npt_sum<<<1, kBlock, 0, s->ctx->stream>>>(a.partials, a.nb, cols, s->ctl);
    ATVH_HIP(hipGetLastError());
}

// ---- the chain and barostat arithmetic (host, FP64, reference order) -----

// NoseHooverNPT::build (nose_hoover_npt.cpp:101-190)
void npt_build(atvh_system* s, NptState& q) {
    q.kT = kB_ev_per_k * q.T_target;
    const double tau_T2 = q.tau_T * q.tau_T;
    const double tau_P2 = q.tau_P * q.tau_P;
    q.Q_t[0] = static_cast<double>(s->ndof) * q.kT * tau_T2;
    for (int k = 1; k < kNptMaxChain; ++k)
        q.Q_t[k] = q.kT * tau_T2;
    q.Q_b[0] = q.kT * tau_P2;
    for (int k = 1; k < kNptMaxChain; ++k)
        q.Q_b[k] = q.kT * tau_P2;
    q.W = (static_cast<double>(s->ndof) + 3.0) * q.kT * tau_P2;
    q.tdrag = 1.0 - q.drag * s->dt / q.tau_T;
    q.pdrag = 1.0 - q.drag * s->dt / q.tau_P;
    if (!q.built) {
        for (int k = 0; k < kNptMaxChain; ++k)
            q.eta_t[k] = q.eta_t_dot[k] = q.eta_b[k] = q.eta_b_dot[k] = 0.0;
        for (int d = 0; d < 3; ++d)
            q.eps_dot[d] = 0.0;
        // equipartition KE for step 0's first chain update
        q.ke_global = 0.5 * static_cast<double>(s->ndof) * q.kT;
        q.pending = 1.0;
        q.built = true;
    }
}

// compute_temp_target (:194-203) and the per-step chain masses (:838-842)
void temp_target(atvh_system* s, NptState& q) {
    double delta = 0.0;
    if (q.nlast > q.nfirst) {
        delta = static_cast<double>(q.step - q.nfirst) / static_cast<double>(q.nlast - q.nfirst);
        delta = std::clamp(delta, 0.0, 1.0);
    }
    const double T_live = q.T_target + delta * (q.T_stop - q.T_target);
    q.kT = kB_ev_per_k * T_live;
    const double tau_T2 = q.tau_T * q.tau_T;
    q.Q_t[0] = static_cast<double>(s->ndof) * q.kT * tau_T2;
    for (int k = 1; k < kNptMaxChain; ++k)
        q.Q_t[k] = q.kT * tau_T2;
}

// nhc_atoms_half (:269-340), split where the velocities are scaled: the
// backward sweep returns the scale, the forward sweep runs after it.
double atoms_chain_backward(const atvh_system* s, NptState& q, double half_dt) {
    const double quarter_dt = 0.5 * half_dt, eighth_dt = 0.5 * quarter_dt;
    const int M = std::min(q.mtchain, kNptMaxChain);
    double* ed = q.eta_t_dot;
    for (int k = M - 1; k >= 0; --k) {
        const double expfac = (k < M - 1) ? std::exp(-ed[k + 1] * eighth_dt) : 1.0;
        ed[k] *= expfac;
        const double G = k == 0 ? (2.0 * q.ke_global - static_cast<double>(s->ndof) * q.kT) / q.Q_t[0]
                                : (q.Q_t[k - 1] * ed[k - 1] * ed[k - 1] - q.kT) / q.Q_t[k];
        ed[k] += G * quarter_dt;
        ed[k] *= q.tdrag;
        ed[k] *= expfac;
    }
    const double scale = std::exp(-ed[0] * half_dt);
    q.ke_global *= scale * scale;
    for (int k = 0; k < M; ++k)
        q.eta_t[k] += ed[k] * half_dt;
    return scale;
}

void atoms_chain_forward(const atvh_system* s, NptState& q, double half_dt) {
    const double quarter_dt = 0.5 * half_dt, eighth_dt = 0.5 * quarter_dt;
    const int M = std::min(q.mtchain, kNptMaxChain);
    double* ed = q.eta_t_dot;
    for (int k = 0; k < M; ++k) {
        const double expfac = (k < M - 1) ? std::exp(-ed[k + 1] * eighth_dt) : 1.0;
        ed[k] *= expfac;
        const double G = k == 0 ? (2.0 * q.ke_global - static_cast<double>(s->ndof) * q.kT) / q.Q_t[0]
                                : (q.Q_t[k - 1] * ed[k - 1] * ed[k - 1] - q.kT) / q.Q_t[k];
        ed[k] += G * quarter_dt;
        ed[k] *= expfac;
    }
}

// nhc_baro_half (:346-421)
void baro_chain(NptState& q, double half_dt) {
    const double quarter_dt = 0.5 * half_dt, eighth_dt = 0.5 * quarter_dt;
    const int M = std::min(q.mpchain, kNptMaxChain);
    const double W_axis = q.W / 3.0;
    double* ed = q.eta_b_dot;
    auto G = [&](int k) -> double {
        if (k == 0) {
            double two_ke = 0.0;
            for (int d = 0; d < 3; ++d)
                if (q.couple[d])
                    two_ke += W_axis * q.eps_dot[d] * q.eps_dot[d];
            return (two_ke - q.kT) / q.Q_b[0];
        }
        return (q.Q_b[k - 1] * ed[k - 1] * ed[k - 1] - q.kT) / q.Q_b[k];
    };
    for (int k = M - 1; k >= 0; --k) {
        const double expfac = (k < M - 1) ? std::exp(-ed[k + 1] * eighth_dt) : 1.0;
        ed[k] *= expfac;
        ed[k] += G(k) * quarter_dt;
        ed[k] *= q.pdrag;
        ed[k] *= expfac;
    }
    const double baro_scale = std::exp(-ed[0] * half_dt);
    for (int d = 0; d < 3; ++d)
        if (q.couple[d])
            q.eps_dot[d] *= baro_scale;
    for (int k = 0; k < M; ++k)
        q.eta_b[k] += ed[k] * half_dt;
    for (int k = 0; k < M; ++k) {
        const double expfac = (k < M - 1) ? std::exp(-ed[k + 1] * eighth_dt) : 1.0;
        ed[k] *= expfac;
        ed[k] += G(k) * quarter_dt;
        ed[k] *= expfac;
    }
}

// update_eps_dot (:459-581), without the couple-xy branch; ke_axis is
// 0.5 mvv2e sum m v_a^2 of the current velocities
void update_eps_dot(const atvh_system* s, NptState& q, const double ke_axis[3], double half_dt) {
    const double V = s->box[0] * s->box[1] * s->box[2];
    if (V <= 0.0)
        return;
    double diag[3] = {q.vir[0] + 2.0 * ke_axis[0], q.vir[1] + 2.0 * ke_axis[1], q.vir[2] + 2.0 * ke_axis[2]};
    if (q.pdim == 3) {
        const double ave = (1.0 / 3.0) * (diag[0] + diag[1] + diag[2]);
        diag[0] = diag[1] = diag[2] = ave;
    }
    double mtk_active_ke = 0.0;
    for (int d = 0; d < 3; ++d)
        if (q.couple[d])
            mtk_active_ke += ke_axis[d];
    const int64_t N_atoms = (s->ndof + 3) / 3;
    const double mtk_term =
        (q.pdim > 0) ? 6.0 * mtk_active_ke / (static_cast<double>(q.pdim) * N_atoms) : 0.0;
    for (int d = 0; d < 3; ++d) {
        if (!q.couple[d])
            continue;
        const double P_axis = diag[d] / V;
        const double G_eps = (3.0 * V * (P_axis - q.P_target) + mtk_term) / q.W;
        q.eps_dot[d] += G_eps * half_dt;
    }
}

// scale_v_by_baro's factors (:591-620)
void baro_v_scales(const atvh_system* s, const NptState& q, double half_dt, double out[3]) {
    double sum_eps_dot = 0.0;
    for (int d = 0; d < 3; ++d)
        if (q.couple[d])
            sum_eps_dot += q.eps_dot[d];
    const int64_t N_atoms = (s->ndof + 3) / 3;
    const double mtk_term2 = (q.pdim > 0) ? sum_eps_dot / (static_cast<double>(q.pdim) * N_atoms) : 0.0;
    for (int d = 0; d < 3; ++d)
        out[d] = std::exp(-(q.eps_dot[d] + mtk_term2) * half_dt);
}

// extended_hamiltonian (:785-820)
double hamiltonian(const atvh_system* s, const NptState& q, double pe) {
    double H = q.ke_global + pe;
    const double W_axis = q.W / 3.0;
    for (int d = 0; d < 3; ++d)
        if (q.couple[d])
            H += 0.5 * W_axis * q.eps_dot[d] * q.eps_dot[d];
    for (int k = 0; k < q.mtchain; ++k)
        H += 0.5 * q.Q_t[k] * q.eta_t_dot[k] * q.eta_t_dot[k];
    for (int k = 0; k < q.mpchain; ++k)
        H += 0.5 * q.Q_b[k] * q.eta_b_dot[k] * q.eta_b_dot[k];
    H += static_cast<double>(s->ndof) * q.kT * q.eta_t[0];
    for (int k = 1; k < q.mtchain; ++k)
        H += q.kT * q.eta_t[k];
    for (int k = 0; k < q.mpchain; ++k)
        H += q.kT * q.eta_b[k];
    H += q.P_target * (s->box[0] * s->box[1] * s->box[2]);
    return H;
}

constexpr double kHalfMvv2e = 0.5 * mass_v2_to_energy;

// initial_integrate P1-P6 (NoseHooverNPT::launch, :826-856)
void npt_initial(atvh_system* s, NptState& q) {
    const double half_dt = 0.5 * s->dt;
    temp_target(s, q);
    baro_chain(q, half_dt);                                 // P1
    const double scale = atoms_chain_backward(s, q, half_dt); // P2: backward sweep
    NptArgs a = npt_args(s);
    a.s0 = q.pending;
    a.s1 = scale;
    q.pending = 1.0;
    {
        ScopedTimer tm(s->ctx, "integrate");
        k_npt_chain_scale<<<a.nb, kBlock, 0, s->ctx->stream>>>(a);
        ATVH_HIP(hipGetLastError());
    }
    reduce(s, a, 3);
    DevCtl h{};
    read_ctl(s, h);
    atoms_chain_forward(s, q, half_dt); // P2: forward sweep
    const double ke_axis[3] = {kHalfMvv2e * h.red[0], kHalfMvv2e * h.red[1], kHalfMvv2e * h.red[2]};
    update_eps_dot(s, q, ke_axis, half_dt); // P3 (the last force evaluation's virial)
    baro_v_scales(s, q, half_dt, a.sa);     // P4 (its KE bookkeeping is overwritten at end_of_step)
    for (int d = 0; d < 3; ++d) {           // P6
        const double eps_dt = std::clamp(q.eps_dot[d] * s->dt, -kEpsDtMaxFraction, kEpsDtMaxFraction);
        a.a[d] = std::exp(eps_dt);
        a.drift[d] = expm1c(eps_dt * 0.5);
    }
    for (int d = 0; d < 3; ++d)
        s->box[d] *= a.a[d];
    a.lx = s->box[0];
    a.ly = s->box[1];
    a.lz = s->box[2];
    {
        ScopedTimer tm(s->ctx, "integrate");
        k_npt_drift<<<a.nb, kBlock, 0, s->ctx->stream>>>(a);
        ATVH_HIP(hipGetLastError());
    }
}

// decide + neighbor (Runtime::step_verlet_ with NeighborList::decide_reneighbor):
// `ago` and `nrebuilds` are the control block's counters, kept on the host
// for the length of a run.
void npt_decide_rebuild(atvh_system* s, int& ago, int& nrebuilds) {
    ago += 1;
    const bool delay_ok = ago >= s->delay;
    const bool every_ok = (ago % std::max(1, s->every)) == 0;
    if (!(delay_ok && every_ok))
        return;
    const bool stale = s->box[0] != s->nl_box[0] || s->box[1] != s->nl_box[1] || s->box[2] != s->nl_box[2];
    bool rebuild = !s->check || stale;
    if (!rebuild) {
        DevCtl t{};
        read_ctl(s, t);
        rebuild = t.disp_flag != 0;
    }
    if (!rebuild)
        return;
    if (stale)
        nl_setup_grid(s); // CellList::compute_grid + build_stencil for the new box
    nl_build(s, true);    // records the box it was built in
    ago = 0;
    nrebuilds += 1;
}

// final_integrate P7 + end_of_step P8-P11 (:857-888); returns the step's PE
double npt_final_end(atvh_system* s, NptState& q) {
    const double half_dt = 0.5 * s->dt;
    NptArgs a = npt_args(s);
    baro_v_scales(s, q, half_dt, a.sa); // P8's factors (eps_dot unchanged since P3)
    {
        ScopedTimer tm(s->ctx, "integrate");
        k_npt_kick_scale<<<a.nb, kBlock, 0, s->ctx->stream>>>(a);
        ATVH_HIP(hipGetLastError());
    }
    reduce(s, a, 5);
    DevCtl h{};
    read_ctl(s, h);
    for (int d = 0; d < 3; ++d)
        q.vir[d] = h.vir[d];
    q.ke_global = kHalfMvv2e * h.red[3]; // compute_global_KE after P7
    if (a.sa[0] == a.sa[1] && a.sa[1] == a.sa[2])
        q.ke_global *= a.sa[0] * a.sa[0];
    else
        q.ke_global = kHalfMvv2e * h.red[4];
    const double ke_axis[3] = {kHalfMvv2e * h.red[0], kHalfMvv2e * h.red[1], kHalfMvv2e * h.red[2]};
    update_eps_dot(s, q, ke_axis, half_dt);                   // P9
    q.pending = atoms_chain_backward(s, q, half_dt);          // P10 (v scaled lazily)
    atoms_chain_forward(s, q, half_dt);
    baro_chain(q, half_dt);                                   // P11
    return h.pe;
}

} // namespace

void npt_free(atvh_system* s) {
    delete s->npt;
    s->npt = nullptr;
}

void npt_run(atvh_system* s, int nsteps, int precision, double* pe, double* ke) {
    NptState& q = *s->npt;
    if (s->pme)
        throw Error(ATVH_E_STATE, "NoseHooverNPT: KSpacePME on a moving box is not supported");
    DevCtl h{};
    read_ctl(s, h);
    if (!q.built)
        for (int d = 0; d < 3; ++d)
            q.vir[d] = h.vir[d]; // the frame's stress: the last evaluation that requested it
    npt_build(s, q);
    q.H.assign(nsteps, 0.0);
    q.V.assign(nsteps, 0.0);
    int ago = h.ago, nrebuilds = h.nrebuilds;
    for (int step = 0; step < nsteps; ++step) {
        npt_initial(s, q);
        npt_decide_rebuild(s, ago, nrebuilds);
        // pre_force + force: the virial is always requested (requires_system_stress)
        pair_enqueue(s, precision, true, true);
        energy_enqueue_finalize(s, nullptr, step);
        const double e = npt_final_end(s, q);
        if (pe)
            pe[step] = e;
        if (ke)
            ke[step] = q.ke_global;
        q.H[step] = hamiltonian(s, q, e);
        q.V[step] = s->box[0] * s->box[1] * s->box[2];
        q.step += 1;
    }
    if (q.pending != 1.0) {
        NptArgs a = npt_args(s);
        a.s0 = q.pending;
        k_npt_scale<<<a.nb, kBlock, 0, s->ctx->stream>>>(a);
        ATVH_HIP(hipGetLastError());
        q.pending = 1.0;
    }
    // the decide counters and the end-of-step KE live in the control block
    read_ctl(s, h);
    h.ago = ago;
    h.nrebuilds = nrebuilds;
    h.ke = q.ke_global;
    ATVH_HIP(hipMemcpy(s->ctl, &h, sizeof(h), hipMemcpyHostToDevice));
}

} // namespace atvh

using namespace atvh;

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

int atvh_npt_configure(atvh_system* s, double dt, double T_target, double tau_T, double P_target, double tau_P,
                       double drag, int mtchain, int mpchain, const int couple[3], double T_stop) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    if (!(dt > 0.0))
        throw Error(ATVH_E_ARG, "NoseHooverNPT: dt must be positive");
    if (!(T_target > 0.0) || !(tau_T > 0.0) || !(tau_P > 0.0))
        throw Error(ATVH_E_ARG, "NoseHooverNPT: T_target, tau_T and tau_P must be positive");
    if (mtchain < 1 || mpchain < 1 || mtchain > kNptMaxChain || mpchain > kNptMaxChain)
        throw Error(ATVH_E_ARG, "NoseHooverNPT: chain lengths must be in [1, 8]");
    auto* q = new NptState();
    q->T_target = T_target;
    q->T_stop = T_stop < 0.0 ? T_target : T_stop; // -1 folds to T_target (no ramp)
    q->tau_T = tau_T;
    q->P_target = P_target;
    q->tau_P = tau_P;
    q->drag = drag;
    q->mtchain = mtchain;
    q->mpchain = mpchain;
    q->pdim = 0;
    for (int d = 0; d < 3; ++d) {
        q->couple[d] = couple ? (couple[d] != 0) : 1;
        q->pdim += q->couple[d];
    }
    npt_free(s);
    s->npt = q;
    // replaces VelocityVerlet + NoseHoover
    s->dt = dt;
    s->thermostat = 0;
    s->T_target = T_target;
    s->tau_T = tau_T;
    s->ndof = 3 * s->n - 3; // publish_global_dof (src/cpu/driver/dof.cpp:49)
    s->integ_configured = true;
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_npt_set_run_span(atvh_system* s, int64_t nfirst, int64_t nlast) {
    if (!s)
        return ATVH_E_ARG;
    if (!s->npt) {
        s->ctx->err = "atvh_npt_set_run_span: no NoseHooverNPT configured";
        return ATVH_E_STATE;
    }
    s->npt->nfirst = nfirst;
    s->npt->nlast = nlast;
    return ATVH_OK;
}

int atvh_npt_thermo(atvh_system* s, double* H, double* V, int nsteps) {
    if (!s)
        return ATVH_E_ARG;
    if (!s->npt || nsteps > static_cast<int>(s->npt->H.size())) {
        s->ctx->err = "atvh_npt_thermo: more steps requested than recorded";
        return ATVH_E_STATE;
    }
    for (int k = 0; k < nsteps; ++k) {
        if (H)
            H[k] = s->npt->H[k];
        if (V)
            V[k] = s->npt->V[k];
    }
    return ATVH_OK;
}

int atvh_npt_state(atvh_system* s, double* eps_dot3, double* eta_t, double* eta_t_dot, double* eta_b,
                   double* eta_b_dot) {
    if (!s)
        return ATVH_E_ARG;
    if (!s->npt) {
        s->ctx->err = "atvh_npt_state: no NoseHooverNPT configured";
        return ATVH_E_STATE;
    }
    const NptState& q = *s->npt;
    for (int d = 0; d < 3 && eps_dot3; ++d)
        eps_dot3[d] = q.eps_dot[d];
    for (int k = 0; k < q.mtchain; ++k) {
        if (eta_t)
            eta_t[k] = q.eta_t[k];
        if (eta_t_dot)
            eta_t_dot[k] = q.eta_t_dot[k];
    }
    for (int k = 0; k < q.mpchain; ++k) {
        if (eta_b)
            eta_b[k] = q.eta_b[k];
        if (eta_b_dot)
            eta_b_dot[k] = q.eta_b_dot[k];
    }
    return ATVH_OK;
}

} // extern "C"
