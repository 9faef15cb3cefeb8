// atvh_internal.hpp — shared definitions of libatv_hip.so (gfx950).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "atvh.h"

namespace atvh {

// include/atomiverse/core/units.h:138,165-166,173 (metal units)
constexpr double kB_ev_per_k = 8.617333262e-5;
constexpr double force_to_accel = 1.0 / 103.6427;
constexpr double mass_v2_to_energy = 103.6427;
constexpr double coulomb = 14.399645;

constexpr int kMaxTypes = 16;
constexpr int kBlock = 256;       // threads per workgroup (4 wave64)
constexpr int kMaxGrid = 8192;    // cap for grid-stride kernels that may be predicated off
constexpr int kPartialStride = 8; // doubles per block partial (pe, vir[6], spare)

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define ATVH_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::atvh::Error(ATVH_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device-resident control block of one system. Written only by device code
// during a run; the host reads it at synchronisation points.
struct DevCtl {
    double xi;        // Nose-Hoover thermostat variable
    double nh_scale;  // NH velocity scale not yet applied to v (lazy)
    double pe;        // last potential energy
    double vir[6];    // last pair virial (Voigt xx yy zz xy xz yz)
    double ke;        // kinetic energy at end of the last step
    double ke_half;   // kinetic energy after the last kick (NH input)
    int disp_flag;    // some atom moved more than skin/2 since the last build
    int rebuild;      // rebuild decided for the current step
    int ago;          // steps since the last rebuild (LAMMPS `ago`)
    int nrebuilds;    // total builds including the initial one
    int overflow;     // neighbor capacity exceeded (count of atoms)
    int max_nbr;      // max neighbors of the last build
    int always;       // constant 1: predicate for unconditional launches
    int halo_overflow; // bricks whose halo exceeded the LDS capacity
    long long nbr_total; // full-list entries of the last build
    double e_kspace[3];  // last k-space energy parts: reciprocal, self, net charge
    int rec_step;        // next thermo record index of a batched run
    double nh_Q;         // Nose-Hoover mass Q = Ndof kB T tau^2
    double nh_ndof_kT;   // Ndof kB T
    double T_target;     // thermostat temperature (moves with replica exchanges)
    double stress_kin[6]; // last kinetic stress sum m v_a v_b (eV, Voigt xx yy zz xy xz yz)
    int steps_done;       // steps of the current run completed (advanced by the end-of-step kernel)
    double red[6];        // fixed-order sums of the last NoseHooverNPT reduction
};

// A rebuild that overflowed the neighbor capacity or a brick halo leaves the
// system halted: the integration kernels of every later phase leave at
// their first instruction, so no step integrates forces from a truncated
// list. The host regrows the list at the halted step and resumes there.
__device__ __forceinline__ bool halted(const DevCtl* c) { return (c->overflow | c->halo_overflow) != 0; }

struct CoulState; // PairLJCutCoulLong tables (coul_pme.hip)
struct PmeState;  // KSpacePME mesh, plans and tables (coul_pme.hip)
struct NptState;  // NoseHooverNPT chains, barostat and records (npt.hip)

struct TimerSlot {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0.0;
    int64_t launches = 0;
};

} // namespace atvh

struct atvh_ctx {
    int device = 0;
    int live_systems = 0; // systems created on this context and not yet destroyed
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    bool timing = false;
    std::string timing_only; // empty: every kernel family
    std::map<std::string, atvh::TimerSlot> timers;
    std::vector<hipEvent_t> event_pool;
};

struct atvh_system {
    atvh_ctx* ctx = nullptr;
    int64_t n = 0;
    int64_t npad = 0; // column stride of the ELL neighbor list
    double box[3] = {0, 0, 0};

    // per-atom device arrays, cell-sorted order
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *fx = nullptr, *fy = nullptr, *fz = nullptr;
    double* q = nullptr;
    int* type = nullptr;
    int* tag = nullptr; // frame row of the atom now at this slot
    // permutation scratch (same layout)
    double *ax = nullptr, *ay = nullptr, *az = nullptr;
    double *avx = nullptr, *avy = nullptr, *avz = nullptr;
    double *afx = nullptr, *afy = nullptr, *afz = nullptr;
    double* aq = nullptr;
    int* atype = nullptr;
    int* atag = nullptr;
    // x_save for the displacement check
    double *xs = nullptr, *ys = nullptr, *zs = nullptr;

    int ntypes = 1;
    double mass_t[atvh::kMaxTypes] = {1.0};
    bool has_charge = false;

    // neighbor list
    bool nl_configured = false;
    bool nl_built = false;
    double cutoff = 0, skin = 0, cutoff_skin = 0;
    double nl_box[3] = {0, 0, 0}; // box of the last list build (a different box makes it stale)
    int every = 1, delay = 10, check = 1;
    int nc[3] = {0, 0, 0};
    double cs[3] = {0, 0, 0};
    int reach[3] = {0, 0, 0};
    int ncells = 0;
    bool use_half = true;
    int nstencil = 0;
    int* d_stencil = nullptr; // int4 {di,dj,dk,kind}
    int ncols = 0;
    int* d_cols = nullptr;    // int4 {di,dj,kmax,0}: the stencil as z-columns
    int* cell_of = nullptr;   // [n] cell of each slot (after sort: of the sorted slot)
    int* acell_of = nullptr;
    int* cell_count = nullptr;
    int* cell_start = nullptr; // [ncells+1]
    int* cell_fill = nullptr;
    int* perm = nullptr;
    int* scan_sums = nullptr;
    int nscan_blocks = 0;
    int* nbr = nullptr;      // small-grid mode: ELL, column-major: nbr[k*npad + i]
    int* numneigh = nullptr; // [n]
    int cap = 0;
    // brick mode (every axis has >= 2*reach+1 cells)
    bool brick_mode = false;
    int brick_pick = 0;     // index of the chosen brick candidate
    int brick_pick_min = 0; // smallest candidate index allowed (raised on halo overflow)
    int bdim[3] = {0, 0, 0};
    int nb[3] = {0, 0, 0};
    int nbricks = 0;
    int* cell_rank = nullptr;       // reference cell index -> layout rank
    int* brick_cell_base = nullptr; // [nbricks+1]
    uint16_t* nbr16 = nullptr;      // per brick, 8-wide groups: [a0*cap + (g*nat + li)*8 + u]
    uint16_t* self16 = nullptr;     // [n] own row in the brick's halo
    int* halo_cells = nullptr;      // [nbricks*kHaloCellCap] static: cell rank | image code << 26
    int* halo_rows = nullptr;       // [nbricks*kHaloCap] global slot | image code << 26
    int* halo_off = nullptr;        // [nbricks*(kHaloCellCap+1)] halo row offset per halo cell
    int* halo_n = nullptr;          // [nbricks] rows in each halo
    int2* brick_range = nullptr;    // [nbricks] (first slot, atom count)

    // pair style: 0 = PairLJCut, 1 = PairLJCutCoulLong
    bool pair_configured = false;
    int pair_style = 0;
    atvh::CoulState* coul = nullptr;
    // k-space (KSpacePME), evaluated after the pair style in the force phase
    atvh::PmeState* pme = nullptr;
    // charge sums of the uploaded frame (ascending atom order)
    double q_sum = 0, q_sq = 0, q_abs = 0, q_max = 0;
    double lj_eps = 1, lj_sigma = 1, lj_cutoff = 0;
    int lj_ntypes = 0;
    std::vector<double> lj_eps_tab, lj_sigma_tab;
    double* d_ljtab = nullptr; // per type pair: {lj1=24eps*s6*2? see pair kernel}, packed

    // integrator
    bool integ_configured = false;
    double dt = 0;
    int thermostat = 0;
    double T_target = 0, tau_T = 0;
    int64_t ndof = 0;

    // reductions
    double* partials = nullptr;
    int npartials = 0;
    atvh::DevCtl* ctl = nullptr;

    // home atoms a list build takes per brick (0: the LDS-bound maximum)
    int brick_atom_limit = 0;

    // MTK NoseHooverNPT in place of VelocityVerlet + NoseHoover (npt.hip)
    atvh::NptState* npt = nullptr;

    // thermo record of the last md_run
    double* d_thermo = nullptr;
    int thermo_cap = 0;
    int thermo_n = 0;
};

// One stage of a batched launch: a device array of per-system argument
// blocks and the prefix of their block counts (nsys + 1 entries).
struct BatchStage {
    void* args = nullptr;
    int* base = nullptr;
    int nblocks = 0;
    int nmembers = 0; // entries of args (base holds nmembers + 1): the member search's range
    int* owner = nullptr; // [nblocks] member of every block (one load instead of a search)
};

struct atvh_batch {
    atvh_ctx* ctx = nullptr;
    std::vector<atvh_system*> sys;
    std::map<std::string, BatchStage> stages;
    atvh::DevCtl** d_ctls = nullptr; // member control blocks
    int* nl_nb = nullptr;            // [kNlStages][nsys] block counts of the rebuild stages
    int* nl_act = nullptr;           // [kNlStages][nsys+1] this step's prefix over rebuilding members
    unsigned char* nl_flags = nullptr; // [nsys] this step's rebuild flags
    int* nl_next = nullptr;            // work counter of the pair-build stage (k_batched_queue)
    bool typed = false;
    // members whose force call is not the batched PairLJCut stage (charged
    // LJ + Coulomb with or without PME): queued one by one inside the step
    std::vector<atvh_system*> solo;
    bool use_graph = false;
    hipGraphExec_t graph[2] = {nullptr, nullptr}; // [energy step?]
    int graph_prec[2] = {-1, -1};
    int thermo_cap = 0;
    int* d_slot_member = nullptr; // replica exchange: member holding ladder slot k
    int* d_accept = nullptr;      // accepted exchanges of the last round
    int* d_halted = nullptr;      // [nsys] members a rebuild overflow halted in the last run
};

namespace atvh {

inline int grid_for(int64_t n, int block = kBlock) {
    int64_t g = (n + block - 1) / block;
    if (g < 1)
        g = 1;
    return static_cast<int>(g);
}
inline int grid_capped(int64_t n, int block = kBlock) {
    int g = grid_for(n, block);
    return g > kMaxGrid ? kMaxGrid : g;
}

// Timing helpers (HIP events on the context stream, around one kernel family).
struct ScopedTimer {
    atvh_ctx* ctx;
    const char* name;
    hipEvent_t start = nullptr;
    ScopedTimer(atvh_ctx* c, const char* nm);
    ~ScopedTimer();
};
hipEvent_t take_event(atvh_ctx* c);
bool timing_on(const atvh_ctx* c, const char* nm);

// One kernel launch on stream st; when the context times the family `name`,
// the kernel itself carries the start and stop events (hipExtLaunchKernelGGL),
// so the interval is the kernel's own and no marker is queued beside it.
template <class... KArgs, class... Args>
void launch_timed(atvh_ctx* c, const char* name, void (*kernel)(KArgs...), dim3 grid, dim3 block,
                  hipStream_t st, Args... args) {
    if (!timing_on(c, name)) {
        hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
        return;
    }
    hipEvent_t start = take_event(c), stop = take_event(c);
    hipExtLaunchKernelGGL(kernel, grid, block, 0, st, start, stop, 0, args...);
    c->timers[name].pending.emplace_back(start, stop);
}

template <class T>
void dev_alloc(T** p, size_t count) {
    if (count == 0)
        count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess)
        throw Error(ATVH_E_NOMEM, std::string("hipMalloc(") + std::to_string(count * sizeof(T)) +
                                      " B): " + hipGetErrorString(e));
}
template <class T>
void dev_free(T*& p) {
    if (p) {
        (void)hipFree(p);
        p = nullptr;
    }
}

// neighbor.hip
constexpr int kNlStages = 7; // kernels of one brick-mode rebuild
void nl_setup_grid(atvh_system* s);
void nl_enqueue_rebuild(atvh_system* s, const int* pred, bool permute_forces);
void nl_enqueue_decide(atvh_system* s);
void nl_free(atvh_system* s);
void nl_realloc_capacity(atvh_system* s, int cap);
void nl_check_ctl(const DevCtl& h);
// List build at the current positions, growing the capacity or stepping to a
// smaller brick shape until it fits. keep_counts: a regrow inside a run (the
// rebuild counter keeps the rebuild the device already counted).
void nl_build(atvh_system* s, bool keep_counts);

// pair_lj.hip
int pair_grid(const atvh_system* s);
void pair_enqueue(atvh_system* s, int precision, bool eflag, bool vflag);
void pair_free(atvh_system* s);

// coul_pme.hip
void ljc_enqueue(atvh_system* s, int precision, bool eflag, bool vflag);
void pme_enqueue(atvh_system* s, int precision, bool ev);
int pme_partial_blocks(const atvh_system* s);
// true when the PME buffers and transform plans of this precision exist (or
// no PME is configured), so a step allocates nothing
bool pme_ready(const atvh_system* s, int precision);
int pme_grid_blocks(const atvh_system* s, int precision); // energy partial blocks of one evaluation
void pme_make_buffers(atvh_system* s, int precision);
// makes and warms every buffer and plan an evaluation of this precision needs
void pme_prepare(atvh_system* s, int precision);
const double* pme_partials(const atvh_system* s);
void pme_energy_consts(const atvh_system* s, double& escale, double& e_self, double& e_net);
// device bytes one FFT plan pair took (largest seen; 64 MiB before the first)
int64_t fft_pair_bytes();
void coul_free(atvh_system* s);

// batched launches (batch.hip + the module builders)
template <class Args>
void batch_stage(atvh_batch* b, const std::string& name, const std::vector<Args>& args,
                 const std::vector<int>& blocks);
void batch_free_stages(atvh_batch* b);
void integ_batch_build(atvh_batch* b);
void integ_batch_launch(atvh_batch* b, const char* stage);
void nl_batch_build(atvh_batch* b);
void nl_batch_launch_rebuild(atvh_batch* b);
void pair_batch_build(atvh_batch* b);
void pair_batch_launch(atvh_batch* b, int precision, bool ef);

// device bytes held per component (atvh_system_memory)
int64_t nl_device_bytes(const atvh_system* s);
int64_t pair_device_bytes(const atvh_system* s);
int64_t coul_device_bytes(const atvh_system* s);
int64_t pme_device_bytes(const atvh_system* s);

// npt.hip
// nsteps of the fused loop with NoseHooverNPT as the integrator; per-step
// potential/kinetic energy go to pe/ke (host arrays, may be null)
void npt_run(atvh_system* s, int nsteps, int precision, double* pe, double* ke);
void npt_free(atvh_system* s);

// integrate.hip
void integ_enqueue_initial(atvh_system* s);
// Queue steps [first, nsteps) of a fused run; resume: the first step starts
// at its force phase (its initial integrate, decide and rebuild already ran).
// Step k's potential and kinetic energy go to pe[k] and ke[k].
struct ThermoRec {
    double* pe;
    double* ke;
};
ThermoRec thermo_rec(atvh_system* s); // the system's own per-step record
void md_enqueue_steps(atvh_system* s, int first, int nsteps, int precision, int energy_every, bool resume,
                      ThermoRec rec);
// Regrow the list of a system halted by an overflow and finish its run alone
// (bit-identical to the batched step); returns after the stream drains.
void md_recover_and_finish(atvh_system* s, int nsteps, int precision, int energy_every);
// Looks up this device's copy of the allocation-free entry point's static
// scratch array (once per device, with the device current).
void noalloc_bind(int device);
void integ_enqueue_final(atvh_system* s);
void energy_enqueue_finalize(atvh_system* s, double* record_pe, int step, bool with_pair = true);

} // namespace atvh

namespace atvh {

template <class Args>
void batch_stage(atvh_batch* b, const std::string& name, const std::vector<Args>& args,
                 const std::vector<int>& blocks) {
    BatchStage& st = b->stages[name];
    dev_free(st.args);
    dev_free(st.base);
    dev_free(st.owner);
    std::vector<int> base(blocks.size() + 1, 0);
    for (size_t k = 0; k < blocks.size(); ++k)
        base[k + 1] = base[k] + blocks[k];
    std::vector<int> owner(std::max(base.back(), 1), 0);
    for (size_t k = 0; k < blocks.size(); ++k)
        for (int b = base[k]; b < base[k + 1]; ++b)
            owner[b] = static_cast<int>(k);
    dev_alloc(&st.owner, owner.size());
    ATVH_HIP(hipMemcpy(st.owner, owner.data(), owner.size() * sizeof(int), hipMemcpyHostToDevice));
    Args* d = nullptr;
    dev_alloc(&d, args.size());
    st.args = d;
    dev_alloc(&st.base, base.size());
    ATVH_HIP(hipMemcpy(d, args.data(), args.size() * sizeof(Args), hipMemcpyHostToDevice));
    ATVH_HIP(hipMemcpy(st.base, base.data(), base.size() * sizeof(int), hipMemcpyHostToDevice));
    st.nblocks = base.back();
    st.nmembers = static_cast<int>(args.size());
}

} // namespace atvh
