/* atvh.h — C ABI of libatv_hip.so, the MI355X (gfx950) classical-MD backend.
 *
 * The hip:: C++ components (include/atomiverse/hip/) sit on top of this ABI,
 * exactly where the reference's cuda:: components call their kernels
 * directly. Conventions:
 *   - every function returns ATVH_OK (0) or a negative ATVH_E_* code and
 *     leaves a message for atvh_last_error(); no C++ exception crosses it;
 *   - host buffers are caller-owned and read/written only during the call;
 *     device buffers are library-owned;
 *   - all work of a context is ordered on the context's one HIP stream; a
 *     context is driven by one host thread at a time;
 *   - per-atom host arrays are in the caller's row order ("frame order");
 *     the library keeps atoms cell-sorted on the device and maps back by tag.
 * Units are the reference's metal units (Å, eV, fs, amu, K, e).
 */
#ifndef ATVH_H
#define ATVH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATVH_OK 0
#define ATVH_E_ARG (-1)      /* invalid argument */
#define ATVH_E_HIP (-2)      /* HIP runtime error */
#define ATVH_E_STATE (-3)    /* call out of order (e.g. no neighbor list) */
#define ATVH_E_CAPACITY (-4) /* a device capacity was exceeded */
#define ATVH_E_NOMEM (-5)    /* device allocation failed */

#define ATVH_PREC_FP64 0  /* plain: FP64 throughout */
#define ATVH_PREC_MIXED 1 /* FP32 pair/PME math; FP64 integration + accumulation */

#define ATVH_PHASE_INITIAL 0 /* md::Phase::initial_integrate */
#define ATVH_PHASE_FINAL 1   /* md::Phase::final_integrate */

typedef struct atvh_ctx atvh_ctx;
typedef struct atvh_system atvh_system;
typedef struct atvh_batch atvh_batch;
typedef struct atvh_server atvh_server;

/* ---- context --------------------------------------------------------------
 * Replaces cuda::Backend::init's stream set-up (src/cuda/driver/driver.cu:25-32)
 * and the ATV_CUDA_CHECK error path (include/atomiverse/cuda/device/error.cuh).
 */
int atvh_ctx_create(int device, atvh_ctx** out);
/* Fails with ATVH_E_STATE (and frees nothing) while systems created on the
 * context are still alive. */
int atvh_ctx_destroy(atvh_ctx* ctx);
const char* atvh_last_error(const atvh_ctx* ctx);
/* Use an external stream (hipStream_t passed as void*); NULL restores the
 * context's own stream. */
int atvh_ctx_set_stream(atvh_ctx* ctx, void* stream);
void* atvh_ctx_stream(atvh_ctx* ctx);
int atvh_ctx_synchronize(atvh_ctx* ctx);
/* Per-kernel timing on the context stream with HIP events (off by default).
 * name = kernel family ("pair", "integrate", "nlist", "pme", ...). */
int atvh_ctx_enable_timing(atvh_ctx* ctx, int enable);
/* Restrict the timing events to one kernel family (NULL: every family).
 * Every event pair records two markers into the stream, so a throughput
 * measurement times only the kernel it needs. */
int atvh_ctx_timing_only(atvh_ctx* ctx, const char* name);
int atvh_ctx_kernel_time(atvh_ctx* ctx, const char* name, double* total_ms,
                         int64_t* launches);
int atvh_ctx_reset_timing(atvh_ctx* ctx);
/* Free and total device memory of the context's device (hipMemGetInfo). */
int atvh_ctx_mem_info(atvh_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);

/* ---- frame (device state) ---------------------------------------------------
 * Replaces cuda::Frame + cpu::to_cuda / cuda::to_cpu
 * (include/atomiverse/cuda/frame/frame.cuh:71-200;
 *  src/cuda/frame/transfer.cu:145-475). Orthorhombic box [0,L).
 * Positions and velocities are FP64 in both precision modes.
 */
int atvh_system_create(atvh_ctx* ctx, int64_t n, const double box[3],
                       atvh_system** out);
int atvh_system_destroy(atvh_system* sys);
int64_t atvh_system_natoms(const atvh_system* sys);
/* Upload frame-order arrays. vx..vz, fx..fz, type, charge may be NULL
 * (zero / type 0 / neutral). mass[n] is required; atoms of one type must
 * share one mass. */
int atvh_frame_upload(atvh_system* sys, const double* x, const double* y,
                      const double* z, const double* vx, const double* vy,
                      const double* vz, const double* fx, const double* fy,
                      const double* fz, const double* mass, const int* type,
                      const double* charge);
/* Download into frame order; any pointer may be NULL. Positions come back
 * wrapped into [0, L) exactly as wrap_orthorhombic leaves them. */
int atvh_frame_download(atvh_system* sys, double* x, double* y, double* z,
                        double* vx, double* vy, double* vz, double* fx,
                        double* fy, double* fz);
int atvh_system_box(const atvh_system* sys, double box[3]);
/* Device bytes a system holds for one part: the frame columns and
 * reductions, the neighbor list, the pair style's tables, the PME mesh and
 * plans, the integrator's records. Backs Component::memory_bytes
 * (include/atomiverse/cuda/component/component.cuh:202). */
enum { ATVH_MEM_FRAME = 0, ATVH_MEM_NLIST = 1, ATVH_MEM_PAIR = 2, ATVH_MEM_PME = 3, ATVH_MEM_INTEGRATOR = 4 };
int atvh_system_memory(const atvh_system* sys, int part, int64_t* bytes);

/* ---- neighbor list ------------------------------------------------------------
 * Replaces cuda::NeighborList (include/atomiverse/cuda/neighbor/neighbor_list.cuh:93-229)
 * with the reference CPU semantics: CellList::CellList/compute_grid/
 * build_stencil (src/cpu/neighbor/cell_list.cpp:24-232), counting-sort
 * binning (:234-314), the half-list pair predicate of refresh_half
 * (:539-738) and the LAMMPS decide cadence (neighbor_list.h:79-141).
 * skin < 0 means 0.3*cutoff.
 */
int atvh_nlist_configure(atvh_system* sys, double cutoff, double skin,
                         int every, int delay, int check);
/* Unconditional build (NeighborList::build at component build time). */
int atvh_nlist_build(atvh_system* sys);
/* LAMMPS Neighbor::decide on the device (++ago, displacement check) followed
 * by the rebuild it decides; synchronises once to read the overflow probe
 * (see atvh_nlist_set_capacity). */
int atvh_nlist_decide_and_rebuild(atvh_system* sys);
/* Brick layout of the list (this backend's tiling of the reference cell
 * grid): the candidate index in use (0 = largest brick, 11 = one cell; -1
 * for the small-grid layout) and its size in cells per axis. */
int atvh_nlist_brick_shape(atvh_system* sys, int* candidate, int* bdim);
/* Use candidate k or the next smaller valid one; the list must be built
 * again (atvh_nlist_build). */
int atvh_nlist_set_brick_shape(atvh_system* sys, int candidate);
/* The fullest halo and the fullest brick (home atoms) of the current list;
 * zeros for the small-grid layout. Synchronises. */
int atvh_nlist_brick_fill(atvh_system* sys, int* max_halo, int* max_atoms);
/* Home atoms a list build takes per brick, at most 320 (0 restores 320). A
 * rebuild that finds a fuller brick overflows as one past the LDS halo
 * capacity does: inside atvh_md_run or atvh_batch_run the system halts at
 * that step's force phase, the brick shape steps down and the run resumes
 * (see atvh_nlist_set_capacity). The current list stays valid while none of
 * its bricks is over the new limit. Tests lower it to provoke that path;
 * the reference's cluster lists have no such bound. */
int atvh_nlist_set_brick_limit(atvh_system* sys, int max_atoms);
/* Autotuner: time the current brick shape and the next two smaller ones on
 * this system (one rebuild + `trials` pair launches each, HIP events), keep
 * the fastest and rebuild the list with it. The frame's forces are kept;
 * the rebuild counter restarts. *chosen = candidate (-1: nothing to tune),
 * *best_us = its pair + rebuild/8 time in microseconds. */
int atvh_nlist_autotune(atvh_system* sys, int precision, int trials,
                        int* chosen, double* best_us);
/* Per-atom list capacity (entries, rounded up to a multiple of 8) for the
 * next build; the list must be built again. Builds and runs grow it when a
 * rebuild needs more: the reference sizes its pair arrays from the density
 * and regrows them inside the rebuild (include/atomiverse/cuda/neighbor/
 * neighbor_list.cuh:119, src/cuda/neighbor/neighbor_list.cu:544-553). A
 * rebuild inside atvh_md_run or atvh_batch_run that overflows halts that
 * system at the step's force phase (no later phase integrates anything);
 * the call then regrows the list at those positions and resumes, so the
 * trajectory equals an uninterrupted run with a larger capacity.
 * atvh_nlist_decide_and_rebuild reads the overflow back (one stream
 * synchronisation) and regrows at once. */
int atvh_nlist_set_capacity(atvh_system* sys, int capacity);
/* Host read of the rebuild count (synchronises the stream). */
int atvh_nlist_stats(atvh_system* sys, int64_t* nrebuilds, int64_t* npairs_half,
                     int* max_neighbors, int* capacity);
/* Export the current pair set as unordered tag pairs (a < b, a and b are
 * frame rows), sorted ascending by (a, b). Synchronises. If cap is too small
 * returns ATVH_E_CAPACITY with *npairs = the needed count. */
int atvh_nlist_export_pairs(atvh_system* sys, int64_t* pairs, int64_t cap,
                            int64_t* npairs);

/* ---- pair styles ----------------------------------------------------------------
 * PairLJCut (src/cpu/potentials/pair_lj_cut.cpp:31-42, 257-298, 300-830;
 * CUDA counterpart src/cuda/potentials/pair_lj_cut.cu:89-512).
 * Single-type: ntypes = 0 and (eps, sigma). Per-type: ntypes > 0 with
 * ntypes*ntypes eps/sigma tables (Lorentz-Berthelot already filled).
 */
int atvh_pair_lj_configure(atvh_system* sys, double eps, double sigma,
                           double cutoff, int ntypes, const double* eps_tab,
                           const double* sigma_tab);
/* PairLJCutCoulLong(coeffs, alpha, cutoff)
 * (src/cpu/potentials/pair_lj_cut_coul_long.cpp:34-70, compute :276-571;
 * CUDA counterpart cuda::CoulLong, src/cuda/potentials/coul_long.cu).
 * The coefficient table is ntypes x ntypes, row-major: entries with
 * set_tab != 0 are given (both orders, as the constructor's `put` stores
 * them), the rest are Lorentz-Berthelot mixed from set diagonals or zero.
 * Real-space Coulomb from the ErfcTable of nbin bins (10000 in the
 * reference). Replaces any PairLJCut of the system. */
int atvh_pair_ljcoul_configure(atvh_system* sys, int ntypes,
                               const double* eps_tab, const double* sigma_tab,
                               const int* set_tab, double alpha, double cutoff,
                               int nbin);
/* One force-phase evaluation: the pair style overwrites the forces, then the
 * k-space solver (if configured) adds its forces (ZeroForces folded in).
 * eflag/vflag request the energy and 6-component virial. */
int atvh_pair_compute(atvh_system* sys, int precision, int eflag, int vflag);

/* ---- k-space -----------------------------------------------------------------------
 * KSpacePME(alpha, grid_spacing, order) (src/cpu/longrange/pme.cpp:31-39,
 * compute :147-253; CUDA counterpart cuda::KSpacePME,
 * src/cuda/longrange/pme.cu). Orders 3..8. Mesh K_d = good_fft_size(
 * ceil(L_d / spacing)) (pme.cpp:66-74). Needs per-atom charges in the frame.
 * Once configured it runs in every force phase after the pair style. */
int atvh_pme_configure(atvh_system* sys, double alpha, double grid_spacing,
                       int order);
/* KSpacePME alone into zeroed forces; energy/virial are the k-space terms. */
int atvh_pme_compute(atvh_system* sys, int precision, int eflag, int vflag);
/* Mesh size K[3] (may be null) and the last k-space energy parts
 * e_parts[3] = {reciprocal, self, net charge} (may be null; synchronises). */
int atvh_pme_info(atvh_system* sys, int* K, double* e_parts);
/* Last computed potential energy (eV) and Voigt virial (eV); synchronises. */
int atvh_energy(atvh_system* sys, double* pe, double* virial6, double* ke);

/* ---- stress -------------------------------------------------------------------------
 * The Virial observer: replaces cpu::md::Virial::launch
 * (src/cpu/observe/virial_stress.cpp:14-56, contract virial_stress.h:16-54) and
 * the CUDA kinetic-stress producer with cuda::Virial
 * (src/cuda/observe/stress.cu:1-56). Queue the kinetic stress
 * sum_i m_i v_ia v_ib * mass_v2_to_energy (eV, extensive, Voigt xx yy zz xy
 * xz yz) of the current velocities; no synchronisation. */
int atvh_kinetic_stress(atvh_system* sys);
/* The last queued kinetic stress, the virial of the last force evaluation
 * that requested it (pair + k-space, eV, extensive) and the pressure
 * P = (tr K + tr W) / (3V) in eV/Å^3 (compute_pressure convention, positive
 * is compressive; virial_stress.h:30-45). Any pointer may be NULL;
 * synchronises. */
int atvh_stress(atvh_system* sys, double* kinetic6, double* virial6, double* pressure);

/* ---- integration ------------------------------------------------------------------
 * VelocityVerlet (src/cpu/dynamics/velocity_verlet.cpp:24-68) and single-chain
 * NoseHoover (src/cpu/dynamics/nose_hoover.cpp:31-170); CUDA counterparts
 * src/cuda/dynamics/velocity_verlet.cu:22-90, nose_hoover.cu:50-150.
 * thermostat: 0 = NVE, 1 = Nose-Hoover (T in K, tau in fs). ndof <= 0 -> 3N-3.
 */
int atvh_integrator_configure(atvh_system* sys, double dt, int thermostat,
                              double T_target, double tau_T, int64_t ndof);
/* One integration phase (ATVH_PHASE_INITIAL: kick + drift + wrap + NH;
 * ATVH_PHASE_FINAL: kick + NH). */
int atvh_step_vvnh(atvh_system* sys, int phase);
/* Nose-Hoover state (xi) read/write; synchronises. */
int atvh_thermostat_state(atvh_system* sys, double* xi_out, const double* xi_in);

/* ---- MTK NoseHooverNPT ---------------------------------------------------------------
 * Replaces cuda::md::NoseHooverNPT(T_target, tau_T, P_target, tau_P, dt, drag,
 * mtchain, mpchain, couple_axis, T_stop)
 * (include/atomiverse/cuda/dynamics/nose_hoover_npt.cuh:56-94; algorithm of
 * src/cpu/dynamics/nose_hoover_npt.cpp). It is the system's integrator in
 * place of VelocityVerlet + NoseHoover (atvh_integrator_configure switches
 * back); atvh_md_run then steps it: P1-P6 at initial_integrate, P7 at
 * final_integrate, P8-P11 at end_of_step, the virial requested every step,
 * the box rescaled every step and the cell grid rebuilt for the new box at
 * each list rebuild. Units: K, fs, eV/Å^3 (positive pressure compressive).
 * couple[3] (NULL = all axes): per-axis barostat flags; T_stop < 0: no ramp.
 * Not supported with KSpacePME (E_STATE at the run), in a batch, or in
 * atvh_step_vvnh. */
int atvh_npt_configure(atvh_system* sys, double dt, double T_target, double tau_T, double P_target,
                       double tau_P, double drag, int mtchain, int mpchain, const int couple[3],
                       double T_stop);
/* NoseHooverNPT::set_run_span (nose_hoover_npt.cuh): the temperature ramp's
 * span in cumulative step indices. */
int atvh_npt_set_run_span(atvh_system* sys, int64_t nfirst, int64_t nlast);
/* Per-step MTK conserved quantity H [eV] (frame.scalars->conserved) and box
 * volume [Å^3] of the last atvh_md_run. */
int atvh_npt_thermo(atvh_system* sys, double* H, double* V, int nsteps);
/* Barostat strain rates eps_dot[3] [1/fs] and chain positions/momenta
 * (mtchain / mpchain entries each; any pointer may be NULL). */
int atvh_npt_state(atvh_system* sys, double* eps_dot3, double* eta_t, double* eta_t_dot, double* eta_b,
                   double* eta_b_dot);

/* ---- fused MD loop -----------------------------------------------------------------
 * nsteps of Runtime::step_verlet_ (src/cpu/driver/runtime.cpp:452-502) queued
 * on the stream with no host synchronisation: initial integrate, device-side
 * decide + rebuild, force, final integrate, kinetic energy.
 * energy_every > 0 requests PE/virial on every energy_every-th step and the
 * last step (LAMMPS eflag cadence); 0 = only the last step.
 */
int atvh_md_run(atvh_system* sys, int nsteps, int precision, int energy_every);
/* Device-recorded per-step potential/kinetic energy of the last atvh_md_run
 * (entries for steps without an energy request hold NaN). Synchronises. */
int atvh_md_thermo(atvh_system* sys, double* pe, double* ke, int nsteps);

/* Allocation-free variant of atvh_md_run. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same steps, kernels and results as atvh_md_run; the per-step
 * energy record lives in a static device array of 65,536 steps per GPU
 * (1 MiB, held until the process ends) instead of a buffer grown per call,
 * and is copied to pe[0..nsteps) / ke[0..nsteps) (host arrays, either may be
 * NULL; NaN on steps without an energy request).
 * - nsteps above 65,536: ATVH_E_ARG.
 * - The list must be built (atvh_nlist_build or an earlier run) and, with
 *   PME, one evaluation of this precision must have run, since those make
 *   their buffers; kernel timing must be off. Otherwise ATVH_E_STATE.
 * - A rebuild that overflows the list capacity is not regrown (that would
 *   allocate): ATVH_E_CAPACITY, the system is left halted mid-step and its
 *   frame must be uploaded again.
 * - Every call on one GPU shares that GPU's array: calls that overlap on one
 *   GPU (two contexts, two host threads) race. atvh_md_thermo does not see
 *   this variant's records. */
int atvh_md_run_noalloc(atvh_system* sys, int nsteps, int precision, int energy_every, double* pe,
                        double* ke);

/* Measurement aid: a plain 16-byte-vector copy of `bytes` between two
 * device buffers, launched iters times on the context stream; *avg_ms =
 * mean kernel time. Its byte counters calibrate the HBM traffic counters. */
int atvh_util_stream_copy(atvh_ctx* ctx, int64_t bytes, int iters,
                          double* avg_ms);

/* ---- synthetic inputs (host only) ---------------------------------------------------
 * Lattice generators for the benchmark workloads and InitVelocity::build
 * (src/cpu/dynamics/init.cpp:107-215): mt19937 + normal(0,1)/sqrt(m), COM
 * removal, rescale to T with ndof. Arrays may be NULL to query the count.
 */
/* FCC: 4 sites per cell, cells x-fastest, basis (0,0,0)(.5,.5,0)(.5,0,.5)(0,.5,.5);
 * stops after nmax sites when nmax >= 0 (trimmed lattice). Returns the count. */
int64_t atvh_util_fcc_lattice(int nx, int ny, int nz, double a, int64_t nmax, double* x,
                              double* y, double* z);
/* Rock salt on a simple-cubic grid: type = (ix+iy+iz)&1, charge +q0 / -q0. */
int64_t atvh_util_rocksalt_lattice(int nx, int ny, int nz, double spacing, double* x,
                                   double* y, double* z, int* type, double* charge, double q0);
int atvh_util_init_velocity(int64_t n, const double* mass, double T, uint32_t seed,
                            int64_t ndof, double* vx, double* vy, double* vz);

/* ---- batches of independent systems ------------------------------------------------
 * Replaces the replica loop of Remd::propagate_ (src/cpu/path/remd.cpp:98-103:
 * one cpu::md::MD per replica, run one after another). Members are ordinary
 * systems of one context, each uploaded and configured through the calls
 * above (NeighborList in the brick layout, PairLJCut, VV +/- NH; atom counts,
 * boxes and temperatures may all differ). atvh_batch_run advances every
 * member by the same number of steps with one launch per stage over all of
 * them; each member's trajectory is bit-identical to atvh_md_run on it alone.
 * Per-member results: atvh_md_thermo, atvh_frame_download, atvh_energy.
 * The batch references its members (destroy it before them).
 */
int atvh_batch_create(atvh_ctx* ctx, atvh_system* const* systems, int nsys,
                      atvh_batch** out);
int atvh_batch_destroy(atvh_batch* batch);
/* Capture the batched step once into a hipGraph and replay it (default off). */
int atvh_batch_set_graph(atvh_batch* batch, int enable);
/* energy_every as in atvh_md_run. Synchronises at the end. A member whose
 * rebuild overflows halts at that step, regrows its list and finishes the
 * run alone (bit-identical to the batched step). */
int atvh_batch_run(atvh_batch* batch, int nsteps, int precision,
                   int energy_every);
/* One replica-exchange round on the device (Remd::exchange_round_,
 * src/cpu/path/remd.cpp:105-151): members are copies of one system (same
 * atom count) with Nose-Hoover thermostats, created in increasing ladder
 * temperature; the members holding adjacent ladder slots (k, k+1), k even
 * for even rounds and odd for odd rounds, exchange with probability
 * min(1, exp((1/kT_k - 1/kT_k+1)(U_k - U_k+1))) using the potential energies
 * of the last batched step and a counter-based uniform u(seed, round, k).
 * An accepted exchange moves the thermostats (T, Q, xi) between the two
 * members and rescales each one's velocities by sqrt(T_new/T_old); the
 * configurations stay in their systems. *accepted = accepted swaps. */
int atvh_batch_exchange(atvh_batch* batch, int64_t round, uint64_t seed,
                        int* accepted);
/* The member holding each ladder slot, and each member's current
 * thermostat temperature (either pointer may be NULL; synchronises). */
int atvh_batch_ladder(atvh_batch* batch, int* member_of_slot,
                      double* T_of_member);

/* ---- serving: MD segments with continuous batching ---------------------------------
 * No counterpart in the reference (its run() is continuable and frames move
 * through cpu::to_cuda / cuda::to_cpu, include/atomiverse/cuda/dynamics/
 * md.cuh:130-178, src/cuda/frame/transfer.cu:145-475); this is the serving
 * layer over that boundary. A segment is one system and a step count.
 * Submitted segments are admitted in FIFO order while their estimated device
 * footprint fits the HBM budget, stay resident until done, and advance
 * together one chunk per atvh_server_advance (every brick-layout segment,
 * LJ or charged with k-space, through one batched step; small-grid segments
 * through their own fused loop). A segment's result does not depend on what
 * it ran beside.
 */
#define ATVH_PAIR_LJ 0
#define ATVH_PAIR_LJCOUL 1

typedef struct atvh_segment {
    int64_t n;
    double box[3];
    /* frame-order host arrays, copied at submission; v*, charge, type may be NULL */
    const double *x, *y, *z, *vx, *vy, *vz, *mass, *charge;
    const int* type;
    double cutoff, skin;            /* NeighborList(cutoff, skin, 1, 0, yes) */
    int pair_style;                 /* ATVH_PAIR_LJ or ATVH_PAIR_LJCOUL */
    double eps, sigma;              /* PairLJCut */
    int ntypes;                     /* PairLJCutCoulLong coefficient table */
    const double *eps_tab, *sigma_tab;
    const int* set_tab;
    double alpha;                   /* Ewald splitting (LJCOUL) */
    double pme_spacing;             /* KSpacePME grid spacing; pme_order 0 = none */
    int pme_order;
    double dt;                      /* VelocityVerlet */
    int thermostat;                 /* 0 NVE, 1 Nose-Hoover */
    double T, tau;
    int64_t nsteps;
} atvh_segment;

/* Admission estimate of a segment's device bytes. */
int64_t atvh_segment_bytes(const atvh_segment* seg);
int atvh_server_create(atvh_ctx* ctx, int64_t hbm_budget_bytes,
                       int chunk_steps, int precision, atvh_server** out);
int atvh_server_destroy(atvh_server* server);
/* Queue a segment; *id identifies it. ATVH_E_CAPACITY if it alone exceeds
 * the budget. */
int atvh_server_submit(atvh_server* server, const atvh_segment* seg,
                       int64_t* id);
/* Admit what fits, advance every resident segment by one chunk (never past
 * its own end), retire the finished ones (device memory released).
 * Synchronises. A segment whose admission or run fails is retired at once
 * (device memory released) and reported by poll like a finished one; the
 * other segments go on. */
int atvh_server_advance(atvh_server* server, int* active, int* queued);
/* Ids finished since the last poll (up to cap). */
int atvh_server_poll(atvh_server* server, int64_t* ids, int cap, int* n);
/* Final positions (wrapped) and velocities of a finished segment in frame
 * order (any pointer may be NULL) and its last potential and kinetic energy;
 * forgets the segment. For a failed segment: returns the error code it
 * failed with (atvh_last_error names the cause) and forgets it. */
int atvh_server_fetch(atvh_server* server, int64_t id, double* x, double* y,
                      double* z, double* vx, double* vy, double* vz,
                      double* pe, double* ke);
int atvh_server_usage(atvh_server* server, int64_t* bytes_in_use,
                      int64_t* budget);

#ifdef __cplusplus
}
#endif

#endif /* ATVH_H */
