/* atv_oracle.h — CPU restatement of the reference classical-MD path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP backend:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and never as the thing measured or shipped. The product path
 * (atomiverse_amd/csrc, libatv_hip.so) neither links nor calls it.
 *
 * Every function restates one reference function operation for operation; the
 * citation is on each definition in atv_oracle.cpp. Plain C ABI so ctypes can
 * bind it; no exceptions cross it (errors are negative return codes).
 */
#ifndef ATV_ORACLE_H
#define ATV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry / inputs ---------------------------------------------------- */

/* FCC lattice, nx*ny*nz conventional cells of edge a, 4 sites per cell,
 * sites emitted cell by cell (x fastest over cells), basis order
 * (0,0,0) (½,½,0) (½,0,½) (0,½,½). Returns the atom count. */
int64_t orc_fcc_lattice(int nx, int ny, int nz, double a, double* x, double* y,
                        double* z);

/* InitVelocity::build: mt19937(seed) + normal_distribution(0,1), vx,vy,vz per
 * atom scaled by 1/sqrt(m); COM removal; rescale to T [K] with ndof. */
void orc_init_velocity(int64_t n, const double* mass, double T, uint32_t seed,
                       int64_t ndof, double* vx, double* vy, double* vz);

/* ---- cell list / Verlet list (CellList) ----------------------------------- */

/* compute_grid + build_stencil for a single-rank orthorhombic box.
 * out_nc[3], out_cs[3], out_reach[3]; returns number of half-stencil offsets,
 * *out_nfull gets the full-stencil count. */
int orc_cell_grid(const double box[3], double cutoff_skin, int* out_nc,
                  double* out_cs, int* out_reach, int* out_nfull);

/* bin_atoms: atom_cell[n], cell_count[ncells], cell_offset[ncells],
 * binned[n] (counting sort, ascending i inside a cell). */
int orc_bin_atoms(int64_t n, const double* x, const double* y, const double* z,
                  const double box[3], double cutoff_skin, int* atom_cell,
                  int* cell_count, int* cell_offset, int* binned);

/* refresh_half (serial shift-form semantics; nthreads>1 computes the same CSR
 * with rows split across threads). Returns total pairs, or -(needed) if
 * `cap` is too small (no partial output then). */
int64_t orc_build_half(int64_t n, const double* x, const double* y,
                       const double* z, const double box[3], double cutoff_skin,
                       int nthreads, int* numneigh, int64_t* firstneigh,
                       int* neighbors, int64_t cap);

/* needs_rebuild: max min-image |x - x_save|^2 * 4 > skin^2. Returns 0/1 and
 * writes the max squared displacement. */
int orc_needs_rebuild(int64_t n, const double* x, const double* y,
                      const double* z, const double* xs, const double* ys,
                      const double* zs, const double box[3], double skin,
                      double* max_disp_sq);

/* ---- pair styles ----------------------------------------------------------- */

/* PairLJCut::compute on a half CSR. Forces are ACCUMULATED into f*, energy
 * and the 6-component Voigt virial (xx yy zz xy xz yz) are accumulated into
 * *energy and virial[6]. ntypes == 0 -> single (eps, sigma); ntypes > 0 ->
 * per-type eps_tab/sigma_tab (ntypes*ntypes, Lorentz-Berthelot already
 * applied by the caller or by orc_lj_mix). */
void orc_pair_lj(int64_t n, const double* x, const double* y, const double* z,
                 const double box[3], const int* numneigh,
                 const int64_t* firstneigh, const int* neighbors, double eps,
                 double sigma, double cutoff, int ntypes, const int* type,
                 const double* eps_tab, const double* sigma_tab, int nthreads,
                 double* fx, double* fy, double* fz, double* energy,
                 double* virial);

/* Lorentz-Berthelot fill of the off-diagonal of an ntypes x ntypes table
 * whose diagonal is set (PairLJCut::finalize_type_table_). */
void orc_lj_mix(int ntypes, double* eps_tab, double* sigma_tab);


/* ---- long-range electrostatics (atv_pme.cpp) ------------------------------ */

/* good_fft_size: smallest m >= max(n, 4) factoring into {2,3,5,7}. */
int orc_good_fft_size(int n);
/* bspline: cardinal B-spline weights theta[order] and derivatives. */
void orc_bspline(double frac, int order, double* theta, double* dtheta);
/* spline_moduli: Euler moduli mod[K]. */
void orc_spline_moduli(int K, int order, double* mod);
/* EwaldSum::recommend_alpha */
double orc_ewald_alpha(double cutoff, double accuracy);
/* KSpacePME mesh dimensions for an orthorhombic box and grid spacing. */
void orc_pme_grid(const double box[3], double spacing, int order, int K[3]);
/* Unnormalised 3-D complex DFT, data interleaved (re, im), layout
 * k2 + K2 (k1 + K1 k0); sign -1 forward, +1 backward. */
void orc_fft3d(int K0, int K1, int K2, double* data, int sign);
/* KSpacePME::compute (orthorhombic, one rank). Accumulates forces, energy
 * (self + recip + net) and the reciprocal virial; parts[3] (optional) gets
 * e_recip, e_self, e_net. Returns 0, or -1 on bad arguments. */
int orc_pme(int64_t n, const double* x, const double* y, const double* z,
            const double* q, const double box[3], double alpha, const int K[3],
            int order, double* fx, double* fy, double* fz, double* energy,
            double* virial, double* parts);
/* ErfcTable::build: e_tab[nbin+1], f_tab[nbin+1]. */
void orc_erfc_table(double alpha, double cutoff, int nbin, double* e_tab,
                    double* f_tab);
/* PairLJCutCoulLong::finalize_type_table_: mixes unset entries in place and
 * fills lj1..lj4 and the energy shift (ntypes*ntypes each). */
void orc_ljcoul_tables(int ntypes, double cutoff, double* eps, double* sigma,
                       const int* set, double* lj1, double* lj2, double* lj3,
                       double* lj4, double* shift);
/* PairLJCutCoulLong::compute on a half CSR (accumulating). */
void orc_pair_lj_coul_long(int64_t n, const double* x, const double* y,
                           const double* z, const double* q, const int* type,
                           const double box[3], const int* numneigh,
                           const int64_t* firstneigh, const int* neighbors,
                           int ntypes, const double* lj1, const double* lj2,
                           const double* lj3, const double* lj4,
                           const double* shift, double cutoff,
                           const double* e_tab, const double* f_tab, int nbin,
                           int nthreads, double* fx, double* fy, double* fz,
                           double* energy, double* virial);
/* CoulLong::compute (closed-form erfc real space, accumulating). */
void orc_coul_long(int64_t n, const double* x, const double* y, const double* z,
                   const double* q, const double box[3], const int* numneigh,
                   const int64_t* firstneigh, const int* neighbors,
                   double alpha, double cutoff, double* fx, double* fy,
                   double* fz, double* energy, double* virial);

/* Force field of a charged run for orc_md_run_ff: PairLJCutCoulLong with
 * per-type tables (lj1..lj4 and shift from orc_ljcoul_tables) and the erfc table,
 * followed by KSpacePME when pme_order > 0 (the reference's evaluator
 * order: pair first, then k-space). */
typedef struct orc_md_ff {
    int ntypes;
    const int* type;
    const double* q;
    const double *lj1, *lj2, *lj3, *lj4, *shift;
    const double *e_tab, *f_tab;
    int nbin;
    double alpha;
    int pme_K[3];
    int pme_order;
} orc_md_ff;

/* ---- integrators ----------------------------------------------------------- */

typedef struct orc_md_params {
    double dt;       /* fs */
    double cutoff;   /* Å (LJ and list) */
    double skin;     /* Å */
    int every, delay, check;
    double eps, sigma;      /* single-type LJ */
    int nh;                 /* 1 = single-chain Nose-Hoover NVT */
    double T_target, tau_T; /* K, fs */
    int64_t ndof;           /* <= 0 -> 3N - 3 */
    int nthreads;
    int stress;             /* compute the pair virial */
} orc_md_params;

typedef struct orc_md_state {
    double xi;       /* NH thermostat variable */
    int64_t step;    /* cumulative step index */
    int64_t nrebuilds;
    int64_t ago;
} orc_md_state;

/* cpu::md::MD::run with VelocityVerlet (+ NoseHoover) + NeighborList +
 * PairLJCut, phase order of Runtime::step_verlet_. Forces f* are the frame's
 * (carried across calls; zero for a fresh frame). `state` is in/out: a
 * zero-initialised state means "build components" (initial list build).
 * Per-step potential and kinetic energy are written to pe[nsteps], ke[nsteps]
 * when non-null. virial[6] gets the last step's pair virial if non-null. */
int orc_md_run(int64_t n, double* x, double* y, double* z, double* vx,
               double* vy, double* vz, double* fx, double* fy, double* fz,
               const double* mass, const double box[3],
               const orc_md_params* p, orc_md_state* state, int nsteps,
               double* pe, double* ke, double* virial);

/* orc_md_run with the charged force field `ff` in place of single-type LJ
 * (p->eps / p->sigma unused; p->cutoff is the real-space cutoff). */
int orc_md_run_ff(int64_t n, double* x, double* y, double* z, double* vx,
                  double* vy, double* vz, double* fx, double* fy, double* fz,
                  const double* mass, const double box[3],
                  const orc_md_params* p, const orc_md_ff* ff,
                  orc_md_state* state, int nsteps, double* pe, double* ke,
                  double* virial);

/* ---- MTK NoseHooverNPT ----------------------------------------------------- */

#define ORC_NPT_MAXCHAIN 8

/* cpu::md::NoseHooverNPT constructor arguments (nose_hoover_npt.h:183-188):
 * T_target [K] (ramp start), tau_T [fs], P_target [eV/Å^3] (positive is
 * compressive), tau_P [fs], drag, chain lengths, per-axis barostat flags,
 * T_stop [K] (< 0: no ramp) and the ramp's run span (set_run_span). */
typedef struct orc_npt_params {
    double T_target, tau_T, P_target, tau_P, drag;
    int mtchain, mpchain;
    int couple[3];
    double T_stop;
    int64_t nfirst, nlast;
} orc_npt_params;

/* Extended-Lagrangian state carried across calls; zero-initialised means
 * "build the integrator" (masses, equipartition KE). vir holds the virial
 * diagonal of the last force evaluation (the frame's kStress[0..2]; zero for
 * a fresh frame). */
typedef struct orc_npt_state {
    double eta_t[ORC_NPT_MAXCHAIN], eta_t_dot[ORC_NPT_MAXCHAIN];
    double eta_b[ORC_NPT_MAXCHAIN], eta_b_dot[ORC_NPT_MAXCHAIN];
    double eps_dot[3];
    double ke_global;
    double vir[3];
    int built;
} orc_npt_state;

/* cpu::md::MD::run with NoseHooverNPT (replacing VelocityVerlet + NoseHoover)
 * + NeighborList + PairLJCut (single type, p->eps / p->sigma), phase order of
 * Runtime::step_verlet_. box[3] is in/out (the barostat rescales it). Per
 * step: pe, ke (the post-NPT kinetic energy), H (the MTK conserved
 * quantity) and V (the box volume) when non-null. p->nh and p->T_target /
 * p->tau_T are unused (q carries the thermostat). */
int orc_npt_run(int64_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                double* fx, double* fy, double* fz, const double* mass, double box[3],
                const orc_md_params* p, const orc_npt_params* q, orc_md_state* state,
                orc_npt_state* nstate, int nsteps, double* pe, double* ke, double* H,
                double* V);

/* KineticEnergy::launch: 0.5 * mvv2e * sum m v^2 */
double orc_kinetic_energy(int64_t n, const double* mass, const double* vx,
                          const double* vy, const double* vz);

/* cpu::md::Virial::launch: mvv2e * sum m v_a v_b, Voigt xx yy zz xy xz yz */
void orc_kinetic_stress(int64_t n, const double* mass, const double* vx,
                        const double* vy, const double* vz, double out[6]);

#ifdef __cplusplus
}
#endif

#endif
