// atomiverse/hip/md.hpp — the hip:: MD backend: Driver/Component/Frame API over
// the C ABI of libatv_hip.so (include/atvh.h). C++20, header-only, no HIP headers.
//
// Mirrors the reference's cuda:: backend so user code switches by namespace:
//   cuda::md::MD            (include/atomiverse/cuda/dynamics/md.cuh:118-385)
//   cuda::Component         (include/atomiverse/cuda/component/component.cuh:41-203)
//   cuda::NeighborList      (include/atomiverse/cuda/neighbor/neighbor_list.cuh:93-229)
//   cuda::PairLJCut         (include/atomiverse/cuda/potentials/pair_lj_cut.cuh:25)
//   cuda::md::VelocityVerlet / NoseHoover (include/atomiverse/cuda/dynamics/
//                            velocity_verlet.cuh:44, nose_hoover.cuh:31)
// Phases and slot order follow ATV::md::Phase (include/atomiverse/core/stage.h:24-35);
// within a phase components run in ascending `order`, insertion order breaking
// ties (PlanBuilder's stable sort, include/atomiverse/core/plan.h).
//
// Integration note: inside the reference tree define ATV_HIP_USE_REFERENCE_CORE
// and include atomiverse/core/stage.h first; this header then uses the core's
// ATV::md::Phase instead of declaring its own.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "atvh.h"

namespace ATV {
#ifndef ATV_HIP_USE_REFERENCE_CORE
namespace md {
/// The MD step's phases, in slot order (core/stage.h:24-35).
enum class Phase {
    initial_integrate,
    post_integrate,
    pre_exchange,
    pre_neighbor,
    neighbor,
    pre_force,
    force,
    post_force,
    final_integrate,
    end_of_step,
    COUNT
};
} // namespace md

/// Role bits of a component (core/role.h).
enum class RoleFlag : uint32_t {
    None = 0,
    FrameImmutable = 1u << 0, ///< does not modify x / v / f / box
    Deferred = 1u << 1,       ///< optional; may run off the main loop
    ExternalIO = 1u << 2,     ///< writes outside the process
};
using Role = uint32_t;
constexpr Role role_bit(RoleFlag f) noexcept { return static_cast<Role>(f); }

/// What a termination criterion says after an iteration, by severity
/// (core/verdict.h): Stop > Converged > Continue.
enum class Verdict { Continue = 0, Converged = 1, Stop = 2 };
constexpr Verdict worse_of(Verdict a, Verdict b) {
    return static_cast<int>(a) >= static_cast<int>(b) ? a : b;
}

/// Stream a component's launches go to (core/binding.h).
enum class Scheduling { blocking, deferrable };
#endif

namespace hip {

struct device_tag {};

/// Runtime precision choice (the reference fixes it at compile time with
/// ATV_CUDA_PREC_MODE; here it is per driver): plain FP64, or FP32 pair math
/// with FP64 integration and accumulation.
enum class Precision { fp64 = ATVH_PREC_FP64, mixed = ATVH_PREC_MIXED };

/// Host-side frame at the boundary: the columns of cpu::Frame this path reads
/// and writes (kAtom_x..z, v, f, mass, type, charge; kEnergy, kEnergyPair,
/// kKinetic, kStress), in the caller's row order. Units: Å, Å/fs, eV/Å, amu, e.
struct HostFrame {
    std::vector<double> x, y, z, vx, vy, vz, fx, fy, fz, mass, charge;
    std::vector<int> type;
    std::array<double, 3> box{0.0, 0.0, 0.0};
    double energy = 0.0, energy_pair = 0.0, kinetic = 0.0;
    std::array<double, 6> stress{0, 0, 0, 0, 0, 0};
    std::size_t nlocal() const { return x.size(); }
    void resize(std::size_t n) {
        for (auto* v : {&x, &y, &z, &vx, &vy, &vz, &fx, &fy, &fz})
            v->assign(n, 0.0);
        mass.assign(n, 1.0);
        charge.clear();
        type.assign(n, 0);
    }
};

inline void check(int rc, atvh_ctx* ctx, const char* what) {
    if (rc != ATVH_OK)
        throw std::runtime_error(std::string(what) + ": " + (ctx ? atvh_last_error(ctx) : "error"));
}

/// Execution context (cuda::Context's role): the device context, the resident
/// system, the run's precision and which consumers need the virial.
class Context {
  public:
    explicit Context(int device = 0) {
        check(atvh_ctx_create(device, &ctx_), nullptr, "atvh_ctx_create");
    }
    ~Context() {
        if (sys_)
            atvh_system_destroy(sys_);
        if (ctx_)
            atvh_ctx_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    atvh_ctx* ctx() const { return ctx_; }
    atvh_system* sys() const { return sys_; }
    Precision precision = Precision::fp64;
    bool stress_required = false;
    bool energy_this_step = false;

    /// What this loop's components measured in the last iteration
    /// (cuda::LoopMetrics; the MD fills `energy` when a criterion or thermo
    /// output reads it).
    struct LoopMetrics {
        double energy = 0.0;  ///< potential + kinetic [eV]
        double potential = 0.0, kinetic = 0.0;
        int step = -1;        ///< the iteration the values belong to
    } metrics;
    /// Set by a component to end the run after this iteration
    /// (IterationStatus::stop_requested).
    bool stop_requested = false;

    void materialize(const HostFrame& f) {
        if (sys_) {
            atvh_system_destroy(sys_);
            sys_ = nullptr;
        }
        const int64_t n = static_cast<int64_t>(f.nlocal());
        check(atvh_system_create(ctx_, n, f.box.data(), &sys_), ctx_, "atvh_system_create");
        auto opt = [&](const std::vector<double>& v) { return v.size() == f.nlocal() ? v.data() : nullptr; };
        check(atvh_frame_upload(sys_, f.x.data(), f.y.data(), f.z.data(), opt(f.vx), opt(f.vy), opt(f.vz),
                                opt(f.fx), opt(f.fy), opt(f.fz), f.mass.data(),
                                f.type.size() == f.nlocal() ? f.type.data() : nullptr, opt(f.charge)),
              ctx_, "atvh_frame_upload");
    }
    void retrieve(HostFrame& f) const {
        check(atvh_frame_download(sys_, f.x.data(), f.y.data(), f.z.data(), f.vx.data(), f.vy.data(),
                                  f.vz.data(), f.fx.data(), f.fy.data(), f.fz.data()),
              ctx_, "atvh_frame_download");
        double pe = 0, ke = 0;
        check(atvh_energy(sys_, &pe, f.stress.data(), &ke), ctx_, "atvh_energy");
        f.energy = pe + 0.0;
        f.energy_pair = pe;
        f.kinetic = ke;
        // a barostat moves the box
        check(atvh_system_box(sys_, f.box.data()), ctx_, "atvh_system_box");
    }

  private:
    atvh_ctx* ctx_ = nullptr;
    atvh_system* sys_ = nullptr;
};

struct BuildContext {
    Context& runtime;
    HostFrame& frame;
};
using DestroyContext = BuildContext;

struct LaunchContext {
    Context& runtime;
    HostFrame& frame;
    ATV::md::Phase stage;
    int step;
    bool at(ATV::md::Phase p) const { return stage == p; }
};

struct Binding {
    ATV::md::Phase phase;
    int order;
};

class Component;

/// Opt-in neighbor protocols (cuda/component/neighbor_protocol.cuh): a
/// consumer reads the published pair list up to its cutoff; a provider
/// publishes one up to its reach (cutoff + skin).
class NeighborConsumer {
  public:
    virtual ~NeighborConsumer() = default;
    virtual double neighbor_cutoff() const = 0;
};
class NeighborProvider {
  public:
    virtual ~NeighborProvider() = default;
    virtual double neighbor_reach() const = 0;
};

/// cuda::Component's lifecycle (include/atomiverse/cuda/component/
/// component.cuh:41-203): build once, launch per step (device work only, no
/// allocation or host synchronisation), destroy; capability queries for
/// install_defaults and the plan checks.
class Component {
  public:
    virtual ~Component() = default;
    virtual std::string name() const = 0;
    virtual Role role() const { return 0; }
    virtual bool requires_system_stress() const { return false; }
    virtual bool provides_system_stress() const { return false; }
    virtual bool provides_energy() const { return false; }
    virtual bool provides_force() const { return false; }
    virtual void set_stress_required(bool) {}
    virtual const NeighborConsumer* as_neighbor_consumer() const { return nullptr; }
    virtual const NeighborProvider* as_neighbor_provider() const { return nullptr; }
    virtual Scheduling scheduling() const { return Scheduling::blocking; }
    virtual int frequency() const { return 0; }
    virtual std::vector<Binding> default_bindings() const { return {}; }
    virtual void build(const BuildContext&) {}
    virtual void launch(const LaunchContext&) = 0;
    virtual void destroy(const DestroyContext&) {}
    /// Device bytes this component's state occupies (0 when it owns none).
    virtual std::size_t memory_bytes() const { return 0; }

    // backend queries (not in the reference's interface)
    bool is_neighbor_provider() const { return as_neighbor_provider() != nullptr; }
    virtual bool is_integrator() const { return false; }
    /// Runs on the host only (reads values already brought back); such a
    /// component may join the fused device loop, which calls it after each
    /// step it is due.
    virtual bool host_only() const { return false; }

    /// The context the component was built on (set by the driver before
    /// build); memory queries read the resident system through it.
    void attach_(const Context* rt) { rt_ = rt; }

  protected:
    const Context* rt_ = nullptr;
    std::size_t system_bytes_(int part) const {
        int64_t b = 0;
        if (rt_ && rt_->sys())
            atvh_system_memory(rt_->sys(), part, &b);
        return static_cast<std::size_t>(b);
    }
};

/// NeighborList(cutoff, skin = -1 (0.3*cutoff), every = 1, delay = 10, check = true):
/// counting-sort cell binning + Verlet list, LAMMPS decide cadence, device-side.
class NeighborList : public Component, public NeighborProvider {
  public:
    explicit NeighborList(double cutoff, double skin = -1.0, int every = 1, int delay = 10, bool check = true)
        : cutoff_(cutoff), skin_(skin), every_(every), delay_(delay), check_(check) {}
    std::string name() const override { return "neighbor_list"; }
    const NeighborProvider* as_neighbor_provider() const override { return this; }
    double neighbor_reach() const override { return cutoff_ + skin(); }
    std::size_t memory_bytes() const override { return system_bytes_(ATVH_MEM_NLIST); }
    void build(const BuildContext& c) override {
        check(atvh_nlist_configure(c.runtime.sys(), cutoff_, skin_, every_, delay_, check_ ? 1 : 0),
              c.runtime.ctx(), "NeighborList::build");
        check(atvh_nlist_build(c.runtime.sys()), c.runtime.ctx(), "NeighborList::build");
    }
    void launch(const LaunchContext& c) override {
        check(atvh_nlist_decide_and_rebuild(c.runtime.sys()), c.runtime.ctx(), "NeighborList::launch");
    }
    int n_rebuilds(const Context& rt) const {
        int64_t nr = 0;
        check(atvh_nlist_stats(rt.sys(), &nr, nullptr, nullptr, nullptr), rt.ctx(), "NeighborList::n_rebuilds");
        return static_cast<int>(nr);
    }
    double cutoff() const { return cutoff_; }
    double skin() const { return skin_ < 0 ? 0.3 * cutoff_ : skin_; }
    int every() const { return every_; }
    int delay() const { return delay_; }
    bool check_enabled() const { return check_; }

  private:
    double cutoff_, skin_;
    int every_, delay_;
    bool check_;
};

/// PairLJCut(eps, sigma, cutoff), or per-type tables (Lorentz-Berthelot filled).
class PairLJCut : public Component {
  public:
    PairLJCut(double epsilon, double sigma, double cutoff) : eps_(epsilon), sigma_(sigma), cutoff_(cutoff) {}
    PairLJCut(int ntypes, std::vector<double> eps_tab, std::vector<double> sigma_tab, double cutoff)
        : cutoff_(cutoff), ntypes_(ntypes), eps_tab_(std::move(eps_tab)), sigma_tab_(std::move(sigma_tab)) {
        if (ntypes <= 0 || eps_tab_.size() != size_t(ntypes * ntypes) || sigma_tab_.size() != eps_tab_.size())
            throw std::invalid_argument("PairLJCut: the per-type form needs ntypes^2 eps and sigma entries");
    }
    std::string name() const override { return "lj"; }
    bool provides_force() const override { return true; }
    bool provides_energy() const override { return true; }
    const NeighborConsumer* as_neighbor_consumer() const override { return &consumer_; }
    std::size_t memory_bytes() const override { return system_bytes_(ATVH_MEM_PAIR); }
    void set_stress_required(bool v) override { stress_ = v; }
    void build(const BuildContext& c) override {
        check(atvh_pair_lj_configure(c.runtime.sys(), eps_, sigma_, cutoff_, ntypes_,
                                     ntypes_ ? eps_tab_.data() : nullptr, ntypes_ ? sigma_tab_.data() : nullptr),
              c.runtime.ctx(), "PairLJCut::build");
    }
    void launch(const LaunchContext& c) override {
        const int ef = c.runtime.energy_this_step ? 1 : 0;
        const int vf = (stress_ || c.runtime.stress_required) && ef ? 1 : 0;
        check(atvh_pair_compute(c.runtime.sys(), static_cast<int>(c.runtime.precision), ef, vf), c.runtime.ctx(),
              "PairLJCut::launch");
    }
    double pair_cutoff() const { return cutoff_; }
    double epsilon() const { return eps_; }
    double sigma() const { return sigma_; }
    int ntypes() const { return ntypes_; }

  private:
    struct Consumer : NeighborConsumer {
        const PairLJCut* p;
        explicit Consumer(const PairLJCut* q) : p(q) {}
        double neighbor_cutoff() const override { return p->cutoff_; }
    };
    double eps_ = 0, sigma_ = 0, cutoff_;
    int ntypes_ = 0;
    std::vector<double> eps_tab_, sigma_tab_;
    bool stress_ = false;
    Consumer consumer_{this};

  public:
    PairLJCut(const PairLJCut& o)
        : Component(o), eps_(o.eps_), sigma_(o.sigma_), cutoff_(o.cutoff_), ntypes_(o.ntypes_),
          eps_tab_(o.eps_tab_), sigma_tab_(o.sigma_tab_), stress_(o.stress_), consumer_(this) {}
    PairLJCut(PairLJCut&& o) noexcept
        : Component(o), eps_(o.eps_), sigma_(o.sigma_), cutoff_(o.cutoff_), ntypes_(o.ntypes_),
          eps_tab_(std::move(o.eps_tab_)), sigma_tab_(std::move(o.sigma_tab_)), stress_(o.stress_),
          consumer_(this) {}
};

/// One coefficient row of PairLJCutCoulLong (types, epsilon [eV], sigma [Å]).
struct LJTypeCoeff {
    int type_i = 0, type_j = 0;
    double epsilon = 0.0, sigma = 1.0;
};

/// PairLJCutCoulLong(coeffs, alpha, cutoff): per-type shifted LJ + real-space
/// Ewald Coulomb from the erfc table (src/cpu/potentials/
/// pair_lj_cut_coul_long.cpp:34-70); unset off-diagonal pairs are
/// Lorentz-Berthelot mixed from set diagonals. With a KSpacePME in the plan
/// its reciprocal part runs in the same force-phase call.
class PairLJCutCoulLong : public Component {
  public:
    PairLJCutCoulLong(std::vector<LJTypeCoeff> coeffs, double alpha, double cutoff)
        : coeffs_(std::move(coeffs)), alpha_(alpha), cutoff_(cutoff) {
        if (coeffs_.empty())
            throw std::invalid_argument("PairLJCutCoulLong: empty coeff table");
        if (!(alpha > 0.0))
            throw std::invalid_argument("PairLJCutCoulLong: alpha must be > 0");
        if (!(cutoff > 0.0))
            throw std::invalid_argument("PairLJCutCoulLong: cutoff must be > 0");
        for (const auto& c : coeffs_) {
            if (c.type_i < 0 || c.type_j < 0)
                throw std::invalid_argument("PairLJCutCoulLong: negative type");
            if (!(c.sigma > 0.0) || !(c.epsilon >= 0.0))
                throw std::invalid_argument("PairLJCutCoulLong: bad epsilon/sigma");
            ntypes_ = std::max(ntypes_, std::max(c.type_i, c.type_j) + 1);
        }
    }
    std::string name() const override { return "lj/cut/coul/long"; }
    bool provides_force() const override { return true; }
    bool provides_energy() const override { return true; }
    const NeighborConsumer* as_neighbor_consumer() const override { return &consumer_; }
    std::size_t memory_bytes() const override { return system_bytes_(ATVH_MEM_PAIR); }
    void set_stress_required(bool v) override { stress_ = v; }
    void build(const BuildContext& c) override {
        const size_t n2 = size_t(ntypes_) * ntypes_;
        std::vector<double> eps(n2, 0.0), sig(n2, 0.0);
        std::vector<int> set(n2, 0);
        for (const auto& cf : coeffs_)
            for (auto [a, b] : {std::pair{cf.type_i, cf.type_j}, std::pair{cf.type_j, cf.type_i}}) {
                eps[size_t(a) * ntypes_ + b] = cf.epsilon;
                sig[size_t(a) * ntypes_ + b] = cf.sigma;
                set[size_t(a) * ntypes_ + b] = 1;
            }
        check(atvh_pair_ljcoul_configure(c.runtime.sys(), ntypes_, eps.data(), sig.data(), set.data(), alpha_,
                                         cutoff_, 10000),
              c.runtime.ctx(), "PairLJCutCoulLong::build");
    }
    void launch(const LaunchContext& c) override {
        const int ef = c.runtime.energy_this_step ? 1 : 0;
        const int vf = (stress_ || c.runtime.stress_required) && ef ? 1 : 0;
        check(atvh_pair_compute(c.runtime.sys(), static_cast<int>(c.runtime.precision), ef, vf), c.runtime.ctx(),
              "PairLJCutCoulLong::launch");
    }
    double pair_cutoff() const { return cutoff_; }
    double alpha() const { return alpha_; }
    int ntypes() const { return ntypes_; }

  private:
    struct Consumer : NeighborConsumer {
        const PairLJCutCoulLong* p;
        explicit Consumer(const PairLJCutCoulLong* q) : p(q) {}
        double neighbor_cutoff() const override { return p->cutoff_; }
    };
    std::vector<LJTypeCoeff> coeffs_;
    double alpha_, cutoff_;
    int ntypes_ = 0;
    bool stress_ = false;
    Consumer consumer_{this};

  public:
    PairLJCutCoulLong(const PairLJCutCoulLong& o)
        : Component(o), coeffs_(o.coeffs_), alpha_(o.alpha_), cutoff_(o.cutoff_), ntypes_(o.ntypes_),
          stress_(o.stress_), consumer_(this) {}
    PairLJCutCoulLong(PairLJCutCoulLong&& o) noexcept
        : Component(o), coeffs_(std::move(o.coeffs_)), alpha_(o.alpha_), cutoff_(o.cutoff_), ntypes_(o.ntypes_),
          stress_(o.stress_), consumer_(this) {}
};

/// CoulLong(alpha, cutoff): real-space Ewald Coulomb alone
/// (src/cpu/potentials/coul_long.cpp), here the LJ-free PairLJCutCoulLong of
/// one type (the erfc table stands in for the closed form, within the
/// table's interpolation error).
class CoulLong : public PairLJCutCoulLong {
  public:
    CoulLong(double alpha, double cutoff) : PairLJCutCoulLong({LJTypeCoeff{0, 0, 0.0, 1.0}}, alpha, cutoff) {}
    std::string name() const override { return "coul/long"; }
};

/// KSpacePME(alpha, grid_spacing, order) (src/cpu/longrange/pme.cpp:31-253).
/// With a pair style in the plan the k-space part runs inside that style's
/// force-phase call, after it (the reference's evaluator order); alone it
/// evaluates the reciprocal sum into zeroed forces.
class KSpacePME : public Component {
  public:
    KSpacePME(double alpha, double grid_spacing, int order = 5)
        : alpha_(alpha), spacing_(grid_spacing), order_(order) {
        if (alpha <= 0.0)
            throw std::invalid_argument("KSpacePME: alpha must be > 0");
        if (grid_spacing <= 0.0)
            throw std::invalid_argument("KSpacePME: grid_spacing must be > 0");
        if (order < 3)
            throw std::invalid_argument("KSpacePME: order must be >= 3");
    }
    std::string name() const override { return "pme"; }
    bool provides_force() const override { return true; }
    bool provides_energy() const override { return true; }
    std::size_t memory_bytes() const override { return system_bytes_(ATVH_MEM_PME); }
    void set_stress_required(bool v) override { stress_ = v; }
    void build(const BuildContext& c) override {
        check(atvh_pme_configure(c.runtime.sys(), alpha_, spacing_, order_), c.runtime.ctx(), "KSpacePME::build");
    }
    void launch(const LaunchContext& c) override {
        if (!standalone_)
            return;
        const int ef = c.runtime.energy_this_step ? 1 : 0;
        const int vf = (stress_ || c.runtime.stress_required) && ef ? 1 : 0;
        check(atvh_pme_compute(c.runtime.sys(), static_cast<int>(c.runtime.precision), ef, vf), c.runtime.ctx(),
              "KSpacePME::launch");
    }
    void set_standalone(bool v) { standalone_ = v; }
    double alpha() const { return alpha_; }
    int order() const { return order_; }

  private:
    double alpha_, spacing_;
    int order_;
    bool stress_ = false;
    bool standalone_ = false;
};

namespace md {

/// VelocityVerlet(dt): initial_integrate and final_integrate at order 0.
class VelocityVerlet : public Component {
  public:
    explicit VelocityVerlet(double dt) : dt_(dt) {}
    std::string name() const override { return "velocity_verlet"; }
    bool is_integrator() const override { return true; }
    std::vector<Binding> default_bindings() const override {
        return {{ATV::md::Phase::initial_integrate, 0}, {ATV::md::Phase::final_integrate, 0}};
    }
    void build(const BuildContext& c) override {
        check(atvh_integrator_configure(c.runtime.sys(), dt_, thermostat_, T_, tau_, 0), c.runtime.ctx(),
              "VelocityVerlet::build");
    }
    void launch(const LaunchContext& c) override {
        const int phase = c.at(ATV::md::Phase::initial_integrate) ? ATVH_PHASE_INITIAL : ATVH_PHASE_FINAL;
        check(atvh_step_vvnh(c.runtime.sys(), phase), c.runtime.ctx(), "VelocityVerlet::launch");
    }
    double dt() const { return dt_; }
    // set by a NoseHoover bound in the same driver (the two share one kernel)
    void couple_thermostat(double T, double tau) {
        thermostat_ = 1;
        T_ = T;
        tau_ = tau;
    }

  private:
    double dt_;
    int thermostat_ = 0;
    double T_ = 0, tau_ = 0;
};

/// NoseHoover(T_target [K], tau [fs], dt [fs]): single chain, after the kick of
/// each integrate phase (order +100). On this backend the update runs inside the
/// velocity-Verlet kernels of the same phases, so launch() has nothing to add.
class NoseHoover : public Component {
  public:
    NoseHoover(double T_target, double tau, double dt) : T_(T_target), tau_(tau), dt_(dt) {}
    std::string name() const override { return "nose_hoover"; }
    std::vector<Binding> default_bindings() const override {
        return {{ATV::md::Phase::initial_integrate, 100}, {ATV::md::Phase::final_integrate, 100}};
    }
    void launch(const LaunchContext&) override {}
    double T_target() const { return T_; }
    double tau() const { return tau_; }
    double dt() const { return dt_; }
    double xi(const Context& rt) const {
        double v = 0;
        check(atvh_thermostat_state(rt.sys(), &v, nullptr), rt.ctx(), "NoseHoover::xi");
        return v;
    }

  private:
    double T_, tau_, dt_;
};

/// cpu::md::Virial (src/cpu/observe/virial_stress.cpp:14-56; cuda::Virial,
/// src/cuda/observe/stress.cu): at post_force (-50) the kinetic stress
/// sum m v_a v_b of the current velocities, queued on the device; mounting it
/// makes the force phase compute the virial (set_stress_required). Mount with
/// sim.on(Phase::post_force).add_ref(virial, -50) to read it afterwards.
class Virial : public Component {
  public:
    std::string name() const override { return "virial"; }
    Role role() const override { return role_bit(RoleFlag::FrameImmutable); }
    bool provides_system_stress() const override { return true; }
    std::vector<Binding> default_bindings() const override { return {{ATV::md::Phase::post_force, -50}}; }
    void launch(const LaunchContext& c) override {
        check(atvh_kinetic_stress(c.runtime.sys()), c.runtime.ctx(), "Virial::launch");
    }
    /// Kinetic stress and virial (eV, extensive, Voigt xx yy zz xy xz yz) and
    /// the pressure (tr K + tr W) / (3V) in eV/Å^3; synchronises.
    void read(const Context& rt, double kinetic[6], double virial[6], double* pressure) const {
        check(atvh_stress(rt.sys(), kinetic, virial, pressure), rt.ctx(), "Virial::read");
    }
};

/// KineticEnergy at end_of_step (-100): reduced by the final-integrate kernel.
class KineticEnergy : public Component {
  public:
    std::string name() const override { return "kinetic_energy"; }
    void launch(const LaunchContext&) override {}
};

/// The zeroing infrastructure cuda::md::MD binds in its constructor
/// (md.cuh:123-127): forces, energy and stress at pre_force, kinetic energy
/// at end_of_step. On this backend the force kernels overwrite their outputs
/// and the kick kernels write fresh kinetic partials, so there is nothing to
/// launch; the components keep the plan's shape (and its orders) the same.
class ZeroForces : public Component {
  public:
    std::string name() const override { return "zero_forces"; }
    void launch(const LaunchContext&) override {}
};
class ZeroEnergy : public Component {
  public:
    std::string name() const override { return "zero_energy"; }
    void launch(const LaunchContext&) override {}
};
class ZeroStress : public Component {
  public:
    std::string name() const override { return "zero_stress"; }
    void launch(const LaunchContext&) override {}
};
class ZeroKinetic : public Component {
  public:
    std::string name() const override { return "zero_kinetic"; }
    void launch(const LaunchContext&) override {}
};

/// MTK NoseHooverNPT(T_target [K], tau_T [fs], P_target [eV/Å^3], tau_P [fs],
/// dt [fs], drag, mtchain, mpchain, couple_axis, T_stop)
/// (include/atomiverse/cuda/dynamics/nose_hoover_npt.cuh:56-94): replaces
/// VelocityVerlet + NoseHoover; bound at initial_integrate 0,
/// final_integrate 0 and end_of_step -40; reads the system stress, so the
/// MD injects a Virial when none is mounted. It steps inside the fused
/// device loop (atvh_md_run), which runs its three phases per step.
class NoseHooverNPT : public Component {
  public:
    NoseHooverNPT(double T_target, double tau_T, double P_target, double tau_P, double dt, double drag = 0.0,
                  int mtchain = 3, int mpchain = 3, std::array<bool, 3> couple_axis = {true, true, true},
                  double T_stop = -1.0)
        : T_(T_target), tau_T_(tau_T), P_(P_target), tau_P_(tau_P), dt_(dt), drag_(drag), mt_(mtchain),
          mp_(mpchain), couple_(couple_axis), T_stop_(T_stop) {}
    NoseHooverNPT& set_run_span(std::int64_t nfirst, std::int64_t nlast) {
        nfirst_ = nfirst;
        nlast_ = nlast;
        return *this;
    }
    std::string name() const override { return "nose_hoover_npt"; }
    bool is_integrator() const override { return true; }
    bool requires_system_stress() const override { return true; }
    std::vector<Binding> default_bindings() const override {
        return {{ATV::md::Phase::initial_integrate, 0},
                {ATV::md::Phase::final_integrate, 0},
                {ATV::md::Phase::end_of_step, -40}};
    }
    void build(const BuildContext& c) override {
        const int cp[3] = {couple_[0] ? 1 : 0, couple_[1] ? 1 : 0, couple_[2] ? 1 : 0};
        check(atvh_npt_configure(c.runtime.sys(), dt_, T_, tau_T_, P_, tau_P_, drag_, mt_, mp_, cp, T_stop_),
              c.runtime.ctx(), "NoseHooverNPT::build");
        if (nlast_ > nfirst_)
            check(atvh_npt_set_run_span(c.runtime.sys(), nfirst_, nlast_), c.runtime.ctx(),
                  "NoseHooverNPT::build");
    }
    void launch(const LaunchContext&) override {
        throw std::logic_error("NoseHooverNPT steps inside the fused device loop (hip::md::MD::set_fused(true))");
    }
    double dt() const { return dt_; }
    /// Box volume [Å^3] and MTK conserved quantity [eV] after the last step.
    double volume(const Context& rt) const {
        double b[3] = {0, 0, 0};
        check(atvh_system_box(rt.sys(), b), rt.ctx(), "NoseHooverNPT::volume");
        return b[0] * b[1] * b[2];
    }

  private:
    double T_, tau_T_, P_, tau_P_, dt_, drag_;
    int mt_, mp_;
    std::array<bool, 3> couple_;
    double T_stop_;
    std::int64_t nfirst_ = 0, nlast_ = 0;
};

/// cuda::Thermo (set_thermo): every `every` steps (0 = every step) a row
/// {step, potential, kinetic, temperature} read back from the device at
/// end_of_step +20. Host-only: it reads what the step already reduced.
class Thermo : public Component {
  public:
    struct Row {
        int step;
        double potential, kinetic, temperature;
    };
    explicit Thermo(int every = 0) : every_(every < 1 ? 1 : every) {}
    std::string name() const override { return "thermo"; }
    Role role() const override { return role_bit(RoleFlag::FrameImmutable); }
    bool host_only() const override { return true; }
    int frequency() const override { return every_; }
    std::vector<Binding> default_bindings() const override { return {{ATV::md::Phase::end_of_step, 20}}; }
    void launch(const LaunchContext& c) override {
        if ((c.step + 1) % every_ != 0)
            return;
        double pe = 0, ke = 0;
        check(atvh_energy(c.runtime.sys(), &pe, nullptr, &ke), c.runtime.ctx(), "Thermo::launch");
        const double kB = 8.617333262e-5; // units.h metal kB [eV/K]
        const double ndof = 3.0 * static_cast<double>(atvh_system_natoms(c.runtime.sys())) - 3.0;
        rows_.push_back({c.step, pe, ke, ndof > 0 ? 2.0 * ke / (ndof * kB) : 0.0});
    }
    int every() const { return every_; }
    const std::vector<Row>& rows() const { return rows_; }

  private:
    int every_;
    std::vector<Row> rows_;
};

/// What a criterion may look at after an iteration (cuda::LoopState).
struct LoopState {
    int iteration;
    const Context& runtime;
    const Context::LoopMetrics& metrics;
};

/// A termination policy asked once after every iteration
/// (include/atomiverse/cuda/driver/criterion.cuh).
class Criterion {
  public:
    virtual ~Criterion() = default;
    virtual const char* name() const = 0;
    virtual Verdict after_iteration(const LoopState& state) = 0;
};

/// Abandon a run whose energy has gone non-finite (criterion.cuh:166-179):
/// reads only the energy the loop already brought back to the host.
class StopOnNaN : public Criterion {
  public:
    const char* name() const override { return "stop_on_nan"; }
    Verdict after_iteration(const LoopState& s) override {
        return std::isfinite(s.metrics.energy) ? Verdict::Continue : Verdict::Stop;
    }
};

/// Column keys of a cpu::Frame-like host frame (core/frame_schema.h), for
/// MD::run on the reference's frame type: the adapter reads them with
/// `frame.span(key)`, `frame.box.length(d)` and writes energies with
/// `frame.meta.get(key)`. Inside the reference tree the schema's own keys
/// are passed.
template <class Key, class MetaKey, class StressKey>
struct FrameKeys {
    Key x, y, z, vx, vy, vz, fx, fy, fz, mass, type, charge;
    MetaKey energy, kinetic;
    StressKey stress;
    bool has_charge = false;
};

/// cuda::md::MD with the same setters and continuable run().
class MD {
  public:
    /// Binds the zeroing and kinetic-energy infrastructure (md.cuh:123-127).
    explicit MD(int device = 0) : rt_(std::make_unique<Context>(device)) {
        on(ATV::md::Phase::pre_force).add(ZeroForces{}, -1000);
        on(ATV::md::Phase::pre_force).add(ZeroEnergy{}, -999);
        on(ATV::md::Phase::pre_force).add(ZeroStress{}, -998);
        on(ATV::md::Phase::end_of_step).add(ZeroKinetic{}, -1000);
        on(ATV::md::Phase::end_of_step).add(KineticEnergy{}, -100);
    }

    struct PhaseAdder {
        MD& md;
        ATV::md::Phase phase;
        template <class C>
        MD& add(C&& comp, int order = 0) {
            md.add_owned_(phase, std::make_unique<std::decay_t<C>>(std::forward<C>(comp)), order);
            return md;
        }
        template <class C>
        MD& add_ref(C& comp, int order = 0) {
            md.add_borrowed_(phase, &comp, order);
            return md;
        }
    };
    PhaseAdder on(ATV::md::Phase p) { return PhaseAdder{*this, p}; }

    /// Singular by name (driver.cuh:650-664): a repeat call with the same
    /// component name is ignored, a different one throws.
    template <class N>
    MD& set_neighbor(N&& nl) {
        if (singular_(neighbor_set_, nl.name(), "set_neighbor"))
            on(ATV::md::Phase::neighbor).add(std::forward<N>(nl), -10);
        return *this;
    }
    template <class E>
    MD& set_evaluator(E&& e) {
        return on(ATV::md::Phase::force).add(std::forward<E>(e), 0);
    }
    template <class I>
    MD& set_integrator(I&& integ) {
        if (singular_(integrator_set_, integ.name(), "set_integrator"))
            bind_defaults_(std::forward<I>(integ));
        return *this;
    }
    template <class T>
    MD& set_thermostat(T&& t) {
        bind_defaults_(std::forward<T>(t));
        return *this;
    }
    template <class B>
    MD& set_barostat(B&& b) {
        bind_defaults_(std::forward<B>(b));
        return *this;
    }
    template <class O>
    MD& set_observer(O&& o) {
        bind_defaults_(std::forward<O>(o));
        return *this;
    }
    /// Thermo rows every `every` steps (cuda::md::MD::set_thermo, md.cuh:
    /// 249-263): a repeat with the same cadence is ignored, another throws.
    MD& set_thermo(int every) {
        if (thermo_every_) {
            if (*thermo_every_ == every)
                return *this;
            throw std::invalid_argument("set_thermo already configured with " + std::to_string(*thermo_every_) +
                                        ", rejected re-call with " + std::to_string(every));
        }
        thermo_every_ = every;
        auto t = std::make_unique<Thermo>(every);
        thermo_ = t.get();
        ops_.push_back(Op{ATV::md::Phase::end_of_step, 20, t.get()});
        owned_.push_back(std::move(t));
        built_ = false;
        return *this;
    }
    const Thermo* thermo() const { return thermo_; }

    /// Attach a termination policy (Driver::add_criterion); plural, composed
    /// by severity.
    template <class H>
    MD& add_criterion(H&& h) {
        using D = std::decay_t<H>;
        static_assert(std::is_base_of_v<Criterion, D>, "add_criterion requires a hip::md::Criterion subclass");
        handlers_.push_back(std::make_unique<D>(std::forward<H>(h)));
        return *this;
    }
    bool has_handler() const { return !handlers_.empty(); }
    bool converged() const { return converged_; }
    bool stopped() const { return stopped_; }
    const std::string& stop_reason() const { return reason_; }

    MD& set_precision(Precision p) {
        rt_->precision = p;
        return *this;
    }
    MD& set_stress_required(bool v) {
        rt_->stress_required = v;
        return *this;
    }
    /// Use the fused device loop (atvh_md_run) when the plan is the standard
    /// integrator + NeighborList + pair set (host-only observers allowed);
    /// on by default.
    MD& set_fused(bool v) {
        fused_allowed_ = v;
        return *this;
    }

    /// install_defaults + validate_components + plan compile (Driver::build).
    void build() {
        install_defaults_();
        std::stable_sort(ops_.begin(), ops_.end(), [](const Op& a, const Op& b) {
            return a.phase != b.phase ? static_cast<int>(a.phase) < static_cast<int>(b.phase) : a.order < b.order;
        });
        bool has_int = false, has_force = false;
        for (auto& op : ops_) {
            has_int |= op.comp->is_integrator();
            has_force |= op.comp->provides_force();
        }
        if (!has_force)
            throw std::invalid_argument("hip::md::MD: no force evaluator; add one with set_evaluator()");
        if (!has_int)
            throw std::invalid_argument("hip::md::MD: no integrator; add one with set_integrator()");
        // couple a Nose-Hoover thermostat into the velocity-Verlet kernels
        VelocityVerlet* vv = nullptr;
        NoseHoover* nh = nullptr;
        bool npt = false;
        for (auto& op : ops_) {
            if (auto* p = dynamic_cast<VelocityVerlet*>(op.comp))
                vv = p;
            if (auto* p = dynamic_cast<NoseHoover*>(op.comp))
                nh = p;
            npt |= dynamic_cast<NoseHooverNPT*>(op.comp) != nullptr;
        }
        if (nh && vv)
            vv->couple_thermostat(nh->T_target(), nh->tau());
        // k-space runs inside a pair style's force call when one is present
        bool has_pair = false;
        for (auto& op : ops_)
            has_pair |= dynamic_cast<PairLJCut*>(op.comp) != nullptr ||
                        dynamic_cast<PairLJCutCoulLong*>(op.comp) != nullptr;
        for (auto& op : ops_)
            if (auto* k = dynamic_cast<KSpacePME*>(op.comp))
                k->set_standalone(!has_pair);
        fused_ = fused_allowed_;
        for (auto& op : ops_) {
            const bool known =
                dynamic_cast<VelocityVerlet*>(op.comp) || dynamic_cast<NoseHoover*>(op.comp) ||
                dynamic_cast<NoseHooverNPT*>(op.comp) || dynamic_cast<NeighborList*>(op.comp) ||
                dynamic_cast<PairLJCut*>(op.comp) || dynamic_cast<PairLJCutCoulLong*>(op.comp) ||
                dynamic_cast<KineticEnergy*>(op.comp) || dynamic_cast<KSpacePME*>(op.comp) ||
                dynamic_cast<ZeroForces*>(op.comp) || dynamic_cast<ZeroEnergy*>(op.comp) ||
                dynamic_cast<ZeroStress*>(op.comp) || dynamic_cast<ZeroKinetic*>(op.comp) ||
                // the fused loop computes the virial whenever the NPT reads it
                (npt && dynamic_cast<Virial*>(op.comp)) || op.comp->host_only();
            fused_ = fused_ && known;
        }
        if (fused_) // the fused loop evaluates forces through a pair style
            fused_ = has_pair;
        if (npt && !fused_)
            throw std::invalid_argument("hip::md::MD: NoseHooverNPT steps in the fused device loop only");
        built_ = true;
    }

    /// Advance n_steps; continuable on the same frame (cuda::md::MD::run,
    /// md.cuh:155-178). Attached criteria are polled after every step and
    /// any verdict but Continue ends the run; returns the steps taken.
    int run(HostFrame& frame, int n_steps) {
        if (!built_)
            build();
        if (live_ && frame_ != &frame)
            teardown_();
        if (!live_) {
            rt_->materialize(frame);
            BuildContext bc{*rt_, frame};
            for (auto& op : ops_) {
                op.comp->attach_(rt_.get());
                if (!op.built) {
                    op.comp->build(bc);
                    op.built = true;
                }
            }
            frame_ = &frame;
            step_ = 0;
            live_ = true;
        }
        converged_ = stopped_ = false;
        reason_.clear();
        int done = 0;
        const bool per_step = !handlers_.empty();
        while (done < n_steps) {
            rt_->stop_requested = false;
            int chunk = per_step ? 1 : n_steps - done;
            if (fused_) {
                // host-only observers run after each step they are due: end the
                // fused chunk there
                for (auto& op : ops_)
                    if (op.comp->host_only() && op.comp->frequency() > 0) {
                        const int f = op.comp->frequency();
                        chunk = std::min(chunk, f - (step_ % f));
                    }
                check(atvh_md_run(rt_->sys(), chunk, static_cast<int>(rt_->precision), 0), rt_->ctx(),
                      "hip::md::MD::run");
                step_ += chunk;
                done += chunk;
                publish_metrics_();
                for (auto& op : ops_)
                    if (op.comp->host_only()) {
                        LaunchContext lc{*rt_, frame, op.phase, step_ - 1};
                        op.comp->launch(lc);
                    }
            } else {
                for (int s = 0; s < chunk; ++s, ++step_, ++done) {
                    rt_->energy_this_step = per_step || (done + 1 == n_steps) || thermo_due_(step_);
                    for (auto& op : ops_) {
                        if (op.comp->is_neighbor_provider() && op.phase != ATV::md::Phase::neighbor)
                            continue;
                        if (op.comp->host_only() && op.phase == ATV::md::Phase::end_of_step)
                            continue; // after the metrics below
                        LaunchContext lc{*rt_, frame, op.phase, step_};
                        op.comp->launch(lc);
                    }
                    if (per_step || thermo_due_(step_))
                        publish_metrics_();
                    for (auto& op : ops_)
                        if (op.comp->host_only() && op.phase == ATV::md::Phase::end_of_step) {
                            LaunchContext lc{*rt_, frame, op.phase, step_};
                            op.comp->launch(lc);
                        }
                }
            }
            if (per_step && poll_(step_ - 1) != Verdict::Continue)
                break;
        }
        rt_->retrieve(frame);
        return done;
    }

    /// run() on a cpu::Frame-like frame (the reference's `run(cpu::Frame&, n)`,
    /// md.cuh:155): its columns are read through `keys`, advanced on the
    /// device and written back with the energies and the virial.
    template <class F, class K, class MK, class SK>
    int run(F& frame, int n_steps, const FrameKeys<K, MK, SK>& keys) {
        const std::size_t n = frame.nlocal();
        if (!adapted_ || adapted_src_ != static_cast<const void*>(&frame)) {
            adapted_ = std::make_unique<HostFrame>();
            adapted_src_ = &frame;
        }
        HostFrame& h = *adapted_;
        h.resize(n);
        auto in = [&](std::vector<double>& dst, K key) {
            const auto* p = frame.span(key);
            for (std::size_t i = 0; i < n; ++i)
                dst[i] = static_cast<double>(p[i]);
        };
        for (auto [dst, key] : {std::pair{&h.x, keys.x}, {&h.y, keys.y}, {&h.z, keys.z}, {&h.vx, keys.vx},
                                {&h.vy, keys.vy}, {&h.vz, keys.vz}, {&h.fx, keys.fx}, {&h.fy, keys.fy},
                                {&h.fz, keys.fz}, {&h.mass, keys.mass}})
            in(*dst, key);
        const auto* tp = frame.span(keys.type);
        for (std::size_t i = 0; i < n; ++i)
            h.type[i] = static_cast<int>(tp[i]);
        if (keys.has_charge) {
            h.charge.resize(n);
            in(h.charge, keys.charge);
        }
        for (int d = 0; d < 3; ++d)
            h.box[d] = frame.box.length(d);
        const int done = run(h, n_steps);
        auto out = [&](const std::vector<double>& src, K key) {
            auto* p = frame.span(key);
            for (std::size_t i = 0; i < n; ++i)
                p[i] = src[i];
        };
        for (auto [src, key] : {std::pair{&h.x, keys.x}, {&h.y, keys.y}, {&h.z, keys.z}, {&h.vx, keys.vx},
                                {&h.vy, keys.vy}, {&h.vz, keys.vz}, {&h.fx, keys.fx}, {&h.fy, keys.fy},
                                {&h.fz, keys.fz}})
            out(*src, key);
        frame.meta.get(keys.energy) = h.energy;
        frame.meta.get(keys.kinetic) = h.kinetic;
        auto& st = frame.meta.get(keys.stress);
        for (int k = 0; k < 6; ++k)
            st[k] = h.stress[k];
        double b[3];
        check(atvh_system_box(rt_->sys(), b), rt_->ctx(), "hip::md::MD::run");
        // a barostat moved the box: Box::scale_diagonal (cpu/frame/box.h:215)
        frame.box.scale_diagonal(b[0] / frame.box.length(0), b[1] / frame.box.length(1), b[2] / frame.box.length(2));
        return done;
    }

    Context& context() { return *rt_; }
    int step() const { return step_; }
    bool fused() const { return fused_; }
    /// Mounted components in plan order (for inspection).
    std::vector<const Component*> components() const {
        std::vector<const Component*> v;
        for (auto& op : ops_)
            v.push_back(op.comp);
        return v;
    }

    ~MD() { teardown_(); }

  private:
    struct Op {
        ATV::md::Phase phase;
        int order;
        Component* comp;
        bool built = false;
    };
    void add_owned_(ATV::md::Phase p, std::unique_ptr<Component> c, int order) {
        ops_.push_back(Op{p, order, c.get()});
        owned_.push_back(std::move(c));
        built_ = false;
    }
    void add_borrowed_(ATV::md::Phase p, Component* c, int order) {
        ops_.push_back(Op{p, order, c});
        built_ = false;
    }
    template <class C>
    void bind_defaults_(C&& comp) {
        auto up = std::make_unique<std::decay_t<C>>(std::forward<C>(comp));
        const auto bindings = up->default_bindings();
        if (bindings.empty())
            throw std::runtime_error("Component '" + up->name() + "' has no default bindings. Use .on(phase).add().");
        for (const Binding& b : bindings)
            ops_.push_back(Op{b.phase, b.order, up.get()});
        owned_.push_back(std::move(up));
        built_ = false;
    }
    static bool singular_(std::optional<std::string>& slot, const std::string& n, const char* setter) {
        if (slot) {
            if (*slot == n)
                return false;
            throw std::invalid_argument(std::string(setter) + " already configured with '" + *slot +
                                        "', rejected re-call with '" + n + "'");
        }
        slot = n;
        return true;
    }
    // install_defaults (md.cuh:285-318): a Virial when a component reads the
    // system stress and none provides it; every virial producer told whether
    // its virial is consumed; a neighbor consumer needs a provider.
    void install_defaults_() {
        bool needs = false, has_virial = false, consumer = false, provider = false;
        std::string first_consumer;
        for (auto& op : ops_) {
            needs |= op.comp->requires_system_stress();
            has_virial |= op.comp->provides_system_stress();
            if (op.comp->as_neighbor_consumer() && !consumer) {
                consumer = true;
                first_consumer = op.comp->name();
            }
            provider |= op.comp->as_neighbor_provider() != nullptr;
        }
        if (needs && !has_virial) {
            on(ATV::md::Phase::post_force).add(Virial{}, -50);
            has_virial = true;
        }
        if (has_virial)
            rt_->stress_required = true;
        for (auto& op : ops_)
            op.comp->set_stress_required(rt_->stress_required);
        if (consumer && !provider)
            throw std::invalid_argument("hip::md::MD: '" + first_consumer +
                                        "' needs a neighbor list. Add `sim.set_neighbor(hip::NeighborList{cutoff})`.");
        if (!provider)
            throw std::invalid_argument(
                "hip::md::MD: the pair evaluator needs a neighbor list. Add `sim.set_neighbor(hip::NeighborList{cutoff})`.");
    }
    bool thermo_due_(int step) const { return thermo_ && (step + 1) % thermo_->every() == 0; }
    void publish_metrics_() {
        double pe = 0, ke = 0;
        check(atvh_energy(rt_->sys(), &pe, nullptr, &ke), rt_->ctx(), "hip::md::MD::run");
        rt_->metrics.potential = pe;
        rt_->metrics.kinetic = ke;
        rt_->metrics.energy = pe + ke;
        rt_->metrics.step = step_ - 1;
    }
    // poll_handlers_ (driver.cuh:397-416): the most severe verdict wins; a
    // component's stop request counts as Stop
    Verdict poll_(int iteration) {
        const LoopState st{iteration, *rt_, rt_->metrics};
        Verdict v = rt_->stop_requested ? Verdict::Stop : Verdict::Continue;
        for (auto& h : handlers_) {
            const Verdict hv = h->after_iteration(st);
            if (hv != Verdict::Continue && worse_of(v, hv) == hv && v != hv)
                reason_ = h->name();
            v = worse_of(v, hv);
        }
        converged_ = v == Verdict::Converged;
        stopped_ = v == Verdict::Stop;
        return v;
    }
    void teardown_() {
        if (!live_)
            return;
        BuildContext bc{*rt_, *frame_};
        for (auto& op : ops_) {
            if (op.built) {
                op.comp->destroy(bc);
                op.built = false;
            }
        }
        live_ = false;
        frame_ = nullptr;
    }

    std::unique_ptr<Context> rt_;
    std::vector<Op> ops_;
    std::vector<std::unique_ptr<Component>> owned_;
    std::vector<std::unique_ptr<Criterion>> handlers_;
    std::optional<std::string> neighbor_set_, integrator_set_;
    std::optional<int> thermo_every_;
    Thermo* thermo_ = nullptr;
    bool built_ = false, live_ = false, fused_ = false, fused_allowed_ = true;
    bool converged_ = false, stopped_ = false;
    std::string reason_;
    HostFrame* frame_ = nullptr;
    std::unique_ptr<HostFrame> adapted_;
    const void* adapted_src_ = nullptr;
    int step_ = 0;
};

} // namespace md
} // namespace hip
} // namespace ATV
