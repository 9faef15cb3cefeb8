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
    }

  private:
    atvh_ctx* ctx_ = nullptr;
    atvh_system* sys_ = nullptr;
};

struct BuildContext {
    Context& runtime;
    HostFrame& frame;
};

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

/// cuda::Component's lifecycle: build once, launch per step (device work only,
/// no allocation or host synchronisation), destroy.
class Component {
  public:
    virtual ~Component() = default;
    virtual std::string name() const = 0;
    virtual std::vector<Binding> default_bindings() const { return {}; }
    virtual void build(const BuildContext&) {}
    virtual void launch(const LaunchContext&) = 0;
    virtual void destroy(const BuildContext&) {}
    virtual void set_stress_required(bool) {}
    virtual bool is_neighbor_provider() const { return false; }
    virtual bool is_integrator() const { return false; }
    virtual bool provides_force() const { return false; }
};

/// NeighborList(cutoff, skin = -1 (0.3*cutoff), every = 1, delay = 10, check = true):
/// counting-sort cell binning + Verlet list, LAMMPS decide cadence, device-side.
class NeighborList : public Component {
  public:
    explicit NeighborList(double cutoff, double skin = -1.0, int every = 1, int delay = 10, bool check = true)
        : cutoff_(cutoff), skin_(skin), every_(every), delay_(delay), check_(check) {}
    std::string name() const override { return "neighbor_list"; }
    bool is_neighbor_provider() const override { return true; }
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
    double eps_ = 0, sigma_ = 0, cutoff_;
    int ntypes_ = 0;
    std::vector<double> eps_tab_, sigma_tab_;
    bool stress_ = false;
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
    std::vector<LJTypeCoeff> coeffs_;
    double alpha_, cutoff_;
    int ntypes_ = 0;
    bool stress_ = false;
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
    bool provides_system_stress() const { return true; }
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

/// cuda::md::MD with the same setters and continuable run().
class MD {
  public:
    explicit MD(int device = 0) : rt_(std::make_unique<Context>(device)) {
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

    template <class N>
    MD& set_neighbor(N&& nl) {
        if (neighbor_set_)
            throw std::invalid_argument("set_neighbor already configured");
        neighbor_set_ = true;
        return on(ATV::md::Phase::neighbor).add(std::forward<N>(nl), -10);
    }
    template <class E>
    MD& set_evaluator(E&& e) {
        return on(ATV::md::Phase::force).add(std::forward<E>(e), 0);
    }
    template <class I>
    MD& set_integrator(I&& integ) {
        if (integrator_set_)
            throw std::invalid_argument("set_integrator already configured");
        integrator_set_ = true;
        bind_defaults_(std::forward<I>(integ));
        return *this;
    }
    template <class T>
    MD& set_thermostat(T&& t) {
        bind_defaults_(std::forward<T>(t));
        return *this;
    }
    MD& set_precision(Precision p) {
        rt_->precision = p;
        return *this;
    }
    MD& set_stress_required(bool v) {
        rt_->stress_required = v;
        return *this;
    }
    /// Use the fused device loop (atvh_md_run) when the plan is the standard
    /// VV(+NH) + NeighborList + pair set; on by default.
    MD& set_fused(bool v) {
        fused_allowed_ = v;
        return *this;
    }

    /// Validate the plan (install_defaults / validate_components).
    void build() {
        std::stable_sort(ops_.begin(), ops_.end(), [](const Op& a, const Op& b) {
            return a.phase != b.phase ? static_cast<int>(a.phase) < static_cast<int>(b.phase) : a.order < b.order;
        });
        bool has_nl = false, has_int = false, has_force = false;
        for (auto& op : ops_) {
            has_nl |= op.comp->is_neighbor_provider();
            has_int |= op.comp->is_integrator();
            has_force |= op.comp->provides_force();
        }
        if (!has_force)
            throw std::invalid_argument("hip::md::MD: no force evaluator; add one with set_evaluator()");
        if (!has_nl)
            throw std::invalid_argument(
                "hip::md::MD: the pair evaluator needs a neighbor list. Add `sim.set_neighbor(hip::NeighborList{cutoff})`.");
        if (!has_int)
            throw std::invalid_argument("hip::md::MD: no integrator; add one with set_integrator()");
        // couple a Nose-Hoover thermostat into the velocity-Verlet kernels
        VelocityVerlet* vv = nullptr;
        NoseHoover* nh = nullptr;
        for (auto& op : ops_) {
            if (auto* p = dynamic_cast<VelocityVerlet*>(op.comp))
                vv = p;
            if (auto* p = dynamic_cast<NoseHoover*>(op.comp))
                nh = p;
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
        // a mounted Virial observer consumes the force phase's virial
        // (the reference's install_defaults virial gating)
        for (auto& op : ops_)
            if (dynamic_cast<Virial*>(op.comp))
                rt_->stress_required = true;
        for (auto& op : ops_)
            op.comp->set_stress_required(rt_->stress_required);
        fused_ = fused_allowed_;
        for (auto& op : ops_) {
            const bool known = dynamic_cast<VelocityVerlet*>(op.comp) || dynamic_cast<NoseHoover*>(op.comp) ||
                               dynamic_cast<NeighborList*>(op.comp) || dynamic_cast<PairLJCut*>(op.comp) ||
                               dynamic_cast<PairLJCutCoulLong*>(op.comp) || dynamic_cast<KineticEnergy*>(op.comp) ||
                               dynamic_cast<KSpacePME*>(op.comp);
            fused_ = fused_ && known;
        }
        if (fused_) { // the fused loop evaluates forces through a pair style
            bool pair = false;
            for (auto& op : ops_)
                pair |= dynamic_cast<PairLJCut*>(op.comp) != nullptr ||
                        dynamic_cast<PairLJCutCoulLong*>(op.comp) != nullptr;
            fused_ = pair;
        }
        built_ = true;
    }

    /// Advance n_steps; continuable on the same frame (cuda::md::MD::run).
    int run(HostFrame& frame, int n_steps) {
        if (!built_)
            build();
        if (live_ && frame_ != &frame) {
            teardown_();
        }
        if (!live_) {
            rt_->materialize(frame);
            BuildContext bc{*rt_, frame};
            for (auto& op : ops_)
                if (!op.built) {
                    op.comp->build(bc);
                    op.built = true;
                }
            frame_ = &frame;
            step_ = 0;
            live_ = true;
        }
        if (fused_) {
            check(atvh_md_run(rt_->sys(), n_steps, static_cast<int>(rt_->precision), 0), rt_->ctx(),
                  "hip::md::MD::run");
            step_ += n_steps;
        } else {
            for (int s = 0; s < n_steps; ++s, ++step_) {
                rt_->energy_this_step = (s == n_steps - 1);
                for (auto& op : ops_) {
                    if (op.comp->is_neighbor_provider() && op.phase != ATV::md::Phase::neighbor)
                        continue;
                    LaunchContext lc{*rt_, frame, op.phase, step_};
                    op.comp->launch(lc);
                }
            }
        }
        rt_->retrieve(frame);
        return n_steps;
    }

    Context& context() { return *rt_; }
    int step() const { return step_; }
    bool fused() const { return fused_; }

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
        for (const Binding& b : up->default_bindings())
            ops_.push_back(Op{b.phase, b.order, up.get()});
        owned_.push_back(std::move(up));
        built_ = false;
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
    bool neighbor_set_ = false, integrator_set_ = false;
    bool built_ = false, live_ = false, fused_ = false, fused_allowed_ = true;
    HostFrame* frame_ = nullptr;
    int step_ = 0;
};

} // namespace md
} // namespace hip
} // namespace ATV
