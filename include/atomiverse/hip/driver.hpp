// atomiverse/hip/driver.hpp — the hip:: backend behind the reference's Driver
// template: backend_traits<hip::device_tag>, hip::Backend with cuda::Backend's
// method set, and (outside the reference tree) a Driver<PhaseEnum, DeviceTag>
// with the contract of include/atomiverse/core/driver_template.h:167-290.
//
// Reference counterparts:
//   backend_traits<cuda::device_tag>  include/atomiverse/cuda/driver/backend_traits.cuh:14-17
//   cuda::Backend                     src/cuda/driver/driver.cu:25-196
//   Driver<md::Phase, cuda::device_tag> (explicit instantiation)
//                                     src/cuda/driver/driver.cu:202-204
// Inside the reference tree (ATV_HIP_USE_REFERENCE_CORE) the core's own Driver
// template, PhaseBuilder and Plan are used with the traits below; the frame
// the core hands Backend::materialize is then converted through
// hip::frame_columns<F> (specialised beside cpu::Frame with the schema keys).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_set>
#include <utility>
#include <vector>

#include "atomiverse/hip/md.hpp"

namespace ATV {

#ifndef ATV_HIP_USE_REFERENCE_CORE
template <class DeviceTag>
struct backend_traits;

/// Compiled execution plan (core/plan.h): one slot per phase, ops in
/// ascending order, insertion order breaking ties.
template <class Component>
struct Plan {
    struct Op {
        Component* component = nullptr;
        int order = 0;
        int frequency = 0;
        Scheduling scheduling = Scheduling::blocking;
    };
    struct Slot {
        int phase = 0;
        std::vector<Op> ops;
    };
    std::vector<Slot> slots;
};

/// Number of slots of a phase enum that ends with COUNT (core/stage.h).
template <class PhaseEnum>
constexpr int slot_count() {
    return static_cast<int>(PhaseEnum::COUNT);
}
#endif

namespace hip {

/// cuda::Backend's method set over the C ABI (src/cuda/driver/driver.cu:25-196):
/// init, load a plan, materialize the frame on the device, build the
/// components, step / run, destroy, retrieve, sync, shutdown. One stream (the
/// context's); deferrable components run on it in order.
class Backend {
  public:
    using Component = hip::Component;

    void init() {
        if (!rt_)
            rt_ = std::make_unique<Context>(device_);
    }
    void set_device(int d) { device_ = d; }
    void load(const Plan<Component>& plan) { plan_ = &plan; }
    void materialize(const HostFrame& frame) {
        init();
        frame_ = frame;
        rt_->materialize(frame_);
    }
    /// Builds every mounted component once (a component bound to several
    /// phases is built on its first slot).
    void build_components() {
        BuildContext bc{*rt_, frame_};
        std::unordered_set<Component*> seen;
        for (auto& slot : plan_->slots)
            for (auto& op : slot.ops)
                if (seen.insert(op.component).second) {
                    op.component->attach_(rt_.get());
                    op.component->build(bc);
                }
        // the force phase produces the virial when a stress consumer is mounted
        bool stress = false;
        for (auto& slot : plan_->slots)
            for (auto& op : slot.ops)
                stress |= op.component->requires_system_stress() || op.component->provides_system_stress();
        rt_->stress_required = rt_->stress_required || stress;
        for (auto* c : seen)
            c->set_stress_required(rt_->stress_required);
        check(atvh_ctx_synchronize(rt_->ctx()), rt_->ctx(), "hip::Backend::build_components");
    }
    /// One step: the plan's slots in phase order, each slot's ops in order
    /// (cuda::Backend::step); ops with a frequency run on its multiples.
    void step(int step_idx) {
        for (auto& slot : plan_->slots)
            for (auto& op : slot.ops) {
                if (op.frequency > 0 && step_idx % op.frequency != 0)
                    continue;
                LaunchContext lc{*rt_, frame_, static_cast<ATV::md::Phase>(slot.phase), step_idx};
                op.component->launch(lc);
            }
    }
    void run(int n_steps) {
        for (int s = 0; s < n_steps; ++s) {
            rt_->energy_this_step = s == n_steps - 1; // the energy the frame carries back
            step(s);
        }
    }
    void destroy_components() {
        DestroyContext dc{*rt_, frame_};
        for (int s = static_cast<int>(plan_->slots.size()) - 1; s >= 0; --s) {
            auto& ops = plan_->slots[static_cast<std::size_t>(s)].ops;
            for (int o = static_cast<int>(ops.size()) - 1; o >= 0; --o)
                ops[static_cast<std::size_t>(o)].component->destroy(dc);
        }
    }
    void retrieve(HostFrame& frame) {
        rt_->retrieve(frame_);
        frame = frame_;
        double b[3];
        check(atvh_system_box(rt_->sys(), b), rt_->ctx(), "hip::Backend::retrieve");
        frame.box = {b[0], b[1], b[2]};
    }
    void sync() {
        if (rt_)
            check(atvh_ctx_synchronize(rt_->ctx()), rt_->ctx(), "hip::Backend::sync");
    }
    void shutdown() {
        rt_.reset();
        plan_ = nullptr;
    }
    Context& context() { return *rt_; }
    const HostFrame& frame() const { return frame_; }

  private:
    int device_ = 0;
    std::unique_ptr<Context> rt_;
    const Plan<Component>* plan_ = nullptr;
    HostFrame frame_;
};

} // namespace hip

/// The hip:: specialisation (backend_traits.cuh:14-17).
template <>
struct backend_traits<hip::device_tag> {
    using component_type = hip::Component;
    using backend_type = hip::Backend;
};

#ifndef ATV_HIP_USE_REFERENCE_CORE

template <typename PhaseEnum, typename DeviceTag>
class Driver;

/// Fluent phase binding (driver_template.h:56-160): add() owns an rvalue,
/// add_ref() borrows; implicit lvalue add is rejected.
template <typename PhaseEnum, typename DeviceTag>
class PhaseBuilder {
    using Component = typename backend_traits<DeviceTag>::component_type;

  public:
    template <typename Comp, std::enable_if_t<std::is_base_of_v<Component, std::remove_reference_t<Comp>> &&
                                                  !std::is_lvalue_reference_v<Comp>,
                                              int> = 0>
    PhaseBuilder& add(Comp&& comp, int order = 0) {
        using C = std::remove_reference_t<Comp>;
        auto ptr = std::make_unique<C>(std::forward<Comp>(comp));
        bindings_->push_back({static_cast<int>(phase_), order, ptr.get()});
        owned_->push_back(std::move(ptr));
        return *this;
    }
    template <typename Comp, std::enable_if_t<std::is_base_of_v<Component, Comp>, int> = 0>
    PhaseBuilder& add(std::unique_ptr<Comp> comp, int order = 0) {
        if (!comp)
            throw std::invalid_argument("PhaseBuilder::add: null component");
        bindings_->push_back({static_cast<int>(phase_), order, comp.get()});
        owned_->push_back(std::move(comp));
        return *this;
    }
    template <typename Comp, std::enable_if_t<std::is_base_of_v<Component, Comp>, int> = 0>
    PhaseBuilder& add_ref(Comp& comp, int order = 0) {
        bindings_->push_back({static_cast<int>(phase_), order, &comp});
        return *this;
    }
    template <typename Comp>
    PhaseBuilder& add(Comp&, int = 0) = delete;

  private:
    friend class Driver<PhaseEnum, DeviceTag>;
    struct Pending {
        int phase;
        int order;
        Component* component;
    };
    PhaseBuilder(std::vector<Pending>* b, std::vector<std::unique_ptr<Component>>* o, PhaseEnum p)
        : bindings_(b), owned_(o), phase_(p) {}
    std::vector<Pending>* bindings_;
    std::vector<std::unique_ptr<Component>>* owned_;
    PhaseEnum phase_;
};

/// Driver<PhaseEnum, DeviceTag> (driver_template.h:167-290 with the run()
/// of driver_template_impl.h): mount table, plan compile, and the
/// init / load / materialize / build / run / destroy / retrieve / shutdown
/// lifecycle of one run() call.
template <typename PhaseEnum, typename DeviceTag>
class Driver {
  protected:
    using traits = backend_traits<DeviceTag>;
    using Component = typename traits::component_type;
    using BackendT = typename traits::backend_type;
    using Pending = typename PhaseBuilder<PhaseEnum, DeviceTag>::Pending;

  public:
    Driver() = default;
    ~Driver() { shutdown(); }
    Driver(const Driver&) = delete;
    Driver& operator=(const Driver&) = delete;

    PhaseBuilder<PhaseEnum, DeviceTag> on(PhaseEnum phase) {
        return PhaseBuilder<PhaseEnum, DeviceTag>(&pending_, &owned_components_, phase);
    }
    const Plan<Component>& plan() const { return plan_; }
    const std::vector<Pending>& mounted() const { return pending_; }

    /// Compiles the plan: one slot per phase, ops sorted by order with a
    /// stable sort (PlanBuilder, core/plan.h).
    void build() {
        plan_.slots.assign(static_cast<std::size_t>(slot_count<PhaseEnum>()), {});
        for (int p = 0; p < slot_count<PhaseEnum>(); ++p)
            plan_.slots[static_cast<std::size_t>(p)].phase = p;
        std::vector<Pending> sorted = pending_;
        std::stable_sort(sorted.begin(), sorted.end(),
                         [](const Pending& a, const Pending& b) { return a.order < b.order; });
        for (const auto& pb : sorted) {
            typename Plan<Component>::Op op;
            op.component = pb.component;
            op.order = pb.order;
            op.frequency = pb.component->frequency();
            op.scheduling = pb.component->scheduling();
            plan_.slots[static_cast<std::size_t>(pb.phase)].ops.push_back(op);
        }
        built_ = true;
    }

    /// The full lifecycle (driver_template_impl.h:18-46): the frame goes to
    /// the device, the plan runs n_steps, the frame comes back.
    void run(hip::HostFrame& frame, int n_steps) {
        if (!built_)
            build();
        if (!backend_)
            backend_ = std::make_unique<BackendT>();
        backend_->init();
        backend_->load(plan_);
        backend_->materialize(frame);
        backend_->build_components();
        backend_->run(n_steps);
        backend_->destroy_components();
        backend_->retrieve(frame);
        backend_->shutdown();
    }

    void shutdown() {
        if (backend_) {
            backend_->shutdown();
            backend_.reset();
        }
    }

  protected:
    std::vector<Pending> pending_;
    std::vector<std::unique_ptr<Component>> owned_components_;
    Plan<Component> plan_;
    std::unique_ptr<BackendT> backend_;
    bool built_ = false;
};

#endif // ATV_HIP_USE_REFERENCE_CORE

// the explicit instantiation (driver.cu:202-204) lives in
// atomiverse_amd/cxx/driver.cpp
extern template class Driver<md::Phase, hip::device_tag>;

} // namespace ATV
