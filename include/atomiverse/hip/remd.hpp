// atomiverse/hip/remd.hpp — replica-exchange MD on the hip:: backend.
//
// Mirrors cpu::path::Remd (include/atomiverse/cpu/path/remd.h,
// src/cpu/path/remd.cpp): a strictly increasing temperature ladder, one
// thermostatted replica per temperature, `steps_per_exchange` MD steps
// between exchange rounds on alternating even/odd adjacent pairs
// (Sugita-Okamoto: P = min(1, exp[(beta_i - beta_j)(U_i - U_j)])), and the
// lowest-temperature replica collapsed back into the frame at the end.
//
// MI355X form: all replicas live on one device and advance through one
// batched step (atvh_batch_run: one launch per stage for the whole ladder),
// and the exchange round runs on the device (atvh_batch_exchange). The force
// program is given as components by value (this backend batches
// NeighborList + PairLJCut replicas) instead of a builder callback. Replicas
// are thermostatted with single-chain Nose-Hoover at their ladder
// temperature (the reference attaches CSVR; both are canonical).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <vector>

#include "atomiverse/hip/md.hpp"

namespace ATV {
namespace hip {
namespace path {

class Remd {
  public:
    std::vector<double> temperatures; ///< strictly increasing ladder [K], size >= 2
    double dt = 1.0;                  ///< MD timestep [fs]
    int steps_per_exchange = 10;      ///< MD steps between exchange rounds
    std::uint64_t seed = 1;           ///< master seed (velocities and exchanges)
    double thermostat_tau = 100.0;    ///< Nose-Hoover coupling time [fs]

    explicit Remd(std::vector<double> temps, double dt_ = 1.0, int steps = 10, std::uint64_t seed_ = 1)
        : temperatures(std::move(temps)), dt(dt_), steps_per_exchange(steps), seed(seed_) {
        if (temperatures.size() < 2)
            throw std::invalid_argument("Remd: need >= 2 temperatures");
        for (size_t i = 1; i < temperatures.size(); ++i)
            if (!(temperatures[i] > temperatures[i - 1]))
                throw std::invalid_argument("Remd: temperatures must be strictly increasing");
        if (!(temperatures[0] > 0.0))
            throw std::invalid_argument("Remd: temperatures must be positive");
        if (steps_per_exchange < 1)
            throw std::invalid_argument("Remd: steps_per_exchange must be >= 1");
    }
    ~Remd() { release_(); }
    Remd(const Remd&) = delete;
    Remd& operator=(const Remd&) = delete;

    void set_neighbor(NeighborList nl) { nl_ = std::make_unique<NeighborList>(nl); }
    void set_evaluator(PairLJCut pair) { pair_ = std::make_unique<PairLJCut>(pair); }
    void set_precision(Precision p) { prec_ = p; }

    /// Replicate the frame once per temperature (velocities drawn per replica
    /// with seed ^ (i + 1), as the reference's InitVelocity), run n_exchanges
    /// rounds, and leave the lowest-temperature replica in `frame`.
    void run(HostFrame& frame, int n_exchanges) {
        if (n_exchanges < 1)
            throw std::invalid_argument("Remd::run: n_exchanges must be >= 1");
        if (!nl_ || !pair_)
            throw std::invalid_argument("Remd: set_neighbor and set_evaluator before run");
        if (!(thermostat_tau > 0.0))
            throw std::invalid_argument("Remd: thermostat_tau must be > 0 fs");
        build_(frame);
        const int n = static_cast<int>(sys_.size());
        for (int r = 0; r < n_exchanges; ++r) {
            check(atvh_batch_run(batch_, steps_per_exchange, static_cast<int>(prec_), 0), ctx_, "Remd::propagate");
            // Remd::run alternates even and odd edges from r = 0 on every call
            // (exchange_round_(r % 2 == 0)); the round index handed to the
            // device is 2 * (rounds so far) + parity, so its parity is this
            // call's and its uniform stream never repeats across calls
            int acc = 0;
            const int64_t parity = r & 1;
            check(atvh_batch_exchange(batch_, 2 * round_ + parity, seed, &acc), ctx_, "Remd::exchange");
            attempts_ += (n - static_cast<int>(parity)) / 2;
            accepts_ += acc;
            ++round_;
        }
        // collapse to the lowest-temperature slot
        std::vector<int> slot(n);
        check(atvh_batch_ladder(batch_, slot.data(), nullptr), ctx_, "Remd::collapse");
        atvh_system* s0 = sys_[slot[0]];
        check(atvh_frame_download(s0, frame.x.data(), frame.y.data(), frame.z.data(), frame.vx.data(),
                                  frame.vy.data(), frame.vz.data(), frame.fx.data(), frame.fy.data(),
                                  frame.fz.data()),
              ctx_, "Remd::collapse");
        double pe = 0, ke = 0;
        check(atvh_energy(s0, &pe, frame.stress.data(), &ke), ctx_, "Remd::collapse");
        frame.energy = pe;
        frame.energy_pair = pe;
        frame.kinetic = ke;
    }
    int exchange_attempts() const { return attempts_; }
    int exchange_accepts() const { return accepts_; }

  private:
    void build_(const HostFrame& frame) {
        if (batch_)
            return;
        check(atvh_ctx_create(0, &ctx_), nullptr, "Remd: atvh_ctx_create");
        const int64_t n = static_cast<int64_t>(frame.nlocal());
        for (size_t i = 0; i < temperatures.size(); ++i) {
            atvh_system* s = nullptr;
            check(atvh_system_create(ctx_, n, frame.box.data(), &s), ctx_, "Remd: system");
            sys_.push_back(s);
            std::vector<double> vx(n), vy(n), vz(n);
            check(atvh_util_init_velocity(n, frame.mass.data(), temperatures[i],
                                          static_cast<uint32_t>(seed ^ (i + 1u)), 3 * n - 3, vx.data(), vy.data(),
                                          vz.data()),
                  ctx_, "Remd: InitVelocity");
            check(atvh_frame_upload(s, frame.x.data(), frame.y.data(), frame.z.data(), vx.data(), vy.data(),
                                    vz.data(), nullptr, nullptr, nullptr, frame.mass.data(),
                                    frame.type.size() == frame.nlocal() ? frame.type.data() : nullptr, nullptr),
                  ctx_, "Remd: upload");
            check(atvh_nlist_configure(s, nl_->cutoff(), nl_->skin(), nl_->every(), nl_->delay(),
                                       nl_->check_enabled() ? 1 : 0),
                  ctx_, "Remd: NeighborList");
            check(atvh_pair_lj_configure(s, pair_->epsilon(), pair_->sigma(), pair_->pair_cutoff(), 0, nullptr,
                                         nullptr),
                  ctx_, "Remd: PairLJCut");
            check(atvh_integrator_configure(s, dt, 1, temperatures[i], thermostat_tau, 0), ctx_, "Remd: integrator");
            check(atvh_nlist_build(s), ctx_, "Remd: list build");
        }
        check(atvh_batch_create(ctx_, sys_.data(), static_cast<int>(sys_.size()), &batch_), ctx_, "Remd: batch");
        atvh_batch_set_graph(batch_, 1);
    }
    void release_() {
        if (batch_)
            atvh_batch_destroy(batch_);
        for (atvh_system* s : sys_)
            atvh_system_destroy(s);
        if (ctx_)
            atvh_ctx_destroy(ctx_);
        batch_ = nullptr;
        sys_.clear();
        ctx_ = nullptr;
    }

    std::unique_ptr<NeighborList> nl_;
    std::unique_ptr<PairLJCut> pair_;
    Precision prec_ = Precision::fp64;
    atvh_ctx* ctx_ = nullptr;
    std::vector<atvh_system*> sys_;
    atvh_batch* batch_ = nullptr;
    int64_t round_ = 0;
    int attempts_ = 0;
    int accepts_ = 0;
};

} // namespace path
} // namespace hip
} // namespace ATV
