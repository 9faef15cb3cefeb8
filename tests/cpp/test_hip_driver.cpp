// The hip:: backend behind the reference's Driver contract
// (include/atomiverse/core/driver_template.h:167-290, src/cuda/driver/driver.cu):
//  1. Driver<md::Phase, hip::device_tag> (explicit instantiation, hip::Backend)
//     steps a 4000-atom FCC LJ fluid 150 NVE steps bit-identically to
//     hip::md::MD on the same frame.
//  2. A criterion stops the run: StopOnNaN after a component poisons the
//     loop's energy metric on step 37 (host-side; the device never sees a
//     NaN) ends both the fused and the per-phase run after 38 steps.
//  3. install_defaults: NoseHooverNPT requires the system stress, so a Virial
//     is injected; the 216-atom npt_lj_mtk regression run through MD ends at
//     the reference CPU's published volume (269.425432 A^3, to its six
//     printed decimals).
//  4. set_thermo rows, set_neighbor singular-by-name, memory_bytes, role and
//     the neighbor protocol accessors.
//  5. run() on a cpu::Frame-like frame (span / box / meta column surface)
//     equals run() on a HostFrame.
// Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>

#include "atomiverse/hip/driver.hpp"

using namespace ATV;
using P = ATV::md::Phase;

static const double kB = 8.617333262e-5;
static const double kTau = std::sqrt(103.6427);

#define REQUIRE(c, ...)                                                                                     \
    do {                                                                                                    \
        if (!(c)) {                                                                                         \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                                       \
            std::fprintf(stderr, __VA_ARGS__);                                                              \
            std::fprintf(stderr, "\n");                                                                     \
            std::exit(1);                                                                                   \
        }                                                                                                   \
    } while (0)

static hip::HostFrame fcc(int nc) {
    const double a = std::cbrt(4.0 / 0.8442);
    hip::HostFrame f;
    const int64_t n = atvh_util_fcc_lattice(nc, nc, nc, a, -1, nullptr, nullptr, nullptr);
    f.resize(static_cast<size_t>(n));
    atvh_util_fcc_lattice(nc, nc, nc, a, -1, f.x.data(), f.y.data(), f.z.data());
    f.box = {nc * a, nc * a, nc * a};
    atvh_util_init_velocity(n, f.mass.data(), 0.72 / kB, 87287u, 3 * n - 3, f.vx.data(), f.vy.data(), f.vz.data());
    return f;
}

static bool same(const hip::HostFrame& a, const hip::HostFrame& b) {
    for (size_t i = 0; i < a.nlocal(); ++i)
        if (a.x[i] != b.x[i] || a.y[i] != b.y[i] || a.z[i] != b.z[i] || a.vx[i] != b.vx[i] || a.vz[i] != b.vz[i] ||
            a.fy[i] != b.fy[i])
            return false;
    return true;
}

// Poisons the loop's energy metric at one step (a host-only observer).
class PoisonEnergy : public hip::Component {
  public:
    explicit PoisonEnergy(int at) : at_(at) {}
    std::string name() const override { return "poison_energy"; }
    bool host_only() const override { return true; }
    std::vector<hip::Binding> default_bindings() const override { return {{P::end_of_step, 50}}; }
    void launch(const hip::LaunchContext& c) override {
        if (c.step == at_)
            c.runtime.metrics.energy = std::nan("");
    }

  private:
    int at_;
};

// ---- a cpu::Frame-like host frame: span(key), box.length / scale_diagonal, meta.get
enum class Col { x, y, z, vx, vy, vz, fx, fy, fz, mass, type };
enum class Meta { energy, kinetic };
enum class Stress { stress };
struct MockBox {
    double L[3];
    double length(int d) const { return L[d]; }
    void scale_diagonal(double a, double b, double c) {
        L[0] *= a;
        L[1] *= b;
        L[2] *= c;
    }
};
struct MockMeta {
    double e = 0, k = 0;
    std::array<double, 6> st{};
    double& get(Meta m) { return m == Meta::energy ? e : k; }
    std::array<double, 6>& get(Stress) { return st; }
};
struct MockFrame {
    std::map<Col, std::vector<double>> cols;
    MockBox box;
    MockMeta meta;
    size_t nlocal() const { return cols.at(Col::x).size(); }
    double* span(Col c) { return cols.at(c).data(); }
};

int main() {
    // 1. Driver template vs hip::md::MD
    {
        hip::HostFrame a = fcc(10), b = fcc(10);
        const double dt = 0.005 * kTau;
        Driver<P, hip::device_tag> drv;
        hip::md::VelocityVerlet vv{dt};
        drv.on(P::neighbor).add(hip::NeighborList{2.5, 0.3, 1, 0, true}, -10);
        drv.on(P::force).add(hip::PairLJCut{1.0, 1.0, 2.5}, 0);
        drv.on(P::initial_integrate).add_ref(vv, 0);
        drv.on(P::final_integrate).add_ref(vv, 0);
        drv.build();
        REQUIRE(drv.plan().slots.size() == static_cast<size_t>(P::COUNT), "plan slots");
        drv.run(a, 150);

        hip::md::MD sim;
        sim.set_neighbor(hip::NeighborList{2.5, 0.3, 1, 0, true});
        sim.set_evaluator(hip::PairLJCut{1.0, 1.0, 2.5});
        sim.set_integrator(hip::md::VelocityVerlet{dt});
        sim.build();
        REQUIRE(sim.fused(), "MD should take the fused loop");
        sim.run(b, 150);
        REQUIRE(same(a, b), "Driver<md::Phase, hip::device_tag> and hip::md::MD differ");
        REQUIRE(a.energy == b.energy && a.kinetic == b.kinetic, "energies differ %.17g %.17g", a.energy, b.energy);
        std::printf("Driver<md::Phase, hip::device_tag> == hip::md::MD (150 steps, %zu atoms)\n", a.nlocal());
    }
    // 2. StopOnNaN
    for (bool fused : {true, false}) {
        hip::HostFrame f = fcc(6);
        hip::md::MD sim;
        sim.set_neighbor(hip::NeighborList{2.5, 0.3, 1, 0, true});
        sim.set_evaluator(hip::PairLJCut{1.0, 1.0, 2.5});
        sim.set_integrator(hip::md::VelocityVerlet{0.005 * kTau});
        sim.set_observer(PoisonEnergy{37});
        sim.add_criterion(hip::md::StopOnNaN{});
        sim.set_fused(fused);
        sim.build();
        REQUIRE(sim.fused() == fused, "fused mode");
        const int done = sim.run(f, 100);
        REQUIRE(done == 38 && sim.step() == 38, "stopped after %d steps", done);
        REQUIRE(sim.stopped() && !sim.converged() && sim.stop_reason() == "stop_on_nan", "stop state");
        REQUIRE(std::isfinite(f.energy), "the frame itself stays finite");
        std::printf("StopOnNaN ended the %s run after %d of 100 steps\n", fused ? "fused" : "per-phase", done);
    }
    // 3. NoseHooverNPT, Virial injection, the reference regression's volume
    {
        const int ns = 6, n = ns * ns * ns;
        const double L = std::cbrt(n / 0.8), sp = L / ns;
        hip::HostFrame f;
        f.resize(n);
        int idx = 0;
        for (int ix = 0; ix < ns; ++ix)
            for (int iy = 0; iy < ns; ++iy)
                for (int iz = 0; iz < ns; ++iz, ++idx) {
                    f.x[idx] = (ix + 0.5) * sp;
                    f.y[idx] = (iy + 0.5) * sp;
                    f.z[idx] = (iz + 0.5) * sp;
                    f.vx[idx] = 0.02 * ((ix + iy + iz) % 3 - 1);
                    f.vy[idx] = 0.02 * ((ix + 2 * iy) % 3 - 1);
                    f.vz[idx] = 0.02 * ((2 * ix + iz) % 3 - 1);
                }
        f.box = {L, L, L};
        hip::md::MD sim;
        sim.set_neighbor(hip::NeighborList{2.5, 0.3});
        sim.set_evaluator(hip::PairLJCut{1.0, 1.0, 2.5});
        sim.set_integrator(hip::md::NoseHooverNPT{1.0 / kB, 0.1, 0.0, 1.0, 0.005});
        sim.build();
        bool virial = false;
        for (const auto* c : sim.components())
            virial |= c->name() == "virial";
        REQUIRE(virial, "no Virial was injected for NoseHooverNPT");
        sim.run(f, 80);
        const double V = f.box[0] * f.box[1] * f.box[2];
        REQUIRE(std::fabs(V - 269.425432) <= 5e-7, "NPT volume %.9f", V);
        std::printf("hip::md::NoseHooverNPT through MD: V %.6f (reference CPU 269.425432), Virial injected\n", V);
    }
    // 4. thermo, singular setters, memory, roles, neighbor protocol
    {
        hip::HostFrame f = fcc(6);
        hip::md::MD sim;
        hip::NeighborList nl{2.5, 0.3, 1, 0, true};
        sim.set_neighbor(hip::NeighborList{2.5, 0.3, 1, 0, true});
        sim.set_neighbor(hip::NeighborList{2.5, 0.3, 1, 0, true}); // same name: ignored
        sim.set_evaluator(hip::PairLJCut{1.0, 1.0, 2.5});
        sim.set_integrator(hip::md::VelocityVerlet{0.005 * kTau});
        sim.set_thermo(10);
        sim.set_thermo(10);
        bool threw = false;
        try {
            sim.set_thermo(5);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        REQUIRE(threw, "set_thermo with another cadence must throw");
        sim.build();
        int providers = 0, consumers = 0;
        size_t nl_bytes = 0, lj_bytes = 0;
        sim.run(f, 50);
        for (const auto* c : sim.components()) {
            providers += c->as_neighbor_provider() != nullptr;
            consumers += c->as_neighbor_consumer() != nullptr;
            if (c->name() == "neighbor_list")
                nl_bytes = c->memory_bytes();
            if (c->name() == "lj") {
                lj_bytes = c->memory_bytes();
                REQUIRE(c->as_neighbor_consumer()->neighbor_cutoff() == 2.5, "consumer cutoff");
            }
        }
        REQUIRE(providers == 1 && consumers == 1, "protocols %d %d", providers, consumers);
        // single-type PairLJCut keeps its coefficients in kernel arguments: no device table
        REQUIRE(nl_bytes > f.nlocal() * 2 * 32 && lj_bytes == 0, "memory %zu %zu", nl_bytes, lj_bytes);
        REQUIRE(sim.thermo() && sim.thermo()->rows().size() == 5, "thermo rows %zu",
                sim.thermo() ? sim.thermo()->rows().size() : 0);
        const auto& r = sim.thermo()->rows().back();
        REQUIRE(r.step == 49 && r.potential == f.energy && r.kinetic == f.kinetic, "thermo row");
        REQUIRE(hip::md::Thermo{}.role() == role_bit(RoleFlag::FrameImmutable), "role");
        std::printf("set_thermo: %zu rows, T %.1f K; NeighborList %zu B, PairLJCut %zu B on the device\n",
                    sim.thermo()->rows().size(), r.temperature, nl_bytes, lj_bytes);
    }
    // 5. cpu::Frame-like frame
    {
        hip::HostFrame h = fcc(6);
        MockFrame m;
        const size_t n = h.nlocal();
        for (auto [c, v] : {std::pair{Col::x, &h.x}, {Col::y, &h.y}, {Col::z, &h.z}, {Col::vx, &h.vx},
                            {Col::vy, &h.vy}, {Col::vz, &h.vz}, {Col::fx, &h.fx}, {Col::fy, &h.fy},
                            {Col::fz, &h.fz}, {Col::mass, &h.mass}})
            m.cols[c] = *v;
        m.cols[Col::type].assign(n, 0.0);
        m.box = MockBox{{h.box[0], h.box[1], h.box[2]}};
        hip::md::FrameKeys<Col, Meta, Stress> keys{Col::x,  Col::y,  Col::z,  Col::vx,   Col::vy,   Col::vz,
                                               Col::fx, Col::fy, Col::fz, Col::mass, Col::type, Col::x,
                                               Meta::energy, Meta::kinetic, Stress::stress, false};
        auto drive = [](auto& frame, auto... k) {
            hip::md::MD sim;
            sim.set_neighbor(hip::NeighborList{2.5, 0.3, 1, 0, true});
            sim.set_evaluator(hip::PairLJCut{1.0, 1.0, 2.5});
            sim.set_integrator(hip::md::VelocityVerlet{0.005 * kTau});
            sim.set_thermostat(hip::md::NoseHoover{0.72 / kB, 0.5 * kTau, 0.005 * kTau});
            sim.run(frame, 40, k...);
            sim.run(frame, 40, k...);
        };
        drive(h);
        drive(m, keys);
        for (size_t i = 0; i < n; ++i)
            REQUIRE(m.cols[Col::x][i] == h.x[i] && m.cols[Col::vy][i] == h.vy[i] && m.cols[Col::fz][i] == h.fz[i],
                    "cpu::Frame-like run differs at %zu", i);
        REQUIRE(m.meta.e == h.energy && m.meta.k == h.kinetic, "frame energies");
        std::printf("run(cpu::Frame-like, keys) == run(HostFrame) (80 NVT steps in two calls)\n");
    }
    std::printf("hip driver boundary: all checks passed\n");
    return 0;
}
