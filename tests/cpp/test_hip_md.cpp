// Drives the hip:: C++ components (include/atomiverse/hip/md.hpp) the way user
// code drives cuda::md::MD: a 4000-atom FCC LJ fluid (rho* = 0.8442, T* = 0.72,
// rc = 2.5, skin 0.3, dt* = 0.005 in metal units) for 200 NVE steps through the
// fused device loop and through the per-phase component walk; the two must give
// bit-identical trajectories and conserve energy. Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "atomiverse/hip/md.hpp"
#include "atomiverse/hip/remd.hpp"

using namespace ATV;

static hip::HostFrame make_frame() {
    const double a = std::cbrt(4.0 / 0.8442);
    const int nc = 10;
    hip::HostFrame f;
    const int64_t n = atvh_util_fcc_lattice(nc, nc, nc, a, -1, nullptr, nullptr, nullptr);
    f.resize(static_cast<size_t>(n));
    atvh_util_fcc_lattice(nc, nc, nc, a, -1, f.x.data(), f.y.data(), f.z.data());
    f.box = {nc * a, nc * a, nc * a};
    const double kB = 8.617333262e-5;
    atvh_util_init_velocity(n, f.mass.data(), 0.72 / kB, 87287u, 3 * n - 3, f.vx.data(), f.vy.data(), f.vz.data());
    return f;
}

static void run(hip::HostFrame& f, bool fused, int steps, double& e_first, double& e_last) {
    const double tau = std::sqrt(103.6427);
    const double dt = 0.005 * tau;
    hip::md::MD sim;
    sim.set_neighbor(hip::NeighborList{2.5, 0.3, 1, 0, true});
    sim.set_evaluator(hip::PairLJCut{1.0, 1.0, 2.5});
    sim.set_integrator(hip::md::VelocityVerlet{dt});
    sim.set_fused(fused);
    sim.build();
    if (sim.fused() != fused) {
        std::fprintf(stderr, "unexpected fused mode\n");
        std::exit(2);
    }
    sim.run(f, 1); // first step starts from f = 0 (fresh frame)
    sim.run(f, 1);
    e_first = f.energy + f.kinetic;
    sim.run(f, steps - 2);
    e_last = f.energy + f.kinetic;
}

// Charged binary melt (rock salt 10^3, rho* = 0.9, T* = 1.2, +-1 reduced
// charges, sigma_A 1.0 / sigma_B 0.8) with PairLJCutCoulLong + KSpacePME under
// Nose-Hoover, configured as the reference's MD would be.
static hip::HostFrame make_melt() {
    const double a = std::cbrt(1.0 / 0.9);
    const int nc = 10;
    hip::HostFrame f;
    const int64_t n = atvh_util_rocksalt_lattice(nc, nc, nc, a, nullptr, nullptr, nullptr, nullptr, nullptr, 0.0);
    f.resize(static_cast<size_t>(n));
    f.charge.assign(static_cast<size_t>(n), 0.0);
    atvh_util_rocksalt_lattice(nc, nc, nc, a, f.x.data(), f.y.data(), f.z.data(), f.type.data(), f.charge.data(),
                               1.0 / std::sqrt(14.399645));
    f.box = {nc * a, nc * a, nc * a};
    const double kB = 8.617333262e-5;
    atvh_util_init_velocity(n, f.mass.data(), 1.2 / kB, 87289u, 3 * n - 3, f.vx.data(), f.vy.data(), f.vz.data());
    return f;
}

static void run_melt(hip::HostFrame& f, bool fused, int steps) {
    const double tau = std::sqrt(103.6427);
    const double dt = 0.005 * tau;
    const double alpha = (1.35 - 0.15 * std::log(1e-5)) / 3.0;
    hip::md::MD sim;
    sim.set_neighbor(hip::NeighborList{3.0, 0.3, 1, 0, true});
    sim.set_evaluator(hip::PairLJCutCoulLong{{{0, 0, 1.0, 1.0}, {1, 1, 1.0, 0.8}}, alpha, 3.0});
    sim.set_evaluator(hip::KSpacePME{alpha, 0.65, 4});
    sim.set_integrator(hip::md::VelocityVerlet{dt});
    sim.set_thermostat(hip::md::NoseHoover{1.2 / 8.617333262e-5, 0.5 * tau, dt});
    sim.set_precision(hip::Precision::mixed);
    sim.set_fused(fused);
    sim.build();
    if (sim.fused() != fused) {
        std::fprintf(stderr, "unexpected fused mode (charged)\n");
        std::exit(2);
    }
    sim.run(f, steps);
}

int main() {
    hip::HostFrame a = make_frame(), b = make_frame();
    double a0, a1, b0, b1;
    run(a, true, 200, a0, a1);
    run(b, false, 200, b0, b1);
    const size_t n = a.nlocal();
    for (size_t i = 0; i < n; ++i) {
        if (a.x[i] != b.x[i] || a.vy[i] != b.vy[i] || a.fz[i] != b.fz[i]) {
            std::fprintf(stderr, "fused and per-phase runs differ at atom %zu\n", i);
            return 1;
        }
    }
    const double drift = std::fabs(a1 - a0) / double(n);
    std::printf("hip::md::MD: %zu atoms, E0/N %.6f, |dE|/N %.3e, per-phase == fused\n", n, a0 / n, drift);
    if (!(drift < 1e-2) || !std::isfinite(a1))
        return 1;
    {
        // the Virial observer on the per-phase walk: pressure of the LJ fluid
        hip::HostFrame v = make_frame();
        const double tau = std::sqrt(103.6427);
        hip::md::MD sim;
        hip::md::Virial virial;
        sim.set_neighbor(hip::NeighborList{2.5, 0.3, 1, 0, true});
        sim.set_evaluator(hip::PairLJCut{1.0, 1.0, 2.5});
        sim.set_integrator(hip::md::VelocityVerlet{0.005 * tau});
        sim.on(ATV::md::Phase::post_force).add_ref(virial, -50);
        sim.build();
        if (sim.fused()) {
            std::fprintf(stderr, "virial: the observer needs the per-phase walk\n");
            return 1;
        }
        sim.run(v, 20);
        double kin[6], vir[6], p = 0.0;
        virial.read(sim.context(), kin, vir, &p);
        const double vol = v.box[0] * v.box[1] * v.box[2];
        const double want = (kin[0] + kin[1] + kin[2] + vir[0] + vir[1] + vir[2]) / (3.0 * vol);
        if (!std::isfinite(p) || std::fabs(p - want) > 1e-12 * std::fabs(want) || !(kin[0] > 0.0) || vir[0] == 0.0) {
            std::fprintf(stderr, "virial: bad pressure %g (want %g)\n", p, want);
            return 1;
        }
        std::printf("hip::md::Virial: P %.6f eV/A^3, tr K %.3f eV, tr W %.3f eV\n", p, kin[0] + kin[1] + kin[2],
                    vir[0] + vir[1] + vir[2]);
    }
    hip::HostFrame c = make_melt(), d = make_melt();
    run_melt(c, true, 60);
    run_melt(d, false, 60);
    for (size_t i = 0; i < c.nlocal(); ++i) {
        if (c.x[i] != d.x[i] || c.vz[i] != d.vz[i] || c.fy[i] != d.fy[i]) {
            std::fprintf(stderr, "charged: fused and per-phase runs differ at atom %zu\n", i);
            return 1;
        }
    }
    if (!std::isfinite(c.energy) || !(c.energy < 0.0)) {
        std::fprintf(stderr, "charged: bad energy %g\n", c.energy);
        return 1;
    }
    std::printf("hip::md::MD charged melt: %zu ions, PE/N %.6f, per-phase == fused (PME)\n", c.nlocal(),
                c.energy / c.nlocal());
    // replica exchange over a 4-temperature ladder (cpu::path::Remd's interface)
    {
        const double a = std::cbrt(4.0 / 0.8442);
        const int nc = 6;
        hip::HostFrame f;
        const int64_t n = atvh_util_fcc_lattice(nc, nc, nc, a, -1, nullptr, nullptr, nullptr);
        f.resize(static_cast<size_t>(n));
        atvh_util_fcc_lattice(nc, nc, nc, a, -1, f.x.data(), f.y.data(), f.z.data());
        f.box = {nc * a, nc * a, nc * a};
        const double kB = 8.617333262e-5, tau = std::sqrt(103.6427);
        hip::path::Remd remd({0.70 / kB, 0.74 / kB, 0.78 / kB, 0.82 / kB}, 0.005 * tau, 25, 7);
        remd.thermostat_tau = 0.5 * tau;
        remd.set_neighbor(hip::NeighborList{2.5, 0.3, 1, 0, true});
        remd.set_evaluator(hip::PairLJCut{1.0, 1.0, 2.5});
        remd.set_precision(hip::Precision::mixed);
        remd.run(f, 8);
        if (remd.exchange_attempts() != 12 || !std::isfinite(f.energy) || !(f.energy < 0.0)) {
            std::fprintf(stderr, "remd: attempts %d energy %g\n", remd.exchange_attempts(), f.energy);
            return 1;
        }
        std::printf("hip::path::Remd: %d attempts, %d accepted, lowest-T replica PE/N %.6f\n",
                    remd.exchange_attempts(), remd.exchange_accepts(), f.energy / f.nlocal());
    }
    // a pair evaluator without a neighbor list is rejected, as cuda::md::MD does
    try {
        hip::md::MD bad;
        bad.set_evaluator(hip::PairLJCut{1.0, 1.0, 2.5});
        bad.set_integrator(hip::md::VelocityVerlet{1.0});
        bad.build();
        return 1;
    } catch (const std::invalid_argument&) {
    }
    return 0;
}
