"""Pins of the oracle's kinetic stress (cpu::md::Virial::launch,
src/cpu/observe/virial_stress.cpp:14-56): its trace is twice the kinetic
energy of the oracle's KineticEnergy, a closed form for a uniform flow, and
relabelling the axes relabels the tensor."""
import numpy as np
import pytest

MVV2E = 103.6427  # units::mass_v2_to_energy (include/atomiverse/core/units.h:166)


def _sample(n=2000, seed=3):
    rng = np.random.default_rng(seed)
    m = rng.uniform(1.0, 40.0, n)
    v = rng.normal(0.0, 0.01, (3, n))
    return m, v[0].copy(), v[1].copy(), v[2].copy()


def test_trace_is_twice_the_kinetic_energy(orc):
    m, vx, vy, vz = _sample()
    k = orc.kinetic_stress(m, vx, vy, vz)
    ke = orc.kinetic_energy(m, vx, vy, vz)
    assert k[0] + k[1] + k[2] == pytest.approx(2.0 * ke, rel=1e-13)


def test_uniform_flow_closed_form(orc):
    n, mass, u = 500, 12.0, 0.02
    m = np.full(n, mass)
    vx = np.full(n, u)
    zero = np.zeros(n)
    k = orc.kinetic_stress(m, vx, zero, zero)
    want = np.zeros(6)
    want[0] = n * mass * u * u * MVV2E
    np.testing.assert_allclose(k, want, rtol=1e-13, atol=0.0)
    # flow along (1, 1, 0)/sqrt(2): xx = yy = xy, the rest zero
    w = np.full(n, u / np.sqrt(2.0))
    k = orc.kinetic_stress(m, w, w, zero)
    assert k[0] == pytest.approx(k[1], rel=1e-15) and k[0] == pytest.approx(k[3], rel=1e-15)
    assert k[0] == pytest.approx(0.5 * n * mass * u * u * MVV2E, rel=1e-12)
    assert k[2] == 0.0 and k[4] == 0.0 and k[5] == 0.0


def test_axis_relabelling(orc):
    m, vx, vy, vz = _sample(seed=11)
    k = orc.kinetic_stress(m, vx, vy, vz)
    s = orc.kinetic_stress(m, vy, vx, vz)  # swap x and y
    np.testing.assert_allclose(s[[1, 0, 2, 3, 5, 4]], k, rtol=1e-15, atol=0.0)
