"""The oracle's integrators and MD loop, pinned by closed forms (one VV step,
one NH update done by hand), InitVelocity's invariants, the LAMMPS decide
cadence and the reference's NVE energy gates.

Reference: velocity_verlet.cpp:24-68, nose_hoover.cpp:31-170, init.cpp:29-215,
runtime.cpp:452-502, tests/cpu/dynamics/test_neighbor_decide.cpp:31-80,
regressions/md/lj_argon/REFERENCE.md:73-90.
"""
import numpy as np
import pytest

from conftest import jittered_lattice

KB = 8.617333262e-5
MVV2E = 103.6427
TAU = np.sqrt(MVV2E)


def params(orc, **kw):
    p = dict(dt=0.005 * TAU, cutoff=2.5, skin=0.3, every=1, delay=0, check=1, eps=1.0, sigma=1.0, nh=0,
             T_target=0.72 / KB, tau_T=0.5 * TAU, ndof=0, nthreads=1, stress=1)
    p.update(kw)
    return orc.MDParams(**p)


def state(x, y, z, vx, vy, vz, mass):
    n = x.size
    return dict(x=x.copy(), y=y.copy(), z=z.copy(), vx=vx.copy(), vy=vy.copy(), vz=vz.copy(),
                fx=np.zeros(n), fy=np.zeros(n), fz=np.zeros(n), mass=mass.copy())


def test_init_velocity_invariants(orc):
    mass = np.where(np.arange(500) % 3 == 0, 2.0, 1.0)
    T = 300.0
    vx, vy, vz = orc.init_velocity(mass, T, 1234)
    p = np.array([np.sum(mass * v) for v in (vx, vy, vz)])
    assert np.max(np.abs(p)) < 1e-12
    ke = orc.kinetic_energy(mass, vx, vy, vz)
    assert ke == pytest.approx(0.5 * MVV2E * np.sum(mass * (vx * vx + vy * vy + vz * vz)), rel=1e-13)
    assert 2 * ke / ((3 * 500 - 3) * KB) == pytest.approx(T, rel=1e-12)
    # the stream is the seed's: same seed same draw, other seed other draw
    assert np.array_equal(orc.init_velocity(mass, T, 1234)[0], vx)
    assert not np.array_equal(orc.init_velocity(mass, T, 1235)[0], vx)


def test_one_velocity_verlet_step_by_hand(orc):
    """Step 1 from rest forces (fresh frame f = 0): x1 = wrap(x0 + dt v0),
    f1 = F(x1), v1 = v0 + dtf/m f1 — evaluated here with numpy."""
    x, y, z, box = jittered_lattice(4, 1.9)
    n = x.size
    mass = np.ones(n)
    rng = np.random.default_rng(4)
    vx, vy, vz = (rng.normal(0, 0.02, n) for _ in range(3))
    s = state(x, y, z, vx, vy, vz, mass)
    dt = 0.5
    p = params(orc, dt=dt, cutoff=2.5, skin=0.4)
    orc.md_run(s, box, p, 1)
    x1 = x + dt * vx
    x1 -= box[0] * np.floor(x1 / box[0])
    y1 = y + dt * vy
    y1 -= box[1] * np.floor(y1 / box[1])
    z1 = z + dt * vz
    z1 -= box[2] * np.floor(z1 / box[2])
    np.testing.assert_array_equal(s["x"], x1)
    half = orc.build_half(x1, y1, z1, box, 2.9)
    fx, fy, fz, e, _ = orc.pair_lj(x1, y1, z1, box, half, 1.0, 1.0, 2.5)
    dtf = 0.5 * dt * (1.0 / 103.6427)
    np.testing.assert_allclose(s["vx"], vx + dtf / mass * fx, rtol=0, atol=1e-15)
    np.testing.assert_allclose(s["fz"], fz, rtol=0, atol=1e-13)


def test_nose_hoover_update_by_hand(orc):
    """xi += dt/2 (2KE - Ndof kB T)/Q after each kick; v *= exp(-dt/2 xi);
    Q = Ndof kB T tau^2 (nose_hoover.cpp:60-61,131-168)."""
    x, y, z, box = jittered_lattice(4, 1.9)
    n = x.size
    mass = np.ones(n)
    T = 200.0
    vx, vy, vz = orc.init_velocity(mass, 2 * T, 9)
    s = state(x, y, z, vx, vy, vz, mass)
    dt, tau = 0.2, 10.0
    p = params(orc, dt=dt, nh=1, T_target=T, tau_T=tau, cutoff=2.5, skin=0.4)
    pe, ke, vir, st = orc.md_run(s, box, p, 1)
    ndof = 3 * n - 3
    Q = ndof * KB * T * tau * tau
    dtf = 0.5 * dt / 103.6427
    # initial half: kick with f = 0, drift
    v = np.stack([vx, vy, vz])
    k1 = 0.5 * MVV2E * np.sum(mass * v * v)
    xi = (dt * 0.5) * (2 * k1 - ndof * KB * T) / Q
    v = v * np.exp(-dt * 0.5 * xi)
    xx = [c + dt * vv for c, vv in zip((x, y, z), np.stack([vx, vy, vz]))]
    xx = [c - L * np.floor(c / L) for c, L in zip(xx, box)]
    half = orc.build_half(*xx, box, 2.9)
    f = np.stack(orc.pair_lj(*xx, box, half, 1.0, 1.0, 2.5)[:3])
    v = v + dtf * f
    k2 = 0.5 * MVV2E * np.sum(mass * v * v)
    xi = xi + (dt * 0.5) * (2 * k2 - ndof * KB * T) / Q
    v = v * np.exp(-dt * 0.5 * xi)
    assert st.xi == pytest.approx(xi, rel=1e-12)
    np.testing.assert_allclose(np.stack([s["vx"], s["vy"], s["vz"]]), v, rtol=1e-12, atol=1e-15)
    assert ke[0] == pytest.approx(0.5 * MVV2E * np.sum(mass * v * v), rel=1e-12)


def test_decide_cadence_delay_ten_check_no(orc):
    """test_neighbor_decide.cpp:31-54: every=1, delay=10, check=no rebuilds on
    the 10th decide after a build; 25 steps -> builds at 0, 10, 20."""
    x, y, z, box = jittered_lattice(4, 1.9)
    n = x.size
    s = state(x, y, z, np.zeros(n), np.zeros(n), np.zeros(n), np.ones(n))
    p = params(orc, dt=0.001, every=1, delay=10, check=0, skin=0.4)
    _, _, _, st = orc.md_run(s, box, p, 9)
    assert st.nrebuilds == 1
    s = state(x, y, z, np.zeros(n), np.zeros(n), np.zeros(n), np.ones(n))
    _, _, _, st = orc.md_run(s, box, p, 10)
    assert st.nrebuilds == 2
    s = state(x, y, z, np.zeros(n), np.zeros(n), np.zeros(n), np.ones(n))
    _, _, _, st = orc.md_run(s, box, p, 25)
    assert st.nrebuilds == 3


def test_nve_energy_gates_small_lj(orc):
    """REFERENCE.md gates of the LJ regression: max|E(t)-E(0)|/(N eps) <= 1e-2 and
    second-half drift <= 4x first-half drift; here on 864 atoms (6^3 FCC cells)
    at the configured rho*=0.8442, T*=0.72, dt*=0.005 for 600 steps."""
    a = (4 / 0.8442) ** (1 / 3)
    x, y, z = orc.fcc_lattice(6, 6, 6, a)
    n = x.size
    mass = np.ones(n)
    vx, vy, vz = orc.init_velocity(mass, 0.72 / KB, 87287)
    box = np.full(3, 6 * a)
    s = state(x, y, z, vx, vy, vz, mass)
    pe, ke, _, st = orc.md_run(s, box, params(orc, nthreads=4), 600)
    E = pe + ke
    # the first step starts from f = 0 (fresh frame): measure from step 1
    E = E[1:]
    assert np.max(np.abs(E - E[0])) / n <= 1e-2
    h = E.size // 2
    d1 = abs(E[h] - E[0])
    d2 = abs(E[-1] - E[h])
    assert d2 <= 4 * d1 + 1e-6 * n
    assert st.nrebuilds > 10
