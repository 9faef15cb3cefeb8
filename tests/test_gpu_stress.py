"""The Virial observer on the device (cpu::md::Virial, src/cpu/observe/
virial_stress.cpp:14-56): the kinetic stress of the current velocities,
pending Nose-Hoover scale included, equals the oracle's on the downloaded
velocities; the pressure is (tr K + tr W) / (3V) with the virial of the last
force evaluation, which equals the oracle's; an ideal gas (no pair within the
cutoff) has P V = 2/3 KE."""
import numpy as np
import pytest

import atomiverse_amd as av
from atomiverse_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_kinetic_stress_and_pressure_c0(hip_ctx, orc):
    w = W.lj_fcc(10, seed=87287, thermostat=1, name="lj_fcc_4000_nvt")
    s = av.System(hip_ctx, w.n, w.box)
    W.upload(s, w)
    W.configure(s, w)
    s.nlist_build()
    s.run(40, av.PREC_FP64)  # Nose-Hoover leaves a pending velocity scale
    s.kinetic_stress()
    d = s.download()
    m = np.asarray(w.mass, dtype=np.float64)
    want_k = orc.kinetic_stress(m, d["vx"], d["vy"], d["vz"])
    s.compute(av.PREC_FP64, True, True)
    kin, vir, p = s.stress()
    # fixed-order block sums against a serial sum: rounding only
    np.testing.assert_allclose(kin, want_k, rtol=1e-12, atol=1e-12 * np.abs(want_k[:3]).max())
    pe, vir_e, ke = s.energy()
    np.testing.assert_array_equal(vir, vir_e)
    half = orc.build_half(d["x"], d["y"], d["z"], w.box, 2.8, nthreads=8)
    _, _, _, _, ovir = orc.pair_lj(d["x"], d["y"], d["z"], w.box, half, 1.0, 1.0, 2.5, nthreads=8)
    np.testing.assert_allclose(vir, ovir, rtol=1e-11)
    vol = float(np.prod(w.box))
    assert p == pytest.approx((want_k[:3].sum() + ovir[:3].sum()) / (3.0 * vol), rel=1e-11)
    # trace of the kinetic stress is twice the kinetic energy of the same velocities
    assert kin[:3].sum() == pytest.approx(2.0 * orc.kinetic_energy(m, d["vx"], d["vy"], d["vz"]), rel=1e-12)


def test_ideal_gas_pressure(hip_ctx):
    """Atoms farther apart than the cutoff: W = 0 and P = N kB T_kin / V with
    T_kin from the same kinetic energy (3/2 N kB T = KE)."""
    n_side, a = 6, 4.0
    g = (np.arange(n_side) + 0.5) * a
    x, y, z = [c.ravel().copy() for c in np.meshgrid(g, g, g, indexing="ij")]
    n = x.size
    box = np.full(3, n_side * a)
    rng = np.random.default_rng(5)
    v = rng.normal(0.0, 0.005, (3, n))
    s = av.System(hip_ctx, n, box)
    s.upload(x, y, z, np.full(n, 39.948), v[0].copy(), v[1].copy(), v[2].copy())
    s.nlist_configure(2.5, 0.3, 1, 0, True)
    s.pair_lj(0.0104, 3.4 / 1.12, 2.5)
    s.integrator(1.0, 0, 0.0, 0.0, 0)
    s.nlist_build()
    s.compute(av.PREC_FP64, True, True)
    s.kinetic_stress()
    kin, vir, p = s.stress()
    assert np.all(vir == 0.0)
    ke = 0.5 * 103.6427 * 39.948 * np.sum(v * v)
    assert kin[:3].sum() == pytest.approx(2.0 * ke, rel=1e-12)
    assert p * box.prod() == pytest.approx(2.0 / 3.0 * ke, rel=1e-12)
