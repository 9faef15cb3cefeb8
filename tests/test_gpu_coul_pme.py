"""Charged systems on the HIP backend against the CPU oracle, through the C ABI:
PairLJCutCoulLong (LJ + tabulated erfc Coulomb) and KSpacePME.

Bars (each written in its test):
  * FP64 pair forces: rounding only, max |dF| <= 1e-12 max |F|;
  * FP64 PME: the oracle transforms with its own DFT and spreads in FP64 in
    atom order, the device with an exact fixed-point grid and hipFFT, so the
    bar is 1e-9 of max |F| for forces and 1e-9 relative for energies (FP32
    transforms of power-of-two meshes are the device's own kernels);
  * mixed precision (FP32 pair / spread / FFT / gather, FP64 accumulation):
    per-atom force RMS relative error <= 1e-5 against the FP64 oracle;
  * physics: the NaCl Madelung energy within 5e-4 (test_pme_coul.cpp:157-175).
"""
import math

import numpy as np
import pytest

import atomiverse_amd as av
from atomiverse_amd import workloads as W

pytestmark = pytest.mark.gpu

COUL = W.QQR2E


def snapshot(nside, jitter=0.06, seed=3):
    """Charged-melt workload with the lattice sites displaced (a thermalised-
    like configuration: the perfect lattice has near-zero forces)."""
    w = W.charged_melt(nside)
    rng = np.random.default_rng(seed)
    # wrapped into [0, L): the reference builds its list from wrapped positions
    w.x = np.mod(w.x + rng.normal(0, jitter, w.n), w.box[0])
    w.y = np.mod(w.y + rng.normal(0, jitter, w.n), w.box[1])
    w.z = np.mod(w.z + rng.normal(0, jitter, w.n), w.box[2])
    return w


def oracle_tables(orc, w):
    return orc.LJCoulTables(w.eps_tab, w.sigma_tab, w.set_tab, w.cutoff, w.alpha)


def make(ctx, w, pme=True):
    s = av.System(ctx, w.n, w.box)
    W.upload(s, w)
    s.nlist_configure(w.cutoff, w.skin, 1, 0, True)
    s.pair_ljcoul(w.eps_tab, w.sigma_tab, w.set_tab, w.alpha, w.cutoff)
    if pme:
        s.pme(w.alpha, w.pme_spacing, w.pme_order)
    s.integrator(w.dt, w.thermostat, w.T, w.tau_T, 0)
    return s


def forces(d):
    return np.stack([d["fx"], d["fy"], d["fz"]])


def rms_rel(F, G):
    return float(np.sqrt(np.sum((F - G) ** 2) / np.sum(G ** 2)))


@pytest.mark.parametrize("nside", [12, 6])  # brick kernels / small-grid kernels
def test_ljcoul_pair_matches_oracle(hip_ctx, orc, nside):
    w = snapshot(nside)
    s = make(hip_ctx, w, pme=False)
    s.nlist_build()
    s.compute(av.PREC_FP64, True, True)
    g = s.download()
    pe, vir, _ = s.energy()
    half = orc.build_half(w.x, w.y, w.z, w.box, w.cutoff + w.skin, nthreads=8)
    fx, fy, fz, e, ovir = orc.pair_lj_coul_long(w.x, w.y, w.z, w.charge, w.type_, w.box, half,
                                                oracle_tables(orc, w), nthreads=8)
    F = np.stack([fx, fy, fz])
    assert np.max(np.abs(forces(g) - F)) <= 1e-12 * np.max(np.abs(F))
    assert pe == pytest.approx(e, rel=1e-12)
    np.testing.assert_allclose(vir, ovir, rtol=1e-10, atol=1e-10 * np.abs(ovir[:3]).max())
    s.compute(av.PREC_MIXED, True, True)
    m = s.download()
    assert rms_rel(forces(m), F) <= 1e-5
    pe_m, vir_m, _ = s.energy()
    assert pe_m == pytest.approx(e, rel=1e-5)
    assert np.linalg.norm(vir_m - ovir) <= 1e-5 * np.linalg.norm(ovir)


@pytest.mark.parametrize("order", [4, 5, 6])
def test_pme_matches_oracle(hip_ctx, orc, order):
    w = snapshot(12, seed=order)
    w.pme_order = order
    w.pme_spacing = 0.55
    s = make(hip_ctx, w)
    K, _ = s.pme_info()
    Ko = orc.pme_grid(w.box, w.pme_spacing, order)
    np.testing.assert_array_equal(K, Ko)
    gx, gy, gz, e, ovir, parts = orc.pme(w.x, w.y, w.z, w.charge, w.box, w.alpha, Ko, order)
    F = np.stack([gx, gy, gz])
    s.pme_compute(av.PREC_FP64, True, True)
    d = s.download()
    pe, vir, _ = s.energy()
    _, gparts = s.pme_info()
    assert np.max(np.abs(forces(d) - F)) <= 1e-9 * np.max(np.abs(F))
    assert pe == pytest.approx(e, rel=1e-9)
    assert gparts[0] == pytest.approx(parts[0], rel=1e-9)
    assert gparts[1] == pytest.approx(parts[1], rel=1e-14)
    assert gparts[2] == pytest.approx(parts[2], abs=1e-12)
    np.testing.assert_allclose(vir, ovir, rtol=1e-8, atol=1e-9 * np.abs(ovir[:3]).max())
    s.pme_compute(av.PREC_MIXED, True, True)
    m = s.download()
    assert rms_rel(forces(m), F) <= 1e-5
    pe_m, vir_m, _ = s.energy()
    assert pe_m == pytest.approx(e, rel=1e-5)
    assert np.linalg.norm(vir_m - ovir) <= 1e-5 * np.linalg.norm(ovir)


@pytest.mark.parametrize("K,order,stretch", [(32, 4, 1), (32, 5, 1), (32, 8, 1), (64, 6, 1), (32, 4, 2)])
def test_native_transform_meshes_match_oracle(hip_ctx, orc, K, order, stretch):
    """Power-of-two meshes transform in FP32 with the library's own kernels
    (five passes over the half spectrum, the influence function applied
    between the forward and inverse x-lines; FP64 stays on hipFFT); the bars
    are those of the other meshes above. stretch 2 doubles the box along z,
    so the mesh is K x K x 2K and every axis runs a different line length.
    The run switches precision back and forth, so the energy partials of
    both transform paths are read with their own block counts."""
    w = snapshot(12, seed=40 + order)
    w.pme_order = order
    if stretch != 1:
        w.box = np.array([w.box[0], w.box[1], stretch * w.box[2]])
        w.z = w.z * stretch
    w.pme_spacing = w.box[0] / K * (1.0 + 1e-9)
    s = make(hip_ctx, w)
    Kd, _ = s.pme_info()
    Ko = orc.pme_grid(w.box, w.pme_spacing, order)
    np.testing.assert_array_equal(Kd, Ko)
    np.testing.assert_array_equal(Kd, [K, K, stretch * K])
    gx, gy, gz, e, ovir, parts = orc.pme(w.x, w.y, w.z, w.charge, w.box, w.alpha, Ko, order)
    F = np.stack([gx, gy, gz])
    s.pme_compute(av.PREC_FP64, True, True)
    d = s.download()
    pe, vir, _ = s.energy()
    _, gparts = s.pme_info()
    assert np.max(np.abs(forces(d) - F)) <= 1e-9 * np.max(np.abs(F))
    assert pe == pytest.approx(e, rel=1e-9)
    assert gparts[0] == pytest.approx(parts[0], rel=1e-9)
    np.testing.assert_allclose(vir, ovir, rtol=1e-8, atol=1e-9 * np.abs(ovir[:3]).max())
    s.pme_compute(av.PREC_MIXED, True, True)
    m = s.download()
    assert rms_rel(forces(m), F) <= 1e-5
    pe_m, vir_m, _ = s.energy()
    assert pe_m == pytest.approx(e, rel=1e-5)
    assert np.linalg.norm(vir_m - ovir) <= 1e-5 * np.linalg.norm(ovir)
    # evaluations repeat bit for bit, also after the other precision ran
    s.pme_compute(av.PREC_FP64, True, True)
    assert s.energy()[0] == pe
    s.pme_compute(av.PREC_MIXED, True, True)
    np.testing.assert_array_equal(forces(s.download()), forces(m))
    assert s.energy()[0] == pe_m


def test_pme_is_bitwise_reproducible(hip_ctx):
    """Fixed-point spreading makes the charge grid independent of the order of
    the atomic adds: repeated evaluations agree bit for bit."""
    w = snapshot(12, seed=11)
    s = make(hip_ctx, w)
    for prec in (av.PREC_FP64, av.PREC_MIXED):
        outs = []
        for _ in range(3):
            s.pme_compute(prec, True, True)
            outs.append((forces(s.download()), s.energy()[0]))
        for F, e in outs[1:]:
            np.testing.assert_array_equal(F, outs[0][0])
            assert e == outs[0][1]


@pytest.mark.parametrize("nside", [12, 24])
def test_pme_spread_groupings_agree_bitwise(hip_ctx, nside):
    """Each spread workgroup sums its atoms' weights in an LDS box of integer
    points and adds the box to the integer grid. Before any list build a
    workgroup takes 256 consecutive slots (the largest box, 6144 points);
    after the build it takes one brick, with a box sized from the brick's
    extent plus the skin. Integer sums do not depend on the grouping, so the
    two give the same forces and energies bit for bit, in both precisions."""
    w = snapshot(nside, seed=5)
    s = make(hip_ctx, w)
    before = {}
    for prec in (av.PREC_FP64, av.PREC_MIXED):
        s.pme_compute(prec, True, True)
        before[prec] = (forces(s.download()), s.energy()[0], s.pme_info()[1])
    s.nlist_build()
    cand, bdim = s.brick_shape()
    assert cand >= 0, "the brick layout is the case under test"
    d = s.download()
    for k in ("x", "y", "z"):  # the build's wrap leaves wrapped positions as they are
        np.testing.assert_array_equal(d[k], getattr(w, k))
    for prec in (av.PREC_FP64, av.PREC_MIXED):
        s.pme_compute(prec, True, True)
        F, e, parts = forces(s.download()), s.energy()[0], s.pme_info()[1]
        np.testing.assert_array_equal(F, before[prec][0])
        assert e == before[prec][1]
        assert parts == before[prec][2]


def test_madelung_nacl_on_device(hip_ctx):
    """test_pme_coul.cpp:157-175 with the device pair style (LJ off) + PME."""
    d, n = 2.82, 4
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x, y, z = (a.reshape(-1).astype(np.float64) * d for a in (i, j, k))
    q = np.where((i + j + k).reshape(-1) % 2 == 0, 1.0, -1.0)
    box = np.full(3, n * d)
    cutoff = 1.9 * d
    alpha = 3.5 / cutoff
    s = av.System(hip_ctx, x.size, box)
    s.upload(x, y, z, np.ones(x.size), charge=q)
    s.nlist_configure(cutoff, 0.3, 1, 0, True)
    s.pair_ljcoul(np.zeros((1, 1)), np.ones((1, 1)), np.ones((1, 1), np.int32), alpha, cutoff)
    s.pme(alpha, 0.5, 6)
    s.nlist_build()
    want = -0.5 * x.size * 1.747564594633182 * COUL / d
    for prec in (av.PREC_FP64, av.PREC_MIXED):
        s.compute(prec, True, True)
        pe, _, _ = s.energy()
        assert abs(pe / want - 1.0) <= 5e-4


def _oracle_state(w):
    n = w.n
    return dict(x=w.x.copy(), y=w.y.copy(), z=w.z.copy(), vx=w.vx.copy(), vy=w.vy.copy(), vz=w.vz.copy(),
                fx=np.zeros(n), fy=np.zeros(n), fz=np.zeros(n), mass=w.mass.copy())


def test_charged_trajectory_matches_oracle(hip_ctx, orc):
    """40 NVT steps of the full loop (VV+NH, decide/rebuild, LJ+Coulomb, PME)
    from identical inputs: positions within 1e-8 Å, energies to 1e-9."""
    w = W.charged_melt(10)
    s = make(hip_ctx, w)
    nsteps = 40
    s.run(nsteps, av.PREC_FP64, energy_every=1)
    g = s.download()
    pe_g, ke_g = s.thermo(nsteps)
    o = _oracle_state(w)
    p = orc.MDParams(dt=w.dt, cutoff=w.cutoff, skin=w.skin, every=1, delay=0, check=1, eps=0, sigma=0, nh=1,
                     T_target=w.T, tau_T=w.tau_T, ndof=0, nthreads=8, stress=1)
    K = orc.pme_grid(w.box, w.pme_spacing, w.pme_order)
    pe_o, ke_o, _, st = orc.md_run_ff(o, w.box, p, oracle_tables(orc, w), w.type_, w.charge, K, w.pme_order,
                                      nsteps)
    for k in ("x", "y", "z"):
        d = g[k] - o[k]
        d -= w.box[0] * np.rint(d / w.box[0])
        assert np.max(np.abs(d)) < 1e-8, k
    np.testing.assert_allclose(pe_g, pe_o, rtol=1e-9)
    np.testing.assert_allclose(ke_g, ke_o, rtol=1e-9)
    assert s.thermostat_xi() == pytest.approx(st.xi, rel=1e-7, abs=1e-12)
    assert s.nlist_stats()["nrebuilds"] == st.nrebuilds


def test_charged_nve_energy_gates(hip_ctx):
    """NVE, 1000 steps, both precisions: max|E-E0|/N <= 1e-2 eV and the late
    drift no more than 4x the early drift (the LJ regression gates)."""
    for prec in (av.PREC_FP64, av.PREC_MIXED):
        w = W.charged_melt(10, thermostat=0)
        s = make(hip_ctx, w)
        s.run(1000, prec, energy_every=1)
        pe, ke = s.thermo(1000)
        E = (pe + ke)[1:]
        assert np.all(np.isfinite(E))
        assert np.max(np.abs(E - E[0])) / w.n <= 1e-2
        h = E.size // 2
        assert abs(E[-1] - E[h]) <= 4 * abs(E[h] - E[0]) + 1e-6 * w.n


def test_charged_errors(hip_ctx):
    w = snapshot(6)
    s = av.System(hip_ctx, w.n, w.box)
    s.upload(w.x, w.y, w.z, w.mass)  # no charges
    s.nlist_configure(w.cutoff, w.skin, 1, 0, True)
    with pytest.raises(av.HipError) as e:
        s.pme(0.0, 0.5, 4)
    assert e.value.code == av._abi.E_ARG
    with pytest.raises(av.HipError):
        s.pme(1.0, 0.5, 2)
    with pytest.raises(av.HipError):
        s.pme(1.0, 0.5, 9)
    with pytest.raises(av.HipError):
        s.pme(1.0, -0.5, 4)
    with pytest.raises(av.HipError):
        s.pair_ljcoul(w.eps_tab, w.sigma_tab, w.set_tab, -1.0, w.cutoff)
    s.pme(1.0, 0.5, 4)
    with pytest.raises(av.HipError) as e:
        s.pme_compute()
    assert e.value.code == av._abi.E_STATE
