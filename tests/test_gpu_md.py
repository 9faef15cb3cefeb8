"""HIP backend parity against the CPU oracle, through the C ABI.

Bars (the tolerances are written in each test):
  * neighbor lists: bit-exact as unordered tag-pair sets on identical positions;
  * FP64 forces/energy/virial: rounding only (summation order, FMA);
  * mixed precision: per-atom force RMS relative error <= 1e-5;
  * trajectories: agreement over a short horizon; NVE gates of the reference.
"""
import numpy as np
import pytest

import atomiverse_amd as av
from atomiverse_amd import workloads as W
from conftest import jittered_lattice

pytestmark = pytest.mark.gpu

KB = W.KB


def make_system(ctx, w):
    s = av.System(ctx, w.n, w.box)
    W.upload(s, w)
    W.configure(s, w)
    return s


def oracle_pairs(orc, x, y, z, box, rcs):
    nn, first, nb = orc.build_half(x, y, z, box, rcs, nthreads=8)
    return orc.pair_keys(orc.half_pairs(nn, first, nb), x.size)


def gpu_pairs(s):
    p = s.export_pairs()
    return np.sort(p[:, 0] * s.n + p[:, 1])


CASE_SEEDS = {"jittered": 1100, "small_grid": 1101, "random_dense": 1102, "anisotropic": 1103}


def random_case(name):
    rng = np.random.default_rng(CASE_SEEDS[name])
    if name == "jittered":
        x, y, z, box = jittered_lattice(4, 1.9)
        return x, y, z, box, 2.6, 0.4
    if name == "small_grid":  # 4 bins < 2*reach+1: min-image fallback
        L = 6.0
        x, y, z = (rng.uniform(0, L, 120) for _ in range(3))
        return x, y, z, np.full(3, L), 2.5, 0.3
    if name == "random_dense":
        L = 12.0
        x, y, z = (rng.uniform(0, L, 1500) for _ in range(3))
        return x, y, z, np.full(3, L), 2.5, 0.3
    if name == "anisotropic":
        box = np.array([9.0, 14.0, 21.0])
        x, y, z = (rng.uniform(0, L, 2000) for L in box)
        return x, y, z, box, 2.5, 0.3
    raise KeyError(name)


@pytest.mark.parametrize("case", ["jittered", "small_grid", "random_dense", "anisotropic"])
def test_pair_set_bit_exact_small(hip_ctx, orc, case):
    x, y, z, box, rc, skin = random_case(case)
    s = av.System(hip_ctx, x.size, box)
    s.upload(x, y, z, np.ones(x.size))
    s.nlist_configure(rc, skin, 1, 0, True)
    s.nlist_build()
    got = gpu_pairs(s)
    want = oracle_pairs(orc, x, y, z, box, rc + skin)
    assert want.size > 0
    np.testing.assert_array_equal(got, want)
    st = s.nlist_stats()
    assert st["npairs_half"] == want.size


def test_pair_set_and_forces_c0_snapshot(hip_ctx, orc):
    """4000-atom FCC fluid: lattice pair set, then 300 GPU steps (FP64), then
    the pair set and FP64 forces/energy/virial on the GPU's own positions."""
    w = W.lj_fcc_4000()
    s = make_system(hip_ctx, w)
    s.nlist_build()
    np.testing.assert_array_equal(gpu_pairs(s), oracle_pairs(orc, w.x, w.y, w.z, w.box, 2.8))
    s.run(300, av.PREC_FP64)
    st = s.nlist_stats()
    assert st["nrebuilds"] > 5
    d = s.download()
    s.nlist_build()  # identical snapshot: rebuild from the positions just read
    want = oracle_pairs(orc, d["x"], d["y"], d["z"], w.box, 2.8)
    np.testing.assert_array_equal(gpu_pairs(s), want)
    # FP64 pair forces on this snapshot
    s.compute(av.PREC_FP64, True, True)
    g = s.download()
    pe, vir, _ = s.energy()
    half = orc.build_half(d["x"], d["y"], d["z"], w.box, 2.8, nthreads=8)
    fx, fy, fz, e, ovir = orc.pair_lj(d["x"], d["y"], d["z"], w.box, half, 1.0, 1.0, 2.5, nthreads=8)
    F = np.stack([fx, fy, fz])
    G = np.stack([g["fx"], g["fy"], g["fz"]])
    # FP64 bar: rounding only — 1e-12 of the largest force component
    assert np.max(np.abs(G - F)) <= 1e-12 * np.max(np.abs(F))
    assert pe == pytest.approx(e, rel=1e-12)
    np.testing.assert_allclose(vir, ovir, rtol=1e-11)
    # mixed precision on the same snapshot: force RMS relative error <= 1e-5
    s.compute(av.PREC_MIXED, True, True)
    m = s.download()
    Mf = np.stack([m["fx"], m["fy"], m["fz"]])
    rms = np.sqrt(np.sum((Mf - F) ** 2) / np.sum(F ** 2))
    assert rms <= 1e-5, rms
    pe_m, vir_m, _ = s.energy()
    assert pe_m == pytest.approx(e, rel=1e-6)
    # virial tensor: 1e-5 relative in Frobenius norm (off-diagonals are ~1e-3 of the trace)
    assert np.linalg.norm(vir_m - ovir) <= 1e-5 * np.linalg.norm(ovir)


def test_two_atoms_at_sigma_give_minus_vshift(hip_ctx):
    """test_pair_lj_cut.cpp:22-62 on the device."""
    eps, sigma, rc = 0.5, 3.0, 8.0
    s = av.System(hip_ctx, 2, np.full(3, 50.0))
    s.upload(np.array([0.0, sigma]), np.zeros(2), np.zeros(2), np.ones(2))
    s.nlist_configure(rc, 0.5, 1, 10, True)
    s.pair_lj(eps, sigma, rc)
    s.nlist_build()
    for prec in (av.PREC_FP64, av.PREC_MIXED):
        s.compute(prec, True, True)
        pe, vir, _ = s.energy()
        sr6 = (sigma / rc) ** 6
        assert pe == pytest.approx(-4 * eps * (sr6 * sr6 - sr6), rel=1e-6 if prec else 1e-14)
        f = s.download()
        assert f["fx"][1] == pytest.approx(24 * eps / sigma, rel=1e-6 if prec else 1e-14)


def test_mixed_cutoff_decision_matches_fp64(hip_ctx):
    """Dimers whose separation lies within 1e-7 relative of the cutoff, on
    both sides: the FP32 r^2 of the mixed kernel cannot resolve them, so the
    kernel takes the FP64 test there and every dimer interacts (or not)
    exactly as in the plain path; the forces then agree to FP32 rounding."""
    rc, nside, spacing = 2.5, 6, 6.5
    k = np.arange(nside ** 3) - nside ** 3 // 2
    r = np.sqrt(rc * rc * (1.0 + 2e-9 * k))  # both sides of rc, none exactly at it
    g = np.stack(np.unravel_index(np.arange(nside ** 3), (nside,) * 3), axis=1) * spacing + 1.0
    x = np.concatenate([g[:, 0], g[:, 0] + r])
    y = np.concatenate([g[:, 1], g[:, 1] + 0.0])
    z = np.concatenate([g[:, 2], g[:, 2] + 0.0])
    box = np.full(3, nside * spacing)
    n = x.size
    s = av.System(hip_ctx, n, box)
    s.upload(x, y, z, np.ones(n))
    s.nlist_configure(rc, 0.3, 1, 0, True)
    s.pair_lj(1.0, 1.0, rc)
    s.nlist_build()
    assert s.brick_shape()[0] >= 0  # the brick kernels
    out = {}
    for prec in (av.PREC_FP64, av.PREC_MIXED):
        s.compute(prec, True, True)
        out[prec] = s.download()["fx"][nside ** 3:]
    f64, f32 = out[av.PREC_FP64], out[av.PREC_MIXED]
    inside = r * r < rc * rc
    assert inside.any() and (~inside).any()
    assert np.array_equal(f64 != 0.0, inside)
    assert np.array_equal(f32 != 0.0, inside), np.nonzero((f32 != 0.0) != inside)
    m = inside
    assert np.max(np.abs(f32[m] - f64[m]) / np.abs(f64[m])) < 1e-5
    s.close()


def test_binary_mixture_per_type(hip_ctx, orc):
    x, y, z, box = jittered_lattice(6, 1.2)
    n = x.size
    t = (np.arange(n) % 2).astype(np.int32)
    em, sm = orc.lj_mix(np.array([1.0, 0, 0, 1.0]), np.array([1.0, 0, 0, 0.8]))
    s = av.System(hip_ctx, n, box)
    s.upload(x, y, z, np.where(t == 0, 1.0, 2.0), type_=t)
    s.nlist_configure(2.5, 0.3, 1, 0, True)
    s.pair_lj(0.0, 0.0, 2.5, em, sm)
    s.nlist_build()
    s.compute(av.PREC_FP64, True, True)
    g = s.download()
    pe, vir, _ = s.energy()
    half = orc.build_half(x, y, z, box, 2.8)
    fx, fy, fz, e, ovir = orc.pair_lj(x, y, z, box, half, 0, 0, 2.5, type_=t, eps_tab=em, sigma_tab=sm)
    F = np.stack([fx, fy, fz])
    assert np.max(np.abs(np.stack([g["fx"], g["fy"], g["fz"]]) - F)) <= 1e-12 * np.max(np.abs(F))
    assert pe == pytest.approx(e, rel=1e-12)
    np.testing.assert_allclose(vir, ovir, rtol=1e-11)


def _oracle_state(w):
    n = w.n
    return dict(x=w.x.copy(), y=w.y.copy(), z=w.z.copy(), vx=w.vx.copy(), vy=w.vy.copy(), vz=w.vz.copy(),
                fx=np.zeros(n), fy=np.zeros(n), fz=np.zeros(n), mass=w.mass.copy())


def _oracle_params(orc, w, nh):
    return orc.MDParams(dt=w.dt, cutoff=w.cutoff, skin=w.skin, every=w.every, delay=w.delay, check=w.check,
                        eps=w.eps, sigma=w.sigma, nh=nh, T_target=w.T, tau_T=w.tau_T, ndof=0, nthreads=8,
                        stress=1)


@pytest.mark.parametrize("nh", [0, 1])
def test_trajectory_matches_oracle_short_horizon(hip_ctx, orc, nh):
    """40 steps of the full loop (VV [+NH], decide/rebuild, LJ) from identical
    inputs: positions within 1e-9 Å, velocities, energies and xi to rounding."""
    w = W.lj_fcc_4000()
    w.thermostat = nh
    s = make_system(hip_ctx, w)
    nsteps = 40
    s.run(nsteps, av.PREC_FP64, energy_every=1)
    g = s.download()
    pe_g, ke_g = s.thermo(nsteps)
    o = _oracle_state(w)
    pe_o, ke_o, _, st = orc.md_run(o, w.box, _oracle_params(orc, w, nh), nsteps)
    for k in ("x", "y", "z"):
        d = g[k] - o[k]
        d -= w.box[0] * np.rint(d / w.box[0])
        assert np.max(np.abs(d)) < 1e-9, k
    for k in ("vx", "vy", "vz"):
        assert np.max(np.abs(g[k] - o[k])) < 1e-9, k
    np.testing.assert_allclose(pe_g, pe_o, rtol=1e-10)
    np.testing.assert_allclose(ke_g, ke_o, rtol=1e-10)
    if nh:
        assert s.thermostat_xi() == pytest.approx(st.xi, rel=1e-8, abs=1e-14)
    assert s.nlist_stats()["nrebuilds"] == st.nrebuilds


def test_nve_energy_gates_c0(hip_ctx):
    """The reference LJ regression gates on the 4000-atom NVE config, 1000 steps,
    both precisions: max|E-E0|/(N eps) <= 1e-2; late drift <= 4x early drift."""
    for prec in (av.PREC_FP64, av.PREC_MIXED):
        w = W.lj_fcc_4000()
        s = make_system(hip_ctx, w)
        s.run(1000, prec, energy_every=1)
        pe, ke = s.thermo(1000)
        E = (pe + ke)[1:]
        assert np.all(np.isfinite(E))
        assert np.max(np.abs(E - E[0])) / w.n <= 1e-2
        h = E.size // 2
        assert abs(E[-1] - E[h]) <= 4 * abs(E[h] - E[0]) + 1e-6 * w.n


def test_decide_cadence_delay_ten_check_no(hip_ctx):
    """test_neighbor_decide.cpp:31-54 on the device: every=1, delay=10,
    check=no -> the 10th decide rebuilds."""
    x, y, z, box = jittered_lattice(4, 1.9)
    s = av.System(hip_ctx, x.size, box)
    s.upload(x, y, z, np.ones(x.size))
    s.nlist_configure(2.0, 0.4, 1, 10, False)
    s.nlist_build()
    assert s.nlist_stats()["nrebuilds"] == 1
    for k in range(1, 10):
        s.nlist_decide_and_rebuild()
        assert s.nlist_stats()["nrebuilds"] == 1, k
    s.nlist_decide_and_rebuild()
    assert s.nlist_stats()["nrebuilds"] == 2


def test_deterministic_bitwise(hip_ctx):
    outs = []
    for _ in range(2):
        w = W.lj_fcc_4000()
        w.thermostat = 1
        s = make_system(hip_ctx, w)
        s.run(120, av.PREC_MIXED, energy_every=10)
        outs.append(s.download())
    for k in outs[0]:
        np.testing.assert_array_equal(outs[0][k], outs[1][k])


def test_errors_are_reported(hip_ctx):
    s = av.System(hip_ctx, 10, np.full(3, 10.0))
    s.upload(np.linspace(0, 9, 10), np.zeros(10), np.zeros(10), np.ones(10))
    with pytest.raises(av.HipError) as e:
        s.compute()
    assert e.value.code == av._abi.E_STATE
    with pytest.raises(av.HipError):
        s.upload(np.zeros(10), np.zeros(10), np.zeros(10), np.zeros(10))  # zero mass
    with pytest.raises(av.HipError):
        s.nlist_configure(-1.0)
