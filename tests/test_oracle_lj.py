"""The oracle's PairLJCut, pinned by the reference's known answers, brute force,
Newton's third law, finite differences and an FCC lattice sum.

Reference: src/cpu/potentials/pair_lj_cut.cpp:31-42,257-298,300-830;
tests/cpu/dynamics/test_pair_lj_cut.cpp:22-62;
tests/cpu/potentials/test_pair_virial.cpp:233-260.
"""
import numpy as np
import pytest

from conftest import brute_lj, cluster, jittered_lattice


def _half_allpairs(n):
    nn = np.arange(n - 1, -1, -1, dtype=np.int32)
    first = np.concatenate([[0], np.cumsum(nn)[:-1]]).astype(np.int64)
    nb = np.concatenate([np.arange(i + 1, n) for i in range(n)]).astype(np.int32)
    return nn, first, nb


def test_two_atoms_at_sigma_give_minus_vshift(orc):
    """test_pair_lj_cut.cpp:22-62: eps=0.5, sigma=3, rc=8, box 50 -> E = -V_shift."""
    eps, sigma, rc = 0.5, 3.0, 8.0
    x = np.array([0.0, sigma])
    y = np.zeros(2)
    z = np.zeros(2)
    box = np.full(3, 50.0)
    half = orc.build_half(x, y, z, box, rc + 0.5)
    fx, fy, fz, e, vir = orc.pair_lj(x, y, z, box, half, eps, sigma, rc)
    sr6 = (sigma / rc) ** 6
    vshift = 4 * eps * (sr6 * sr6 - sr6)
    assert e == pytest.approx(-vshift, abs=1e-15)
    # F(sigma) = 24 eps/sigma (2 - 1) along the separation
    assert fx[1] == pytest.approx(24 * eps / sigma, rel=1e-14)
    assert fx[0] == -fx[1]
    # and through the full MD loop (one step, dt = 0.001 fs, v = 0), as the reference does
    s = dict(x=x.copy(), y=y.copy(), z=z.copy(), vx=np.zeros(2), vy=np.zeros(2), vz=np.zeros(2),
             fx=np.zeros(2), fy=np.zeros(2), fz=np.zeros(2), mass=np.ones(2))
    p = orc.MDParams(dt=0.001, cutoff=rc, skin=0.5, every=1, delay=10, check=1, eps=eps, sigma=sigma,
                     nh=0, T_target=0.0, tau_T=0.0, ndof=0, nthreads=1, stress=1)
    pe, ke, vir2, st = orc.md_run(s, box, p, 1)
    assert pe[0] == pytest.approx(-vshift, abs=1e-6)


def test_virial_matches_pair_sum_on_cluster(orc):
    """test_pair_virial.cpp:233-253 (eps=0.0104, sigma=3.4, rc=14): 1e-11 relative."""
    eps, sig, rc = 0.0104, 3.4, 14.0
    x, y, z, box = cluster()
    half = _half_allpairs(x.size)
    fx, fy, fz, e, vir = orc.pair_lj(x, y, z, box, half, eps, sig, rc)
    bfx, bfy, bfz, be, bvir = brute_lj(x, y, z, box, eps, sig, rc)
    for k in range(6):
        assert vir[k] == pytest.approx(bvir[k], abs=1e-11 * max(1.0, abs(bvir[k])))
    assert abs(bvir[:3].sum()) > 1.0
    assert e == pytest.approx(be, rel=1e-12)
    want = np.stack([bfx, bfy, bfz])
    np.testing.assert_allclose(np.stack([fx, fy, fz]), want, rtol=1e-12, atol=1e-14 * np.abs(want).max())


def test_virial_trace_is_minus_3v_dEdV(orc):
    """test_pair_virial.cpp:255-260: tr(W) = -3V dE/dV by central difference (h=1e-6)."""
    eps, sig, rc = 0.0104, 3.4, 14.0
    h = 1e-6
    half = _half_allpairs(24)

    def energy(lam):
        x, y, z, box = cluster(lam)
        return orc.pair_lj(x, y, z, box, half, eps, sig, rc)[3:]

    e_p, _ = energy(1 + h)
    e_m, _ = energy(1 - h)
    _, w0 = energy(1.0)
    V = lambda lam: (40.0 * lam) ** 3
    dedv = -3 * V(1.0) * (e_p - e_m) / (V(1 + h) - V(1 - h))
    assert w0[:3].sum() == pytest.approx(dedv, rel=1e-6)
    assert abs(dedv) > 1.0


def test_forces_are_minus_energy_gradient_and_sum_to_zero(orc):
    x, y, z, box = jittered_lattice(4, 1.9)
    rng = np.random.default_rng(1)
    x = x + rng.normal(0, 0.05, x.size)
    rc = 2.5
    half = orc.build_half(x, y, z, box, rc + 0.5)
    fx, fy, fz, e, _ = orc.pair_lj(x, y, z, box, half, 1.0, 1.0, rc)
    assert abs(fx.sum()) < 1e-10 and abs(fy.sum()) < 1e-10 and abs(fz.sum()) < 1e-10
    h = 1e-6
    for i in (0, 17, 42):
        for comp, f in ((x, fx), (y, fy), (z, fz)):
            old = comp[i]
            comp[i] = old + h
            ep = orc.pair_lj(x, y, z, box, half, 1.0, 1.0, rc)[3]
            comp[i] = old - h
            em = orc.pair_lj(x, y, z, box, half, 1.0, 1.0, rc)[3]
            comp[i] = old
            # the shifted potential jumps at rc only by the force discontinuity; pairs
            # here sit away from rc so the central difference is clean
            assert -(ep - em) / (2 * h) == pytest.approx(f[i], rel=1e-5, abs=1e-6)


def test_fcc_lattice_sum_closed_form(orc):
    """Perfect FCC at rho*=0.8442: zero force on every atom and E/N equal to the
    shell sum 0.5 * sum_k n_k V(r_k) over the shells inside rc = 2.5
    (r_k^2 = k a^2/2, n_k = 12, 6, 24, 12)."""
    a = (4 / 0.8442) ** (1 / 3)
    x, y, z = orc.fcc_lattice(6, 6, 6, a)
    box = np.full(3, 6 * a)
    rc = 2.5
    half = orc.build_half(x, y, z, box, rc + 0.3)
    fx, fy, fz, e, vir = orc.pair_lj(x, y, z, box, half, 1.0, 1.0, rc)
    assert np.max(np.abs(np.stack([fx, fy, fz]))) < 1e-10
    vs = 4 * (rc ** -12 - rc ** -6)
    shells = [(1, 12), (2, 6), (3, 24), (4, 12)]
    want = 0.5 * sum(nk * (4 * ((k * a * a / 2) ** -6 - (k * a * a / 2) ** -3) - vs) for k, nk in shells)
    assert np.sqrt(5 * a * a / 2) > rc
    assert e / x.size == pytest.approx(want, rel=1e-12)
    # isotropic: off-diagonal virial vanishes, diagonal equal
    assert abs(vir[3]) < 1e-8 and abs(vir[4]) < 1e-8 and abs(vir[5]) < 1e-8
    assert vir[0] == pytest.approx(vir[1], rel=1e-12) and vir[0] == pytest.approx(vir[2], rel=1e-12)


def test_per_type_table_and_lorentz_berthelot(orc):
    et = np.array([[1.0, 0.0], [0.0, 0.25]])
    st = np.array([[1.0, 0.0], [0.0, 0.8]])
    em, sm = orc.lj_mix(et, st)
    assert em.reshape(2, 2)[0, 1] == pytest.approx(0.5) and sm.reshape(2, 2)[1, 0] == pytest.approx(0.9)
    # identical types reproduce the single-type path bit for bit (same arithmetic)
    x, y, z, box = jittered_lattice(4, 1.9)
    half = orc.build_half(x, y, z, box, 2.8)
    t = (np.arange(x.size) % 2).astype(np.int32)
    one = orc.pair_lj(x, y, z, box, half, 1.0, 1.0, 2.5)
    two = orc.pair_lj(x, y, z, box, half, 0, 0, 2.5, type_=t, eps_tab=np.ones(4), sigma_tab=np.ones(4))
    for u, v in zip(one[:3], two[:3]):
        np.testing.assert_array_equal(u, v)
    assert one[3] == two[3]
    # a genuine binary mixture against brute force with per-pair coefficients
    x2 = x + 0.03 * np.sin(np.arange(x.size))
    half2 = orc.build_half(x2, y, z, box, 2.8)
    got = orc.pair_lj(x2, y, z, box, half2, 0, 0, 2.5, type_=t, eps_tab=em, sigma_tab=sm)
    # brute force: split by type pair
    n = x.size
    fxb = np.zeros(n)
    eb = 0.0
    emm = em.reshape(2, 2)
    smm = sm.reshape(2, 2)
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            d = np.array([x2[i] - x2[j], y[i] - y[j], z[i] - z[j]])
            d -= box * np.rint(d / box)
            r2 = d @ d
            if r2 >= 2.5 ** 2:
                continue
            e_, s_ = emm[t[i], t[j]], smm[t[i], t[j]]
            sr6 = (s_ * s_ / r2) ** 3
            fxb[i] += 24 * e_ / r2 * (2 * sr6 * sr6 - sr6) * d[0]
            sc6 = (s_ * s_ / 6.25) ** 3
            eb += 0.5 * (4 * e_ * (sr6 * sr6 - sr6) - 4 * e_ * (sc6 * sc6 - sc6))
    np.testing.assert_allclose(got[0], fxb, rtol=0, atol=1e-11)
    assert got[3] == pytest.approx(eb, rel=1e-12)
