"""Pins of the oracle's long-range electrostatics (no GPU).

Known answers restated from the reference's own tests
(tests/cpu/longrange/test_pme_math.cpp, test_pme_coul.cpp,
tests/cpu/potentials/test_pair_lj_cut_coul_long.cpp) plus what the
mathematics fixes independently of the oracle's code: the closed-form
cardinal B-spline, numpy's FFT, a direct Ewald sum written here from the
textbook formula, the Wigner constant of a point charge in a neutralising
background, finite differences of the energy for forces and strain.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

COUL = 14.399645
MADELUNG_NACL = 1.747564594633182
WIGNER_SC = 2.837297479480620  # E = -WIGNER_SC q^2 / (2L) per charge, sc lattice + background


# ---------------------------------------------------------------- PME math

def test_good_fft_size_pins():
    """test_pme_math.cpp:19-34"""
    assert [O.good_fft_size(n) for n in (1, 2, 3, 4)] == [4, 4, 4, 4]
    assert [O.good_fft_size(n) for n in (11, 13, 15, 16, 23, 49)] == [12, 14, 15, 16, 24, 49]
    # every result factors into {2,3,5,7} and nothing smaller does
    for n in range(4, 300):
        m = O.good_fft_size(n)
        r = m
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        assert r == 1 and m >= n
        for k in range(n, m):
            r = k
            for p in (2, 3, 5, 7):
                while r % p == 0:
                    r //= p
            assert r != 1


def _cardinal_bspline(u, n):
    """M_n(u) = 1/(n-1)! sum_k (-1)^k C(n,k) (u-k)_+^(n-1), support [0, n]."""
    from fractions import Fraction
    u = Fraction(u)  # exact: the alternating sum cancels heavily at order 8
    s = Fraction(0)
    for k in range(n + 1):
        if u - k > 0:
            s += (-1) ** k * math.comb(n, k) * (u - k) ** (n - 1)
    return float(s / math.factorial(n - 1))


@pytest.mark.parametrize("n", [3, 4, 5, 6, 8])
def test_bspline_matches_closed_form(n):
    """theta[k] = M_n(frac + n-1-k), dtheta[k] = M_n'(.) = M_{n-1}(.) - M_{n-1}(. - 1)."""
    for frac in (0.0, 0.1, 0.33, 0.5, 0.75, 0.999):
        th, dth = O.bspline(frac, n)
        assert abs(th.sum() - 1.0) <= 1e-12  # test_pme_math.cpp:36-49
        assert abs(dth.sum()) <= 1e-12
        for k in range(n):
            u = frac + n - 1 - k
            assert th[k] == pytest.approx(_cardinal_bspline(u, n), abs=1e-13)
            d = _cardinal_bspline(u, n - 1) - _cardinal_bspline(u - 1, n - 1)
            assert dth[k] == pytest.approx(d, abs=1e-13)


def test_bspline_order4_essmann_appendix():
    """test_pme_math.cpp:51-61"""
    th, _ = O.bspline(0.0, 4)
    np.testing.assert_allclose(th, [1 / 6, 4 / 6, 1 / 6, 0.0], atol=1e-14)


def test_spline_moduli_pins_and_dft():
    """test_pme_math.cpp:63-86, plus the moduli as numpy's DFT of M_n(k+1)."""
    m = O.spline_moduli(8, 4)
    assert np.all(m > 0)
    assert m[0] == pytest.approx(1.0, abs=1e-12)
    assert m[4] == pytest.approx(9.0, abs=1e-12)
    m16 = O.spline_moduli(16, 6)
    th, _ = O.bspline(0.0, 6)
    assert m16[0] == pytest.approx(1.0 / th[:5].sum() ** 2, abs=1e-12)
    for K, n in ((16, 6), (24, 4), (35, 5), (128, 4)):
        vals = np.array([_cardinal_bspline(k + 1, n) for k in range(n - 1)] + [0.0] * (K - n + 1))
        b = np.fft.ifft(vals) * K  # sum_k v_k exp(+2 pi i m k / K)
        want = 1.0 / np.abs(b) ** 2
        got = O.spline_moduli(K, n)
        finite = np.abs(b) ** 2 > 1e-12
        np.testing.assert_allclose(got[finite], want[finite], rtol=1e-11)


@pytest.mark.parametrize("shape", [(6, 10, 7), (12, 11, 13), (8, 8, 8), (4, 210, 3)])
def test_fft3d_matches_numpy(shape):
    rng = np.random.default_rng(7)
    a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    scale = np.abs(a).sum()
    np.testing.assert_allclose(O.fft3d(a, -1), np.fft.fftn(a), atol=1e-13 * scale)
    np.testing.assert_allclose(O.fft3d(a, +1), np.fft.ifftn(a) * a.size, atol=1e-13 * scale)


def test_ewald_alpha_and_mesh_of_the_charged_melt():
    """ewald_sum.h:58-60 at tol 1e-5, r_c 3.0; pme.cpp:66-74 at L 82.85953, spacing 0.65."""
    assert O.ewald_alpha(3.0, 1e-5) == pytest.approx(1.025646, abs=1e-6)
    np.testing.assert_array_equal(O.pme_grid([82.85953] * 3, 0.65, 4), [128, 128, 128])
    np.testing.assert_array_equal(O.pme_grid([2.0, 8.0, 11.28], 0.5, 6), [7, 16, 24])


# ---------------------------------------------------------------- Ewald helpers

def rocksalt(d, n):
    """make_rocksalt (test_pme_coul.cpp:32-52): site (i,j,k) at d*(i,j,k), charge +-1."""
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x, y, z = (a.reshape(-1) * d for a in (i, j, k))
    q = np.where((i + j + k).reshape(-1) % 2 == 0, 1.0, -1.0)
    return x, y, z, q, np.full(3, n * d)


def ewald_direct_recip(x, y, z, q, box, alpha, tol=1e-17):
    """Reciprocal-space Ewald sum evaluated mode by mode (no mesh):
    E = k_e/(2 pi V) sum_{m!=0} exp(-pi^2 m^2/alpha^2)/m^2 |S(m)|^2, S = sum q e^{2 pi i m.r},
    forces -dE/dr and the virial W_ab = sum_m E_m (delta_ab - 2 (1/m^2 + pi^2/alpha^2) m_a m_b)."""
    V = np.prod(box)
    hmax = [int(math.ceil(alpha * L * math.sqrt(-math.log(tol)) / math.pi)) for L in box]
    h = [np.arange(-hm, hm + 1) for hm in hmax]
    H0, H1, H2 = np.meshgrid(*h, indexing="ij")
    m = np.stack([H0.reshape(-1) / box[0], H1.reshape(-1) / box[1], H2.reshape(-1) / box[2]], 1)
    m = m[np.any(m != 0, axis=1)]
    m2 = (m ** 2).sum(1)
    g = np.exp(-math.pi ** 2 * m2 / alpha ** 2) / m2
    keep = g > 1e-300
    m, m2, g = m[keep], m2[keep], g[keep]
    r = np.stack([x, y, z], 1)
    phase = np.exp(2j * math.pi * (m @ r.T))  # (M, N)
    S = phase @ q
    escale = COUL / (2 * math.pi * V)
    E = escale * np.sum(g * np.abs(S) ** 2)
    im = np.imag(np.conj(S)[:, None] * phase)  # (M, N)
    F = escale * 4 * math.pi * q[:, None] * ((g[:, None] * im).T @ m)
    Em = escale * g * np.abs(S) ** 2
    vt = -2.0 * (1.0 / m2 + math.pi ** 2 / alpha ** 2)
    W = np.array([np.sum(Em * (1 + vt * m[:, 0] ** 2)), np.sum(Em * (1 + vt * m[:, 1] ** 2)),
                  np.sum(Em * (1 + vt * m[:, 2] ** 2)), np.sum(Em * vt * m[:, 0] * m[:, 1]),
                  np.sum(Em * vt * m[:, 0] * m[:, 2]), np.sum(Em * vt * m[:, 1] * m[:, 2])])
    return E, F, W


def real_space_direct(x, y, z, q, box, alpha, cutoff):
    """sum_{i<j} k_e q_i q_j erfc(alpha r)/r over minimum images inside the cutoff."""
    from scipy.special import erfc
    r = np.stack([x, y, z], 1)
    E = 0.0
    F = np.zeros_like(r)
    for i in range(len(q)):
        d = r[i] - r[i + 1:]
        d -= box * np.rint(d / box)
        rr = np.sqrt((d ** 2).sum(1))
        ok = rr < cutoff
        rr, dd, qj = rr[ok], d[ok], q[i + 1:][ok]
        E += np.sum(COUL * q[i] * qj * erfc(alpha * rr) / rr)
        fp = COUL * q[i] * qj * (erfc(alpha * rr) + 2 / math.sqrt(math.pi) * alpha * rr *
                                 np.exp(-(alpha * rr) ** 2)) / rr ** 3
        fv = fp[:, None] * dd
        F[i] += fv.sum(0)
        np.subtract.at(F, np.nonzero(ok)[0] + i + 1, fv)
    return E, F


def electrostatic_energy(x, y, z, q, box, alpha, cutoff, spacing, order):
    """CoulLong + KSpacePME on the all-pairs half list (test_pme_coul.cpp:89-110)."""
    half = O.all_pairs_half(x.size)
    fx, fy, fz, e, vir = O.coul_long(x, y, z, q, box, half, alpha, cutoff)
    K = O.pme_grid(box, spacing, order)
    gx, gy, gz, ek, virk, parts = O.pme(x, y, z, q, box, alpha, K, order)
    return e + ek, np.stack([fx + gx, fy + gy, fz + gz], 1), vir + virk, parts


# ---------------------------------------------------------------- PME physics

def test_madelung_nacl():
    """test_pme_coul.cpp:157-175: within 5e-4 of the NaCl Madelung energy."""
    d, n = 2.82, 4
    x, y, z, q, box = rocksalt(d, n)
    cutoff = 1.9 * d
    alpha = 3.5 / cutoff
    e, F, _, _ = electrostatic_energy(x, y, z, q, box, alpha, cutoff, 0.5, 6)
    want = -0.5 * x.size * MADELUNG_NACL * COUL / d
    assert abs(e / want - 1.0) <= 5e-4
    assert np.abs(F).max() < 1e-6  # perfect lattice: zero force by symmetry


def test_real_space_matches_direct_sum():
    rng = np.random.default_rng(3)
    box = np.array([9.0, 10.0, 11.0])
    x, y, z = (rng.uniform(0, L, 40) for L in box)
    q = rng.normal(size=40)
    half = O.all_pairs_half(40)
    fx, fy, fz, e, _ = O.coul_long(x, y, z, q, box, half, 0.4, 4.4)
    E, F = real_space_direct(x, y, z, q, box, 0.4, 4.4)
    assert e == pytest.approx(E, rel=1e-12)
    np.testing.assert_allclose(np.stack([fx, fy, fz], 1), F, atol=1e-12 * np.abs(F).max())


def test_pme_agrees_with_direct_ewald():
    """test_pme_coul.cpp:177-192 (5e-4 on the total), plus reciprocal forces
    and virial against the mode-by-mode Ewald sum on a random neutral cell."""
    d, n = 2.82, 4
    x, y, z, q, box = rocksalt(d, n)
    cutoff = 1.9 * d
    alpha = 3.5 / cutoff
    E_pme, _, _, parts = electrostatic_energy(x, y, z, q, box, alpha, cutoff, 0.5, 6)
    E_real, _ = real_space_direct(x, y, z, q, box, alpha, cutoff)
    E_rec, _, _ = ewald_direct_recip(x, y, z, q, box, alpha)
    e_self = -COUL * alpha / math.sqrt(math.pi) * np.sum(q * q)
    assert parts[1] == pytest.approx(e_self, rel=1e-14)
    assert parts[2] == 0.0
    E_ew = E_real + E_rec + e_self
    assert abs(E_pme / E_ew - 1.0) <= 5e-4

    rng = np.random.default_rng(11)
    box = np.array([10.0, 11.0, 12.0])
    N = 60
    x, y, z = (rng.uniform(0, L, N) for L in box)
    q = rng.normal(size=N)
    q -= q.mean()
    alpha = 0.5
    K = O.pme_grid(box, 0.3, 6)
    fx, fy, fz, e, vir, parts = O.pme(x, y, z, q, box, alpha, K, 6)
    E, F, W = ewald_direct_recip(x, y, z, q, box, alpha)
    assert parts[0] == pytest.approx(E, rel=1e-5)
    G = np.stack([fx, fy, fz], 1)
    assert np.sqrt(np.sum((G - F) ** 2) / np.sum(F ** 2)) <= 1e-4
    np.testing.assert_allclose(vir, W, atol=1e-4 * np.abs(W[:3]).max())


def test_single_charge_in_background_is_the_wigner_energy():
    """One charge, no pairs: e_self + e_recip + e_net = -xi q^2 k_e / (2L)
    (simple-cubic lattice of charges in a neutralising background)."""
    L = 10.0
    box = np.full(3, L)
    for pos in ((1.3, 7.2, 4.4), (0.0, 0.0, 0.0)):
        x, y, z = (np.array([c]) for c in pos)
        q = np.array([0.7])
        K = O.pme_grid(box, 0.2, 8)
        _, _, _, e, _, parts = O.pme(x, y, z, q, box, 0.6, K, 8)
        want = -WIGNER_SC * 0.49 * COUL / (2 * L)
        assert e == pytest.approx(want, rel=1e-8)
        # the mode-by-mode Ewald sum used as a checker above gives it to rounding
        E, _, _ = ewald_direct_recip(x, y, z, q, box, 0.6)
        assert E + parts[1] + parts[2] == pytest.approx(want, rel=1e-13)
        assert parts[2] == pytest.approx(-COUL * math.pi / (2 * L ** 3 * 0.36) * 0.49, rel=1e-14)


def test_finite_difference_forces():
    """test_pme_coul.cpp:194-241: |F - (-dE/dx)| <= 2e-3 with h = 1e-5."""
    x = np.array([1.0, 3.2, 1.3, 2.7])
    y = np.array([1.0, 1.4, 3.1, 2.2])
    z = np.array([1.0, 1.1, 2.4, 3.6])
    q = np.array([0.6, -0.4, 0.5, -0.7])
    box = np.full(3, 8.0)
    cutoff, alpha = 3.9, 0.45
    _, F, _, _ = electrostatic_energy(x, y, z, q, box, alpha, cutoff, 0.7, 6)
    h = 1e-5
    for a in range(4):
        for ax, arr in enumerate((x, y, z)):
            o = arr[a]
            arr[a] = o + h
            ep = electrostatic_energy(x, y, z, q, box, alpha, cutoff, 0.7, 6)[0]
            arr[a] = o - h
            em = electrostatic_energy(x, y, z, q, box, alpha, cutoff, 0.7, 6)[0]
            arr[a] = o
            assert abs(F[a, ax] + (ep - em) / (2 * h)) <= 2e-3


def test_stress_agrees_with_ewald_on_rock_salt():
    """test_pme_coul.cpp:424-450: 5e-3 * max(1, |W|) per Voigt component."""
    d, n = 2.82, 4
    x, y, z, q, box = rocksalt(d, n)
    cutoff = 1.9 * d
    alpha = 3.5 / cutoff
    _, _, W_pme, _ = electrostatic_energy(x, y, z, q, box, alpha, cutoff, 0.5, 6)
    half = O.all_pairs_half(x.size)
    _, _, _, _, W_real = O.coul_long(x, y, z, q, box, half, alpha, cutoff)
    _, _, W_rec = ewald_direct_recip(x, y, z, q, box, alpha)
    W_ew = W_real + W_rec
    for s in range(6):
        assert abs(W_pme[s] - W_ew[s]) <= 5e-3 * max(1.0, abs(W_ew[s]))


def test_reciprocal_virial_is_minus_strain_derivative():
    """test_pme_coul.cpp (FiniteStrainDiagonalMatchesVirial): V_aa = -dE/d eps_a
    at fixed fractional coordinates, reciprocal part only (neutral cell)."""
    rng = np.random.default_rng(5)
    box = np.array([8.0, 9.0, 10.0])
    N = 30
    r = np.stack([rng.uniform(0, L, N) for L in box], 1)
    q = rng.normal(size=N)
    q -= q.mean()
    alpha, order = 0.55, 6
    K = O.pme_grid(box, 0.3, order)

    def erec(b):
        rr = r * (b / box)
        return O.pme(rr[:, 0], rr[:, 1], rr[:, 2], q, b, alpha, K, order)[5][0]

    _, _, _, _, vir, _ = O.pme(r[:, 0], r[:, 1], r[:, 2], q, box, alpha, K, order)
    h = 1e-5
    for a in range(3):
        bp, bm = box.copy(), box.copy()
        bp[a] *= 1 + h
        bm[a] *= 1 - h
        dE = (erec(bp) - erec(bm)) / (2 * h)
        assert vir[a] == pytest.approx(-dE, rel=1e-5, abs=1e-8)


# ---------------------------------------------------------------- erfc table / LJ+Coulomb pair

def test_erfc_table_matches_direct_erfc():
    """test_pair_lj_cut_coul_long.cpp:20-38 (5e-3 relative), plus the exact
    table nodes: entry i is erfc(alpha r)/r at r^2 = i dr2."""
    from scipy.special import erfc
    a, cut, nb = 0.3, 10.0, 5000
    e_tab, f_tab = O.erfc_table(a, cut, nb)
    dr2 = cut * cut / nb
    for r in (0.5, 1.0, 2.5, 5.0, 9.5):
        r2 = r * r
        i = int(r2 / dr2)
        t = r2 / dr2 - i
        e = e_tab[i] + t * (e_tab[i + 1] - e_tab[i])
        f = f_tab[i] + t * (f_tab[i + 1] - f_tab[i])
        g = a * r
        e_ref = erfc(g) / r
        f_ref = (erfc(g) + 2 / math.sqrt(math.pi) * g * math.exp(-g * g)) / r ** 3
        assert abs(e - e_ref) <= 5e-3 * abs(e_ref)
        assert abs(f - f_ref) <= 5e-3 * abs(f_ref)
    for i in (1, 17, 2500, 5000):
        r = math.sqrt(i * dr2)
        assert e_tab[i] == pytest.approx(erfc(a * r) / r, rel=1e-14)
    assert e_tab[0] == e_tab[1] and f_tab[0] == f_tab[1]


def _melt_tables(cutoff=3.0, alpha=1.025646):
    eps = np.array([[1.0, 0.0], [0.0, 1.0]])
    sig = np.array([[1.0, 0.0], [0.0, 0.8]])
    st = np.array([[1, 0], [0, 1]])
    return O.LJCoulTables(eps, sig, st, cutoff, alpha)


def test_lorentz_berthelot_and_lj_tables():
    """finalize_type_table_ (pair_lj_cut_coul_long.cpp:132-186) for sigma_A 1.0, sigma_B 0.8."""
    t = _melt_tables()
    assert t.sigma[1] == pytest.approx(0.9) and t.eps[1] == pytest.approx(1.0)
    s6 = 0.9 ** 6
    assert t.lj1[1] == pytest.approx(48 * s6 * s6)
    assert t.lj4[1] == pytest.approx(4 * s6)
    rc6 = 3.0 ** 6
    assert t.shift[1] == pytest.approx(4 * (s6 * s6 / rc6 ** 2 - s6 / rc6), rel=1e-14)
    # unset diagonal -> eps 0, sigma 1, no LJ
    u = O.LJCoulTables(np.array([[1.0, 0.0], [0.0, 0.0]]), np.array([[1.0, 0.0], [0.0, 0.0]]),
                       np.array([[1, 0], [0, 0]]), 3.0, 1.0)
    assert u.lj1[1] == 0.0 and u.lj1[3] == 0.0 and u.sigma[3] == 1.0


def _melt_snapshot(nside=6, seed=9):
    rng = np.random.default_rng(seed)
    a = 1.03574
    i, j, k = np.meshgrid(np.arange(nside), np.arange(nside), np.arange(nside), indexing="ij")
    L = nside * a
    x, y, z = (np.mod(c.reshape(-1) * a + rng.normal(0, 0.05, nside ** 3), L) for c in (i, j, k))
    t = ((i + j + k).reshape(-1) % 2).astype(np.int32)
    q = np.where(t == 0, 1.0, -1.0) / math.sqrt(COUL)
    return x, y, z, q, t, np.full(3, nside * a)


def test_lj_coul_long_reduces_to_lj_and_to_coul_long():
    """Charges off -> PairLJCut with the mixed table; LJ off -> CoulLong to
    the table's interpolation error (1e-5 relative)."""
    x, y, z, q, t, box = _melt_snapshot()
    tab = _melt_tables()
    half = O.build_half(x, y, z, box, 3.3)
    fx, fy, fz, e, vir = O.pair_lj_coul_long(x, y, z, np.zeros_like(q), t, box, half, tab)
    gx, gy, gz, e2, vir2 = O.pair_lj(x, y, z, box, half, 0, 0, 3.0, type_=t, eps_tab=tab.eps,
                                     sigma_tab=tab.sigma)
    F, G = np.stack([fx, fy, fz]), np.stack([gx, gy, gz])
    assert np.abs(F - G).max() <= 1e-12 * np.abs(G).max()
    assert e == pytest.approx(e2, rel=1e-12)
    np.testing.assert_allclose(vir, vir2, rtol=1e-11)

    zero = O.LJCoulTables(np.zeros((2, 2)), np.ones((2, 2)), np.ones((2, 2)), 3.0, 1.025646)
    fx, fy, fz, e, vir = O.pair_lj_coul_long(x, y, z, q, t, box, half, zero)
    gx, gy, gz, e2, vir2 = O.coul_long(x, y, z, q, box, half, 1.025646, 3.0)
    F, G = np.stack([fx, fy, fz]), np.stack([gx, gy, gz])
    assert np.sqrt(np.sum((F - G) ** 2) / np.sum(G ** 2)) <= 1e-5
    assert e == pytest.approx(e2, rel=1e-5)


def test_charged_melt_md_conserves_energy():
    """orc_md_run_ff (LJ+Coulomb+PME, NVE): total energy stays within the LJ
    regression gate, and the pair/PME composition is what the force phase
    sums (pair first, then k-space)."""
    x, y, z, q, t, box = _melt_snapshot(6, seed=4)
    n = x.size
    tab = _melt_tables()
    K = O.pme_grid(box, 0.4, 4)
    mass = np.ones(n)
    kB = 8.617333262e-5
    vx, vy, vz = O.init_velocity(mass, 1.2 / kB, 87289)
    s = dict(x=x, y=y, z=z, vx=vx, vy=vy, vz=vz, fx=np.zeros(n), fy=np.zeros(n), fz=np.zeros(n), mass=mass)
    tau = math.sqrt(103.6427)
    p = O.MDParams(dt=0.005 * tau, cutoff=3.0, skin=0.3, every=1, delay=0, check=1, eps=0, sigma=0, nh=0,
                   T_target=0, tau_T=0, ndof=0, nthreads=4, stress=1)
    pe, ke, vir, st = O.md_run_ff(s, box, p, tab, t, q, K, 4, 200)
    E = pe + ke
    assert np.all(np.isfinite(E))
    assert np.max(np.abs(E - E[0])) / n <= 1e-2
    # forces left in the frame are pair + PME at the final positions
    half = O.build_half(s["x"], s["y"], s["z"], box, 3.3)
    fx, fy, fz, e1, _ = O.pair_lj_coul_long(s["x"], s["y"], s["z"], q, t, box, half, tab)
    gx, gy, gz, e2, _, _ = O.pme(s["x"], s["y"], s["z"], q, box, tab.alpha, K, 4)
    np.testing.assert_allclose(s["fx"], fx + gx, atol=1e-10)
    assert pe[-1] == pytest.approx(e1 + e2, rel=1e-10)
