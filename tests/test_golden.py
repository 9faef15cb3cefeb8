"""The oracle reproduces the committed golden fixtures (made by
tests/golden/make_golden.py and cross-checked there against numpy brute
force and a mode-by-mode Ewald sum). Integer results bit-exact; FP64 to
1e-12 relative (summation order of threaded reductions)."""
import numpy as np
import pytest

from golden_io import load

TOL = 1e-12


@pytest.mark.parametrize("name", ["jittered", "small_grid", "anisotropic"])
def test_pair_sets(orc, name):
    g = load(f"pairs_{name}")
    nn, first, nb = orc.build_half(g["x"], g["y"], g["z"], g["box"], float(g["rcs"]))
    np.testing.assert_array_equal(orc.pair_keys(orc.half_pairs(nn, first, nb), g["x"].size), g["keys"])


def test_lj_forces(orc):
    g = load("lj_forces")
    half = orc.build_half(g["x"], g["y"], g["z"], g["box"], float(g["cutoff"] + g["skin"]))
    fx, fy, fz, e, vir = orc.pair_lj(g["x"], g["y"], g["z"], g["box"], half, float(g["eps"]), float(g["sigma"]),
                                     float(g["cutoff"]))
    F = np.stack([fx, fy, fz], 1)
    assert np.abs(F - g["f"]).max() <= TOL * np.abs(g["f"]).max()
    assert e == pytest.approx(float(g["energy"]), rel=TOL)
    np.testing.assert_allclose(vir, g["virial"], rtol=1e-11)


def test_ljcoul_forces(orc):
    g = load("ljcoul_forces")
    tab = orc.LJCoulTables(g["eps"], g["sigma"], g["set"], float(g["cutoff"]), float(g["alpha"]))
    half = orc.build_half(g["x"], g["y"], g["z"], g["box"], float(g["cutoff"] + g["skin"]))
    fx, fy, fz, e, vir = orc.pair_lj_coul_long(g["x"], g["y"], g["z"], g["q"], g["type"], g["box"], half, tab)
    F = np.stack([fx, fy, fz], 1)
    assert np.abs(F - g["f"]).max() <= TOL * np.abs(g["f"]).max()
    assert e == pytest.approx(float(g["energy"]), rel=TOL)


@pytest.mark.parametrize("name", ["pme_order4", "pme_order8"])
def test_pme(orc, name):
    g = load(name)
    fx, fy, fz, e, vir, parts = orc.pme(g["x"], g["y"], g["z"], g["q"], g["box"], float(g["alpha"]), g["K"],
                                        int(g["order"]))
    F = np.stack([fx, fy, fz], 1)
    assert np.abs(F - g["f"]).max() <= TOL * np.abs(g["f"]).max()
    assert e == pytest.approx(float(g["energy"]), rel=TOL)
    np.testing.assert_allclose(vir, g["virial"], rtol=1e-10, atol=1e-12)


def test_md_trajectory(orc):
    g = load("md_nvt")
    n = g["x0"].size
    st = dict(x=g["x0"].copy(), y=g["y0"].copy(), z=g["z0"].copy(), vx=g["vx0"].copy(), vy=g["vy0"].copy(),
              vz=g["vz0"].copy(), fx=np.zeros(n), fy=np.zeros(n), fz=np.zeros(n), mass=np.ones(n))
    tau = np.sqrt(103.6427)
    kb = 8.617333262e-5
    p = orc.MDParams(dt=0.005 * tau, cutoff=2.5, skin=0.3, every=1, delay=0, check=1, eps=1.0, sigma=1.0, nh=1,
                     T_target=0.72 / kb, tau_T=0.5 * tau, ndof=0, nthreads=1, stress=1)
    pe, ke, _, state = orc.md_run(st, g["box"], p, g["pe"].size)
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(st[k], g[k])
    np.testing.assert_array_equal(pe, g["pe"])
    assert state.nrebuilds == int(g["nrebuilds"])
