"""The HIP path against the committed golden fixtures, through the C ABI:
pair sets bit-exact; FP64 pair forces to 1e-12 of max|F|; FP64 PME to 1e-9;
mixed precision to force RMS 1e-5; a 60-step NVT trajectory to 1e-9 Å."""
import numpy as np
import pytest

import atomiverse_amd as av
from golden_io import load

pytestmark = pytest.mark.gpu


def rms(F, G):
    return float(np.sqrt(np.sum((F - G) ** 2) / np.sum(G ** 2)))


def forces(d):
    return np.stack([d["fx"], d["fy"], d["fz"]], 1)


@pytest.mark.parametrize("name", ["jittered", "small_grid", "anisotropic"])
def test_pair_sets(hip_ctx, name):
    g = load(f"pairs_{name}")
    n = g["x"].size
    s = av.System(hip_ctx, n, g["box"])
    s.upload(g["x"], g["y"], g["z"], np.ones(n))
    s.nlist_configure(float(g["rcs"]), 0.0, 1, 0, True)
    s.nlist_build()
    p = s.export_pairs()
    np.testing.assert_array_equal(np.sort(p[:, 0] * n + p[:, 1]), g["keys"])


def test_lj_forces(hip_ctx):
    g = load("lj_forces")
    n = g["x"].size
    s = av.System(hip_ctx, n, g["box"])
    s.upload(g["x"], g["y"], g["z"], np.ones(n))
    s.nlist_configure(float(g["cutoff"]), float(g["skin"]), 1, 0, True)
    s.pair_lj(float(g["eps"]), float(g["sigma"]), float(g["cutoff"]))
    s.nlist_build()
    s.compute(av.PREC_FP64, True, True)
    F = forces(s.download())
    pe, vir, _ = s.energy()
    assert np.abs(F - g["f"]).max() <= 1e-12 * np.abs(g["f"]).max()
    assert pe == pytest.approx(float(g["energy"]), rel=1e-12)
    np.testing.assert_allclose(vir, g["virial"], rtol=1e-10, atol=1e-10)
    s.compute(av.PREC_MIXED, True, True)
    assert rms(forces(s.download()), g["f"]) <= 1e-5


def test_ljcoul_forces(hip_ctx):
    g = load("ljcoul_forces")
    n = g["x"].size
    s = av.System(hip_ctx, n, g["box"])
    s.upload(g["x"], g["y"], g["z"], np.ones(n), type_=g["type"], charge=g["q"])
    s.nlist_configure(float(g["cutoff"]), float(g["skin"]), 1, 0, True)
    s.pair_ljcoul(g["eps"], g["sigma"], g["set"], float(g["alpha"]), float(g["cutoff"]))
    s.nlist_build()
    s.compute(av.PREC_FP64, True, True)
    F = forces(s.download())
    pe, vir, _ = s.energy()
    assert np.abs(F - g["f"]).max() <= 1e-12 * np.abs(g["f"]).max()
    assert pe == pytest.approx(float(g["energy"]), rel=1e-12)
    s.compute(av.PREC_MIXED, True, True)
    assert rms(forces(s.download()), g["f"]) <= 1e-5


@pytest.mark.parametrize("name", ["pme_order4", "pme_order8"])
def test_pme(hip_ctx, name):
    g = load(name)
    n = g["x"].size
    s = av.System(hip_ctx, n, g["box"])
    s.upload(g["x"], g["y"], g["z"], np.ones(n), type_=g["type"], charge=g["q"])
    s.pme(float(g["alpha"]), float(g["spacing"]), int(g["order"]))
    K, _ = s.pme_info()
    np.testing.assert_array_equal(K, g["K"])
    s.pme_compute(av.PREC_FP64, True, True)
    F = forces(s.download())
    pe, vir, _ = s.energy()
    assert np.abs(F - g["f"]).max() <= 1e-9 * np.abs(g["f"]).max()
    assert pe == pytest.approx(float(g["energy"]), rel=1e-9)
    np.testing.assert_allclose(vir, g["virial"], rtol=1e-8, atol=1e-9 * np.abs(g["virial"][:3]).max())
    s.pme_compute(av.PREC_MIXED, True, True)
    assert rms(forces(s.download()), g["f"]) <= 1e-5


def test_md_trajectory(hip_ctx):
    g = load("md_nvt")
    n = g["x0"].size
    tau = np.sqrt(103.6427)
    kb = 8.617333262e-5
    s = av.System(hip_ctx, n, g["box"])
    s.upload(g["x0"], g["y0"], g["z0"], np.ones(n), g["vx0"], g["vy0"], g["vz0"])
    s.nlist_configure(2.5, 0.3, 1, 0, True)
    s.pair_lj(1.0, 1.0, 2.5)
    s.integrator(0.005 * tau, 1, 0.72 / kb, 0.5 * tau, 0)
    nsteps = g["pe"].size
    s.run(nsteps, av.PREC_FP64, energy_every=1)
    d = s.download()
    pe, ke = s.thermo(nsteps)
    L = g["box"][0]
    for k in ("x", "y", "z"):
        dd = d[k] - g[k]
        dd -= L * np.rint(dd / L)
        assert np.abs(dd).max() < 1e-9, k
    for k in ("vx", "vy", "vz"):
        assert np.abs(d[k] - g[k]).max() < 1e-9, k
    np.testing.assert_allclose(pe, g["pe"], rtol=1e-10)
    np.testing.assert_allclose(ke, g["ke"], rtol=1e-10)
    assert s.nlist_stats()["nrebuilds"] == int(g["nrebuilds"])
