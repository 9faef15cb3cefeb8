"""pytest configuration: the `gpu` marker and shared fixtures.

GPU tests call libatv_hip.so through its C ABI and compare against the CPU
oracle (oracle/libatv_oracle.so). They never skip for a missing library: a
GPU run without the HIP backend must fail loudly.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP backend)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def hip_ctx():
    import atomiverse_amd as av
    ctx = av.Context(0)
    yield ctx
    ctx.close()


def jittered_lattice(n_side, spacing):
    """Deterministic jittered simple-cubic lattice (restates the fixture of
    tests/cpu/dynamics/test_full_neighbor_list.cpp:25-45), x-major order."""
    idx = np.arange(n_side ** 3)
    ix, iy, iz = np.unravel_index(idx, (n_side, n_side, n_side))
    j = 0.07 * np.sin(7.0 * idx + 1.0)
    x = (ix + 0.5) * spacing + j
    y = (iy + 0.5) * spacing - j
    z = (iz + 0.5) * spacing + 0.5 * j
    return x, y, z, np.full(3, n_side * spacing)


def cluster(lam=1.0, n=24):
    """24-atom pseudo-random cluster in a 40 Å box (restates
    tests/cpu/potentials/test_pair_virial.cpp:55-80)."""
    a = 1.7 * np.arange(n) + 0.3
    x = lam * (3.5 * np.sin(a) + 1.1 * np.cos(2.9 * a))
    y = lam * (3.5 * np.cos(1.3 * a) + 0.9 * np.sin(3.7 * a))
    z = lam * (3.5 * np.sin(2.1 * a + 0.5) + 1.3 * np.cos(a))
    return x, y, z, np.full(3, 40.0 * lam)


def brute_pairs(x, y, z, box, rc):
    """All unordered pairs with minimum-image r^2 < rc^2 (numpy, O(N^2))."""
    n = x.size
    keys = []
    for i0 in range(0, n, 512):
        sl = slice(i0, min(n, i0 + 512))
        d = []
        for c, L in zip((x, y, z), box):
            dd = c[sl, None] - c[None, :]
            dd -= L * np.rint(dd / L)
            d.append(dd)
        r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
        ii, jj = np.nonzero(r2 < rc * rc)
        ii = ii + i0
        m = ii < jj
        keys.append(ii[m].astype(np.int64) * n + jj[m])
    return np.sort(np.concatenate(keys)) if keys else np.zeros(0, np.int64)


def brute_lj(x, y, z, box, eps, sigma, rc):
    """O(N^2) shifted-LJ forces, energy, Voigt virial (numpy)."""
    n = x.size
    d = []
    for c, L in zip((x, y, z), box):
        dd = c[:, None] - c[None, :]
        dd -= L * np.rint(dd / L)
        d.append(dd)
    r2 = d[0] ** 2 + d[1] ** 2 + d[2] ** 2
    np.fill_diagonal(r2, np.inf)
    m = r2 < rc * rc
    r2s = np.where(m, r2, 1.0)
    sr6 = (sigma ** 2 / r2s) ** 3
    fp = np.where(m, 24 * eps / r2s * (2 * sr6 * sr6 - sr6), 0.0)
    sr6c = (sigma ** 2 / rc ** 2) ** 3
    vsh = 4 * eps * (sr6c ** 2 - sr6c)
    e = 0.5 * np.sum(np.where(m, 4 * eps * (sr6 * sr6 - sr6) - vsh, 0.0))
    f = [np.sum(fp * dd, axis=1) for dd in d]
    vir = [0.5 * np.sum(fp * a * b) for a, b in ((d[0], d[0]), (d[1], d[1]), (d[2], d[2]),
                                                  (d[0], d[1]), (d[0], d[2]), (d[1], d[2]))]
    return f[0], f[1], f[2], e, np.array(vir)
