"""Generate the golden fixtures of this directory (run from the repo root:
`python tests/golden/make_golden.py`).

Each fixture is a small .npz (inputs and expected outputs, no pickles) made
by the CPU oracle and cross-checked here, before it is written, against an
independent computation: numpy brute force for pair sets and pair forces, a
mode-by-mode Ewald sum for PME. The fixtures pin the oracle (tests/
test_golden.py re-runs it on the CPU) and the HIP path
(tests/test_gpu_golden.py) to the same numbers.
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import oracle as O  # noqa: E402
from conftest import brute_lj, brute_pairs, jittered_lattice  # noqa: E402
from test_oracle_pme import ewald_direct_recip, real_space_direct  # noqa: E402

COUL = 14.399645
KB = 8.617333262e-5
TAU = math.sqrt(103.6427)


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print(f"{name}: {os.path.getsize(path)} bytes")


def pair_set_cases():
    rng = np.random.default_rng(20261)
    cases = {}
    x, y, z, box = jittered_lattice(6, 1.55)
    cases["jittered"] = (x, y, z, box, 2.8)
    L = 6.0
    cases["small_grid"] = (rng.uniform(0, L, 150), rng.uniform(0, L, 150), rng.uniform(0, L, 150), np.full(3, L),
                           2.8)
    box = np.array([9.5, 12.0, 16.5])
    cases["anisotropic"] = (rng.uniform(0, box[0], 900), rng.uniform(0, box[1], 900), rng.uniform(0, box[2], 900),
                            box, 2.8)
    for name, (x, y, z, box, rcs) in cases.items():
        nn, first, nb = O.build_half(x, y, z, box, rcs)
        keys = O.pair_keys(O.half_pairs(nn, first, nb), x.size)
        brute = brute_pairs(x, y, z, box, rcs)
        assert np.array_equal(keys, brute), name
        save(f"pairs_{name}", x=x, y=y, z=z, box=box, rcs=np.array(rcs), keys=keys)


def lj_case():
    x, y, z, box = jittered_lattice(7, 1.3)
    rcs = 2.8
    half = O.build_half(x, y, z, box, rcs)
    fx, fy, fz, e, vir = O.pair_lj(x, y, z, box, half, 1.0, 1.0, 2.5)
    bx, by, bz, be, bv = brute_lj(x, y, z, box, 1.0, 1.0, 2.5)
    bf = np.stack([bx, by, bz], 1)
    F = np.stack([fx, fy, fz], 1)
    assert np.abs(F - bf).max() <= 1e-11 * np.abs(bf).max()
    assert abs(e - be) <= 1e-11 * abs(be)
    assert np.abs(vir - bv).max() <= 1e-10 * np.abs(bv).max()
    save("lj_forces", x=x, y=y, z=z, box=box, eps=np.array(1.0), sigma=np.array(1.0), cutoff=np.array(2.5),
         skin=np.array(0.3), f=F, energy=np.array(e), virial=vir)


def melt(nside, seed):
    rng = np.random.default_rng(seed)
    a = (1.0 / 0.9) ** (1.0 / 3.0)
    i, j, k = np.meshgrid(np.arange(nside), np.arange(nside), np.arange(nside), indexing="ij")
    t = ((i + j + k).reshape(-1) % 2).astype(np.int32)
    L = nside * a
    # displaced sites, wrapped into [0, L) as the integrator leaves them
    x, y, z = (np.mod(c.reshape(-1) * a + rng.normal(0, 0.06, nside ** 3), L) for c in (i, j, k))
    q = np.where(t == 0, 1.0, -1.0) / math.sqrt(COUL)
    return x, y, z, q, t, np.full(3, L)


def ljcoul_case():
    x, y, z, q, t, box = melt(8, 31)
    cutoff, alpha = 3.0, O.ewald_alpha(3.0, 1e-5)
    eps = np.array([[1.0, 0.0], [0.0, 1.0]])
    sig = np.array([[1.0, 0.0], [0.0, 0.8]])
    st = np.array([[1, 0], [0, 1]], np.int32)
    tab = O.LJCoulTables(eps, sig, st, cutoff, alpha)
    half = O.build_half(x, y, z, box, cutoff + 0.3)
    fx, fy, fz, e, vir = O.pair_lj_coul_long(x, y, z, q, t, box, half, tab)
    # cross-check: closed-form erfc Coulomb (numpy) + per-type LJ (oracle's pinned PairLJCut)
    ec, Fc = real_space_direct(x, y, z, q, box, alpha, cutoff)
    lx, ly, lz, el, _ = O.pair_lj(x, y, z, box, half, 0, 0, cutoff, type_=t, eps_tab=tab.eps, sigma_tab=tab.sigma)
    F = np.stack([fx, fy, fz], 1)
    ref = Fc + np.stack([lx, ly, lz], 1)
    assert math.sqrt(np.sum((F - ref) ** 2) / np.sum(ref ** 2)) <= 1e-5
    assert abs(e - (ec + el)) <= 1e-5 * abs(ec + el)
    save("ljcoul_forces", x=x, y=y, z=z, q=q, type=t, box=box, eps=eps, sigma=sig, set=st, cutoff=np.array(cutoff),
         alpha=np.array(alpha), skin=np.array(0.3), f=F, energy=np.array(e), virial=vir)


def pme_case():
    """Two meshes of one configuration: order 4 at spacing 0.55 (the charged
    melt's order; mesh error ~3e-2 of the reciprocal part) and order 8 at
    spacing 0.3, which the mode-by-mode Ewald sum pins to 1e-5."""
    x, y, z, q, t, box = melt(8, 32)
    alpha = O.ewald_alpha(3.0, 1e-5)
    er, Fr, Wr = ewald_direct_recip(x, y, z, q, box, alpha)
    for name, spacing, order, tol in (("pme_order4", 0.55, 4, 5e-2), ("pme_order8", 0.3, 8, 1e-5)):
        K = O.pme_grid(box, spacing, order)
        fx, fy, fz, e, vir, parts = O.pme(x, y, z, q, box, alpha, K, order)
        F = np.stack([fx, fy, fz], 1)
        assert abs(parts[0] - er) <= tol * abs(er), name
        assert math.sqrt(np.sum((F - Fr) ** 2) / np.sum(Fr ** 2)) <= tol, name
        assert np.abs(vir - Wr).max() <= tol * np.abs(Wr[:3]).max(), name
        save(name, x=x, y=y, z=z, q=q, type=t, box=box, alpha=np.array(alpha), spacing=np.array(spacing),
             order=np.array(order), K=K, f=F, energy=np.array(e), virial=vir, parts=np.array(parts))


def md_case():
    a = (4.0 / 0.8442) ** (1.0 / 3.0)
    x, y, z = O.fcc_lattice(5, 5, 5, a)
    n = x.size
    box = np.full(3, 5 * a)
    mass = np.ones(n)
    vx, vy, vz = O.init_velocity(mass, 0.72 / KB, 87287)
    st = dict(x=x.copy(), y=y.copy(), z=z.copy(), vx=vx.copy(), vy=vy.copy(), vz=vz.copy(), fx=np.zeros(n),
              fy=np.zeros(n), fz=np.zeros(n), mass=mass)
    p = O.MDParams(dt=0.005 * TAU, cutoff=2.5, skin=0.3, every=1, delay=0, check=1, eps=1.0, sigma=1.0, nh=1,
                   T_target=0.72 / KB, tau_T=0.5 * TAU, ndof=0, nthreads=1, stress=1)
    pe, ke, _, state = O.md_run(st, box, p, 60)
    save("md_nvt", x0=x, y0=y, z0=z, vx0=vx, vy0=vy, vz0=vz, box=box, x=st["x"], y=st["y"], z=st["z"], vx=st["vx"],
         vy=st["vy"], vz=st["vz"], pe=pe, ke=ke, xi=np.array(state.xi), nrebuilds=np.array(state.nrebuilds))


if __name__ == "__main__":
    pair_set_cases()
    lj_case()
    ljcoul_case()
    pme_case()
    md_case()
