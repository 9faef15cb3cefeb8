"""Synthetic benchmark workloads in the reference's metal units.

Reduced Lennard-Jones quantities map to metal units (eV, Å, fs, amu, K) with
eps = 1 eV, sigma = 1 Å, m = 1 amu:
    T   = T* / kB                     (kB = 8.617333262e-5 eV/K)
    tau = sigma sqrt(m/eps) = sqrt(103.6427) fs = 10.1805 fs
    dt  = dt* tau,  tau_NH = tau_NH* tau
    q   = q* / sqrt(14.399645) e      (so q*q*/r* is in units of eps)
The generators are the library's own (atvh_util_*), so the GPU path and the
CPU checker consume bit-identical inputs.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import dptr, iptr

KB = 8.617333262e-5
MVV2E = 103.6427
TAU_FS = math.sqrt(MVV2E)
QQR2E = 14.399645


@dataclass
class LJWorkload:
    name: str
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    mass: np.ndarray
    box: np.ndarray
    cutoff: float = 2.5
    skin: float = 0.3
    dt: float = 0.005 * TAU_FS
    eps: float = 1.0
    sigma: float = 1.0
    T: float = 0.72 / KB
    tau_T: float = 0.5 * TAU_FS
    thermostat: int = 0
    every: int = 1
    delay: int = 0
    check: int = 1
    type_: np.ndarray | None = None
    charge: np.ndarray | None = None
    # charged systems: PairLJCutCoulLong tables and KSpacePME
    pair: str = "lj"
    eps_tab: np.ndarray | None = None
    sigma_tab: np.ndarray | None = None
    set_tab: np.ndarray | None = None
    alpha: float = 0.0
    pme_spacing: float = 0.0
    pme_order: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.x.size


def fcc_lattice(nx, ny, nz, a, nmax=-1):
    lib = _abi.load()
    n = lib.atvh_util_fcc_lattice(nx, ny, nz, a, nmax, None, None, None)
    x, y, z = (np.empty(n) for _ in range(3))
    lib.atvh_util_fcc_lattice(nx, ny, nz, a, nmax, dptr(x), dptr(y), dptr(z))
    return x, y, z


def rocksalt_lattice(nx, ny, nz, spacing, q0):
    lib = _abi.load()
    n = lib.atvh_util_rocksalt_lattice(nx, ny, nz, spacing, None, None, None, None, None, q0)
    x, y, z, q = (np.empty(n) for _ in range(4))
    t = np.empty(n, np.int32)
    lib.atvh_util_rocksalt_lattice(nx, ny, nz, spacing, dptr(x), dptr(y), dptr(z), iptr(t), dptr(q), q0)
    return x, y, z, t, q


def init_velocity(mass, T, seed, ndof=None):
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    n = mass.size
    if ndof is None:
        ndof = 3 * n - 3
    vx, vy, vz = (np.empty(n) for _ in range(3))
    rc = _abi.load().atvh_util_init_velocity(n, dptr(mass), T, seed, ndof, dptr(vx), dptr(vy), dptr(vz))
    if rc != 0:
        raise _abi.HipError(rc, "atvh_util_init_velocity")
    return vx, vy, vz


def lj_fcc(ncell, density=0.8442, T_star=0.72, seed=87287, thermostat=0, name=None):
    """FCC LJ fluid: ncell^3 cells, 4 atoms per cell, seeded MB velocities."""
    a = (4.0 / density) ** (1.0 / 3.0)
    x, y, z = fcc_lattice(ncell, ncell, ncell, a)
    n = x.size
    L = ncell * a
    mass = np.ones(n)
    T = T_star / KB
    vx, vy, vz = init_velocity(mass, T, seed)
    return LJWorkload(name or f"lj_fcc_{n}", x, y, z, vx, vy, vz, mass, np.array([L, L, L]), T=T,
                      thermostat=thermostat, meta=dict(ncell=ncell, density=density, T_star=T_star, seed=seed))


def lj_fcc_4000():
    """4000-atom FCC LJ fluid, NVE (10^3 cells, rho*=0.8442, T*=0.72, seed 87287)."""
    return lj_fcc(10, seed=87287, thermostat=0, name="lj_fcc_4000_nve")


def lj_fcc_2048000(thermostat=1):
    """2,048,000-atom FCC LJ fluid (80^3 cells), Nose-Hoover tau*=0.5, seed 87288."""
    return lj_fcc(80, seed=87288, thermostat=thermostat, name="lj_fcc_2048000_nvt")


def npt_sc_216():
    """The reference's MTK NPT regression input (regressions/md/npt_lj_mtk/
    reg_npt_lj_mtk.cpp:15-40): 6^3 simple-cubic sites at rho = 0.8, unit masses,
    the deterministic velocity pattern 0.02*((i+j+k)%3-1), 0.02*((i+2j)%3-1),
    0.02*((2i+k)%3-1) [Å/fs], NeighborList(2.5, 0.3) with its defaults (every 1,
    delay 10, check). Run: NoseHooverNPT(T = 1/kB, tau_T 0.1, P 0, tau_P 1,
    dt 0.005 fs), 80 steps."""
    n_side = 6
    n = n_side ** 3
    L = (n / 0.8) ** (1.0 / 3.0)
    sp = L / n_side
    ix, iy, iz = np.unravel_index(np.arange(n), (n_side, n_side, n_side))
    x = (ix + 0.5) * sp
    y = (iy + 0.5) * sp
    z = (iz + 0.5) * sp
    vx = 0.02 * ((ix + iy + iz) % 3 - 1)
    vy = 0.02 * ((ix + 2 * iy) % 3 - 1)
    vz = 0.02 * ((2 * ix + iz) % 3 - 1)
    return LJWorkload("npt_sc_216", x.astype(float), y.astype(float), z.astype(float), vx.astype(float),
                      vy.astype(float), vz.astype(float), np.ones(n), np.array([L, L, L]), dt=0.005,
                      T=1.0 / KB, tau_T=0.1, delay=10,
                      meta=dict(P_target=0.0, tau_P=1.0, steps=80, V_reference_cpu=269.425432))


def ewald_alpha(cutoff, accuracy=1e-5):
    """EwaldSum::recommend_alpha (ewald_sum.h:58-60)."""
    return (1.35 - 0.15 * math.log(accuracy)) / cutoff


def charged_melt(nside, density=0.9, T_star=1.2, seed=87289, thermostat=1, cutoff=3.0, skin=0.3,
                 accuracy=1e-5, pme_spacing=0.65, pme_order=4, name=None):
    """Charged LJ binary melt: nside^3 simple-cubic sites (rock salt), alternating
    +-1 reduced charges, species A (sigma 1.0) and B (sigma 0.8), eps 1,
    Lorentz-Berthelot for A-B, unit masses, seeded MB velocities."""
    a = (1.0 / density) ** (1.0 / 3.0)
    x, y, z, t, q = rocksalt_lattice(nside, nside, nside, a, 1.0 / math.sqrt(QQR2E))
    n = x.size
    L = nside * a
    mass = np.ones(n)
    T = T_star / KB
    vx, vy, vz = init_velocity(mass, T, seed)
    return LJWorkload(name or f"charged_melt_{n}", x, y, z, vx, vy, vz, mass, np.array([L, L, L]),
                      cutoff=cutoff, skin=skin, T=T, thermostat=thermostat, type_=t, charge=q, pair="ljcoul",
                      eps_tab=np.array([[1.0, 0.0], [0.0, 1.0]]), sigma_tab=np.array([[1.0, 0.0], [0.0, 0.8]]),
                      set_tab=np.array([[1, 0], [0, 1]], np.int32), alpha=ewald_alpha(cutoff, accuracy),
                      pme_spacing=pme_spacing, pme_order=pme_order,
                      meta=dict(nside=nside, density=density, T_star=T_star, seed=seed, accuracy=accuracy))


def charged_melt_512000():
    """512,000-ion charged LJ melt (80^3 rock salt, rho*=0.9, T*=1.2, seed 87289), NVT
    tau*=0.5, r_c 3.0 + skin 0.3, Ewald tolerance 1e-5, PME 128^3 order 4."""
    return charged_melt(80, name="charged_melt_512000_nvt")


def lj_trimmed(natoms, density=0.8442, T_star=0.72, seed=87287, thermostat=1, name=None):
    """FCC lattice trimmed to `natoms` sites in lattice order inside the box of
    that density: L = (N/rho)^(1/3), ceil((N/4)^(1/3)) cells per axis of edge
    L/cells (the first sites fill the box bottom-up; the rest is void)."""
    L = (natoms / density) ** (1.0 / 3.0)
    nc = math.ceil((natoms / 4.0) ** (1.0 / 3.0) - 1e-12)
    x, y, z = fcc_lattice(nc, nc, nc, L / nc, natoms)
    mass = np.ones(natoms)
    T = T_star / KB
    vx, vy, vz = init_velocity(mass, T, seed)
    return LJWorkload(name or f"lj_trimmed_{natoms}", x, y, z, vx, vy, vz, mass, np.array([L, L, L]), T=T,
                      thermostat=thermostat, meta=dict(density=density, T_star=T_star, seed=seed, cells=nc))


def lj_ladder(nsys=256, nmin=500, nmax=20000, seed=1234, t_lo=0.7, t_hi=1.5):
    """Batched REMD-style ensemble: nsys independent LJ fluids, N log-uniform in
    [nmin, nmax] (seeded), temperatures on the geometric ladder
    T*_k = t_lo (t_hi/t_lo)^(k/(nsys-1)), velocities seeded 87287 + k, NVT."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(nmin), math.log(nmax), nsys)).astype(np.int64)
    out = []
    for k, n in enumerate(sizes):
        t = t_lo * (t_hi / t_lo) ** (k / max(nsys - 1, 1))
        out.append(lj_trimmed(int(n), T_star=t, seed=87287 + k, thermostat=1, name=f"ladder_{k}_{int(n)}"))
    return out


def upload(sysobj, w: LJWorkload):
    sysobj.upload(w.x, w.y, w.z, w.mass, w.vx, w.vy, w.vz, type_=w.type_, charge=w.charge)


def configure(sysobj, w: LJWorkload):
    """NeighborList + PairLJCut (or PairLJCutCoulLong + KSpacePME) + VelocityVerlet(+NoseHoover)."""
    sysobj.nlist_configure(w.cutoff, w.skin, w.every, w.delay, bool(w.check))
    if w.pair == "ljcoul":
        sysobj.pair_ljcoul(w.eps_tab, w.sigma_tab, w.set_tab, w.alpha, w.cutoff)
        if w.pme_order > 0:
            sysobj.pme(w.alpha, w.pme_spacing, w.pme_order)
    else:
        sysobj.pair_lj(w.eps, w.sigma, w.cutoff)
    sysobj.integrator(w.dt, w.thermostat, w.T, w.tau_T, 0)
