"""ctypes binding of the CPU oracle (libatv_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product path never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libatv_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_int64)


class MDParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("cutoff", C.c_double),
        ("skin", C.c_double),
        ("every", C.c_int),
        ("delay", C.c_int),
        ("check", C.c_int),
        ("eps", C.c_double),
        ("sigma", C.c_double),
        ("nh", C.c_int),
        ("T_target", C.c_double),
        ("tau_T", C.c_double),
        ("ndof", C.c_int64),
        ("nthreads", C.c_int),
        ("stress", C.c_int),
    ]


class MDState(C.Structure):
    _fields_ = [("xi", C.c_double), ("step", C.c_int64), ("nrebuilds", C.c_int64), ("ago", C.c_int64)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -ffp-contract=off -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_fcc_lattice.restype = C.c_int64
        L.orc_fcc_lattice.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp]
        L.orc_init_velocity.restype = None
        L.orc_init_velocity.argtypes = [C.c_int64, _dp, C.c_double, C.c_uint32, C.c_int64, _dp, _dp, _dp]
        L.orc_cell_grid.restype = C.c_int
        L.orc_cell_grid.argtypes = [_dp, C.c_double, _ip, _dp, _ip, _ip]
        L.orc_bin_atoms.restype = C.c_int
        L.orc_bin_atoms.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, C.c_double, _ip, _ip, _ip, _ip]
        L.orc_build_half.restype = C.c_int64
        L.orc_build_half.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, C.c_double, C.c_int, _ip, _lp, _ip,
                                     C.c_int64]
        L.orc_needs_rebuild.restype = C.c_int
        L.orc_needs_rebuild.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_pair_lj.restype = None
        L.orc_pair_lj.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _ip, _lp, _ip, C.c_double, C.c_double,
                                  C.c_double, C.c_int, _ip, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_lj_mix.restype = None
        L.orc_lj_mix.argtypes = [C.c_int, _dp, _dp]
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_kinetic_energy.argtypes = [C.c_int64, _dp, _dp, _dp, _dp]
        L.orc_md_run.restype = C.c_int
        L.orc_md_run.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                 C.POINTER(MDParams), C.POINTER(MDState), C.c_int, _dp, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _l(a):
    return a.ctypes.data_as(_lp)


def fcc_lattice(nx, ny, nz, a):
    n = lib().orc_fcc_lattice(nx, ny, nz, a, None, None, None)
    x, y, z = (np.empty(n) for _ in range(3))
    lib().orc_fcc_lattice(nx, ny, nz, a, _d(x), _d(y), _d(z))
    return x, y, z


def init_velocity(mass, T, seed, ndof=None):
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    n = mass.size
    if ndof is None:
        ndof = 3 * n - 3
    vx, vy, vz = (np.empty(n) for _ in range(3))
    lib().orc_init_velocity(n, _d(mass), T, seed, ndof, _d(vx), _d(vy), _d(vz))
    return vx, vy, vz


def cell_grid(box, cutoff_skin):
    box = np.ascontiguousarray(box, dtype=np.float64)
    nc = np.zeros(3, np.int32)
    cs = np.zeros(3)
    reach = np.zeros(3, np.int32)
    nfull = np.zeros(1, np.int32)
    nhalf = lib().orc_cell_grid(_d(box), cutoff_skin, _i(nc), _d(cs), _i(reach), _i(nfull))
    return nc, cs, reach, nhalf, int(nfull[0])


def bin_atoms(x, y, z, box, cutoff_skin):
    box = np.ascontiguousarray(box, dtype=np.float64)
    nc, _, _, _, _ = cell_grid(box, cutoff_skin)
    n = x.size
    ncell = int(np.prod(nc))
    atom_cell = np.zeros(n, np.int32)
    binned = np.zeros(n, np.int32)
    count = np.zeros(ncell, np.int32)
    offset = np.zeros(ncell, np.int32)
    lib().orc_bin_atoms(n, _d(x), _d(y), _d(z), _d(box), cutoff_skin, _i(atom_cell), _i(count),
                        _i(offset), _i(binned))
    return atom_cell, count, offset, binned


def build_half(x, y, z, box, cutoff_skin, nthreads=1):
    """Half CSR (numneigh, firstneigh, neighbors) exactly as CellList::refresh_half."""
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    box = np.ascontiguousarray(box, dtype=np.float64)
    n = x.size
    numneigh = np.zeros(n, np.int32)
    first = np.zeros(n, np.int64)
    cap = max(16, n * 48)
    while True:
        nb = np.zeros(cap, np.int32)
        tot = lib().orc_build_half(n, _d(x), _d(y), _d(z), _d(box), cutoff_skin, nthreads, _i(numneigh),
                                   _l(first), _i(nb), cap)
        if tot >= 0:
            return numneigh, first, nb[:tot].copy()
        cap = -tot


def half_pairs(numneigh, first, neighbors):
    """(P, 2) int64 array of (i, j) for every half-list entry."""
    i = np.repeat(np.arange(numneigh.size, dtype=np.int64), numneigh)
    return np.stack([i, neighbors.astype(np.int64)], axis=1)


def pair_keys(pairs, n):
    """Order-free key per unordered pair: min*n + max, sorted."""
    a = np.minimum(pairs[:, 0], pairs[:, 1]).astype(np.int64)
    b = np.maximum(pairs[:, 0], pairs[:, 1]).astype(np.int64)
    return np.sort(a * n + b)


def needs_rebuild(x, y, z, xs, ys, zs, box, skin):
    box = np.ascontiguousarray(box, dtype=np.float64)
    m = np.zeros(1)
    r = lib().orc_needs_rebuild(x.size, _d(x), _d(y), _d(z), _d(xs), _d(ys), _d(zs), _d(box), skin, _d(m))
    return bool(r), float(m[0])


def pair_lj(x, y, z, box, half, eps, sigma, cutoff, type_=None, eps_tab=None, sigma_tab=None,
            nthreads=1):
    """PairLJCut::compute on a half CSR -> (fx, fy, fz, energy, virial[6])."""
    numneigh, first, nb = half
    box = np.ascontiguousarray(box, dtype=np.float64)
    n = x.size
    fx, fy, fz = (np.zeros(n) for _ in range(3))
    e = np.zeros(1)
    vir = np.zeros(6)
    ntypes = 0
    tp = et = st = None
    if type_ is not None:
        eps_tab = np.ascontiguousarray(eps_tab, dtype=np.float64)
        sigma_tab = np.ascontiguousarray(sigma_tab, dtype=np.float64)
        ntypes = int(round(np.sqrt(eps_tab.size)))
        type_ = np.ascontiguousarray(type_, dtype=np.int32)
        tp, et, st = _i(type_), _d(eps_tab), _d(sigma_tab)
    lib().orc_pair_lj(n, _d(x), _d(y), _d(z), _d(box), _i(numneigh), _l(first), _i(nb), eps, sigma, cutoff,
                      ntypes, tp, et, st, nthreads, _d(fx), _d(fy), _d(fz), _d(e), _d(vir))
    return fx, fy, fz, float(e[0]), vir


def lj_mix(eps_tab, sigma_tab):
    eps_tab = np.ascontiguousarray(eps_tab, dtype=np.float64).copy()
    sigma_tab = np.ascontiguousarray(sigma_tab, dtype=np.float64).copy()
    nt = int(round(np.sqrt(eps_tab.size)))
    lib().orc_lj_mix(nt, _d(eps_tab), _d(sigma_tab))
    return eps_tab, sigma_tab


def kinetic_energy(mass, vx, vy, vz):
    return lib().orc_kinetic_energy(mass.size, _d(mass), _d(vx), _d(vy), _d(vz))


def md_run(state_arrays, box, params: MDParams, nsteps, md_state: MDState | None = None):
    """Run the oracle MD loop in place on dict arrays x,y,z,vx,vy,vz,fx,fy,fz,mass.

    Returns (pe[nsteps], ke[nsteps], virial[6], MDState)."""
    s = state_arrays
    box = np.ascontiguousarray(box, dtype=np.float64)
    if md_state is None:
        md_state = MDState()
    pe = np.zeros(nsteps)
    ke = np.zeros(nsteps)
    vir = np.zeros(6)
    rc = lib().orc_md_run(s["x"].size, _d(s["x"]), _d(s["y"]), _d(s["z"]), _d(s["vx"]), _d(s["vy"]),
                          _d(s["vz"]), _d(s["fx"]), _d(s["fy"]), _d(s["fz"]), _d(s["mass"]), _d(box),
                          C.byref(params), C.byref(md_state), nsteps, _d(pe), _d(ke), _d(vir))
    if rc != 0:
        raise RuntimeError(f"orc_md_run failed ({rc})")
    return pe, ke, vir, md_state
