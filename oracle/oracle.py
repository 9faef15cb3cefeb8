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


class MDForceField(C.Structure):
    _fields_ = [
        ("ntypes", C.c_int),
        ("type", _ip),
        ("q", _dp),
        ("lj1", _dp),
        ("lj2", _dp),
        ("lj3", _dp),
        ("lj4", _dp),
        ("shift", _dp),
        ("e_tab", _dp),
        ("f_tab", _dp),
        ("nbin", C.c_int),
        ("alpha", C.c_double),
        ("pme_K", C.c_int * 3),
        ("pme_order", C.c_int),
    ]


class MDState(C.Structure):
    _fields_ = [("xi", C.c_double), ("step", C.c_int64), ("nrebuilds", C.c_int64), ("ago", C.c_int64)]


NPT_MAXCHAIN = 8


class NPTParams(C.Structure):
    """cpu::md::NoseHooverNPT constructor arguments (nose_hoover_npt.h:183-188)."""
    _fields_ = [
        ("T_target", C.c_double),
        ("tau_T", C.c_double),
        ("P_target", C.c_double),
        ("tau_P", C.c_double),
        ("drag", C.c_double),
        ("mtchain", C.c_int),
        ("mpchain", C.c_int),
        ("couple", C.c_int * 3),
        ("T_stop", C.c_double),
        ("nfirst", C.c_int64),
        ("nlast", C.c_int64),
    ]


class NPTState(C.Structure):
    _fields_ = [
        ("eta_t", C.c_double * NPT_MAXCHAIN),
        ("eta_t_dot", C.c_double * NPT_MAXCHAIN),
        ("eta_b", C.c_double * NPT_MAXCHAIN),
        ("eta_b_dot", C.c_double * NPT_MAXCHAIN),
        ("eps_dot", C.c_double * 3),
        ("ke_global", C.c_double),
        ("vir", C.c_double * 3),
        ("built", C.c_int),
    ]


def npt_params(T_target, tau_T, P_target, tau_P, drag=0.0, mtchain=3, mpchain=3, couple=(1, 1, 1),
               T_stop=-1.0, nfirst=0, nlast=0):
    q = NPTParams(T_target=T_target, tau_T=tau_T, P_target=P_target, tau_P=tau_P, drag=drag,
                  mtchain=mtchain, mpchain=mpchain, T_stop=T_stop, nfirst=nfirst, nlast=nlast)
    for a in range(3):
        q.couple[a] = 1 if couple[a] else 0
    return q


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
        L.orc_kinetic_stress.restype = None
        L.orc_kinetic_stress.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp]
        L.orc_md_run.restype = C.c_int
        L.orc_md_run.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                 C.POINTER(MDParams), C.POINTER(MDState), C.c_int, _dp, _dp, _dp]
        L.orc_md_run_ff.restype = C.c_int
        L.orc_md_run_ff.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                    C.POINTER(MDParams), C.POINTER(MDForceField), C.POINTER(MDState),
                                    C.c_int, _dp, _dp, _dp]
        L.orc_npt_run.restype = C.c_int
        L.orc_npt_run.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                  C.POINTER(MDParams), C.POINTER(NPTParams), C.POINTER(MDState),
                                  C.POINTER(NPTState), C.c_int, _dp, _dp, _dp, _dp]
        L.orc_good_fft_size.restype = C.c_int
        L.orc_good_fft_size.argtypes = [C.c_int]
        L.orc_bspline.restype = None
        L.orc_bspline.argtypes = [C.c_double, C.c_int, _dp, _dp]
        L.orc_spline_moduli.restype = None
        L.orc_spline_moduli.argtypes = [C.c_int, C.c_int, _dp]
        L.orc_ewald_alpha.restype = C.c_double
        L.orc_ewald_alpha.argtypes = [C.c_double, C.c_double]
        L.orc_pme_grid.restype = None
        L.orc_pme_grid.argtypes = [_dp, C.c_double, C.c_int, _ip]
        L.orc_fft3d.restype = None
        L.orc_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int]
        L.orc_pme.restype = C.c_int
        L.orc_pme.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, C.c_double, _ip, C.c_int, _dp, _dp, _dp,
                              _dp, _dp, _dp]
        L.orc_erfc_table.restype = None
        L.orc_erfc_table.argtypes = [C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.orc_ljcoul_tables.restype = None
        L.orc_ljcoul_tables.argtypes = [C.c_int, C.c_double, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _dp]
        L.orc_pair_lj_coul_long.restype = None
        L.orc_pair_lj_coul_long.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _ip, _dp, _ip, _lp, _ip, C.c_int,
                                            _dp, _dp, _dp, _dp, _dp, C.c_double, _dp, _dp, C.c_int, C.c_int,
                                            _dp, _dp, _dp, _dp, _dp]
        L.orc_coul_long.restype = None
        L.orc_coul_long.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _ip, _lp, _ip, C.c_double,
                                    C.c_double, _dp, _dp, _dp, _dp, _dp]
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


def kinetic_stress(mass, vx, vy, vz):
    """cpu::md::Virial's kinetic stress (eV, Voigt xx yy zz xy xz yz)."""
    out = np.zeros(6)
    lib().orc_kinetic_stress(mass.size, _d(mass), _d(vx), _d(vy), _d(vz), _d(out))
    return out


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


def npt_run(state_arrays, box, params: MDParams, npt: NPTParams, nsteps, md_state: MDState | None = None,
            npt_state: NPTState | None = None):
    """MTK NoseHooverNPT MD loop in place on dict arrays (x..fz, mass); `box`
    (float64[3]) is rescaled in place. Returns (pe, ke, H, V) per step and the
    two state structs."""
    s = state_arrays
    if box.dtype != np.float64 or not box.flags.c_contiguous:
        raise ValueError("box must be a contiguous float64 array (updated in place)")
    md_state = md_state or MDState()
    npt_state = npt_state or NPTState()
    out = [np.zeros(nsteps) for _ in range(4)]
    rc = lib().orc_npt_run(s["x"].size, _d(s["x"]), _d(s["y"]), _d(s["z"]), _d(s["vx"]), _d(s["vy"]),
                           _d(s["vz"]), _d(s["fx"]), _d(s["fy"]), _d(s["fz"]), _d(s["mass"]), _d(box),
                           C.byref(params), C.byref(npt), C.byref(md_state), C.byref(npt_state), nsteps,
                           *[_d(a) for a in out])
    if rc != 0:
        raise RuntimeError(f"orc_npt_run failed ({rc})")
    return (*out, md_state, npt_state)


# ---- long-range electrostatics ------------------------------------------------

def good_fft_size(n):
    return lib().orc_good_fft_size(int(n))


def bspline(frac, order):
    th = np.zeros(order)
    dth = np.zeros(order)
    lib().orc_bspline(float(frac), order, _d(th), _d(dth))
    return th, dth


def spline_moduli(K, order):
    m = np.zeros(K)
    lib().orc_spline_moduli(K, order, _d(m))
    return m


def ewald_alpha(cutoff, accuracy=1e-5):
    return lib().orc_ewald_alpha(cutoff, accuracy)


def pme_grid(box, spacing, order):
    box = np.ascontiguousarray(box, dtype=np.float64)
    K = np.zeros(3, np.int32)
    lib().orc_pme_grid(_d(box), spacing, order, _i(K))
    return K


def fft3d(a, sign):
    """Unnormalised 3-D DFT of a complex array (sign -1 forward, +1 backward)."""
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    K0, K1, K2 = a.shape
    lib().orc_fft3d(K0, K1, K2, a.ctypes.data_as(_dp), sign)
    return a


def pme(x, y, z, q, box, alpha, K, order):
    """KSpacePME::compute -> (fx, fy, fz, energy, virial[6], (e_recip, e_self, e_net))."""
    x, y, z, q = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, q))
    box = np.ascontiguousarray(box, dtype=np.float64)
    K = np.ascontiguousarray(K, dtype=np.int32)
    n = x.size
    fx, fy, fz = (np.zeros(n) for _ in range(3))
    e = np.zeros(1)
    vir = np.zeros(6)
    parts = np.zeros(3)
    rc = lib().orc_pme(n, _d(x), _d(y), _d(z), _d(q), _d(box), alpha, _i(K), order, _d(fx), _d(fy), _d(fz),
                       _d(e), _d(vir), _d(parts))
    if rc != 0:
        raise ValueError("orc_pme: bad arguments")
    return fx, fy, fz, float(e[0]), vir, tuple(parts)


def erfc_table(alpha, cutoff, nbin=10000):
    e = np.zeros(nbin + 1)
    f = np.zeros(nbin + 1)
    lib().orc_erfc_table(alpha, cutoff, nbin, _d(e), _d(f))
    return e, f


class LJCoulTables:
    """finalize_type_table_ output for a coefficient table (ntypes x ntypes;
    `set` marks entries given explicitly, the rest are mixed)."""

    def __init__(self, eps, sigma, set_, cutoff, alpha, nbin=10000):
        self.eps = np.ascontiguousarray(eps, dtype=np.float64).reshape(-1).copy()
        self.sigma = np.ascontiguousarray(sigma, dtype=np.float64).reshape(-1).copy()
        self.set = np.ascontiguousarray(set_, dtype=np.int32).reshape(-1)
        self.ntypes = int(round(np.sqrt(self.eps.size)))
        self.lj1, self.lj2, self.lj3, self.lj4, self.shift = (np.zeros(self.eps.size) for _ in range(5))
        lib().orc_ljcoul_tables(self.ntypes, cutoff, _d(self.eps), _d(self.sigma), _i(self.set), _d(self.lj1),
                                _d(self.lj2), _d(self.lj3), _d(self.lj4), _d(self.shift))
        self.cutoff = cutoff
        self.alpha = alpha
        self.nbin = nbin
        self.e_tab, self.f_tab = erfc_table(alpha, cutoff, nbin)


def pair_lj_coul_long(x, y, z, q, type_, box, half, tab: LJCoulTables, nthreads=1):
    """PairLJCutCoulLong::compute on a half CSR -> (fx, fy, fz, energy, virial[6])."""
    numneigh, first, nb = half
    x, y, z, q = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, q))
    type_ = np.ascontiguousarray(type_, dtype=np.int32)
    box = np.ascontiguousarray(box, dtype=np.float64)
    n = x.size
    fx, fy, fz = (np.zeros(n) for _ in range(3))
    e = np.zeros(1)
    vir = np.zeros(6)
    lib().orc_pair_lj_coul_long(n, _d(x), _d(y), _d(z), _d(q), _i(type_), _d(box), _i(numneigh), _l(first),
                                _i(nb), tab.ntypes, _d(tab.lj1), _d(tab.lj2), _d(tab.lj3), _d(tab.lj4),
                                _d(tab.shift), tab.cutoff, _d(tab.e_tab), _d(tab.f_tab), tab.nbin, nthreads,
                                _d(fx), _d(fy), _d(fz), _d(e), _d(vir))
    return fx, fy, fz, float(e[0]), vir


def coul_long(x, y, z, q, box, half, alpha, cutoff):
    """CoulLong::compute (closed-form erfc) -> (fx, fy, fz, energy, virial[6])."""
    numneigh, first, nb = half
    x, y, z, q = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, q))
    box = np.ascontiguousarray(box, dtype=np.float64)
    n = x.size
    fx, fy, fz = (np.zeros(n) for _ in range(3))
    e = np.zeros(1)
    vir = np.zeros(6)
    lib().orc_coul_long(n, _d(x), _d(y), _d(z), _d(q), _d(box), _i(numneigh), _l(first), _i(nb), alpha,
                        cutoff, _d(fx), _d(fy), _d(fz), _d(e), _d(vir))
    return fx, fy, fz, float(e[0]), vir


def all_pairs_half(n):
    """Half CSR holding every pair j > i (the reference tests' build_full_halflist)."""
    numneigh = (n - 1 - np.arange(n)).astype(np.int32)
    first = np.concatenate([[0], np.cumsum(numneigh)[:-1]]).astype(np.int64)
    nb = np.concatenate([np.arange(i + 1, n) for i in range(n)] + [np.zeros(0, np.int64)]).astype(np.int32)
    return numneigh, first, nb


def md_run_ff(state_arrays, box, params: MDParams, tab: LJCoulTables, type_, q, pme_K, pme_order, nsteps,
              md_state: MDState | None = None):
    """orc_md_run_ff: the MD loop with PairLJCutCoulLong (+ KSpacePME when pme_order > 0)."""
    s = state_arrays
    box = np.ascontiguousarray(box, dtype=np.float64)
    type_ = np.ascontiguousarray(type_, dtype=np.int32)
    q = np.ascontiguousarray(q, dtype=np.float64)
    ff = MDForceField()
    ff.ntypes = tab.ntypes
    ff.type = _i(type_)
    ff.q = _d(q)
    ff.lj1, ff.lj2, ff.lj3, ff.lj4, ff.shift = (_d(a) for a in (tab.lj1, tab.lj2, tab.lj3, tab.lj4, tab.shift))
    ff.e_tab, ff.f_tab = _d(tab.e_tab), _d(tab.f_tab)
    ff.nbin = tab.nbin
    ff.alpha = tab.alpha
    for d in range(3):
        ff.pme_K[d] = int(pme_K[d]) if pme_order > 0 else 0
    ff.pme_order = int(pme_order)
    if md_state is None:
        md_state = MDState()
    pe = np.zeros(nsteps)
    ke = np.zeros(nsteps)
    vir = np.zeros(6)
    keep = (type_, q)  # noqa: F841  (alive across the call)
    rc = lib().orc_md_run_ff(s["x"].size, _d(s["x"]), _d(s["y"]), _d(s["z"]), _d(s["vx"]), _d(s["vy"]),
                             _d(s["vz"]), _d(s["fx"]), _d(s["fy"]), _d(s["fz"]), _d(s["mass"]), _d(box),
                             C.byref(params), C.byref(ff), C.byref(md_state), nsteps, _d(pe), _d(ke), _d(vir))
    if rc != 0:
        raise RuntimeError(f"orc_md_run_ff failed ({rc})")
    return pe, ke, vir, md_state
