"""Thin Python handles over the C ABI (test and benchmark plumbing).

The product is the C ABI plus the hip:: C++ components in include/; these
classes only marshal numpy arrays through ctypes.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref

import numpy as np

from . import _abi
from ._abi import dptr, iptr, lptr  # noqa: F401

# Live handles, released in dependency order (batches, systems, then contexts) before
# the interpreter and the HIP runtime shut down.
_live_batches = weakref.WeakSet()
_live_systems = weakref.WeakSet()
_live_contexts = weakref.WeakSet()


@atexit.register
def _release_all():
    for b in list(_live_batches):
        b.close()
    for s in list(_live_systems):
        s.close()
    for c in list(_live_contexts):
        c.close()


def _check(rc, ctx_handle):
    if rc != _abi.OK:
        msg = _abi.load().atvh_last_error(ctx_handle)
        raise _abi.HipError(rc, msg.decode() if msg else "")


class Context:
    """atvh_ctx: one device, one stream."""

    def __init__(self, device: int = 0):
        self.lib = _abi.load()
        h = C.c_void_p()
        rc = self.lib.atvh_ctx_create(device, C.byref(h))
        if rc != _abi.OK:
            raise _abi.HipError(rc, f"atvh_ctx_create(device={device}) failed")
        self.h = h
        _live_contexts.add(self)

    def close(self):
        """Destroy the context; a no-op while systems created on it are alive
        (the last System.close() or the garbage collector finishes it)."""
        if self.h and self.lib.atvh_ctx_destroy(self.h) == _abi.OK:
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(self.lib.atvh_ctx_synchronize(self.h), self.h)

    def enable_timing(self, on=True, only=None):
        """Per-kernel HIP-event timing; `only` restricts it to one kernel family."""
        _check(self.lib.atvh_ctx_timing_only(self.h, None if only is None else only.encode()), self.h)
        _check(self.lib.atvh_ctx_enable_timing(self.h, 1 if on else 0), self.h)

    def reset_timing(self):
        _check(self.lib.atvh_ctx_reset_timing(self.h), self.h)

    def mem_info(self):
        """(free, total) device bytes."""
        f = C.c_int64()
        t = C.c_int64()
        _check(self.lib.atvh_ctx_mem_info(self.h, C.byref(f), C.byref(t)), self.h)
        return f.value, t.value

    def kernel_time(self, name):
        ms = C.c_double()
        n = C.c_int64()
        _check(self.lib.atvh_ctx_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)), self.h)
        return ms.value, n.value


class System:
    """atvh_system: one orthorhombic periodic system resident on the device."""

    def __init__(self, ctx: Context, n: int, box):
        self.ctx = ctx
        self.lib = ctx.lib
        self.n = int(n)
        self.box = np.ascontiguousarray(box, dtype=np.float64)
        h = C.c_void_p()
        _check(self.lib.atvh_system_create(ctx.h, self.n, dptr(self.box), C.byref(h)), ctx.h)
        self.h = h
        _live_systems.add(self)

    def close(self):
        if self.h:
            self.lib.atvh_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        _check(rc, self.ctx.h)

    @staticmethod
    def _f(a, n):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (n,)
        return a

    def upload(self, x, y, z, mass, vx=None, vy=None, vz=None, fx=None, fy=None, fz=None, type_=None,
               charge=None):
        n = self.n
        arrs = [self._f(a, n) for a in (x, y, z, vx, vy, vz, fx, fy, fz)]
        mass = self._f(mass, n)
        t = None if type_ is None else np.ascontiguousarray(type_, dtype=np.int32)
        q = self._f(charge, n)
        self._keep = (arrs, mass, t, q)
        self._ck(self.lib.atvh_frame_upload(self.h, *[dptr(a) for a in arrs], dptr(mass), iptr(t), dptr(q)))

    def download(self):
        n = self.n
        out = {k: np.empty(n) for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")}
        self._ck(self.lib.atvh_frame_download(self.h, *[dptr(out[k]) for k in
                                                        ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")]))
        return out

    # neighbor list
    def nlist_configure(self, cutoff, skin=-1.0, every=1, delay=10, check=True):
        self._ck(self.lib.atvh_nlist_configure(self.h, cutoff, skin, every, delay, 1 if check else 0))

    def nlist_build(self):
        self._ck(self.lib.atvh_nlist_build(self.h))

    def nlist_decide_and_rebuild(self):
        self._ck(self.lib.atvh_nlist_decide_and_rebuild(self.h))

    def brick_shape(self):
        k = C.c_int()
        b = np.zeros(3, np.int32)
        self._ck(self.lib.atvh_nlist_brick_shape(self.h, C.byref(k), iptr(b)))
        return k.value, tuple(int(v) for v in b)

    def set_brick_shape(self, candidate):
        self._ck(self.lib.atvh_nlist_set_brick_shape(self.h, int(candidate)))

    def brick_fill(self):
        """(fullest halo, fullest brick's home atoms) of the current list."""
        mh = C.c_int()
        ma = C.c_int()
        self._ck(self.lib.atvh_nlist_brick_fill(self.h, C.byref(mh), C.byref(ma)))
        return mh.value, ma.value

    def set_brick_limit(self, max_atoms):
        """Home atoms a list build takes per brick (<= 320; 0 restores 320)."""
        self._ck(self.lib.atvh_nlist_set_brick_limit(self.h, int(max_atoms)))

    def autotune(self, precision=_abi.PREC_MIXED, trials=3):
        """Pick the fastest brick shape on this system; returns (candidate, microseconds)."""
        k = C.c_int()
        us = C.c_double()
        self._ck(self.lib.atvh_nlist_autotune(self.h, precision, int(trials), C.byref(k), C.byref(us)))
        return k.value, us.value

    def set_capacity(self, capacity):
        """Per-atom list capacity for the next build (grown on overflow)."""
        self._ck(self.lib.atvh_nlist_set_capacity(self.h, int(capacity)))

    def nlist_stats(self):
        nr = C.c_int64()
        npairs = C.c_int64()
        mx = C.c_int()
        cap = C.c_int()
        self._ck(self.lib.atvh_nlist_stats(self.h, C.byref(nr), C.byref(npairs), C.byref(mx), C.byref(cap)))
        return dict(nrebuilds=nr.value, npairs_half=npairs.value, max_neighbors=mx.value, capacity=cap.value)

    def export_pairs(self):
        """(P, 2) int64 unordered tag pairs (a < b), sorted."""
        cnt = C.c_int64()
        self._ck(self.lib.atvh_nlist_export_pairs(self.h, None, 0, C.byref(cnt)))
        out = np.empty((max(cnt.value, 1), 2), np.int64)
        self._ck(self.lib.atvh_nlist_export_pairs(self.h, lptr(out.reshape(-1)), cnt.value, C.byref(cnt)))
        return out[: cnt.value]

    # pair style
    def pair_lj(self, eps, sigma, cutoff, eps_tab=None, sigma_tab=None):
        if eps_tab is None:
            self._ck(self.lib.atvh_pair_lj_configure(self.h, eps, sigma, cutoff, 0, None, None))
        else:
            et = np.ascontiguousarray(eps_tab, dtype=np.float64).reshape(-1)
            st = np.ascontiguousarray(sigma_tab, dtype=np.float64).reshape(-1)
            nt = int(round(np.sqrt(et.size)))
            self._ck(self.lib.atvh_pair_lj_configure(self.h, eps, sigma, cutoff, nt, dptr(et), dptr(st)))

    def pair_ljcoul(self, eps_tab, sigma_tab, set_tab, alpha, cutoff, nbin=10000):
        """PairLJCutCoulLong: ntypes x ntypes coefficient tables (set_tab marks given entries)."""
        et = np.ascontiguousarray(eps_tab, dtype=np.float64).reshape(-1)
        st = np.ascontiguousarray(sigma_tab, dtype=np.float64).reshape(-1)
        mt = np.ascontiguousarray(set_tab, dtype=np.int32).reshape(-1)
        nt = int(round(np.sqrt(et.size)))
        self._ck(self.lib.atvh_pair_ljcoul_configure(self.h, nt, dptr(et), dptr(st), iptr(mt), alpha, cutoff,
                                                     int(nbin)))

    def pme(self, alpha, spacing, order):
        """KSpacePME(alpha, grid_spacing, order)."""
        self._ck(self.lib.atvh_pme_configure(self.h, alpha, spacing, int(order)))

    def pme_compute(self, precision=_abi.PREC_FP64, eflag=True, vflag=True):
        self._ck(self.lib.atvh_pme_compute(self.h, precision, 1 if eflag else 0, 1 if vflag else 0))

    def pme_info(self):
        """(K[3], (e_recip, e_self, e_net) of the last k-space energy evaluation)."""
        K = np.zeros(3, np.int32)
        parts = np.zeros(3)
        self._ck(self.lib.atvh_pme_info(self.h, iptr(K), dptr(parts)))
        return K, tuple(parts)

    def compute(self, precision=_abi.PREC_FP64, eflag=True, vflag=True):
        self._ck(self.lib.atvh_pair_compute(self.h, precision, 1 if eflag else 0, 1 if vflag else 0))

    def energy(self):
        pe = C.c_double()
        ke = C.c_double()
        vir = np.zeros(6)
        self._ck(self.lib.atvh_energy(self.h, C.byref(pe), dptr(vir), C.byref(ke)))
        return pe.value, vir, ke.value

    def kinetic_stress(self):
        """Queue the Virial observer's kinetic stress of the current velocities."""
        self._ck(self.lib.atvh_kinetic_stress(self.h))

    def stress(self):
        """(kinetic stress [6], virial [6], pressure) in eV, eV and eV/Å^3."""
        kin = np.zeros(6)
        vir = np.zeros(6)
        p = C.c_double()
        self._ck(self.lib.atvh_stress(self.h, dptr(kin), dptr(vir), C.byref(p)))
        return kin, vir, p.value

    # integration
    def integrator(self, dt, thermostat=0, T=0.0, tau=0.0, ndof=0):
        self._ck(self.lib.atvh_integrator_configure(self.h, dt, thermostat, T, tau, ndof))

    def npt(self, dt, T_target, tau_T, P_target, tau_P, drag=0.0, mtchain=3, mpchain=3, couple=(1, 1, 1),
            T_stop=-1.0):
        """MTK NoseHooverNPT as the integrator (replaces VelocityVerlet + NoseHoover)."""
        cp = np.array([1 if c else 0 for c in couple], np.int32)
        self._ck(self.lib.atvh_npt_configure(self.h, dt, T_target, tau_T, P_target, tau_P, drag, mtchain,
                                             mpchain, iptr(cp), T_stop))

    def current_box(self):
        """The box edges now (a barostat moves them)."""
        out = np.zeros(3)
        self._ck(self.lib.atvh_system_box(self.h, dptr(out)))
        return out

    def memory(self, part):
        """Device bytes this system holds for one part (MEM_FRAME, MEM_NLIST,
        MEM_PAIR, MEM_PME, MEM_INTEGRATOR)."""
        out = C.c_int64(0)
        self._ck(self.lib.atvh_system_memory(self.h, int(part), C.byref(out)))
        return out.value

    def npt_run_span(self, nfirst, nlast):
        self._ck(self.lib.atvh_npt_set_run_span(self.h, int(nfirst), int(nlast)))

    def npt_thermo(self, nsteps):
        """Per-step conserved quantity H and box volume of the last run."""
        H = np.empty(nsteps)
        V = np.empty(nsteps)
        self._ck(self.lib.atvh_npt_thermo(self.h, dptr(H), dptr(V), int(nsteps)))
        return H, V

    def npt_state(self, mtchain=3, mpchain=3):
        out = [np.zeros(3), np.zeros(mtchain), np.zeros(mtchain), np.zeros(mpchain), np.zeros(mpchain)]
        self._ck(self.lib.atvh_npt_state(self.h, *[dptr(a) for a in out]))
        return dict(eps_dot=out[0], eta_t=out[1], eta_t_dot=out[2], eta_b=out[3], eta_b_dot=out[4])

    def step_phase(self, phase):
        self._ck(self.lib.atvh_step_vvnh(self.h, phase))

    def thermostat_xi(self, set_to=None):
        out = C.c_double()
        if set_to is None:
            self._ck(self.lib.atvh_thermostat_state(self.h, C.byref(out), None))
        else:
            v = C.c_double(set_to)
            self._ck(self.lib.atvh_thermostat_state(self.h, C.byref(out), C.byref(v)))
        return out.value

    def run(self, nsteps, precision=_abi.PREC_FP64, energy_every=0):
        self._ck(self.lib.atvh_md_run(self.h, int(nsteps), precision, int(energy_every)))

    def thermo(self, nsteps):
        pe = np.empty(nsteps)
        ke = np.empty(nsteps)
        self._ck(self.lib.atvh_md_thermo(self.h, dptr(pe), dptr(ke), int(nsteps)))
        return pe, ke

    def synchronize(self):
        self.ctx.synchronize()


class Batch:
    """atvh_batch: member systems advanced together, one launch per stage."""

    def __init__(self, ctx: Context, systems):
        self.ctx = ctx
        self.lib = ctx.lib
        self.systems = list(systems)
        arr = (C.c_void_p * len(self.systems))(*[s.h.value for s in self.systems])
        h = C.c_void_p()
        _check(self.lib.atvh_batch_create(ctx.h, arr, len(self.systems), C.byref(h)), ctx.h)
        self.h = h
        _live_batches.add(self)

    def close(self):
        if self.h:
            self.lib.atvh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, enable=True):
        _check(self.lib.atvh_batch_set_graph(self.h, 1 if enable else 0), self.ctx.h)

    def run(self, nsteps, precision=_abi.PREC_FP64, energy_every=0):
        _check(self.lib.atvh_batch_run(self.h, int(nsteps), precision, int(energy_every)), self.ctx.h)

    def exchange(self, round_index, seed):
        """One replica-exchange round; returns the number of accepted swaps."""
        acc = C.c_int()
        _check(self.lib.atvh_batch_exchange(self.h, int(round_index), int(seed), C.byref(acc)), self.ctx.h)
        return acc.value

    def ladder(self):
        """(member holding each ladder slot, thermostat T of each member)."""
        n = len(self.systems)
        slots = np.zeros(n, np.int32)
        T = np.zeros(n)
        _check(self.lib.atvh_batch_ladder(self.h, iptr(slots), dptr(T)), self.ctx.h)
        return slots, T


class Segment(C.Structure):
    """atvh_segment (include/atvh.h)."""
    _fields_ = [
        ("n", C.c_int64), ("box", C.c_double * 3),
        ("x", _abi._dp), ("y", _abi._dp), ("z", _abi._dp), ("vx", _abi._dp), ("vy", _abi._dp), ("vz", _abi._dp),
        ("mass", _abi._dp), ("charge", _abi._dp), ("type", _abi._ip),
        ("cutoff", C.c_double), ("skin", C.c_double),
        ("pair_style", C.c_int), ("eps", C.c_double), ("sigma", C.c_double),
        ("ntypes", C.c_int), ("eps_tab", _abi._dp), ("sigma_tab", _abi._dp), ("set_tab", _abi._ip),
        ("alpha", C.c_double), ("pme_spacing", C.c_double), ("pme_order", C.c_int),
        ("dt", C.c_double), ("thermostat", C.c_int), ("T", C.c_double), ("tau", C.c_double),
        ("nsteps", C.c_int64),
    ]


def segment_of(w, nsteps):
    """Segment spec of a workload (keeps its arrays alive on the returned object)."""
    g = Segment()
    g.n = w.n
    for d in range(3):
        g.box[d] = float(w.box[d])
    keep = []

    def d_(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return dptr(a)

    def i_(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return iptr(a)

    g.x, g.y, g.z = d_(w.x), d_(w.y), d_(w.z)
    g.vx, g.vy, g.vz = d_(w.vx), d_(w.vy), d_(w.vz)
    g.mass, g.charge, g.type = d_(w.mass), d_(w.charge), i_(w.type_)
    g.cutoff, g.skin = w.cutoff, w.skin
    if w.pair == "ljcoul":
        g.pair_style = 1
        g.ntypes = int(round(np.sqrt(np.asarray(w.eps_tab).size)))
        g.eps_tab, g.sigma_tab, g.set_tab = d_(w.eps_tab), d_(w.sigma_tab), i_(w.set_tab)
        g.alpha, g.pme_spacing, g.pme_order = w.alpha, w.pme_spacing, w.pme_order
    else:
        g.pair_style = 0
        g.eps, g.sigma = w.eps, w.sigma
    g.dt, g.thermostat, g.T, g.tau = w.dt, w.thermostat, w.T, w.tau_T
    g.nsteps = int(nsteps)
    g._keep = keep
    return g


class Server:
    """atvh_server: HBM-budget admission and continuous batching of MD segments."""

    def __init__(self, ctx: Context, budget_bytes, chunk_steps=100, precision=_abi.PREC_MIXED):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        _check(self.lib.atvh_server_create(ctx.h, int(budget_bytes), int(chunk_steps), precision, C.byref(h)),
               ctx.h)
        self.h = h
        self.n = {}
        _live_batches.add(self)

    def close(self):
        if self.h:
            self.lib.atvh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, w, nsteps):
        g = segment_of(w, nsteps)
        sid = C.c_int64()
        _check(self.lib.atvh_server_submit(self.h, C.byref(g), C.byref(sid)), self.ctx.h)
        self.n[sid.value] = w.n
        return sid.value

    def advance(self):
        a = C.c_int()
        q = C.c_int()
        _check(self.lib.atvh_server_advance(self.h, C.byref(a), C.byref(q)), self.ctx.h)
        return a.value, q.value

    def poll(self):
        ids = np.zeros(1024, np.int64)
        k = C.c_int()
        _check(self.lib.atvh_server_poll(self.h, lptr(ids), ids.size, C.byref(k)), self.ctx.h)
        return [int(v) for v in ids[:k.value]]

    def fetch(self, sid):
        n = self.n.pop(sid)
        out = {k: np.empty(n) for k in ("x", "y", "z", "vx", "vy", "vz")}
        pe = C.c_double()
        ke = C.c_double()
        _check(self.lib.atvh_server_fetch(self.h, sid, *[dptr(out[k]) for k in ("x", "y", "z", "vx", "vy", "vz")],
                                          C.byref(pe), C.byref(ke)), self.ctx.h)
        out["pe"], out["ke"] = pe.value, ke.value
        return out

    def usage(self):
        u = C.c_int64()
        b = C.c_int64()
        _check(self.lib.atvh_server_usage(self.h, C.byref(u), C.byref(b)), self.ctx.h)
        return u.value, b.value


def segment_bytes(w, nsteps=0):
    return _abi.load().atvh_segment_bytes(C.byref(segment_of(w, nsteps)))
