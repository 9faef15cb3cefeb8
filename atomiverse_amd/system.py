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
from ._abi import dptr, iptr, lptr

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

    def enable_timing(self, on=True):
        _check(self.lib.atvh_ctx_enable_timing(self.h, 1 if on else 0), self.h)

    def reset_timing(self):
        _check(self.lib.atvh_ctx_reset_timing(self.h), self.h)

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

    # integration
    def integrator(self, dt, thermostat=0, T=0.0, tau=0.0, ndof=0):
        self._ck(self.lib.atvh_integrator_configure(self.h, dt, thermostat, T, tau, ndof))

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
