"""The allocation-free MD run (atvh_md_run_noalloc) against atvh_md_run,
through the C ABI.

Two calls in a row on one GPU, the second on a smaller system with fewer
steps, each compared bit for bit with atvh_md_run from identical inputs
(positions, velocities, the per-step potential and kinetic energies and
where they are NaN). A charged PME system follows in mixed precision. The
entry point's refusals (no built list, PME not yet evaluated, too many steps,
a list rebuild beyond the capacity) are checked on inputs whose kernels all
stay in bounds: the capacity case is the one the capacity test already runs,
where a row beyond the capacity is counted and not stored.
"""
import ctypes as C

import numpy as np
import pytest

import atomiverse_amd as av
from atomiverse_amd import _abi
from atomiverse_amd import workloads as W

pytestmark = pytest.mark.gpu

MAX_STEPS = 1 << 16


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _err(ctx):
    msg = _abi.load().atvh_last_error(ctx.h)
    return msg.decode() if msg else ""


def _noalloc(s, nsteps, precision, energy_every):
    pe = np.full(nsteps, -1.0)
    ke = np.full(nsteps, -1.0)
    rc = s.lib.atvh_md_run_noalloc(s.h, int(nsteps), precision, int(energy_every), _dp(pe), _dp(ke))
    return rc, pe, ke


def _system(ctx, w, precision=None):
    s = av.System(ctx, w.n, w.box)
    W.upload(s, w)
    W.configure(s, w)
    s.nlist_build()
    if precision is not None:
        s.compute(precision, True, True)  # makes the PME buffers and plans of this precision
    return s


def _same_run(ctx, w, nsteps, precision, energy_every, ready=None):
    print(f"noalloc vs run: {w.name} n={w.n} steps={nsteps} prec={precision} every={energy_every}", flush=True)
    ref = _system(ctx, w, ready)
    ref.run(nsteps, precision, energy_every=energy_every)
    want = ref.download()
    pe_ref, ke_ref = ref.thermo(nsteps)
    ref.close()

    s = _system(ctx, w, ready)
    rc, pe, ke = _noalloc(s, nsteps, precision, energy_every)
    assert rc == _abi.OK, _err(ctx)
    got = s.download()
    s.close()
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(np.isnan(pe), np.isnan(pe_ref))
    m = ~np.isnan(pe_ref)
    assert m.sum() >= 2
    assert np.array_equal(pe[m], pe_ref[m])
    assert np.array_equal(ke, ke_ref)


def test_noalloc_run_equals_run_two_calls_second_smaller(hip_ctx):
    _same_run(hip_ctx, W.lj_fcc(10, thermostat=1, name="lj_fcc_4000_nvt"), 200, av.PREC_FP64, 10)
    _same_run(hip_ctx, W.lj_fcc(7, seed=31, thermostat=0, name="lj_fcc_1372_nve"), 60, av.PREC_MIXED, 7)


def test_noalloc_run_charged_pme(hip_ctx):
    w = W.charged_melt(10, name="charged_melt_1000")
    _same_run(hip_ctx, w, 40, av.PREC_MIXED, 5, ready=av.PREC_MIXED)


def test_noalloc_refusals(hip_ctx):
    w = W.lj_fcc(6, name="lj_fcc_864")
    s = av.System(hip_ctx, w.n, w.box)
    W.upload(s, w)
    W.configure(s, w)
    rc, _, _ = _noalloc(s, 10, av.PREC_FP64, 0)
    assert rc == _abi.E_STATE and "not built" in _err(hip_ctx)
    s.nlist_build()
    rc, _, _ = _noalloc(s, MAX_STEPS + 1, av.PREC_FP64, 0)
    assert rc == _abi.E_ARG
    s.close()

    wc = W.charged_melt(8, name="charged_melt_512")
    c = _system(hip_ctx, wc)
    rc, _, _ = _noalloc(c, 5, av.PREC_MIXED, 0)
    assert rc == _abi.E_STATE and "PME" in _err(hip_ctx)
    c.close()


def test_noalloc_overflow_is_reported_not_regrown(hip_ctx):
    w = W.lj_fcc(10, T_star=1.5, seed=4242, thermostat=1, name="lj_fcc_4000_hot")
    s = _system(hip_ctx, w)
    first = s.nlist_stats()
    s.set_capacity(first["max_neighbors"])
    s.nlist_build()  # the lattice's rows fit exactly; the melt's later rebuilds do not
    rc, _, _ = _noalloc(s, 300, av.PREC_FP64, 50)
    assert rc == _abi.E_CAPACITY, _err(hip_ctx)
    # nothing was regrown: the overflow stays reported until a list build
    with pytest.raises(av.HipError) as e:
        s.nlist_stats()
    assert e.value.code == _abi.E_CAPACITY
    s.close()
