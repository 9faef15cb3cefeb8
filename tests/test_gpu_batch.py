"""Batched multi-system step (atvh_batch_*) through the C ABI.

The bar is batch invariance, bit for bit: every member's positions,
velocities, forces, per-step energies, thermostat state and rebuild count
after a batched run equal those of the same system run alone with
atvh_md_run (the single-system path, itself checked against the oracle).
"""
import numpy as np
import pytest

import atomiverse_amd as av
from atomiverse_amd import workloads as W

pytestmark = pytest.mark.gpu

SIZES = [500, 777, 1500, 2601, 4000, 6100]


def members(ctx, sizes, t_lo=0.7, t_hi=1.5):
    out = []
    for k, n in enumerate(sizes):
        t = t_lo * (t_hi / t_lo) ** (k / max(len(sizes) - 1, 1))
        w = W.lj_trimmed(n, T_star=t, seed=87287 + k, thermostat=1)
        s = av.System(ctx, w.n, w.box)
        W.upload(s, w)
        W.configure(s, w)
        out.append((w, s))
    return out


def snapshot(s, nsteps):
    d = s.download()
    pe, ke = s.thermo(nsteps)
    return d, pe, ke, s.thermostat_xi(), s.nlist_stats()["nrebuilds"]


@pytest.mark.parametrize("prec", [av.PREC_FP64, av.PREC_MIXED])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_members_alone(hip_ctx, prec, graph):
    nsteps = 150
    alone = []
    for w, s in members(hip_ctx, SIZES):
        s.run(nsteps, prec, energy_every=10)
        alone.append(snapshot(s, nsteps))
        s.close()
    ms = members(hip_ctx, SIZES)
    b = av.Batch(hip_ctx, [s for _, s in ms])
    b.set_graph(graph)
    b.run(nsteps, prec, energy_every=10)
    for (w, s), (d0, pe0, ke0, xi0, nr0) in zip(ms, alone):
        d1, pe1, ke1, xi1, nr1 = snapshot(s, nsteps)
        for k in d0:
            np.testing.assert_array_equal(d1[k], d0[k], err_msg=f"{w.name} {k}")
        np.testing.assert_array_equal(pe1, pe0)
        np.testing.assert_array_equal(ke1, ke0)
        assert xi1 == xi0 and nr1 == nr0
        assert nr1 > 1
    b.close()
    for _, s in ms:
        s.close()


def test_batch_continues_and_matches_one_long_run(hip_ctx):
    """Two batched runs of 60 steps equal one of 120 for every member."""
    ms_a = members(hip_ctx, SIZES[:3])
    ms_b = members(hip_ctx, SIZES[:3])
    ba = av.Batch(hip_ctx, [s for _, s in ms_a])
    bb = av.Batch(hip_ctx, [s for _, s in ms_b])
    ba.run(120, av.PREC_MIXED)
    bb.run(60, av.PREC_MIXED)
    bb.run(60, av.PREC_MIXED)
    for (_, sa), (_, sb) in zip(ms_a, ms_b):
        da, db = sa.download(), sb.download()
        for k in da:
            np.testing.assert_array_equal(da[k], db[k])
    ba.close()
    bb.close()


def test_batch_rejects_unsupported_members(hip_ctx):
    w = W.lj_trimmed(500)
    good = av.System(hip_ctx, w.n, w.box)
    W.upload(good, w)
    W.configure(good, w)
    # a box with fewer than 2*reach+1 cells per axis has no brick layout
    tiny = W.lj_fcc(3, thermostat=1)
    small = av.System(hip_ctx, tiny.n, tiny.box)
    W.upload(small, tiny)
    W.configure(small, tiny)
    b = av.Batch(hip_ctx, [good, small])
    with pytest.raises(av.HipError) as e:
        b.run(5)
    assert e.value.code == av._abi.E_ARG
    b.close()
    with pytest.raises(av.HipError):
        av.Batch(hip_ctx, [good, good])
