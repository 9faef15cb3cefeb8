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


def charged_members(ctx, layout="two_lj"):
    """two_lj: two LJ fluids and two charged melts with PME (1000 and 1728
    ions, different meshes). one_lj: one LJ fluid and five charged melts, so
    the batched pair launch's member table (PairLJCut members only) is much
    shorter than the batch: its member search must stay inside that table."""
    if layout == "two_lj":
        ws = [W.lj_trimmed(900, seed=11, thermostat=1),
              W.charged_melt(10, seed=21, name="charged_melt_1000"),
              W.lj_trimmed(2000, seed=12, thermostat=1),
              W.charged_melt(12, seed=22, name="charged_melt_1728")]
    else:
        ws = [W.lj_trimmed(900, seed=11, thermostat=1)]
        ws += [W.charged_melt(side, seed=30 + k, name=f"charged_melt_{side ** 3}_{k}")
               for k, side in enumerate((10, 12, 10, 12, 10))]
    out = []
    for k, w in enumerate(ws):
        s = av.System(ctx, w.n, w.box)
        W.upload(s, w)
        W.configure(s, w)
        out.append((w, s))
    return out


@pytest.mark.parametrize("prec", [av.PREC_FP64, av.PREC_MIXED])
@pytest.mark.parametrize("graph,layout", [(False, "two_lj"), (True, "two_lj"), (False, "one_lj")])
def test_batch_with_pme_members_equals_members_alone(hip_ctx, prec, graph, layout):
    """Charged members (LJ + Coulomb + KSpacePME) advance inside the batched
    step (their force calls queued in the step; a step holding k-space
    members is not captured into a graph even when graphs are on) and equal
    their runs alone bit for bit, energies and k-space included."""
    nsteps = 60
    alone = []
    for w, s in charged_members(hip_ctx, layout):
        s.run(nsteps, prec, energy_every=10)
        alone.append(snapshot(s, nsteps))
        s.close()
    ms = charged_members(hip_ctx, layout)
    b = av.Batch(hip_ctx, [s for _, s in ms])
    b.set_graph(graph)
    b.run(nsteps // 2, prec, energy_every=10)
    b.run(nsteps - nsteps // 2, prec, energy_every=10)
    for (w, s), (d0, pe0, ke0, xi0, nr0) in zip(ms, alone):
        d1, _, _, xi1, nr1 = snapshot(s, nsteps - nsteps // 2)
        for k in d0:
            np.testing.assert_array_equal(d1[k], d0[k], err_msg=f"{w.name} {k}")
        pe1, ke1 = s.thermo(nsteps - nsteps // 2)
        np.testing.assert_array_equal(pe1, pe0[nsteps // 2:])
        np.testing.assert_array_equal(ke1, ke0[nsteps // 2:])
        assert xi1 == xi0 and nr1 == nr0
    b.close()
    for _, s in ms:
        s.close()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_continues_and_matches_one_long_run(hip_ctx, graph):
    """Batched runs of 60, 30 and 30 steps equal one of 120 for every member,
    with and without the step captured into graphs (each run captures its
    own)."""
    ms_a = members(hip_ctx, SIZES[:3])
    ms_b = members(hip_ctx, SIZES[:3])
    ba = av.Batch(hip_ctx, [s for _, s in ms_a])
    bb = av.Batch(hip_ctx, [s for _, s in ms_b])
    bb.set_graph(graph)
    ba.run(120, av.PREC_MIXED)
    bb.run(60, av.PREC_MIXED)
    bb.run(30, av.PREC_MIXED)
    bb.run(30, av.PREC_MIXED)
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


M64 = (1 << 64) - 1


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def test_replica_exchange_on_device(hip_ctx):
    """Remd::exchange_round_ (remd.cpp:105-151) on the device: decisions equal
    the Metropolis rule evaluated on the host from the members' energies with
    the same counter-based uniforms; temperatures stay a permutation of the
    ladder and every member's kinetic energy is rescaled by T_new/T_old."""
    kb = W.KB
    base = W.lj_fcc(7)
    ladder = [0.70, 0.72, 0.74, 0.76, 0.78, 0.80]
    ms = []
    for k, t in enumerate(ladder):
        s = av.System(hip_ctx, base.n, base.box)
        vx, vy, vz = W.init_velocity(base.mass, t / kb, 87287 + k)
        s.upload(base.x, base.y, base.z, base.mass, vx, vy, vz)
        s.nlist_configure(2.5, 0.3, 1, 0, True)
        s.pair_lj(1.0, 1.0, 2.5)
        s.integrator(base.dt, 1, t / kb, base.tau_T, 0)
        ms.append(s)
    b = av.Batch(hip_ctx, ms)
    seed = 4242
    total = 0
    for rnd in range(6):
        b.run(40, av.PREC_MIXED)
        slots, T = b.ladder()
        pe = np.array([s.energy()[0] for s in ms])
        ke = np.array([s.energy()[2] for s in ms])
        # host restatement of the round
        want_slots = slots.copy()
        want_T = T.copy()
        acc = 0
        for k in range(rnd & 1, len(ladder) - 1, 2):
            i, j = want_slots[k], want_slots[k + 1]
            arg = (1 / (kb * want_T[i]) - 1 / (kb * want_T[j])) * (pe[i] - pe[j])
            p = 1.0 if arg >= 0 else np.exp(arg)
            r = _splitmix64(seed ^ _splitmix64(((rnd << 32) ^ k) & M64))
            u = (r >> 11) * (1.0 / 9007199254740992.0)
            if u < p:
                acc += 1
                want_T[i], want_T[j] = want_T[j], want_T[i]
                want_slots[k], want_slots[k + 1] = j, i
        got = b.exchange(rnd, seed)
        total += got
        slots2, T2 = b.ladder()
        assert got == acc
        np.testing.assert_array_equal(slots2, want_slots)
        np.testing.assert_allclose(T2, want_T, rtol=0, atol=0)
        assert sorted(T2) == sorted(t / kb for t in ladder)
        ke2 = np.array([s.energy()[2] for s in ms])
        np.testing.assert_allclose(ke2, ke * T2 / T, rtol=1e-12)
    assert total > 0
    b.close()


def test_batched_autotune_chooses_valid_shapes(hip_ctx):
    """atomiverse_amd.tuning.autotune_batch: one record per size class, every
    member left on a valid brick shape with a built list; with the chosen
    shapes, a batched run still equals each member run alone."""
    from atomiverse_amd.tuning import autotune_batch

    ms = members(hip_ctx, SIZES)
    systems = [s for _, s in ms]
    rep = autotune_batch(hip_ctx, systems, av.PREC_MIXED, nclass=3, steps=4)
    assert 1 <= len(rep["classes"]) <= 3
    assert rep["step_ms_after"] <= rep["step_ms_before"]
    shapes = [s.brick_shape()[0] for s in systems]
    assert all(0 <= k <= 10 for k in shapes)
    for s in systems:
        s.close()
    nsteps = 60
    alone = []
    for (w, s), k in zip(members(hip_ctx, SIZES), shapes):
        s.set_brick_shape(k)
        s.run(nsteps, av.PREC_MIXED)
        alone.append(s.download())
        s.close()
    ms = members(hip_ctx, SIZES)
    for (_, s), k in zip(ms, shapes):
        s.set_brick_shape(k)
        s.nlist_build()
    b = av.Batch(hip_ctx, [s for _, s in ms])
    b.set_graph(True)
    b.run(nsteps, av.PREC_MIXED)
    for (w, s), d0 in zip(ms, alone):
        d1 = s.download()
        for k in d0:
            np.testing.assert_array_equal(d1[k], d0[k], err_msg=f"{w.name} {k}")
        s.close()
    b.close()
