"""Neighbor-capacity overflow inside a run: the system halts at the step whose
rebuild overflowed, the list is regrown at those positions and the run
resumes, so the trajectory equals an uninterrupted run with a larger
capacity (the reference regrows inside the rebuild,
src/cuda/neighbor/neighbor_list.cu:544-553).

The overflow is provoked from the host: the capacity is set to the largest
row of the perfect lattice, and a hot start (T* = 1.5) melts the lattice, so
a later rebuild needs more entries. Every kernel stays within its bounds (a
row beyond the capacity is counted, not stored).

A brick that outgrows the home atoms its list build takes is the same kind
of overflow: the system halts, the brick shape steps down and the run
resumes. After a warm-up, the limit is lowered from the host to the
fluid's fullest brick, so a later rebuild crosses it.
"""
import numpy as np
import pytest

import atomiverse_amd as av
from atomiverse_amd import workloads as W

STEPS = 300


def _hot():
    return W.lj_fcc(10, T_star=1.5, seed=4242, thermostat=1, name="lj_fcc_4000_hot")


def _system(ctx, w, tight):
    s = av.System(ctx, w.n, w.box)
    W.upload(s, w)
    W.configure(s, w)
    s.nlist_build()
    first = s.nlist_stats()
    if tight:
        s.set_capacity(first["max_neighbors"])
    return s, first


@pytest.mark.gpu
def test_run_regrows_list_and_matches_uninterrupted(hip_ctx):
    w = _hot()
    ref, _ = _system(hip_ctx, w, tight=False)
    ref.run(STEPS, av.PREC_FP64, energy_every=50)
    want = ref.download()
    want_stats = ref.nlist_stats()
    pe_ref, ke_ref = ref.thermo(STEPS)

    t, first = _system(hip_ctx, w, tight=True)
    print(f"capacity set to {first['max_neighbors']} (lattice rows)", flush=True)
    t.run(STEPS, av.PREC_FP64, energy_every=50)
    got = t.download()
    st = t.nlist_stats()
    pe, ke = t.thermo(STEPS)
    assert st["capacity"] > first["max_neighbors"], "no rebuild overflowed: the test provoked nothing"
    assert st["max_neighbors"] > first["max_neighbors"]
    assert st["nrebuilds"] == want_stats["nrebuilds"]
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(np.isnan(pe), np.isnan(pe_ref))
    m = ~np.isnan(pe)
    assert np.array_equal(pe[m], pe_ref[m])
    assert np.array_equal(ke, ke_ref)
    ref.close()
    t.close()


@pytest.mark.gpu
def test_batch_member_regrows_and_matches_alone(hip_ctx):
    ws = [_hot(), W.lj_fcc(9, T_star=0.9, seed=77, thermostat=1, name="lj_fcc_2916")]
    alone = []
    for w in ws:
        s, _ = _system(hip_ctx, w, tight=False)
        s.run(STEPS, av.PREC_MIXED, energy_every=0)
        alone.append(s.download())
        s.close()
    members = []
    tight0 = None
    for k, w in enumerate(ws):
        s, first = _system(hip_ctx, w, tight=(k == 0))
        if k == 0:
            tight0 = first["max_neighbors"]
        members.append(s)
    b = av.Batch(hip_ctx, members)
    b.set_graph(True)
    b.run(STEPS, av.PREC_MIXED, energy_every=0)
    assert members[0].nlist_stats()["capacity"] > tight0
    for s, want in zip(members, alone):
        got = s.download()
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            assert np.array_equal(got[k], want[k]), k
    b.close()
    for s in members:
        s.close()


@pytest.mark.gpu
def test_phase_walk_decide_regrows(hip_ctx):
    """atvh_nlist_decide_and_rebuild reads the overflow probe and regrows at once."""
    w = _hot()
    s, first = _system(hip_ctx, w, tight=True)
    s.nlist_build()
    grew = False
    for _ in range(STEPS):
        s.step_phase(av.PHASE_INITIAL)
        s.nlist_decide_and_rebuild()
        s.compute(av.PREC_FP64, False, False)
        s.step_phase(av.PHASE_FINAL)
        if s.nlist_stats()["capacity"] > first["max_neighbors"]:
            grew = True
            break
    assert grew
    got = s.download()
    s.nlist_build()
    p = s.export_pairs()
    from oracle import oracle as O
    nn, fst, nb = O.build_half(got["x"], got["y"], got["z"], w.box, w.cutoff + w.skin)
    want = O.pair_keys(O.half_pairs(nn, fst, nb), w.n)
    assert np.array_equal(np.sort(p[:, 0] * w.n + p[:, 1]), want)
    s.close()



WARM = 100
CHUNK = 100
MAX_CHUNKS = 20


def _brick_limited(ctx, w, prec):
    """A melted fluid whose list build takes no more home atoms per brick than
    its fullest brick holds now. (The lattice's fullest bricks, with lattice
    planes on their faces, are an artefact the melt removes.) A later rebuild
    that finds one atom more in some brick overflows; on the oracle's
    trajectory of this input that happens 7 to 282 steps after the warm-up,
    depending on its length."""
    s, _ = _system(ctx, w, tight=False)
    s.run(WARM, prec, energy_every=0)
    s.nlist_build()
    k0, _ = s.brick_shape()
    _, ma = s.brick_fill()
    s.set_brick_limit(ma)
    return s, k0, ma


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [av.PREC_FP64, av.PREC_MIXED])
def test_brick_overflow_steps_shape_down(hip_ctx, orc, prec):
    """A rebuild that finds a brick over the build's home-atom limit halts the
    system at that step; the run steps the brick shape down, rebuilds at the
    halted positions and resumes. No pair kernel reads the skipped brick's
    list rows (they are a previous build's: the mixed kernel's exact-cutoff
    walk indexes positions through them). The run ends on the reference's
    pair set, and in a batch (graph on) the member equals its run alone bit
    for bit."""
    w = _hot()
    s, k0, ma = _brick_limited(hip_ctx, w, prec)
    print(f"brick shape {k0}, fullest brick {ma} atoms = limit, precision {prec}", flush=True)
    nchunks = 0
    while s.brick_shape()[0] == k0 and nchunks < MAX_CHUNKS:
        s.run(CHUNK, prec, energy_every=50)
        nchunks += 1
    k1, _ = s.brick_shape()
    print(f"shape {k0} -> {k1} within {nchunks * CHUNK} steps", flush=True)
    assert k1 > k0, "no rebuild overflowed a brick: the test provoked nothing"
    want = s.download()
    pe0, ke0 = s.thermo(CHUNK)
    assert np.all(np.isfinite(ke0))
    s.nlist_build()
    half = orc.build_half(want["x"], want["y"], want["z"], w.box, w.cutoff + w.skin)
    p = s.export_pairs()
    np.testing.assert_array_equal(np.sort(p[:, 0] * w.n + p[:, 1]), orc.pair_keys(orc.half_pairs(*half), w.n))
    s.close()

    other = W.lj_fcc(9, T_star=0.9, seed=77, thermostat=1, name="lj_fcc_2916")
    o, _ = _system(hip_ctx, other, tight=False)
    o.run(WARM, prec, energy_every=0)
    for _ in range(nchunks):
        o.run(CHUNK, prec, energy_every=50)
    want_o = o.download()
    o.close()
    m0, _, _ = _brick_limited(hip_ctx, w, prec)
    m1, _ = _system(hip_ctx, other, tight=False)
    m1.run(WARM, prec, energy_every=0)
    b = av.Batch(hip_ctx, [m0, m1])
    b.set_graph(True)
    for _ in range(nchunks):
        b.run(CHUNK, prec, energy_every=50)
    assert m0.brick_shape()[0] == k1
    for m, ref in ((m0, want), (m1, want_o)):
        got = m.download()
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            assert np.array_equal(got[k], ref[k]), k
    pe1, ke1 = m0.thermo(CHUNK)
    assert np.array_equal(ke1, ke0)
    b.close()
    m0.close()
    m1.close()
