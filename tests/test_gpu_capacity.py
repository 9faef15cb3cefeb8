"""Neighbor-capacity overflow inside a run: the system halts at the step whose
rebuild overflowed, the list is regrown at those positions and the run
resumes, so the trajectory equals an uninterrupted run with a larger
capacity (the reference regrows inside the rebuild,
src/cuda/neighbor/neighbor_list.cu:544-553).

The overflow is provoked from the host: the capacity is set to the largest
row of the perfect lattice, and a hot start (T* = 1.5) melts the lattice, so
a later rebuild needs more entries. Every kernel stays within its bounds (a
row beyond the capacity is counted, not stored).
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
