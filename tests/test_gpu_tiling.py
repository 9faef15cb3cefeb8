"""Brick tiling and its autotuner: the neighbor list is the reference's pair
set whatever the brick shape, forces agree to rounding, and the autotuner
keeps the frame's forces and a valid shape."""
import numpy as np
import pytest

import atomiverse_amd as av
from atomiverse_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_every_brick_shape_gives_the_reference_pair_set(hip_ctx, orc):
    w = W.lj_trimmed(9000, seed=5)
    s = av.System(hip_ctx, w.n, w.box)
    W.upload(s, w)
    W.configure(s, w)
    s.nlist_build()
    s.run(50, av.PREC_FP64)  # move off the lattice
    d = s.download()
    want = orc.pair_keys(orc.half_pairs(*orc.build_half(d["x"], d["y"], d["z"], w.box, 2.8)), w.n)
    ref = None
    seen = set()
    for k in range(12):
        s.set_brick_shape(k)
        s.upload(d["x"], d["y"], d["z"], w.mass)
        s.nlist_build()
        kk, bdim = s.brick_shape()
        if kk in seen:
            continue
        seen.add(kk)
        p = s.export_pairs()
        np.testing.assert_array_equal(np.sort(p[:, 0] * w.n + p[:, 1]), want, err_msg=str(bdim))
        s.compute(av.PREC_FP64, True, True)
        F = np.stack([s.download()[c] for c in ("fx", "fy", "fz")])
        if ref is None:
            ref = F
        assert np.abs(F - ref).max() <= 1e-12 * np.abs(ref).max()
    assert len(seen) >= 3


def test_autotune_keeps_forces_and_list(hip_ctx, orc):
    w = W.lj_trimmed(12000, seed=6)
    s = av.System(hip_ctx, w.n, w.box)
    W.upload(s, w)
    W.configure(s, w)
    s.nlist_build()
    s.run(20, av.PREC_MIXED)
    before = s.download()
    k, us = s.autotune(av.PREC_MIXED, trials=3)
    assert 0 <= k <= 11 and us > 0
    after = s.download()
    for c in before:
        np.testing.assert_array_equal(after[c], before[c])
    want = orc.pair_keys(orc.half_pairs(*orc.build_half(after["x"], after["y"], after["z"], w.box, 2.8)), w.n)
    p = s.export_pairs()
    np.testing.assert_array_equal(np.sort(p[:, 0] * w.n + p[:, 1]), want)
    s.run(20, av.PREC_MIXED)  # the tuned layout steps on


def test_dense_input_steps_down_and_runs(hip_ctx, orc):
    """A trimmed FCC lattice packs its atoms at ~1.46x the nominal density in
    part of the box: the largest brick shapes, chosen for uniform density,
    would leave almost no halo or brick headroom there. The list build reads
    back the fullest halo and brick and steps down; the hot ladder run that
    follows keeps every rebuild within capacity, and the list is still the
    reference's pair set."""
    w = W.lj_trimmed(1406, T_star=1.5, seed=9)
    s = av.System(hip_ctx, w.n, w.box)
    W.upload(s, w)
    W.configure(s, w)
    s.set_brick_shape(0)
    s.nlist_build()
    k, bdim = s.brick_shape()
    assert k > 0, bdim  # the 4x5x5 shape does not keep its headroom here
    s.run(200, av.PREC_MIXED)  # raises if any rebuild overflowed a halo or brick
    d = s.download()
    s.nlist_build()
    want = orc.pair_keys(orc.half_pairs(*orc.build_half(d["x"], d["y"], d["z"], w.box, 2.8)), w.n)
    p = s.export_pairs()
    np.testing.assert_array_equal(np.sort(p[:, 0] * w.n + p[:, 1]), want)


def test_long_cutoff_list_paths(hip_ctx, orc):
    """A cutoff of 4.0 + 0.4 at liquid density: more than 256 neighbours per
    atom (list rows stored directly instead of staged per atom) and more than
    768 candidates per cell (the build walks the stencil columns one at a
    time): the pair set is still the reference's and the FP64 forces its
    forces."""
    w = W.lj_trimmed(6000, seed=13)
    s = av.System(hip_ctx, w.n, w.box)
    W.upload(s, w)
    s.nlist_configure(4.0, 0.4, 1, 0, True)
    s.pair_lj(1.0, 1.0, 4.0)
    s.integrator(w.dt, 0, 0.0, 0.0, 0)
    s.nlist_build()
    s.run(30, av.PREC_FP64)  # off the lattice
    d = s.download()
    s.nlist_build()
    st = s.nlist_stats()
    assert st["capacity"] > 256 and st["max_neighbors"] > 256, st
    half = orc.build_half(d["x"], d["y"], d["z"], w.box, 4.4, nthreads=8)
    want = orc.pair_keys(orc.half_pairs(*half), w.n)
    p = s.export_pairs()
    np.testing.assert_array_equal(np.sort(p[:, 0] * w.n + p[:, 1]), want)
    s.compute(av.PREC_FP64, True, True)
    g = s.download()
    fx, fy, fz, e, _ = orc.pair_lj(d["x"], d["y"], d["z"], w.box, half, 1.0, 1.0, 4.0, nthreads=8)
    F = np.stack([fx, fy, fz])
    G = np.stack([g["fx"], g["fy"], g["fz"]])
    assert np.max(np.abs(G - F)) <= 1e-12 * np.max(np.abs(F))
