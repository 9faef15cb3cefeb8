"""Brick-shape autotuning for a batched ensemble.

atvh_nlist_autotune times brick shapes on one system alone. A small system
alone leaves most of the GPU idle, so it favours small bricks (more
workgroups), while in a batch of hundreds of members the GPU is full and the
halo overhead of small bricks dominates. autotune_batch therefore times the
whole batched step: members are split into log-size classes, and for each
class in turn every candidate shape is given to all members of the class and
the batch is timed over a few graph-replayed steps; the fastest shape (the
members' own default included) is kept.
"""
import time

import numpy as np

from . import _abi
from .system import Batch


def size_classes(sizes, nclass):
    """Class index (0..nclass-1) of every size: equal-count classes in log N."""
    logs = np.log(np.asarray(sizes, dtype=np.float64))
    edges = np.quantile(logs, np.linspace(0.0, 1.0, nclass + 1))
    return np.clip(np.searchsorted(edges, logs, side="right") - 1, 0, nclass - 1)


def _timed_step(ctx, systems, precision, steps):
    for s in systems:
        s.nlist_build()  # lists for the shapes now set
    b = Batch(ctx, systems)
    try:
        b.set_graph(True)
        b.run(steps, precision)  # capture and warm
        ctx.synchronize()
        t0 = time.perf_counter()
        b.run(steps, precision)
        ctx.synchronize()
        return (time.perf_counter() - t0) / steps
    finally:
        b.close()


def autotune_batch(ctx, systems, precision=_abi.PREC_MIXED, nclass=8, steps=10, span=(-1, 1, 2)):
    """Choose brick shapes for `systems` (brick-mode members of one batch) by
    timing the batched step. Candidate shapes for a class are the class's
    smallest default candidate index plus each offset in `span` (index 0 is
    the largest brick; a shape whose halo would not fit is skipped by the
    library). Members are advanced by the timing runs. Returns one record per
    class and the step time before and after."""
    sizes = np.array([s.n for s in systems])
    cls = size_classes(sizes, nclass)
    default = [s.brick_shape()[0] for s in systems]
    best_t = _timed_step(ctx, systems, precision, steps)
    t_start = best_t
    report = []
    for c in range(nclass):
        idx = np.nonzero(cls == c)[0]
        if idx.size == 0:
            continue
        k0 = min(default[i] for i in idx)
        chosen = None  # None: the members' own defaults
        for off in span:
            k = k0 + off
            if k < 0 or k > 11:
                continue
            for i in idx:
                systems[i].set_brick_shape(k)
            t = _timed_step(ctx, systems, precision, steps)
            if t < best_t:
                best_t, chosen = t, k
        for i in idx:
            systems[i].set_brick_shape(default[i] if chosen is None else chosen)
        report.append({"class_atoms": [int(sizes[idx].min()), int(sizes[idx].max())], "members": int(idx.size),
                       "candidate": "default" if chosen is None else int(chosen),
                       "brick": list(systems[int(idx[0])].brick_shape()[1])})
    for s in systems:
        s.nlist_build()
    return {"classes": report, "step_ms_before": t_start * 1e3, "step_ms_after": best_t * 1e3}
