#!/usr/bin/env python3
"""Benchmark: atom-steps/s of the 2,048,000-atom LJ fluid (80^3 FCC cells,
rho*=0.8442, T*=0.72, rc=2.5, skin 0.3, Nose-Hoover tau*=0.5) on one MI355X,
mixed precision (headline `value`) beside the plain FP64 path, with the force
RMS error of mixed against FP64, the pair kernel's roofline fraction and the
CPU oracle timed on the host cores.

A step is one full MD step of the fused loop: initial kick/drift/wrap + NH,
device-side neighbor decide (+ rebuild when due), LJ forces, final kick + NH.
Inputs are resident in HBM before the timed region starts.

Usage: python bench.py [--gpus 1] [--steps 500] [--warmup 20]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# HBM bytes per launch of the pair kernel at this workload, from rocprofv3
# FETCH_SIZE / WRITE_SIZE passes scaled by a plain copy's counters (the file
# records the commands and numbers); absent -> traffic null.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_k_lj_brick_mixed_lj_fcc_2048000.txt")
VALU_FILE = os.path.join(ROOT, "profiles", "pmc_sq_k_lj_brick_mixed_lj_fcc_2048000.txt")
CU_COUNT, CLOCK_HZ = 256, 2.4e9  # MI355X: one wave64 VALU instruction per clock per CU


def measured_traffic():
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def valu_issue(avg_s):
    """The pair kernel's VALU issue time per launch against its measured duration:
    the bound that the HBM fraction leaves unexplained."""
    insts = measured_valu_insts()
    if insts is None:
        return None
    issue_s = insts / (CU_COUNT * CLOCK_HZ)
    return {"insts_per_launch": insts, "issue_us": issue_s * 1e6, "frac": issue_s / avg_s,
            "source": os.path.relpath(VALU_FILE, ROOT)}


def measured_valu_insts():
    """VALU wave-instructions per launch of the pair kernel, from the committed
    SQ_INSTS_VALU summary (None if absent)."""
    try:
        for line in open(VALU_FILE):
            if line.startswith("valu_insts_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_breakdown(ctx, s, precision, steps, names):
    """Average time per launch of each kernel family, from a separate pass with
    HIP events around every family (the events add markers to the stream, so
    the throughput of the timed region is measured with the pair kernel's
    events only)."""
    ctx.enable_timing(True)
    ctx.reset_timing()
    s.run(steps, precision, energy_every=0)
    ctx.synchronize()
    ctx.enable_timing(False)
    return {k: ctx.kernel_time(k) for k in names}


def run_gpu(av, W, ctx, precision, steps, warmup):
    w = W.lj_fcc_2048000(thermostat=1)
    s = av.System(ctx, w.n, w.box)
    W.upload(s, w)
    W.configure(s, w)
    s.nlist_build()
    s.run(warmup, precision, energy_every=0)
    ctx.synchronize()
    st0 = s.nlist_stats()
    ctx.enable_timing(True, only="pair")
    ctx.reset_timing()
    ctx.synchronize()
    t0 = time.perf_counter()
    s.run(steps, precision, energy_every=0)
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.enable_timing(False)
    st1 = s.nlist_stats()
    pair = ctx.kernel_time("pair")
    kt = stage_breakdown(ctx, s, precision, min(steps, 100), ("pair", "integrate", "nlist"))
    return w, s, t1 - t0, pair, kt, st0, st1


def run_charged(av, W, ctx, precision, steps, warmup):
    """The 512,000-ion charged melt (LJ + real-space Coulomb + PME 128^3, order 4, NVT)."""
    w = W.charged_melt_512000()
    s = av.System(ctx, w.n, w.box)
    W.upload(s, w)
    W.configure(s, w)
    s.nlist_build()
    s.run(warmup, precision, energy_every=0)
    ctx.synchronize()
    t0 = time.perf_counter()
    s.run(steps, precision, energy_every=0)
    ctx.synchronize()
    t1 = time.perf_counter()
    kt = stage_breakdown(ctx, s, precision, min(steps, 50), ("pair", "pme", "integrate", "nlist"))
    return w, s, t1 - t0, kt


def charged_section(av, W, ctx, steps, warmup, threads):
    out = {}
    sys_keep = {}
    for name, prec in (("fp64", av.PREC_FP64), ("mixed", av.PREC_MIXED)):
        log(f"bench: charged melt {name}, {warmup} warmup + {steps} timed steps")
        w, s, dt, kt = run_charged(av, W, ctx, prec, steps, warmup)
        out[f"{name}_atom_steps_per_s"] = w.n * steps / dt
        out[f"{name}_ms_per_step"] = dt * 1e3 / steps
        out[f"{name}_kernel_avg_us"] = {k: v[0] * 1e3 / max(v[1], 1) for k, v in kt.items()}
        log(f"  {name}: {dt * 1e3 / steps:.3f} ms/step, " + ", ".join(
            f"{k} {v[0] / max(v[1], 1) * 1e3:.1f}us x{v[1]}" for k, v in kt.items()))
        sys_keep[name] = s
    s = sys_keep["mixed"]
    s.compute(av.PREC_FP64, False, False)
    F = s.download()
    s.compute(av.PREC_MIXED, False, False)
    M = s.download()
    F = np.stack([F["fx"], F["fy"], F["fz"]])
    M = np.stack([M["fx"], M["fy"], M["fz"]])
    out["force_rms_rel_err_mixed_vs_fp64"] = float(np.sqrt(np.sum((M - F) ** 2) / np.sum(F ** 2)))
    from oracle import oracle as O
    log("bench: charged melt mixed forces against the oracle")
    out["force_rms_rel_err_mixed_vs_oracle"] = oracle_force_rms(O, s, w, threads)
    out["mixed_over_fp64"] = out["mixed_atom_steps_per_s"] / out["fp64_atom_steps_per_s"]
    out["workload"] = "charged_melt_512000_nvt (pme 128^3 order 4, rc 3.0, alpha 1.025646)"
    out["steps"] = steps
    for s in sys_keep.values():
        s.close()
    return out


def nve_section(av, W, ctx, lj_steps, melt_steps, every=100):
    """NVE side runs from the 2.048M-atom LJ and 512k-ion melt starting states (thermostat off), both
    precisions: total energy per atom recorded every `every` steps, its linear
    drift in eV/atom/ns, and the largest deviation from the first record."""
    out = {}
    for label, make, steps in (("lj_fcc_2048000", lambda: W.lj_fcc_2048000(thermostat=0), lj_steps),
                               ("charged_melt_512000", lambda: W.charged_melt(80, thermostat=0), melt_steps)):
        if steps <= 0:
            continue
        out[label] = {"steps": steps, "record_every": every}
        for name, prec in (("fp64", av.PREC_FP64), ("mixed", av.PREC_MIXED)):
            w = make()
            log(f"bench: NVE side run {label} {name}, {steps} steps")
            s = av.System(ctx, w.n, w.box)
            W.upload(s, w)
            W.configure(s, w)
            s.nlist_build()
            s.run(steps, prec, energy_every=every)
            pe, ke = s.thermo(steps)
            s.close()
            e = (pe + ke) / w.n
            k = np.nonzero(np.isfinite(e))[0]
            t_fs = (k + 1) * w.dt
            slope = float(np.polyfit(t_fs, e[k], 1)[0])  # eV/atom/fs
            out[label][name] = {"drift_ev_per_atom_per_ns": slope * 1e6,
                                "max_abs_dev_ev_per_atom": float(np.max(np.abs(e[k] - e[k[0]]))),
                                "span_fs": float(t_fs[-1] - t_fs[0])}
            log(f"  drift {slope * 1e6:.3e} eV/atom/ns, max |E-E0| {out[label][name]['max_abs_dev_ev_per_atom']:.3e}"
                " eV/atom")
    return out


def ladder_section(av, W, ctx, steps, loop_steps, eager=False):
    """Batched REMD-style ensemble: 256 LJ fluids (N log-uniform in [500, 20000],
    T* ladder 0.7..1.5, NVT, mixed), one batched step per launch stage (hipGraph
    replay) against a loop of plain per-system launches over the same members.
    Brick shapes come from the batched autotuner (atomiverse_amd.tuning), per
    size class; the loop uses the same shapes. eager: the batched steps run
    without graph replay on the library's default shapes (for kernel traces)."""
    from atomiverse_amd.tuning import autotune_batch
    ws = W.lj_ladder(256)
    total = sum(w.n for w in ws)
    shapes = []

    def make():
        out = []
        for w in ws:
            s = av.System(ctx, w.n, w.box)
            W.upload(s, w)
            W.configure(s, w)
            s.nlist_build()
            out.append(s)
        if shapes:
            for s, (k, _) in zip(out, shapes):
                s.set_brick_shape(k)
                s.nlist_build()
        return out

    log(f"bench: ladder batched ({len(ws)} systems, {total} atoms), {steps} steps"
        + (" without graph replay, default brick shapes" if eager else ""))
    ms = make()
    if eager:
        tuned = None
    else:
        tuned = autotune_batch(ctx, ms, av.PREC_MIXED)
        shapes.extend(s.brick_shape() for s in ms)
        log(f"  autotuned brick shapes: step {tuned['step_ms_before']:.3f} -> {tuned['step_ms_after']:.3f} ms; "
            + ", ".join(f"{c['class_atoms']}->{c['brick']}" for c in tuned["classes"]))
    b = av.Batch(ctx, ms)
    b.set_graph(not eager)
    b.run(20, av.PREC_MIXED)  # warm (graph capture, list growth)
    ctx.synchronize()
    t0 = time.perf_counter()
    b.run(steps, av.PREC_MIXED)
    ctx.synchronize()
    dt_b = time.perf_counter() - t0
    # per-family breakdown: a separate eager pass (timed steps are not
    # captured) with HIP events around each launch family of the step
    nb = 40
    ctx.enable_timing(True)
    ctx.reset_timing()
    b.run(nb, av.PREC_MIXED)
    ctx.synchronize()
    ctx.enable_timing(False)
    stages = {k: 1e3 * ctx.kernel_time(k)[0] / nb
              for k in ("batch_step", "batch_integrate", "batch_nlist", "batch_pair")}
    b.close()
    for s in ms:
        s.close()
    log(f"  batched: {dt_b * 1e3 / steps:.3f} ms/step, {total * steps / dt_b:.4e} atom-steps/s")
    log("  eager pass with events, us per step: " + ", ".join(f"{k} {v:.1f}" for k, v in stages.items()))
    log(f"bench: ladder per-system loop, {loop_steps} steps")
    ms = make()
    for s in ms:
        s.run(20, av.PREC_MIXED)
    ctx.synchronize()
    t0 = time.perf_counter()
    for s in ms:
        s.run(loop_steps, av.PREC_MIXED)
    ctx.synchronize()
    dt_l = time.perf_counter() - t0
    for s in ms:
        s.close()
    log(f"  loop: {dt_l * 1e3 / loop_steps:.3f} ms/step, {total * loop_steps / dt_l:.4e} atom-steps/s")
    return {"workload": "lj_ladder_256 (N log-uniform 500..20000, T* 0.7..1.5, NVT, mixed)",
            "atoms": total, "systems": len(ws), "graph_replay": not eager, "autotuned_brick_shapes": tuned,
            "batched_atom_steps_per_s": total * steps / dt_b, "batched_steps": steps,
            "eager_us_per_step_with_events": stages,
            "loop_atom_steps_per_s": total * loop_steps / dt_l, "loop_steps": loop_steps,
            "batched_over_loop": (total * steps / dt_b) / (total * loop_steps / dt_l)}


def serving_requests(W, count, seed=4242):
    """Request set: 70% LJ / 30% charged melt with PME (grid ~1.2 points per
    sigma, order 4), N log-uniform in [1000, 50000], steps uniform in [1000, 10000]."""
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(50000))))
        steps = int(rng.integers(1000, 10001))
        if rng.uniform() < 0.3:
            side = max(10, 2 * int(round(n ** (1.0 / 3.0) / 2)))
            w = W.charged_melt(side, pme_spacing=1.0 / 1.2, seed=9000 + k, name=f"req{k}_charged_{side ** 3}")
        else:
            w = W.lj_trimmed(n, seed=7000 + k, name=f"req{k}_lj_{n}")
        reqs.append((w, steps))
    return reqs


def serving_section(av, W, ctx, count, chunk):
    """Offline batched run of the request set (all submitted at once) against
    the same set arriving as a seeded Poisson stream at the offline service
    rate; throughput in atom-steps/s and completion latency percentiles."""
    reqs = serving_requests(W, count)
    work = sum(w.n * st for w, st in reqs)
    budget = 64 << 30
    log(f"bench: serving, {count} requests, {work:.3e} atom-steps, chunk {chunk}")

    srv = av.Server(ctx, budget, chunk, av.PREC_MIXED)
    t0 = time.perf_counter()
    ids = [srv.submit(w, st) for w, st in reqs]
    left = set(ids)
    while left:
        srv.advance()
        for sid in srv.poll():
            srv.fetch(sid)
            left.discard(sid)
    t_off = time.perf_counter() - t0
    srv.close()
    thr_off = work / t_off
    log(f"  offline: {t_off:.2f} s, {thr_off:.4e} atom-steps/s")

    lam = count / t_off  # arrivals per second = offline service rate
    rng = np.random.default_rng(4243)
    arrive = np.cumsum(rng.exponential(1.0 / lam, count))
    srv = av.Server(ctx, budget, chunk, av.PREC_MIXED)
    start = time.perf_counter()
    k = 0
    arrival_of = {}
    latency = []
    while len(latency) < count:
        now = time.perf_counter() - start
        while k < count and arrive[k] <= now:
            sid = srv.submit(*reqs[k])
            arrival_of[sid] = arrive[k]
            k += 1
        active, queued = srv.advance()
        t = time.perf_counter() - start
        for sid in srv.poll():
            srv.fetch(sid)
            latency.append(t - arrival_of[sid])
        if active == 0 and queued == 0 and k < count:
            time.sleep(max(0.0, arrive[k] - (time.perf_counter() - start)))
    span = (time.perf_counter() - start) - arrive[0]
    srv.close()
    thr_srv = work / span
    lat = np.array(latency)
    log(f"  served: {span:.2f} s, {thr_srv:.4e} atom-steps/s, p50 {np.percentile(lat, 50):.3f} s, "
        f"p99 {np.percentile(lat, 99):.3f} s")
    return {"workload": f"{count} segments (70% LJ / 30% charged+PME, N 1e3..5e4, 1e3..1e4 steps, mixed)",
            "atom_steps": work, "chunk_steps": chunk, "hbm_budget_bytes": budget,
            "offline_atom_steps_per_s": thr_off, "served_atom_steps_per_s": thr_srv,
            "served_over_offline": thr_srv / thr_off, "arrival_rate_per_s": lam,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99))}


def host_cpu():
    """lscpu's model, sockets and cores, and the threads this process may use."""
    info = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k] = v
    except (OSError, ValueError):
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    info["threads_available"] = avail
    return info


def _state(w):
    n = w.n
    return dict(x=w.x.copy(), y=w.y.copy(), z=w.z.copy(), vx=w.vx.copy(), vy=w.vy.copy(), vz=w.vz.copy(),
                fx=np.zeros(n), fy=np.zeros(n), fz=np.zeros(n), mass=w.mass.copy())


def _oracle_params(O, w, threads, nh=None):
    return O.MDParams(dt=w.dt, cutoff=w.cutoff, skin=w.skin, every=w.every, delay=w.delay, check=w.check,
                      eps=w.eps, sigma=w.sigma, nh=w.thermostat if nh is None else nh, T_target=w.T,
                      tau_T=w.tau_T, ndof=0, nthreads=threads, stress=0)


def _time_oracle(O, w, steps, threads, K=None):
    """Oracle MD loop on `threads` host threads: one warm step, then `steps` timed."""
    st = _state(w)
    p = _oracle_params(O, w, threads)
    if w.pair == "ljcoul":
        tab = O.LJCoulTables(w.eps_tab, w.sigma_tab, w.set_tab, w.cutoff, w.alpha)
        run = lambda k: O.md_run_ff(st, w.box, p, tab, w.type_, w.charge, K, w.pme_order, k)  # noqa: E731
    else:
        run = lambda k: O.md_run(st, w.box, p, k)  # noqa: E731
    run(1)
    t0 = time.perf_counter()
    run(steps)
    return time.perf_counter() - t0


def cpu_baseline(W, nsteps, threads):
    """The CPU oracle (FP64 restatement of the reference CPU path) on the host
    cores: bounded samples of the headline workload on all available threads
    (`value`) and on one thread, and of the other configs: the 4000-atom NVE
    fluid, the charged melt with PME, and the 256-system ladder as the
    reference's sequential Remd loop (each replica propagated in turn)."""
    from oracle import oracle as O

    out = {}
    w = W.lj_fcc_2048000(thermostat=1)
    dt = _time_oracle(O, w, nsteps, threads)
    out.update(value=w.n * nsteps / dt, unit="atom-steps/s", cores=threads, kind="port",
               sample=f"{nsteps} NVT steps of the 2,048,000-atom LJ config (FP64, its list builds included), "
                      f"{dt:.1f} s on {threads} host threads")
    rows = {}
    serial_steps = max(2, nsteps // 16)
    dt1 = _time_oracle(O, w, serial_steps, 1)
    rows["lj_fcc_2048000_serial"] = {"atom_steps_per_s": w.n * serial_steps / dt1, "threads": 1,
                                     "steps": serial_steps, "seconds": dt1}
    w0 = W.lj_fcc_4000()
    for th in (1, threads):
        d0 = _time_oracle(O, w0, 1000, th)
        rows[f"lj_fcc_4000_nve_{th}t"] = {"atom_steps_per_s": w0.n * 1000 / d0, "threads": th, "steps": 1000,
                                          "seconds": d0}
    wc = W.charged_melt_512000()
    K = O.pme_grid(wc.box, wc.pme_spacing, wc.pme_order)
    dc = _time_oracle(O, wc, 3, threads, K=K)
    rows["charged_melt_512000"] = {"atom_steps_per_s": wc.n * 3 / dc, "threads": threads, "steps": 3,
                                   "seconds": dc}
    ladder = W.lj_ladder(256)
    t_l = 0.0
    steps_l = 5
    for m in ladder:  # Remd::propagate_ (src/cpu/path/remd.cpp:80-85): one replica after another
        t_l += _time_oracle(O, m, steps_l, threads)
    rows["lj_ladder_256_sequential"] = {"atom_steps_per_s": sum(m.n for m in ladder) * steps_l / t_l,
                                        "threads": threads, "steps": steps_l, "seconds": t_l}
    out["configs"] = rows
    out["host"] = host_cpu()
    return out


def oracle_force_rms(O, s, w, threads):
    """Mixed-precision forces of the device against the oracle's FP64 forces on
    the same (downloaded) positions: sqrt(sum |F - F_ref|^2 / sum |F_ref|^2)."""
    import atomiverse_amd as av
    s.compute(av.PREC_MIXED, False, False)
    d = s.download()
    x, y, z = d["x"], d["y"], d["z"]
    half = O.build_half(x, y, z, w.box, w.cutoff + w.skin, nthreads=threads)
    if w.pair == "ljcoul":
        tab = O.LJCoulTables(w.eps_tab, w.sigma_tab, w.set_tab, w.cutoff, w.alpha)
        fx, fy, fz, _, _ = O.pair_lj_coul_long(x, y, z, w.charge, w.type_, w.box, half, tab, nthreads=threads)
        if w.pme_order > 0:
            K = O.pme_grid(w.box, w.pme_spacing, w.pme_order)
            kx, ky, kz = O.pme(x, y, z, w.charge, w.box, w.alpha, K, w.pme_order)[:3]
            fx, fy, fz = fx + kx, fy + ky, fz + kz
    else:
        fx, fy, fz, _, _ = O.pair_lj(x, y, z, w.box, half, w.eps, w.sigma, w.cutoff)
    F = np.stack([fx, fy, fz])
    M = np.stack([d["fx"], d["fy"], d["fz"]])
    return float(np.sqrt(np.sum((M - F) ** 2) / np.sum(F ** 2)))


def _nve(av, W, ctx, make, steps, every, prec):
    w = make()
    s = av.System(ctx, w.n, w.box)
    W.upload(s, w)
    W.configure(s, w)
    s.nlist_build()
    t0 = time.perf_counter()
    s.run(steps, prec, energy_every=every)
    ctx.synchronize()
    el = time.perf_counter() - t0
    pe, ke = s.thermo(steps)
    e = (pe + ke) / w.n
    k = np.nonzero(np.isfinite(e))[0]
    t_fs = (k + 1) * w.dt
    slope = float(np.polyfit(t_fs, e[k], 1)[0])
    return s, w, el, {"drift_ev_per_atom_per_ns": slope * 1e6,
                      "max_abs_dev_ev_per_atom": float(np.max(np.abs(e[k] - e[k[0]]))),
                      "span_fs": float(t_fs[-1] - t_fs[0])}


def config_rows(av, W, ctx, threads, steps_small=1000):
    """The other configs' precision figures: NVE drift (both precisions) and the
    mixed force RMS against the oracle. The 4000-atom NVE fluid also reports its
    own throughput (1000 steps); for the ladder the smallest, median and
    largest members, for the served stream one LJ and one charged request,
    each as an NVE side run of `steps_small` steps."""
    from oracle import oracle as O
    rows = {}
    cases = [("lj_fcc_4000", lambda: W.lj_fcc_4000())]
    lad = sorted(W.lj_ladder(256), key=lambda m: m.n)
    for tag, m in (("ladder_smallest", lad[0]), ("ladder_median", lad[len(lad) // 2]), ("ladder_largest", lad[-1])):
        cases.append((f"{tag}_{m.n}", lambda m=m: W.lj_trimmed(m.n, T_star=m.meta["T_star"], seed=m.meta["seed"],
                                                               thermostat=0, name=m.name)))
    reqs = serving_requests(W, 24)
    lj = next(w for w, _ in reqs if w.pair == "lj")
    ch = next(w for w, _ in reqs if w.pair == "ljcoul")
    cases.append((f"served_{lj.name}", lambda lj=lj: W.lj_trimmed(lj.n, seed=lj.meta["seed"], thermostat=0,
                                                                   name=lj.name)))
    side = round(ch.n ** (1.0 / 3.0))
    cases.append((f"served_{ch.name}", lambda side=side, ch=ch: W.charged_melt(
        side, pme_spacing=ch.pme_spacing, seed=ch.meta["seed"], thermostat=0, name=ch.name)))
    for label, make in cases:
        row = {}
        for name, prec in (("fp64", av.PREC_FP64), ("mixed", av.PREC_MIXED)):
            s, w, el, drift = _nve(av, W, ctx, make, steps_small, 10, prec)
            row[f"nve_{name}"] = drift
            if name == "fp64" and label == "lj_fcc_4000":
                row["fp64_atom_steps_per_s"] = w.n * steps_small / el
            if name == "mixed":
                row["force_rms_rel_err_mixed_vs_oracle"] = oracle_force_rms(O, s, w, threads)
                row["atoms"] = w.n
            s.close()
        rows[label] = row
        log(f"  {label}: mixed force rms vs oracle {row['force_rms_rel_err_mixed_vs_oracle']:.3e}, "
            f"NVE drift fp64 {row['nve_fp64']['drift_ev_per_atom_per_ns']:.3e} mixed "
            f"{row['nve_mixed']['drift_ev_per_atom_per_ns']:.3e} eV/atom/ns")
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests of the serving measurement (0 skips it)")
    ap.add_argument("--ladder-steps", type=int, default=200,
                    help="timed steps of the batched 256-system ladder; 0 skips it")
    ap.add_argument("--ladder-eager", action="store_true",
                    help="run the ladder's batched steps without graph replay and with the default brick "
                         "shapes (no autotuning, which replays graphs), e.g. under a kernel tracer")
    ap.add_argument("--charged-steps", type=int, default=100,
                    help="timed steps of the charged-melt (PME) side measurement; 0 skips it")
    ap.add_argument("--nve-steps", type=int, default=10000,
                    help="NVE side run length at 2.048M atoms, both precisions (0: skip)")
    ap.add_argument("--nve-charged-steps", type=int, default=2000,
                    help="NVE side run length for the charged melt, both precisions (0: skip)")
    ap.add_argument("--config-steps", type=int, default=1000,
                    help="NVE side-run length of the C0 / ladder-member / served-request rows (0: skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device copy (the reference for byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one MI355X; multi-GPU runs are not part of it "
              "(the single-GPU throughput is the reported value).", file=sys.stderr)
        sys.exit(2)

    import atomiverse_amd as av
    from atomiverse_amd import workloads as W

    ctx = av.Context(0)
    if args.calibrate_copy:
        import ctypes as C
        ms = C.c_double()
        rc = av.load().atvh_util_stream_copy(ctx.h, 1 << 30, 3, C.byref(ms))
        if rc != 0:
            raise av.HipError(rc, "atvh_util_stream_copy")
        log(f"bench: copy calibration, 1 GiB per launch, {ms.value:.3f} ms, {2 * (1 << 30) / ms.value / 1e6:.1f} GB/s")
    res = {}
    for name, prec in (("fp64", av.PREC_FP64), ("mixed", av.PREC_MIXED)):
        log(f"bench: {name} run, {args.warmup} warmup + {args.steps} timed steps")
        w, s, dt, pair, kt, st0, st1 = run_gpu(av, W, ctx, prec, args.steps, args.warmup)
        res[name] = dict(w=w, s=s, dt=dt, pair=pair, kt=kt, st0=st0, st1=st1)
        log(f"  {name}: {dt * 1e3 / args.steps:.3f} ms/step, {w.n * args.steps / dt:.4e} atom-steps/s, "
            f"rebuilds {st1['nrebuilds'] - st0['nrebuilds']}, kernels "
            + ", ".join(f"{k} {v[0] / max(v[1], 1) * 1e3:.1f}us x{v[1]}" for k, v in kt.items()))

    # force RMS error, mixed vs FP64, on the mixed run's final positions
    s = res["mixed"]["s"]
    s.compute(av.PREC_FP64, False, False)
    f64 = s.download()
    s.compute(av.PREC_MIXED, False, False)
    f32 = s.download()
    F = np.stack([f64["fx"], f64["fy"], f64["fz"]])
    M = np.stack([f32["fx"], f32["fy"], f32["fz"]])
    rms = float(np.sqrt(np.sum((M - F) ** 2) / np.sum(F ** 2)))
    from oracle import oracle as O
    threads = max(1, min(host_cpu()["threads_available"], 16))  # the host share of one GPU
    log("bench: mixed forces against the oracle at the headline size")
    rms_orc = oracle_force_rms(O, s, res["mixed"]["w"], threads)

    w = res["mixed"]["w"]
    n = w.n
    steps = args.steps
    value = n * steps / res["mixed"]["dt"]
    fp64_value = n * steps / res["fp64"]["dt"]
    # roofline of the dominant kernel (pair): algorithmic bytes per launch =
    # 24N positions + 8N list offsets/counts + 4 B per half-list entry + 24N forces
    lh = res["mixed"]["st1"]["npairs_half"]
    bytes_alg = 56 * n + 4 * lh
    ms, launches = res["mixed"]["pair"]  # events on the pair kernel's stream, inside the timed region
    avg_s = ms / 1e3 / max(launches, 1)
    achieved = bytes_alg / avg_s / 1e9
    out = {
        "metric": "atom-steps/s, 2.048M-atom LJ, mixed vs FP64 at stated err; p50/p99 latency",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": res["mixed"]["dt"] * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair forces, fp64 integration/accumulation",
        "data": "synthetic (seeded FCC lattice + Maxwell-Boltzmann velocities)",
        "config": {"workload": "lj_fcc_2048000_nvt", "atoms": n, "precision": "mixed", "cutoff": 2.5,
                   "skin": 0.3, "thermostat": "nose-hoover tau*=0.5", "dt_fs": w.dt,
                   "rebuilds_in_timed_steps": res["mixed"]["st1"]["nrebuilds"] - res["mixed"]["st0"]["nrebuilds"],
                   "brick_cells": list(res["mixed"]["s"].brick_shape()[1])},
        "fp64_value": fp64_value,
        "fp64_ms_per_step": res["fp64"]["dt"] * 1e3 / steps,
        "mixed_over_fp64": value / fp64_value,
        "force_rms_rel_err_mixed_vs_fp64": rms,
        "force_rms_rel_err_mixed_vs_oracle": rms_orc,
        "roofline": {"bound": "hbm", "kernel": "k_single<OpLjBrickMixed<false, false, false>>", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "bytes_alg_per_launch": bytes_alg, "avg_launch_us": avg_s * 1e6,
                     "valu_issue": valu_issue(avg_s)},
        "kernel_avg_us": {name: {k: v[0] * 1e3 / max(v[1], 1) for k, v in r["kt"].items()}
                          for name, r in res.items()},
    }
    for r in res.values():
        r["s"].close()
    if args.charged_steps > 0:
        out["charged_melt"] = charged_section(av, W, ctx, args.charged_steps, 10, threads)
    if args.nve_steps > 0 or args.nve_charged_steps > 0:
        out["nve_drift"] = nve_section(av, W, ctx, args.nve_steps, args.nve_charged_steps)
    if args.ladder_steps > 0:
        out["batched_ladder"] = ladder_section(av, W, ctx, args.ladder_steps, max(20, args.ladder_steps // 5),
                                               eager=args.ladder_eager)
    if args.serve_requests > 0:
        out["serving"] = serving_section(av, W, ctx, args.serve_requests, 250)
    if args.config_steps > 0:
        log("bench: per-config NVE drift and mixed force error against the oracle")
        out["config_rows"] = config_rows(av, W, ctx, threads, args.config_steps)
    if not args.no_cpu:
        log(f"bench: cpu baseline ({args.cpu_steps} oracle steps on {threads} threads, then the serial "
            "and per-config samples)")
        out["cpu_baseline"] = cpu_baseline(W, args.cpu_steps, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
