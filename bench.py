"""Benchmark: spin-steps/s at 1M spins on one MI355X.

Workload (the at-scale configuration): 1e6 spins, 2000 z-aligned axons with
radii ~ Gamma(k=3, mean 1 um) packed to an intra-axonal fraction of 0.7
(voxel 107.9 um, periodic), 64-gon open prisms (286k triangles with the
periodic images), permeability 0.01 um/ms, T2 70/90 ms, dt 5 us, PGSE
delta/Delta/TE = 10/40/60 ms (12,000 steps), 64 directions on b = 1/2/3
ms/um^2 shells plus b=0, fp32.  One bench step is one complete acquisition:
every spin walks 12,000 steps once (the gradient-moment accumulation serves
all 65 gradient lines) and the 65 signals are read out in FP64.

Beside the headline line the JSON carries: the batched configuration (256
crossing-bundle substrates advanced together by the server's batched
launches), the served configuration (capacity, latency against solo runs,
and the 50 req/s stream over a fixed window), the waveform-phase path (4e6
spins x 1024 waveforms on the matrix cores), the reference-semantics path,
and CPU baselines (the oracle on this host's cores).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config scale|small]
                       [--cell-um H] [--sort S] [--no-autotune] [--no-cpu-baseline]
                       [--batched-substrates N] [--served-requests N] [--waveform-spins N]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from dsim_amd import autotune as AT  # noqa: E402
from dsim_amd import lib as D  # noqa: E402
from dsim_amd import substrate as S  # noqa: E402

PEAK_FP32 = 157.3e12    # MI355X FP32 vector, FLOP/s
GAMMA = 42576.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def schedule(cfg):
    L = int(cfg.delta / cfg.dt + 0.5)
    B = int(cfg.Delta / cfg.dt + 0.5)
    R = int((cfg.TE - cfg.Delta - cfg.delta) / cfg.dt + 0.5)
    # +G [0,L), 0 [L,B), -G [B,B+L), 0 [B+L, B+L+R)  (diffusion.cpp:564-585)
    return [(1.0, L), (0.0, B - L), (-1.0, L), (0.0, R)]


def flops_per_spin_step(n_tri, n_coll, n_ch=3):
    """Work model: 50 + 7 n_ch + 45 n_tri + 21 n_coll FP32 operations (FMA = 2,
    transcendental = 1) per spin-step: Box-Muller 24, segment 3, step 6, cell
    9, wrap 6, T2 2, moment channels 7 each, Moller-Trumbore 45 per candidate
    whose box overlaps the step's box, reflection 21 per membrane hit."""
    return 50 + 7 * n_ch + 45 * n_tri + 21 * n_coll


def host_cpu():
    """The host's CPU as this process sees it: the model (/proc/cpuinfo),
    logical CPUs (nproc), the CPUs this process may run on (affinity), the
    cgroup's CPU quota, and the threads the CPU baselines use: every CPU the
    process may run on, within the quota."""
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = nproc
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    threads = aff if quota is None else max(1, min(aff, int(quota)))
    return {"model": model, "nproc": nproc, "affinity_cpus": aff, "cgroup_cpu_quota": quota,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"), "threads_used": threads}


def cpu_baseline(env, target_s, threads):
    """The oracle (plain-C restatement, same physical-mode semantics) on a
    bounded sample of the same workload, timed on this host's cores: a short
    calibration run sizes the sample to about target_s seconds."""
    import oracle as O

    grid = env["grid"]
    w = O.World(env["vertices"], env["triangles"], env["tri_info"], grid="aabb", dims=grid["dims"],
                lo=grid["lo"], cell=grid["cell"], eps=grid["eps"], std_dev=env["std_dev"],
                t2=env["t2"], dt=env["cfg"].dt, permeability=env["permeability"], test=1,
                traversal=1, transitions=1, boundary=1)

    def spins(count, seed):
        rng = np.random.default_rng(seed)
        pos = rng.uniform(-1, 1, (count, 3)).astype(np.float32)
        comp, fib, _, _ = w.find_spins(pos, env["xmax"], env["xmax"], env["xmax"], plus_x=True)
        return O.State(pos, comp=comp, fiber=fib, moment=True)

    # work-model counts from a short run with statistics on, then an untimed
    # warm-up and a calibration run without them
    stats = np.zeros(3, np.uint64)
    cal_spins, cal_steps = 20000, 200
    st = spins(cal_spins, 11)
    O.integrate(w, st, 100, rng=O.RNG_PHILOX, key=(5, 6), step0=0, weight=1.0,
                n_threads=threads, stats=stats)
    t0 = time.perf_counter()
    O.integrate(w, st, cal_steps, rng=O.RNG_PHILOX, key=(5, 6), step0=100, weight=1.0,
                n_threads=threads)
    rate0 = cal_spins * cal_steps / max(time.perf_counter() - t0, 1e-6)
    total = rate0 * target_s
    n_spins = int(min(200000, max(cal_spins, total / 2000)))
    steps = int(max(100, min(12000, total / n_spins)))
    st = spins(n_spins, 12)
    t0 = time.perf_counter()
    O.integrate(w, st, steps, rng=O.RNG_PHILOX, key=(5, 6), step0=0, weight=1.0,
                n_threads=threads)
    el = time.perf_counter() - t0
    n_steps, n_cand, n_hit = (int(x) for x in stats)
    return dict(value=n_spins * steps / el, unit="spin-steps/s", cores=threads, kind="port",
                sample=f"{n_spins} spins x {steps} steps of the same substrate and physics "
                       f"(oracle, {threads} OpenMP threads, {el:.1f} s)",
                n_tri=n_cand / max(1, n_steps), n_coll=n_hit / max(1, n_steps))


def _pct(a, q):
    return float(np.percentile(np.asarray(a, np.float64), q)) if len(a) else None


def batched_bench(n_substrates, plain_sample, round_steps=256):
    """The batched configuration: crossing-bundle substrates (axon counts
    50-5000, mean radii 0.5-3 um, two bundles at 0-90 degrees snapped to a
    lattice direction), each with one request of log-uniform 1e4-1e6 spins
    and 6-256 PGSE directions at one b <= 3 ms/um^2 (delta/Delta/TE
    10/40/60 ms), all advanced together by the batched step launches of the
    server; cell sizes from the autotuner.  The plain path runs a random
    sample of the same requests one at a time for the comparison (spin-steps/s
    on both sides: the requests differ in size)."""
    from dsim_amd import batched as B

    cs = B.CrossingSet(n_substrates=n_substrates)
    t0 = time.perf_counter()
    items = list(B.meshes(cs))
    t_mesh = time.perf_counter() - t0
    ntri = sum(len(it[3]) for it in items)
    log(f"batched: {len(items)} crossing meshes ({ntri} triangles) in {t_mesh:.1f} s")
    rs = sorted(items, key=lambda it: float(np.mean(np.concatenate([it[1].a_r, it[1].b_r]))))
    t0 = time.perf_counter()
    ratio, score = B.tune_cell_ratio(cs, [rs[0], rs[len(rs) // 2], rs[-1]], log=log)
    t_tune = time.perf_counter() - t0
    log(f"batched: cell = {ratio} x mean radius ({t_tune:.1f} s)")
    t0 = time.perf_counter()
    envs = []
    for k, it in enumerate(items):
        envs.append(B.make_env(cs, it, ratio))
        items[k] = None  # the host arrays are not needed any more
    t_build = time.perf_counter() - t0
    sub_bytes = sum(e["bytes"] for e in envs)
    log(f"batched: {len(envs)} device substrates in {t_build:.1f} s ({sub_bytes / 1e9:.1f} GB)")
    reqs = B.batched_requests(cs, envs)
    wk = B.work(reqs)
    t0 = time.perf_counter()
    out, span = B.run_server(cs, envs, reqs, round_steps=round_steps)
    wall = time.perf_counter() - t0
    rate = wk / span
    log(f"batched: {len(out)} requests, {sum(q['n'] for q in reqs)} spins, {wk:.3e} spin-steps in "
        f"{span:.2f} s: {rate:.3e} spin-steps/s, {len(out) / span * 3600:.0f} simulations/hour")
    idx = np.sort(np.random.default_rng(3).choice(len(reqs), min(plain_sample, len(reqs)),
                                                  replace=False))
    sample = [reqs[i] for i in idx]
    plain = B.run_plain(cs, envs, sample)
    t_plain = sum(t for t, _ in plain)
    same = all(np.array_equal(out[i]["signals"], sig) for i, (_, sig) in zip(idx, plain))
    plain_rate = B.work(sample) / t_plain
    log(f"batched: plain path on {len(sample)} of the requests {plain_rate:.3e} spin-steps/s; "
        f"batched signals {'bit-identical' if same else 'DIFFER'}")
    res = {"workload": (f"{len(envs)} crossing-bundle substrates ({cs.axons[0]}-{cs.axons[1]} axons, "
                        f"mean radius {cs.radius[0]}-{cs.radius[1]} um, 32-gon, crossing "
                        f"0-90 deg on lattice directions), one request each: 1e4-1e6 spins "
                        f"(log-uniform), 6-256 directions at b 0.5-3 ms/um^2 + b=0, PGSE "
                        f"10/40/60 ms"),
           "substrates": len(envs), "requests": len(reqs), "spins": int(sum(q["n"] for q in reqs)),
           "spin_steps": wk, "seconds": span, "spin_steps_per_s": rate,
           "sims_per_hour": len(out) / span * 3600, "plain_sample": len(sample),
           "plain_spin_steps_per_s": plain_rate, "speedup_vs_plain": rate / plain_rate,
           "bit_identical_to_plain": bool(same), "round_steps": round_steps,
           "cell_ratio": ratio, "cell_ratio_candidates": score,
           "cells_um": [round(min(e["cell_um"] for e in envs), 3),
                        round(max(e["cell_um"] for e in envs), 3)],
           "triangles": ntri, "substrate_bytes": sub_bytes,
           "setup_s": {"meshes": round(t_mesh, 1), "tuning": round(t_tune, 1),
                       "device_builds": round(t_build, 1)}}
    return res, cs, envs


def served_bench(cs, envs, n_requests, round_steps=256, loads=(0.4,), overload_rate=50.0,
                 overload_window=8.0, overload_drain=20.0, overload_active_spins=4 << 20,
                 small_first_blocks=512):
    """The served configuration on resident crossing substrates.  A stream
    scaled to one GPU (requests of 1e4-2e5 spins, protocols as configured):
    capacity with every request queued at once, the same requests one at a
    time through the plain handle, and Poisson streams at a fraction of the
    capacity for the latency percentiles, each request's latency also against
    its own solo (plain) time, under the default round policy and under the
    small-first one (requests of at most small_first_blocks one-wave blocks
    walk in rounds of their own), whose capacity is measured too.  Then the
    configured stream (Poisson 50 req/s, 1e4-2e6 spins) over a fixed arrival
    window: the requests done within the window plus a drain period, their
    latencies, and the sustained rate."""
    from dsim_amd import batched as B

    reqs = B.stream_requests(cs, envs, n_requests, spins=(1e4, 2e5))
    out, span = B.run_server(cs, envs, reqs, round_steps=round_steps)
    cap = len(out) / span
    log(f"served: {len(out)} requests ({B.work(reqs):.3e} spin-steps) in {span:.2f} s: "
        f"{cap * 3600:.0f} simulations/hour")
    plain = B.run_plain(cs, envs, reqs)
    solo = np.array([t for t, _ in plain])
    same = all(np.array_equal(o["signals"], sig) for o, (_, sig) in zip(out, plain))
    log(f"plain: the same {len(reqs)} requests one at a time in {solo.sum():.2f} s "
        f"({len(reqs) / solo.sum() * 3600:.0f} simulations/hour); served signals "
        f"{'bit-identical' if same else 'DIFFER'}")
    by_seed = {q["seed"]: t for q, t in zip(reqs, solo)}
    out_sf, span_sf = B.run_server(cs, envs, reqs, round_steps=round_steps,
                                   small_first_blocks=small_first_blocks)
    cap_sf = len(out_sf) / span_sf
    same_sf = all(np.array_equal(o["signals"], sig) for o, (_, sig) in zip(out_sf, plain))
    log(f"served, small-first rounds (<= {small_first_blocks} blocks): {cap_sf * 3600:.0f} "
        f"simulations/hour; signals {'bit-identical' if same_sf else 'DIFFER'}")
    streams = []
    for load in loads:
        rate = load * cap
        stream = B.stream_requests(cs, envs, n_requests, rate=rate, spins=(1e4, 2e5))
        for policy in (0, small_first_blocks):
            live, _ = B.run_server(cs, envs, stream, realtime=True, round_steps=round_steps,
                                   small_first_blocks=policy)
            lat = np.array([o["latency"] for o in live])
            vs = np.array([o["latency"] / by_seed[o["req"]["seed"]] for o in live])
            small = np.array([o["req"]["n"] <= 64 * small_first_blocks for o in live])
            streams.append({"load": load, "rate_per_s": rate, "requests": len(live),
                            "small_first_blocks": policy,
                            "p50_latency_s": _pct(lat, 50), "p99_latency_s": _pct(lat, 99),
                            "p50_latency_vs_solo": _pct(vs, 50),
                            "p99_latency_vs_solo": _pct(vs, 99),
                            "p99_vs_solo_small_requests": _pct(vs[small], 99),
                            "p99_vs_solo_large_requests": _pct(vs[~small], 99)})
            log(f"stream at {rate:.1f} req/s ({load:.0%} of capacity), small-first {policy}, "
                f"{len(live)} requests: latency p50 {_pct(lat, 50):.3f} s, p99 {_pct(lat, 99):.3f} s; "
                f"vs solo p50 {_pct(vs, 50):.2f}x, p99 {_pct(vs, 99):.2f}x (requests <= "
                f"{64 * small_first_blocks} spins {_pct(vs[small], 99):.2f}x, larger "
                f"{_pct(vs[~small], 99):.2f}x)")
    # the latency floor under load: the stream's smallest request submitted
    # together with its k largest, against its solo time
    small = min(reqs, key=lambda q: q["n"])
    large = sorted(reqs, key=lambda q: -q["n"])[:3]
    corun = []
    for k in (1, 2, 3):
        batch = [dict(q, arrival=0.0) for q in large[:k]] + [dict(small, arrival=0.0)]
        o3, _ = B.run_server(cs, envs, batch, round_steps=round_steps)
        lat = [o["latency"] for o in o3 if o["req"]["seed"] == small["seed"]][0]
        corun.append({"large_requests": k, "large_spins": [q["n"] for q in large[:k]],
                      "small_spins": small["n"], "small_latency_s": lat,
                      "small_latency_vs_solo": lat / by_seed[small["seed"]]})
        log(f"co-run: a {small['n']}-spin request with {k} of {[q['n'] for q in large[:k]]} spins: "
            f"latency {lat:.3f} s, {lat / by_seed[small['seed']]:.2f}x solo")
    # the configured stream: far beyond one GPU's capacity, so a finite window
    n_over = int(overload_rate * overload_window * 1.5) + 10
    over = B.stream_requests(cs, envs, n_over, rate=overload_rate, seed=9)
    sub = [q for q in over if q["arrival"] <= overload_window]
    # admission holds the spins in flight to what fills the device (FIFO):
    # without the limit every queued request shares the rounds and none ends
    # within the window
    o2, span2 = B.run_server(cs, envs, sub, realtime=True, round_steps=round_steps,
                             window_s=overload_window + overload_drain,
                             max_active_spins=overload_active_spins)
    done = [o for o in o2 if o["done"] is not None]
    lat = [o["latency"] for o in done]
    wait = [o["queued"] for o in done]
    svc = [o["service"] for o in done]
    wdone = B.work([o["req"] for o in done])
    T = overload_window + overload_drain
    overload = {"rate_per_s": overload_rate, "arrival_window_s": overload_window,
                "max_active_spins": overload_active_spins,
                "measured_until_s": T, "submitted": len(sub), "completed": len(done),
                "offered_spin_steps_per_s": B.work(sub) / overload_window,
                "sustained_spin_steps_per_s": wdone / T,
                "sustained_sims_per_hour": len(done) / T * 3600,
                "completed_p50_latency_s": _pct(lat, 50), "completed_p99_latency_s": _pct(lat, 99),
                "completed_p50_queue_s": _pct(wait, 50), "completed_p99_queue_s": _pct(wait, 99),
                "completed_p50_service_s": _pct(svc, 50), "completed_p99_service_s": _pct(svc, 99)}
    log(f"configured stream: {len(sub)} requests at {overload_rate:.0f}/s over {overload_window:.0f} s "
        f"(offered {overload['offered_spin_steps_per_s']:.2e} spin-steps/s); {len(done)} done within "
        f"{T:.0f} s (sustained {overload['sustained_spin_steps_per_s']:.2e} spin-steps/s); latency p50 "
        f"{_pct(lat, 50)} s, p99 {_pct(lat, 99)} s")
    return {"workload": (f"{len(envs)} resident crossing substrates (drawn as the batched set); "
                         f"requests: 6-256 directions, b 0.5-5 ms/um^2, delta 5-20 ms, Delta "
                         f"20-60 ms; capacity and latency streams of {n_requests} requests of "
                         f"1e4-2e5 spins, the configured stream of 1e4-2e6 spins at "
                         f"{overload_rate:.0f} req/s"),
            "requests": len(reqs), "substrates": len(envs), "round_steps": round_steps,
            "sims_per_hour": cap * 3600, "plain_sims_per_hour": len(reqs) / solo.sum() * 3600,
            "small_first_sims_per_hour": cap_sf * 3600,
            "small_first_bit_identical": bool(same_sf),
            "speedup_vs_plain": cap / (len(reqs) / solo.sum()),
            "served_spin_steps_per_s": B.work(reqs) / span,
            # the served rate is the device's batched rate on these substrates
            # (the capacity run keeps every block slot busy), so the capacity
            # speedup cannot exceed 1 / this fraction
            "plain_spin_steps_per_s": B.work(reqs) / solo.sum(),
            "plain_rate_fraction_of_served": (B.work(reqs) / solo.sum()) / (B.work(reqs) / span),
            "streams": streams, "small_request_with_large_ones": corun,
            "bit_identical_to_plain": bool(same), "configured_stream": overload}


def reference_mode_bench(n, steps, cpu_seconds, threads, cpu=True):
    """The reference-semantics path a drop-in user runs (spinKernelCpu.cpp
    arithmetic: softsurfer triangle test, first-cell walk over the 50^3 grid,
    clamp box, MWC seeds; k_step on the device) on the small configuration's
    mesh: device rate over `steps` steps of `n` spins, and the oracle's rate in
    the same mode on a sample of the same spins on the host cores."""
    import oracle as O

    cfg = S.SMALL
    env = S.build(cfg, cell_um=0.2, axons=S.pack(cfg))
    V, T, I = env["vertices"], env["triangles"], env["tri_info"]
    sub = D.Substrate(V, T, I)                       # reference 50^3 grid
    pc = np.float32(2 * np.pi * GAMMA * cfg.dt)
    grad = np.array([0.0, 61.7e-9 * env["space_scale"], 0.0], np.float32)
    sim = D.Simulation(sub, n, dt=cfg.dt, std_dev=env["std_dev"], t2=env["t2"],
                       permeability=env["permeability"], mode=D.MODE_REFERENCE,
                       boundary=D.BOUNDARY_CLAMP, rng=D.RNG_MWC, seed=5)
    sim.init_uniform(1.0)
    sim.label(env["xmax"], env["xmax"], env["xmax"])
    stopped = 0

    def walk(k):
        # a step on which the reference's collision loop never ends (a segment
        # on the diagonal two coplanar facet triangles share; about one in
        # 1e9 spin-steps on this mesh) is stopped after 4096 collisions and
        # reported once; the walk goes on
        nonlocal stopped
        sim.integrate(grad, pc, k)
        try:
            sim.sync()
        except D.DsimError as e:
            if e.code != -5:
                raise
            stopped += 1

    walk(64)                                          # warm-up
    sim.step_timing_reset()
    t0 = time.perf_counter()
    walk(steps)
    wall = time.perf_counter() - t0
    kernel_ms, launches = sim.step_timing()
    rate = n * steps / wall
    out = {"spins": n, "steps": steps, "value": rate, "unit": "spin-steps/s",
           "kernel_ms_per_launch": kernel_ms / max(1, launches), "launches": launches,
           "calls_with_a_step_stopped_at_the_collision_guard": stopped,
           "substrate": f"{cfg.name} mesh, reference 50^3 grid, clamp box, MWC seeds"}
    if cpu:
        k = min(n, 20000)
        st = sim.get_state(seeds=True)
        w = O.World(V, T, I, std_dev=env["std_dev"], t2=env["t2"], dt=cfg.dt,
                    permeability=env["permeability"])
        s0 = O.State(st["pos"][:k], comp=st["comp"][:k], fiber=st["fiber"][:k],
                     seed=st["seeds"][:k])
        t0 = time.perf_counter()
        O.integrate(w, s0, 20, rng=O.RNG_MWC, grad=grad, phase_constant=pc, n_threads=threads)
        cal = k * 20 / max(time.perf_counter() - t0, 1e-6)
        cpu_steps = int(max(20, min(steps, cal * cpu_seconds / k)))
        t0 = time.perf_counter()
        O.integrate(w, s0, cpu_steps, rng=O.RNG_MWC, grad=grad, phase_constant=pc,
                    n_threads=threads)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": k * cpu_steps / el, "unit": "spin-steps/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{k} spins x {cpu_steps} steps, same mode and mesh "
                                         f"(oracle, {el:.1f} s)"}
    sim.close()
    sub.close()
    log(f"reference mode: {n} spins x {steps} steps at {rate:.3e} spin-steps/s"
        + (f"; oracle {out['cpu_baseline']['value']:.3e} on {threads} cores" if cpu else ""))
    return out


PEAK_BF16 = 2.5e15      # MI355X dense BF16 matrix, FLOP/s
PEAK_HBM = 8.0e12       # MI355X HBM3E, B/s


def substrate_prep(env):
    """Time to a resident substrate: the host build (multi-threaded C++) and
    the device build of the same records, grid and clearance, each from the
    mesh arrays in host memory; the grids and clearances are compared."""
    g = env["grid"]
    kw = dict(grid="aabb", dims=g["dims"], lo=g["lo"], cell=g["cell"], eps=g["eps"])
    mesh = (env["vertices"], env["triangles"], env["tri_info"])
    D.Substrate(*mesh, device=True, **kw).close()  # first use of the build kernels
    t0 = time.perf_counter()
    h = D.Substrate(*mesh, **kw)
    t1 = time.perf_counter()
    d = D.Substrate(*mesh, device=True, **kw)
    t2 = time.perf_counter()
    sa, ta, ma = h.grid()
    sb, tb, mb = d.grid()
    same = bool(np.array_equal(sa, sb) and np.array_equal(ta, tb) and ma == mb
                and np.array_equal(h.clearance(), d.clearance()))
    h.close()
    d.close()
    return {"triangles": len(env["triangles"]), "cells": int(np.prod(g["dims"])),
            "entries": int(len(ta)), "host_s": round(t1 - t0, 4), "device_s": round(t2 - t1, 4),
            "identical": same}


def waveform_bench(n_spins, n_wave, sample, chunk=64, buffer=2048):
    """Waveform-phase path: n_spins spins on the small substrate, n_wave random
    piecewise-constant waveforms (|G| <= 80 mT/m, a new vector every 10 us) over
    12,000 steps, Kahan-folded every `chunk` steps (GEMM K = 3 * chunk), the
    trajectory walked `buffer` steps per GEMM launch; the phase GEMM against
    the dense BF16 MFMA peak; per-waveform error against FP64 on a
    `sample`-spin run of the same substrate and waveforms."""
    from dsim_amd import phase as PH

    cfg = S.SMALL
    steps = 12000
    t0 = time.perf_counter()
    env = S.build(cfg, cell_um=0.2, axons=S.pack(cfg))
    g = env["grid"]
    sub = D.Substrate(env["vertices"], env["triangles"], env["tri_info"], grid="aabb",
                      dims=g["dims"], lo=g["lo"], cell=g["cell"], eps=g["eps"], device=True)
    pc = np.float32(2 * np.pi * GAMMA * cfg.dt)
    grads = PH.random_waveforms(n_wave, steps, seed=17) * np.float32(1e-9 * env["space_scale"])
    log(f"waveforms: substrate and {n_wave} waveforms in {time.perf_counter() - t0:.1f} s")

    def prepared(n, seed):
        sim = D.Simulation(sub, n, dt=cfg.dt, std_dev=env["std_dev"], t2=env["t2"],
                           permeability=env["permeability"], mode=D.MODE_PHYSICAL,
                           boundary=D.BOUNDARY_PERIODIC, rng=D.RNG_PHILOX, seed=seed)
        sim.init_uniform(1.0)
        sim.label(env["xmax"], env["xmax"], env["xmax"])
        sim.moment_reset()
        return sim

    sim = prepared(n_spins, 3)
    wf = PH.Waveforms(sim, grads, pc, chunk_steps=chunk, buffer_steps=buffer)
    t0 = time.perf_counter()
    wf.run(steps)
    sig = wf.signal()
    wall = time.perf_counter() - t0
    walk_ms, gemm_ms = wf.timing()
    wf.close()
    sim.close()
    flops = 3 * 2.0 * n_spins * n_wave * 3 * steps   # hi*hi + hi*lo + lo*hi
    n_buffers = -(-steps // buffer)
    # per buffer the (phase, compensation) pairs are read and written once (4 B
    # phase + 2 B bf16 compensation each way); the trajectory is read once
    gemm_bytes = n_buffers * (12.0 * n_spins * n_wave) + 4.0 * n_spins * 3 * steps
    log(f"waveforms: {n_spins} spins x {steps} steps x {n_wave} waveforms in {wall:.2f} s "
        f"(walk {walk_ms / 1e3:.2f} s, phase GEMM {gemm_ms / 1e3:.2f} s = "
        f"{flops / (gemm_ms / 1e3) / 1e12:.0f} TFLOP/s); b=0-free signal range "
        f"{sig[:, 0].min():.4f}-{sig[:, 0].max():.4f}")
    # FP64 check on a sample: the same product against a float64 GEMM of the
    # recorded trajectory (numpy, checker only)
    sim = prepared(sample, 4)
    wf = PH.Waveforms(sim, grads, pc, chunk_steps=chunk, buffer_steps=buffer)
    traj = wf.run(steps, trajectory=True)
    phi, _ = wf.phase()
    got = wf.signal()
    st = sim.get_state()
    _, x0 = sim.get_moment()
    wf.close()
    sim.close()
    t0 = time.perf_counter()
    coef = (pc * grads).astype(np.float32).astype(np.float64)          # [steps, n_wave, 3]
    A = traj.transpose(2, 0, 1).reshape(sample, -1).astype(np.float64)
    del traj
    B = np.ascontiguousarray(coef.transpose(0, 2, 1).reshape(-1, n_wave))
    ref = A @ B
    del A
    dphi = np.abs(phi.astype(np.float64) - ref).max(0)
    tot = ref + x0.astype(np.float64) @ coef.sum(0).T
    z = st["mag"].astype(np.float64)[:, None] * np.exp(1j * tot)
    want = np.abs(z.sum(0)) / sample
    dsig = np.abs(got[:, 0] - want)
    log(f"waveforms: FP64 reference in {time.perf_counter() - t0:.1f} s")
    log(f"waveforms vs FP64 ({sample} spins): phase error per waveform max {dphi.max():.2e} rad "
        f"(median {np.median(dphi):.2e}); signal error max {dsig.max():.2e} (median "
        f"{np.median(dsig):.2e})")
    sub.close()
    return {"spins": n_spins, "waveforms": n_wave, "steps": steps, "chunk_steps": chunk,
            "gemm_k": 3 * chunk, "buffer_steps": buffer, "substrate": cfg.name, "seconds": wall,
            "spin_steps_per_s": n_spins * steps / wall, "walk_s": walk_ms / 1e3,
            "phase_gemm_s": gemm_ms / 1e3,
            # algorithmic flops: the three bf16 products (hi*hi, hi*lo, lo*hi)
            # of the spins x (3 * steps) x waveforms GEMM; bytes: the
            # read-modify-write of the (phase, compensation) pairs, 6 B per
            # spin and waveform each way, once per trajectory buffer, plus one
            # read of the trajectory (4 B per spin and k)
            "phase_gemm_roofline": {"bound": "mfma", "achieved": flops / (gemm_ms / 1e3) / 1e12,
                                    "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
                                    "frac": flops / (gemm_ms / 1e3) / PEAK_BF16,
                                    "hbm_gbs": gemm_bytes / (gemm_ms / 1e3) / 1e9,
                                    "hbm_frac": gemm_bytes / (gemm_ms / 1e3) / PEAK_HBM,
                                    "traffic": committed_gemm_traffic(n_spins, n_wave, chunk,
                                                                      buffer),
                                    "traffic_unit": "bytes per full trajectory buffer"},
            "fp64_check": {"spins": sample, "phase_err_max_rad": float(dphi.max()),
                           "phase_err_median_rad": float(np.median(dphi)),
                           "signal_err_max": float(dsig.max()),
                           "signal_err_median": float(np.median(dsig))}}


def footprint_bench(axon_counts, n=1_000_000, cell_um=0.2, sort_interval=1024):
    """The at-scale walk on the same packing statistics in voxels of each axon
    count (the voxel sized for the same intra-axonal fraction): smaller voxels
    hold the whole substrate in cache, so their rate, with the search work
    per step from the walk counters, bounds what moving the triangle records
    nearer the compute units can win.  One acquisition each after a warm-up."""
    import dataclasses

    out = []
    for na in axon_counts:
        cfg = dataclasses.replace(S.SCALE, n_axons=int(na), name=f"scale-{na}axon")
        env = S.build(cfg, cell_um=cell_um, axons=S.pack(cfg))
        sub, sim = AT.make_sim(env, n, sort_interval=sort_interval,
                               steps_per_launch=AT.launch_steps(sort_interval))
        sim.init_uniform(1.0)
        sim.label(env["xmax"], env["xmax"], env["xmax"])
        sched = [(1.0, 2000), (0.0, 6000), (-1.0, 2000), (0.0, 2000)]

        def acq():
            sim.moment_reset()
            for w, k in sched:
                sim.integrate_moment(w, k)

        acq()
        sim.sync()
        sim.step_timing_reset()
        t0 = time.perf_counter()
        acq()
        sim.sync()
        wall = time.perf_counter() - t0
        ms, launches = sim.step_timing()
        sim.walk_stats(True)
        acq()
        sim.sync()
        st = sim.walk_stats()
        sim.walk_stats(False)
        k = st["steps"]
        r = {"axons": int(na), "voxel_um": round(env["axons"].voxel, 2),
             "triangles": len(env["triangles"]), "substrate_bytes": sub.nbytes(),
             "spin_steps_per_s": n * 12000 / wall, "kernel_ms_per_launch": ms / max(1, launches),
             "per_spin_step": {key: st[key] / k for key in ("trips", "plane_tests", "tri_tests",
                                                              "hits", "shortcuts", "cells")}}
        r["per_spin_step"]["trips"] *= 64  # wave trips per lane-step
        log(f"footprint: {na} axons, voxel {r['voxel_um']} um, {r['substrate_bytes'] / 2**20:.1f} "
            f"MiB of substrate: {r['spin_steps_per_s']:.4e} spin-steps/s, plane tests "
            f"{r['per_spin_step']['plane_tests']:.3f} per spin-step")
        out.append(r)
        sim.close()
        sub.close()
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_walk_scale-2000axon-1e6_hbm_traffic.md")


def committed_traffic(cell_um, sort_interval, n):
    """HBM bytes per step launch of k_walk from the committed FETCH_SIZE /
    WRITE_SIZE summary (rocprofv3 --pmc cannot run inside the timed process),
    for the matching grid cell, reorder interval and spin count; else None."""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        if ln.startswith("traffic_bytes_per_launch"):
            f = dict(kv.split("=") for kv in ln.split()[1:])
            if (abs(float(f["cell_um"]) - cell_um) < 1e-9 and int(f["sort"]) == sort_interval
                    and int(f["spins"]) == n):
                return int(f["bytes"])
    return None


GEMM_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_phase_gemm_small-16axon-4e6x1024_hbm_traffic.md")


def committed_gemm_traffic(n_spins, n_wave, chunk, buffer):
    """HBM bytes per full-buffer launch of k_phase_gemm from the committed
    FETCH_SIZE / WRITE_SIZE summary for the matching sizes; else None."""
    try:
        lines = open(GEMM_TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        if ln.startswith("traffic_bytes_per_buffer"):
            f = dict(kv.split("=") for kv in ln.split()[1:])
            if (int(f["spins"]) == n_spins and int(f["waveforms"]) == n_wave
                    and int(f["chunk"]) == chunk and int(f["buffer"]) == buffer):
                return int(f["bytes"])
    return None


def copy_probe(mib):
    """A plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE counter
    readings calibrate the HBM traffic counted for the step kernel."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied {mib} MiB three times")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="scale", choices=["scale", "small"])
    ap.add_argument("--spins", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cell-um", type=float, default=0.0)
    ap.add_argument("--sort", type=int, default=-1)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--batched-substrates", type=int, default=256,
                    help="crossing substrates of the batched measurement (0 skips it and the "
                         "served one)")
    ap.add_argument("--plain-sample", type=int, default=16,
                    help="batched requests also run one at a time through the plain handle")
    ap.add_argument("--served-substrates", type=int, default=64)
    ap.add_argument("--served-requests", type=int, default=200,
                    help="requests of the served capacity and latency streams (0 skips them)")
    ap.add_argument("--round-steps", type=int, default=256,
                    help="steps per batched round of the served form")
    ap.add_argument("--waveform-spins", type=int, default=4_000_000,
                    help="spins of the waveform-phase measurement (0 skips it)")
    ap.add_argument("--waveforms", type=int, default=1024)
    ap.add_argument("--reference-spins", type=int, default=1_000_000,
                    help="spins of the reference-semantics run (0 skips it)")
    ap.add_argument("--chunk-steps", type=int, default=64,
                    help="Kahan interval of the waveform phases in steps (multiple of 32)")
    ap.add_argument("--buffer-steps", type=int, default=2048,
                    help="steps walked per phase-GEMM launch (multiple of --chunk-steps)")
    ap.add_argument("--footprint-axons", default="", metavar="N,N,...",
                    help="only the substrate-footprint measurement at these axon counts")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the simulate path is data-parallel over spins and "
              f"would shard by spin id with one FP64 signal all-reduce; "
              f"--gpus {args.gpus} is not implemented.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.footprint_axons:
        print(json.dumps({"footprint": footprint_bench(
            [int(a) for a in args.footprint_axons.split(",")])}))
        return
    if D.device_count() == 0:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    cfg = S.SCALE if args.config == "scale" else S.SMALL
    n = args.spins or cfg.n_spins
    t0 = time.perf_counter()
    axons = S.pack(cfg)
    log(f"packed {cfg.n_axons} axons: voxel {axons.voxel:.2f} um, intra fraction "
        f"{axons.intra_fraction:.3f} ({time.perf_counter() - t0:.1f} s)")

    tuned = None
    cell_um, sort_interval = args.cell_um or 0.25, 128 if args.sort < 0 else args.sort
    if not args.no_autotune and not args.cell_um:
        t0 = time.perf_counter()
        best, allr = AT.tune(cfg, axons, n, sorts=AT.SORTS if args.sort < 0 else (args.sort,),
                             log=log)
        cell_um, sort_interval = best["cell_um"], best["sort_interval"]
        tuned = dict(cell_um=cell_um, sort_interval=sort_interval,
                     candidates=[{k: r[k] for k in ("cell_um", "sort_interval", "rate")} for r in allr],
                     seconds=round(time.perf_counter() - t0, 1))
        log(f"autotune picked cell {cell_um} um, sort every {sort_interval} steps")

    env = S.build(cfg, cell_um=cell_um, axons=axons)
    g = env["grid"]
    prep = substrate_prep(env)
    log(f"substrate prep ({prep['triangles']} triangles, {prep['cells']} cells): host build "
        f"{prep['host_s']:.3f} s, device build {prep['device_s']:.3f} s, identical "
        f"{prep['identical']}")
    sub, sim = AT.make_sim(env, n, sort_interval=sort_interval,
                           steps_per_launch=AT.launch_steps(sort_interval))
    _, _, maxc = sub.grid()
    sim.init_uniform(1.0)
    sim.sync()
    t0 = time.perf_counter()
    amb = sim.label(env["xmax"], env["xmax"], env["xmax"])   # returns after the device finishes
    label_s = time.perf_counter() - t0
    st = sim.get_state()
    frac_intra = float(np.mean(st["comp"] == 1))
    labelling = {"spins": n, "seconds": round(label_s, 4), "ambiguous": int(amb),
                 "intra_fraction": round(frac_intra, 4)}
    log(f"{len(env['triangles'])} triangles, grid {g['dims']} cell {cell_um} um; labelled {n} "
        f"spins in {label_s:.3f} s: intra {frac_intra:.3f}, ambiguous {amb}")

    prot = env["protocol"]
    grads = (prot[:, 3:6].astype(np.float64) * 1e-9 * env["space_scale"]).astype(np.float32)
    pc = np.float32(2 * np.pi * GAMMA * cfg.dt)
    sched = schedule(cfg)
    steps_per_acq = sum(k for _, k in sched)

    def acquisition():
        sim.moment_reset()
        for wgt, k in sched:
            sim.integrate_moment(wgt, k)
        return sim.signal_moment(grads, pc)

    for _ in range(args.warmup):
        acquisition()
    sim.sync()
    sim.step_timing_reset()
    t0 = time.perf_counter()
    sig = None
    for _ in range(args.steps):
        sig = acquisition()
    sim.sync()
    wall = time.perf_counter() - t0
    step_ms, launches = sim.step_timing()
    spin_steps = n * steps_per_acq * args.steps
    value = spin_steps / wall
    log(f"{args.steps} acquisitions in {wall:.3f} s; step kernels {step_ms / 1e3:.3f} s over "
        f"{launches} launches; b=0 signal {sig[0, 0]:.5f}, b=1 mean {sig[1:22, 0].mean():.5f}")

    cpu = None
    n_tri, n_coll = None, None
    hcpu = host_cpu()
    if not args.no_cpu_baseline:
        threads = hcpu["threads_used"]
        cpu = cpu_baseline(env, args.cpu_seconds, threads)
        cpu["host"] = hcpu
        n_tri, n_coll = cpu.pop("n_tri"), cpu.pop("n_coll")
        log(f"cpu baseline {cpu['value']:.3e} spin-steps/s ({cpu['sample']}); work model "
            f"n_tri {n_tri:.3f}, n_coll {n_coll:.4f}")
    roof = None
    traffic = committed_traffic(cell_um, sort_interval, n)
    if n_tri is not None:
        F = flops_per_spin_step(n_tri, n_coll)
        avg_launch_s = step_ms / 1e3 / max(1, launches)
        achieved = F * n * steps_per_acq * args.steps / (step_ms / 1e3)
        roof = {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": traffic,
                "flops_per_spin_step": F, "n_tri": n_tri, "n_coll": n_coll,
                "avg_launch_ms": avg_launch_s * 1e3,
                "spin_steps_per_launch": n * steps_per_acq * args.steps / max(1, launches)}
    batched = served = None
    if args.batched_substrates:
        sim.close()
        sub.close()
        batched, cs, envs = batched_bench(args.batched_substrates, args.plain_sample,
                                          args.round_steps)
        keep = envs[:args.served_substrates]
        for e in envs[args.served_substrates:]:
            e["sub"].close()
        if args.served_requests:
            served = served_bench(cs, keep, args.served_requests, args.round_steps)
        for e in keep:
            e["sub"].close()
    waves = None
    if args.waveform_spins:
        waves = waveform_bench(args.waveform_spins, args.waveforms, 8192, args.chunk_steps,
                               args.buffer_steps)
    refmode = None
    if args.reference_spins:
        refmode = reference_mode_bench(args.reference_spins, 1000, 5.0, hcpu["threads_used"],
                                       cpu=not args.no_cpu_baseline)
        if refmode.get("cpu_baseline"):
            refmode["gpu_over_cpu"] = refmode["value"] / refmode["cpu_baseline"]["value"]
    out = {
        "metric": "spin-steps/s at 1M spins",
        "value": value,
        "unit": "spin-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg.name, "spins": n, "axons": cfg.n_axons,
                   "triangles": int(len(env["triangles"])), "voxel_um": round(axons.voxel, 2),
                   "intra_fraction": round(axons.intra_fraction, 4),
                   "steps_per_acquisition": steps_per_acq, "gradient_lines": int(len(prot)),
                   "dt_ms": cfg.dt, "permeability_per_hit": env["permeability"],
                   "grid": list(g["dims"]), "cell_um": cell_um, "sort_interval": sort_interval,
                   "max_tris_per_cell": maxc, "parallelism": "none (1 GPU)"},
        "reference_equivalent_value": value * len(prot),
        "autotune": tuned,
        "substrate_prep": prep,
        "labelling": labelling,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "waveforms": waves,
        "reference_mode": refmode,
        "host": dict(hcpu, arch=platform.machine()),
        "gpu_over_cpu": value / cpu["value"] if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
