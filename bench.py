"""Benchmark: spin-steps/s at 1M spins on one MI355X.

Workload (the at-scale configuration): 1e6 spins, 2000 z-aligned axons with
radii ~ Gamma(k=3, mean 1 um) packed to an intra-axonal fraction of 0.7
(voxel 107.9 um, periodic), 64-gon open prisms (286k triangles with the
periodic images), permeability 0.01 um/ms, T2 70/90 ms, dt 5 us, PGSE
delta/Delta/TE = 10/40/60 ms (12,000 steps), 64 directions on b = 1/2/3
ms/um^2 shells plus b=0, fp32.  One bench step is one complete acquisition:
every spin walks 12,000 steps once (the gradient-moment accumulation serves
all 65 gradient lines) and the 65 signals are read out in FP64.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config scale|small]
                       [--cell-um H] [--sort S] [--no-autotune] [--no-cpu-baseline]
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


def served_bench(n_requests, plain_requests, round_steps=256):
    """The served form against the plain one on a stream of requests over
    resident substrates: capacity (every request queued at once) in
    simulations per hour, the same requests one at a time through the plain
    handle, and Poisson streams at 40 % and 70 % of the measured capacity for
    the latency percentiles (each request's latency also against its own
    solo run)."""
    from dsim_amd import served as SVD

    sc = SVD.StreamConfig(n_requests=n_requests)
    t0 = time.perf_counter()
    envs = SVD.make_substrates(sc, log)
    log(f"served: {sc.n_substrates} resident substrates in {time.perf_counter() - t0:.1f} s")
    reqs = SVD.make_requests(sc, envs)
    out, mk = SVD.run_served(envs, reqs, realtime=False, round_steps=round_steps)
    cap = len(out) / mk
    log(f"served: {len(out)} requests ({SVD.work(reqs):.3e} spin-steps) in {mk:.2f} s")
    sub = reqs[:plain_requests]
    plain = SVD.run_plain(envs, sub)
    t_plain = sum(t for t, _ in plain)
    same = all(np.array_equal(o["signals"], sig) for o, (_, sig) in zip(out, plain))
    log(f"plain: {len(sub)} requests one at a time in {t_plain:.2f} s; served signals "
        f"{'bit-identical' if same else 'DIFFER'}")
    solo = np.array([t for t, _ in plain])
    streams = []
    for load in (0.4, 0.7):
        rate = load * cap
        live, _ = SVD.run_served(envs, SVD.make_requests(sc, envs, rate=rate), realtime=True,
                                 round_steps=round_steps)
        lat = np.array([o["latency"] for o in live])
        vs = lat[:len(solo)] / solo
        streams.append({"load": load, "rate_per_s": rate,
                        "p50_latency_s": float(np.percentile(lat, 50)),
                        "p99_latency_s": float(np.percentile(lat, 99)),
                        "p50_latency_vs_solo": float(np.percentile(vs, 50)),
                        "p99_latency_vs_solo": float(np.percentile(vs, 99))})
        log(f"stream at {rate:.1f} req/s ({load:.0%} of capacity): latency p50 "
            f"{np.percentile(lat, 50):.3f} s, p99 {np.percentile(lat, 99):.3f} s; vs solo p50 "
            f"{np.percentile(vs, 50):.2f}x, p99 {np.percentile(vs, 99):.2f}x")
    for e in envs:
        e["sub"].close()
    return {"requests": len(reqs), "substrates": sc.n_substrates, "round_steps": round_steps,
            "sims_per_hour": cap * 3600, "plain_sims_per_hour": len(sub) / t_plain * 3600,
            "speedup_vs_plain": cap / (len(sub) / t_plain),
            "served_spin_steps_per_s": SVD.work(reqs) / mk,
            "streams": streams,
            "bit_identical_to_plain": bool(same),
            "workload": (f"{sc.n_substrates} substrates of {sc.axons[0]}-{sc.axons[1]} axons "
                         f"(mean radius {sc.radius[0]}-{sc.radius[1]} um, {sc.ngon}-gon), "
                         f"requests of {sc.spins[0]:.0e}-{sc.spins[1]:.0e} spins (log-uniform), "
                         f"{sc.lines[0]}-{sc.lines[1]} lines, b {sc.b[0]}-{sc.b[1]} ms/um^2, "
                         f"delta {sc.delta[0]}-{sc.delta[1]} ms, Delta {sc.Delta[0]}-{sc.Delta[1]} ms; "
                         f"plain = first {len(sub)} requests one at a time")}


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
    sim.integrate(grad, pc, 64)                       # warm-up
    sim.sync()
    sim.step_timing_reset()
    t0 = time.perf_counter()
    sim.integrate(grad, pc, steps)
    sim.sync()
    wall = time.perf_counter() - t0
    kernel_ms, launches = sim.step_timing()
    rate = n * steps / wall
    out = {"spins": n, "steps": steps, "value": rate, "unit": "spin-steps/s",
           "kernel_ms_per_launch": kernel_ms / max(1, launches), "launches": launches,
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


def waveform_bench(n_spins, n_wave, sample, chunk=256):
    """Waveform-phase path: n_spins spins on the small substrate, n_wave random
    piecewise-constant waveforms (|G| <= 80 mT/m, a new vector every 10 us) over
    12,000 steps in chunks of `chunk` steps; the phase GEMM against the HBM peak;
    per-waveform error against FP64 on a `sample`-spin run of the same
    substrate and waveforms."""
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
    wf = PH.Waveforms(sim, grads, pc, chunk_steps=chunk)
    t0 = time.perf_counter()
    wf.run(steps)
    sig = wf.signal()
    wall = time.perf_counter() - t0
    walk_ms, gemm_ms = wf.timing()
    wf.close()
    sim.close()
    flops = 3 * 2.0 * n_spins * n_wave * 3 * steps   # hi*hi + hi*lo + lo*hi
    n_chunks = -(-steps // chunk)
    gemm_bytes = n_chunks * (16.0 * n_spins * n_wave) + 4.0 * n_spins * 3 * steps
    log(f"waveforms: {n_spins} spins x {steps} steps x {n_wave} waveforms in {wall:.2f} s "
        f"(walk {walk_ms / 1e3:.2f} s, phase GEMM {gemm_ms / 1e3:.2f} s = "
        f"{flops / (gemm_ms / 1e3) / 1e12:.0f} TFLOP/s); b=0-free signal range "
        f"{sig[:, 0].min():.4f}-{sig[:, 0].max():.4f}")
    # FP64 check on a sample: the same product against a float64 GEMM of the
    # recorded trajectory (numpy, checker only)
    sim = prepared(sample, 4)
    wf = PH.Waveforms(sim, grads, pc, chunk_steps=chunk)
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
            "gemm_k": 3 * chunk, "substrate": cfg.name, "seconds": wall,
            "spin_steps_per_s": n_spins * steps / wall, "walk_s": walk_ms / 1e3,
            "phase_gemm_s": gemm_ms / 1e3,
            # algorithmic bytes: the read-modify-write of the (phase,
            # compensation) pairs, 8 B per spin and waveform each way, once per
            # chunk, plus one read of the chunk's trajectory (4 B per spin and k)
            "phase_gemm_roofline": {"bound": "hbm", "achieved": gemm_bytes / (gemm_ms / 1e3) / 1e9,
                                    "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                                    "frac": gemm_bytes / (gemm_ms / 1e3) / PEAK_HBM,
                                    "traffic": committed_gemm_traffic(n_spins, n_wave, chunk),
                                    "traffic_unit": "bytes per full chunk",
                                    "mfma_tflops": flops / (gemm_ms / 1e3) / 1e12,
                                    "mfma_frac_of_bf16_peak": flops / (gemm_ms / 1e3) / PEAK_BF16},
            "fp64_check": {"spins": sample, "phase_err_max_rad": float(dphi.max()),
                           "phase_err_median_rad": float(np.median(dphi)),
                           "signal_err_max": float(dsig.max()),
                           "signal_err_median": float(np.median(dsig))}}


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


def committed_gemm_traffic(n_spins, n_wave, chunk):
    """HBM bytes per full-chunk launch of k_phase_gemm from the committed
    FETCH_SIZE / WRITE_SIZE summary for the matching sizes; else None."""
    try:
        lines = open(GEMM_TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        if ln.startswith("traffic_bytes_per_chunk"):
            f = dict(kv.split("=") for kv in ln.split()[1:])
            if (int(f["spins"]) == n_spins and int(f["waveforms"]) == n_wave
                    and int(f["chunk"]) == chunk):
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
    ap.add_argument("--served-requests", type=int, default=120,
                    help="requests in the served-form comparison (0 skips it)")
    ap.add_argument("--plain-requests", type=int, default=40)
    ap.add_argument("--round-steps", type=int, default=256,
                    help="steps per batched round of the served form")
    ap.add_argument("--waveform-spins", type=int, default=4_000_000,
                    help="spins of the waveform-phase measurement (0 skips it)")
    ap.add_argument("--waveforms", type=int, default=1024)
    ap.add_argument("--reference-spins", type=int, default=1_000_000,
                    help="spins of the reference-semantics run (0 skips it)")
    ap.add_argument("--chunk-steps", type=int, default=256,
                    help="steps per phase-GEMM chunk of the waveform run (multiple of 32)")
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
        best, allr = AT.tune(cfg, axons, n, sorts=(0, 256, 512, 1024) if args.sort < 0 else (args.sort,),
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
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(env, args.cpu_seconds, threads)
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
    served = None
    if args.served_requests:
        sim.close()
        sub.close()
        served = served_bench(args.served_requests, min(args.plain_requests, args.served_requests),
                              args.round_steps)
    waves = None
    if args.waveform_spins:
        waves = waveform_bench(args.waveform_spins, args.waveforms, 8192, args.chunk_steps)
    refmode = None
    if args.reference_spins:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        refmode = reference_mode_bench(args.reference_spins, 1000, 5.0, threads,
                                       cpu=not args.no_cpu_baseline)
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
        "served": served,
        "waveforms": waves,
        "reference_mode": refmode,
        "host": platform.processor() or platform.machine(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
