"""Served workload: a seeded stream of simulation requests over resident
substrates, run through the server (dsim_amd.serve) and, for comparison, one
at a time through the plain simulation handle.

Each request draws its substrate, spin count, readout lines, b-value and PGSE
timing (delta, Delta) from the stream configuration; the substrates are
z-aligned axon packings whose axon count and mean radius are drawn too.
"""
from __future__ import annotations

import dataclasses
import time
from dataclasses import dataclass

import numpy as np

from . import lib as D
from . import serve as SV
from . import substrate as S

GAMMA = 42576.0


@dataclass
class StreamConfig:
    n_substrates: int = 4
    axons: tuple = (50, 600)        # log-uniform axon count per substrate
    radius: tuple = (0.5, 3.0)      # uniform mean axon radius, um
    ngon: int = 32
    spins: tuple = (1e4, 2e5)       # log-uniform spins per request
    lines: tuple = (6, 64)          # uniform readout lines per request
    b: tuple = (0.5, 5.0)           # uniform b-value, ms/um^2
    delta: tuple = (5.0, 20.0)      # uniform, ms
    Delta: tuple = (20.0, 60.0)     # uniform, ms (at least delta + 1 ms)
    n_requests: int = 120
    D: float = 2.0
    dt: float = 0.005
    kappa: float = 0.01
    t2_intra: float = 70.0
    t2_extra: float = 90.0
    cell_sigma: float = 1.4         # grid cell in units of the step sigma (autotuned)
    seed: int = 5


def make_substrates(sc: StreamConfig, log=None):
    rng = np.random.default_rng(sc.seed)
    envs = []
    for k in range(sc.n_substrates):
        n_ax = int(round(np.exp(rng.uniform(*np.log(sc.axons)))))
        r = float(rng.uniform(*sc.radius))
        cfg = dataclasses.replace(S.SCALE, name=f"served-{k}", n_spins=0, n_axons=n_ax,
                                  mean_radius=r, ngon=sc.ngon, D=sc.D, dt=sc.dt,
                                  t2_intra=sc.t2_intra, t2_extra=sc.t2_extra, kappa=sc.kappa,
                                  seed=sc.seed * 1000 + k)
        sigma_um = np.sqrt(2 * sc.D * sc.dt)
        t0 = time.perf_counter()
        env = S.build(cfg, cell_um=sc.cell_sigma * sigma_um, axons=S.pack(cfg))
        g = env["grid"]
        env["sub"] = D.Substrate(env["vertices"], env["triangles"], env["tri_info"], grid="aabb",
                                 dims=g["dims"], lo=g["lo"], cell=g["cell"], eps=g["eps"],
                                 device=True)
        if log:
            log(f"served substrate {k}: {n_ax} axons, mean radius {r:.2f} um, "
                f"{len(env['triangles'])} triangles, grid {g['dims']} "
                f"({time.perf_counter() - t0:.1f} s)")
        envs.append(env)
    return envs


def make_requests(sc: StreamConfig, envs, rate=None):
    """Requests with arrival times (Poisson at `rate` per second, or all at 0).
    The requests themselves do not depend on the rate."""
    rng = np.random.default_rng(sc.seed + 1)
    arrivals = np.random.default_rng(sc.seed + 2)
    reqs = []
    t = 0.0
    for k in range(sc.n_requests):
        if rate:
            t += arrivals.exponential(1.0 / rate)
        e = int(rng.integers(len(envs)))
        env = envs[e]
        n = int(round(np.exp(rng.uniform(*np.log(sc.spins)))))
        nl = int(rng.integers(sc.lines[0], sc.lines[1] + 1))
        b = float(rng.uniform(*sc.b))
        delta = float(rng.uniform(*sc.delta))
        Delta = float(rng.uniform(max(sc.Delta[0], delta + 1.0), sc.Delta[1]))
        L, B = int(round(delta / sc.dt)), int(round(Delta / sc.dt))
        G = S.gradient_amplitude(b, delta, Delta)
        # readout lines: b=0 plus directions on the sphere, randomly turned
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        dirs = S.fibonacci_directions(nl - 1) @ q.T
        grads = np.vstack([np.zeros((1, 3)), dirs * G]) * 1e-9 * env["space_scale"]
        reqs.append(dict(arrival=t, env=e, n=n, seed=sc.seed * 100000 + k, dt=sc.dt,
                         std_dev=env["std_dev"], t2=env["t2"], permeability=env["permeability"],
                         lobes=SV.pgse_lobes(L, B, 0), grads=grads.astype(np.float32),
                         extent=(env["xmax"],) * 3,
                         pc=np.float32(2 * np.pi * GAMMA * sc.dt), steps=2 * L + (B - L)))
    return reqs


def work(reqs):
    """Simulated spin-steps of the requests."""
    return float(sum(q["n"] * q["steps"] for q in reqs))


def _submit(srv, ids, q):
    return srv.submit(ids[q["env"]], q["n"], seed=q["seed"], dt=q["dt"], std_dev=q["std_dev"],
                      t2=q["t2"], permeability=q["permeability"], lobes=q["lobes"],
                      grads=q["grads"], phase_constant=q["pc"], boundary=D.BOUNDARY_PERIODIC,
                      label_extent=q["extent"])


def run_served(envs, reqs, *, hbm_budget=64 << 30, round_steps=256, realtime=True):
    """Runs the stream through a server with its background loop.  Returns
    per-request dicts (latency, queueing, signals) and the makespan."""
    srv = SV.Server(hbm_budget, round_steps=round_steps, max_active_spins=16 << 20)
    ids = [srv.add_substrate(e["sub"]) for e in envs]
    srv.start()
    t0 = time.perf_counter()
    tickets = []
    for q in sorted(reqs, key=lambda q: q["arrival"]):
        if realtime:
            d = t0 + q["arrival"] - time.perf_counter()
            if d > 0:
                time.sleep(d)
        tickets.append((q, _submit(srv, ids, q)))
    out = []
    for q, t in tickets:
        srv.wait(t, 600.0)
        sig, st = srv.result(t)
        out.append(dict(req=q, signals=sig, latency=st["done_s"] - st["submit_s"],
                        queued=st["admit_s"] - st["submit_s"], submit=st["submit_s"],
                        done=st["done_s"], rounds=st["rounds"]))
    srv.stop()
    srv.close()
    makespan = max(o["done"] for o in out) - min(o["submit"] for o in out)
    return out, makespan


def run_plain(envs, reqs):
    """Each request alone through the plain handle (dsim_integrate_moment per
    lobe), one after the other.  Returns per-request (seconds, signals)."""
    out = []
    for q in reqs:
        env = envs[q["env"]]
        t0 = time.perf_counter()
        sim = D.Simulation(env["sub"], q["n"], dt=q["dt"], std_dev=q["std_dev"], t2=q["t2"],
                           permeability=q["permeability"], mode=D.MODE_PHYSICAL,
                           boundary=D.BOUNDARY_PERIODIC, rng=D.RNG_PHILOX, seed=q["seed"])
        sim.init_uniform(1.0)
        sim.label(*q["extent"])
        sim.moment_reset()
        for w, k in q["lobes"]:
            if k:
                sim.integrate_moment(w, k)
        sig = sim.signal_moment(q["grads"], q["pc"])
        sim.close()
        out.append((time.perf_counter() - t0, sig))
    return out
