"""PGSE through the per-step phase (spinKernelCpu.cpp:490), and what a free
Gaussian walk must give for it.

The walk adds phase_constant * g . x_t after every step t, with g = +G, 0, -G
over the updates of the reference's sequencer (diffusion.cpp:564-585).  For a
free walk x_t = x_0 + xi_1 + ... + xi_t with independent Gaussian steps of
variance sd^2 per axis, and lobes that cancel (sum_t s_t = 0), the phase is
phase_constant * |G| * sum_k xi_k . u F_k with F_k = sum_{t >= k} s_t, so it
is Gaussian with variance v = (phase_constant |G| sd)^2 sum_k F_k^2 and the
signal is exp(-v / 2).  For long lobes this tends to exp(-b D) with the
reference's b (diffusion.cpp:317).
"""
from __future__ import annotations

import numpy as np

GAMMA = 42576.0  # kHz/T (sim.cfg gyroMagneticRatio, spinSystem.cpp:1938)


def step_signs(schedule, steps_per_update):
    """Per-step gradient sign from the per-update schedule."""
    return np.repeat(np.asarray(schedule, np.float64), steps_per_update)


def walk_attenuation(signs, grad_scaled_norm, phase_constant, sd):
    """Signal of a free Gaussian walk under the step signs, and the phase
    variance (both exact for Gaussian steps)."""
    s = np.asarray(signs, np.float64)
    F = np.cumsum(s[::-1])[::-1]
    v = (float(phase_constant) * float(grad_scaled_norm) * float(sd)) ** 2 * float(np.sum(F * F))
    return float(np.exp(-0.5 * v)), v


def mc_sigma(v, n):
    """Standard error of |mean exp(i phase)| over n spins, phase ~ N(0, v)."""
    ec = np.exp(-0.5 * v)
    var = 0.5 * (1.0 + np.exp(-2.0 * v)) - ec * ec
    return float(np.sqrt(max(var, 0.0) / n))


def b_value(G_mT_m, delta, Delta):
    """diffusion.cpp:317 (ms/um^2 for G in mT/m, times in ms)."""
    return (2.0 * 3.1415926535 * 42576.0 * G_mT_m / 1e9 * delta) ** 2 * (Delta - delta / 3.0)


def run_updates(schedule, steps_per_update, advance):
    """Calls advance(sign, steps) once per update, merging equal neighbours
    (the state after the last update is the readout state)."""
    k = 0
    sched = list(schedule)
    while k < len(sched):
        j = k
        while j < len(sched) and sched[j] == sched[k]:
            j += 1
        advance(int(sched[k]), (j - k) * steps_per_update)
        k = j
