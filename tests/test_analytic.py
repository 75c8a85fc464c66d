"""Pins of the closed-form physics references (tests/analytic.py) against
independent limits, so that the GPU physics tests compare with a formula
that is itself checked."""
import numpy as np

import analytic as A


def test_cylinder_gpa_long_pulse_limit_is_neuman():
    # delta >> R^2/D: ln S = -(7/96) (gamma G)^2 R^4 (2 delta - 99 R^2 / (112 D)) / D
    # (Neuman, J Chem Phys 60:4508, 1974)
    for R, G, D in ((1.0, 40.0, 2.0), (0.5, 100.0, 1.0), (2.0, 10.0, 3.0)):
        delta, Delta = 400 * R * R / D, 600 * R * R / D
        got = A.cylinder_gpa_log_signal(G, delta, Delta, D, R)
        g = A.gamma_g(G)
        want = -(7 / 96) * g * g * R ** 4 * (2 * delta - 99 * R * R / (112 * D)) / D
        assert abs(got / want - 1) < 2e-3, (R, G, D, got, want)


def test_cylinder_gpa_approaches_free_diffusion_from_below():
    # restriction only lowers the attenuation, and less so for wider cylinders;
    # short-time surface correction ~ (4/(3 sqrt(pi))) (S/V) sqrt(D Delta) / 2
    G, delta, Delta, D = 40.0, 2.0, 5.0, 2.0
    free = A.free_log_signal(G, delta, Delta, D)
    prev = 0.0
    for R in (10.0, 25.0, 50.0, 100.0, 200.0):
        v = A.cylinder_gpa_log_signal(G, delta, Delta, D, R, n_roots=4000)
        assert free < v < prev, (R, v, free)
        rel = 1 - v / free
        mitra = 4 / (3 * np.sqrt(np.pi)) * (2 / R) * np.sqrt(D * Delta) / 2
        assert rel < 2 * mitra, (R, rel, mitra)
        prev = v


def test_free_signal_matches_reference_b_value_formula():
    # b of diffusion.cpp:317: (2 pi gamma G delta)^2 (Delta - delta/3), in ms/um^2
    G, delta, Delta = 61.7, 10.0, 40.0
    b = (2 * np.pi * 42576.0 * G * 1e-9 * delta) ** 2 * (Delta - delta / 3)
    assert abs(A.free_log_signal(G, delta, Delta, 1.0) + b) < 1e-12
    assert abs(b - 1.0) < 0.01  # 61.7 mT/m at 10/40 ms is the b = 1 ms/um^2 shell


def test_polygon_radius():
    assert abs(A.polygon_area_radius(1.0, 4) - np.sqrt(2 / np.pi)) < 1e-12
    assert 0.999 < A.polygon_area_radius(1.0, 64) < 1.0
