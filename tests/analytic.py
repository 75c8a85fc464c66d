"""Closed-form diffusion-MRI signals used as physics references by the tests.

Units: D in um^2/ms, lengths in um, times in ms, gradients in mT/m;
gamma/2pi = 42576 Hz/mT, so gamma*G = 2*pi*42576*G*1e-9 in 1/(um ms)."""
import numpy as np
from scipy.special import jnp_zeros

GAMMA = 42576.0


def gamma_g(G):
    return 2 * np.pi * GAMMA * np.asarray(G, np.float64) * 1e-9


def free_log_signal(G, delta, Delta, D):
    """PGSE, free diffusion: ln S = -b D, b = (gamma G delta)^2 (Delta - delta/3)."""
    return -(gamma_g(G) * delta) ** 2 * (Delta - delta / 3) * D


def cylinder_gpa_log_signal(G, delta, Delta, D, R, n_roots=200):
    """PGSE, gradient across an impermeable cylinder of radius R, Gaussian phase
    approximation (van Gelderen et al., J Magn Reson B 103:255, 1994):

      ln S = -2 (gamma G)^2 sum_m [2 D a^2 delta - 2 + 2 e^{-D a^2 delta}
              + 2 e^{-D a^2 Delta} - e^{-D a^2 (Delta - delta)}
              - e^{-D a^2 (Delta + delta)}] / [D^2 a^6 (R^2 a^2 - 1)],

    a = x_m / R with x_m the roots of J1'."""
    a2 = (jnp_zeros(1, n_roots) / R) ** 2
    x = D * a2
    num = (2 * x * delta - 2 + 2 * np.exp(-x * delta) + 2 * np.exp(-x * Delta)
           - np.exp(-x * (Delta - delta)) - np.exp(-x * (Delta + delta)))
    g = gamma_g(G)
    return -2 * g * g * np.sum(num / (D * D * a2 ** 3 * (R * R * a2 - 1)))


def polygon_area_radius(R, n):
    """Radius of the circle with the area of the regular n-gon inscribed in a
    circle of radius R."""
    return R * np.sqrt(n / (2 * np.pi) * np.sin(2 * np.pi / n))
