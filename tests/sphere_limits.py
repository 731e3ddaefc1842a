"""The exact added mass of the floating hemisphere (radius 1, infinite depth) at the radiation limits,
for the tests of omega = 0 and omega = infinity.

At omega = 0 the free surface is a rigid lid (d phi / dz = 0 on z = 0) and the potential extends
evenly to the whole sphere; at omega = infinity phi = 0 on z = 0 and it extends oddly. Either way
it solves the exterior Neumann problem of the unit sphere with the extended normal velocity g
(d phi / dr = g on r = 1): with mu = cos(theta) and the azimuthal order m (0 heave, 1 surge with a
cos(phi) factor), g = sum_n g_n P_n^m(mu) and phi = sum_n -g_n / (n + 1) r^-(n + 1) P_n^m(mu).
The added mass is -rho times the integral of n_i phi over the wetted lower half.

Heave: n_3 = mu on the lower half; g = -|mu| (even), mu (odd). Surge: n_1 = sin(theta) cos(phi);
g = sin(theta) (even), -sign(mu) sin(theta) (odd). Heave at infinity and surge at zero are the
translating sphere, whose added mass is half the displaced mass: 0.5 of the hemisphere's m0."""

import math

import numpy as np
import scipy.special as sp


def _gauss(n, a, b):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (b - a) * x + 0.5 * (a + b), 0.5 * (b - a) * w


def hemisphere_added_mass(mode, limit, terms=240, points=400):
    """Added mass / m0 (m0 = 2 pi / 3: rho = 1, radius 1) of mode "heave" or "surge" at limit
    "zero" or "infinity"."""
    m = 0 if mode == "heave" else 1
    lo, wlo = _gauss(points, -1.0, 0.0)  # the kink of the extended velocity at mu = 0 between rules
    hi, whi = _gauss(points, 0.0, 1.0)
    mu, wq = np.concatenate([lo, hi]), np.concatenate([wlo, whi])
    s = np.sqrt(1.0 - mu * mu)
    if m == 0:
        n_i = mu
        g = -np.abs(mu) if limit == "zero" else mu
    else:
        n_i = s
        g = s if limit == "zero" else -np.sign(mu) * s
    phi = np.zeros_like(mu)
    for n in range(m, terms):
        P = sp.lpmv(m, n, mu)
        norm = 2.0 / (2 * n + 1) * (1 if m == 0 else n * (n + 1))  # (n + m)! / (n - m)!, m = 0, 1
        gn = np.sum(wq * g * P) / norm
        phi -= gn / (n + 1) * P
    wet = mu < 0.0
    azimuth = 2.0 * math.pi if m == 0 else math.pi  # the integral of 1 or cos^2 over the azimuth
    return -azimuth * np.sum((wq * n_i * phi)[wet]) / (2.0 * math.pi / 3.0)
