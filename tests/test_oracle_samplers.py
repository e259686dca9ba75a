"""Pin the oracle's Philox samplers add_particles_bessel (particles.cpp:701-726)
and add_particles_everywhere (particles.cpp:686-699) by the distributions the
reference's code defines:

- Bessel: density J0(k r / R) on the disk (k the first zero of J0), so the
  radius has the CDF r J1(k r/R) / (R J1(k)) and the angle is uniform;
- everywhere: x, z uniform in the box, Maxwellian components of thermal speed
  vth, time_to_death = lifetime * Exp(1).

A slip in the rejection test (uni() < J0 instead of >, r^2 for r) or in a
scale fails the Kolmogorov-Smirnov tests below.
"""
import ctypes as C

import numpy as np
from scipy import special, stats

from oracle import oracle as O

K0 = 2.404825557695773


def box(L=2e-2, n=32):
    g = O.grid2(n, n, L / (n - 1), L / (n - 1))
    mask = np.full(n * n, O.FREE, np.uint8)
    return O.make_fields(g, mask, np.zeros(n * n), x_max=L, z_max=L, geometry_empty=1)


def test_bessel_sampler_radial_law():
    f = box()
    n, R, c, vth = 20000, 5e-3, 1e-2, 300.0
    p = O.particles(n)
    O.lib().o_add_particles_bessel_philox(O.dptr(p), n, c, c, R, vth, C.byref(f), 1234, 3)
    assert (p["empty"] == 0).all()
    r = np.hypot(p["x"] - c, p["z"] - c) / R
    assert r.max() <= 1.0
    assert stats.kstest(r, lambda s: s * special.j1(K0 * s) / special.j1(K0)).pvalue > 1e-3
    ang = np.arctan2(p["z"] - c, p["x"] - c)
    assert stats.kstest((ang + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3
    for k in ("vx", "vy", "vz"):
        assert stats.kstest(p[k] / vth, "norm").pvalue > 1e-3


def test_bessel_sampler_drops_samples_outside_the_box():
    f = box()
    n, R = 5000, 5e-3
    p = O.particles(n)
    O.lib().o_add_particles_bessel_philox(O.dptr(p), n, 1e-3, 1e-2, R, 300.0, C.byref(f), 1234, 4)
    live = p["empty"] == 0
    assert 0.1 * n < (~live).sum() < 0.6 * n
    assert p["x"][live].min() >= 0.0


def test_everywhere_sampler_is_uniform_in_the_box():
    L = 2e-2
    f = box(L)
    n, vth, life = 20000, 450.0, 3e-9
    p = O.particles(n)
    O.lib().o_add_particles_everywhere_philox(O.dptr(p), n, vth, life, C.byref(f), 1234, 5)
    assert (p["empty"] == 0).all()
    for k in ("x", "z"):
        assert stats.kstest(p[k] / L, "uniform").pvalue > 1e-3
    for k in ("vx", "vy", "vz"):
        assert stats.kstest(p[k] / vth, "norm").pvalue > 1e-3
    assert stats.kstest(p["time_to_death"] / life, "expon").pvalue > 1e-3
