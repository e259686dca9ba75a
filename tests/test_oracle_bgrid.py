"""Pin the oracle's gridded magnetic field, Fields::B with
magnetic_field_const = 0 (fields.hpp:171-176, Field2D::interpolate,
Field2D.hpp:64-78), as advance_boris uses it per particle
(particles.cpp:946-975):

- a grid holding one value everywhere gives the constant-B trajectories;
- with E = 0 one Boris step turns the velocity about the local B by
  2 atan(q |B(x)| dt / 2m) and keeps the speed; a field linear in x is
  interpolated exactly, so each particle's angle has a closed form.  Br and
  Bz are checked separately: swapping them turns about the wrong axis.
"""
import ctypes as C

import numpy as np

from oracle import oracle as O

QE = O.QE
M_E = 9.11e-31
L = 2e-2


def box(**kw):
    n = 33
    g = O.grid2(n, n, L / (n - 1), L / (n - 1))
    mask = np.full(n * n, O.FREE, np.uint8)
    return O.make_fields(g, mask, np.zeros(n * n), x_max=L, z_max=L, geometry_empty=1, **kw)


def bgrid(fr, fz, n=41, lo=-1e-3, hi=21e-3):
    h = (hi - lo) / (n - 1)
    gB = O.grid2(n, n, h, h, lo, lo)
    r = lo + h * np.arange(n)
    R, Z = np.meshgrid(r, r, indexing="ij")
    return gB, fr(R, Z).ravel(), fz(R, Z).ravel()


def electrons(n, seed):
    rng = np.random.default_rng(seed)
    p = O.particles(n)
    p["x"] = rng.uniform(4e-3, 16e-3, n)
    p["z"] = rng.uniform(4e-3, 16e-3, n)
    for k in ("vx", "vy", "vz"):
        p[k] = rng.normal(size=n) * 3e5
    return p


def run(f, p, steps, dt=1e-11):
    s = O.make_species(-QE, M_E, dt)
    for k in range(steps):
        s.niter = k
        O.lib().o_advance_boris(C.byref(f), C.byref(s), O.dptr(p), len(p), None)


def test_uniform_grid_equals_constant_field():
    Br, Bz = 0.02, 0.07
    fc = box(Br=Br, Bz=Bz)
    fg = box()
    gB, br, bz = bgrid(lambda R, Z: Br + 0 * R, lambda R, Z: Bz + 0 * R)
    O.set_b_grid(fg, gB, br, bz)
    a, b = electrons(300, 1), electrons(300, 1)
    run(fc, a, 40)
    run(fg, b, 40)
    for k in ("x", "z", "vx", "vy", "vz"):
        np.testing.assert_allclose(b[k], a[k], rtol=1e-11, atol=1e-11 * np.abs(a[k]).max())


def _angle(u, w):
    cross = np.linalg.norm(np.cross(u, w), axis=1)
    return np.arctan2(cross, (u * w).sum(1))


def test_local_rotation_angle_follows_the_field():
    dt = 1e-11
    B0, alpha = 0.1, 40.0  # B doubles across the 2.5 cm of the grid
    for comp in ("Bz", "Br"):
        f = box()
        lin = lambda R, Z: B0 * (1 + alpha * R)  # noqa: E731
        zero = lambda R, Z: 0 * R  # noqa: E731
        gB, br, bz = bgrid(lin if comp == "Br" else zero, lin if comp == "Bz" else zero)
        O.set_b_grid(f, gB, br, bz)
        p = electrons(500, 2)
        p0 = p.copy()
        run(f, p, 1, dt)
        Bloc = B0 * (1 + alpha * p0["x"])
        want = 2 * np.arctan(QE * Bloc * dt / (2 * M_E))
        # (x, y, z) ~ (r, theta, z): Bz turns (vx, vy), Br turns (vy, vz)
        if comp == "Bz":
            u, w, fixed = ("vx", "vy"), ("vx", "vy"), "vz"
        else:
            u, w, fixed = ("vy", "vz"), ("vy", "vz"), "vx"
        U = np.stack([p0[u[0]], p0[u[1]], 0 * p0[u[0]]], 1)
        W = np.stack([p[w[0]], p[w[1]], 0 * p[w[0]]], 1)
        np.testing.assert_allclose(_angle(U, W), want, rtol=1e-9)
        np.testing.assert_array_equal(p[fixed], p0[fixed])
        sp0 = np.sqrt(p0["vx"] ** 2 + p0["vy"] ** 2 + p0["vz"] ** 2)
        sp1 = np.sqrt(p["vx"] ** 2 + p["vy"] ** 2 + p["vz"] ** 2)
        np.testing.assert_allclose(sp1, sp0, rtol=1e-13)


def test_field_outside_its_grid_is_reported():
    f = box()
    gB, br, bz = bgrid(lambda R, Z: 0 * R, lambda R, Z: 0.1 + 0 * R, lo=0.0, hi=1e-2)
    O.set_b_grid(f, gB, br, bz)
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    assert O.lib().o_fields_B(C.byref(f), 5e-3, 5e-3, C.byref(a), C.byref(b), C.byref(c)) == 0
    assert b.value == 0.1 and c.value == 0.0
    assert O.lib().o_fields_B(C.byref(f), 15e-3, 5e-3, C.byref(a), C.byref(b), C.byref(c)) == -2
