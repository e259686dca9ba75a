"""Pin the oracle's 3D path (Field3D.hpp:40-163, species3d.cpp:3-94,
fields3d.cpp:35-101) by closed forms: the reference's 3D species does not
compile (species3d.hpp:12-14) and its solver needs UMFPACK, so only Field3D
accumulate/interpolate run against the reference (tests/test_oracle_golden.py)."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

QE = O.QE
M_ION = 40 * 1.66053906660e-27


def box3(n=17, L=1e-2):
    g = O.grid3(n, n, n, L / (n - 1), L / (n - 1), L / (n - 1))
    mask = np.full((n, n, n), O.FREE, np.uint8)
    mask[0, :, :] = mask[-1, :, :] = mask[:, 0, :] = mask[:, -1, :] = mask[:, :, 0] = mask[:, :, -1] = O.FIXED
    return g, mask.ravel()


def nodes(g):
    i, j, k = np.meshgrid(np.arange(g.nx), np.arange(g.ny), np.arange(g.nz), indexing="ij")
    return i * g.dx, j * g.dy, k * g.dz


def one(x, y, z, vx, vy, vz):
    p = O.particles(1)
    p["x"], p["y"], p["z"], p["vx"], p["vy"], p["vz"] = x, y, z, vx, vy, vz
    return p


def test_f3_grad_of_linear_potential_is_exact_including_clamped_edges():
    g, _ = box3(13)
    X, Y, Z = nodes(g)
    u = (2.0 * X - 3.0 * Y + 5.0 * Z).ravel()
    rng = np.random.default_rng(1)
    L = g.xmax
    pts = np.concatenate([rng.uniform(0, L, (300, 3)), [[0, 0, 0], [L, L, L], [0, L, 0.3 * L]]])
    for x, y, z in pts:
        gx, gy, gz = C.c_double(), C.c_double(), C.c_double()
        O.lib().o_f3_grad(C.byref(g), O.dptr(u), x, y, z, C.byref(gx), C.byref(gy), C.byref(gz))
        assert (gx.value, gy.value, gz.value) == pytest.approx((2.0, -3.0, 5.0), rel=1e-9)


def test_f3_grad_of_quadratic_is_the_staggered_difference_at_nodes():
    # u = x^2: at node x_i the x component interpolates the differences at
    # i - 1/2 and i + 1/2 equally: ((x_i^2 - x_{i-1}^2) + (x_{i+1}^2 - x_i^2)) / (2 dx) = 2 x_i
    g, _ = box3(13)
    X, _, _ = nodes(g)
    u = (X ** 2).ravel()
    for i in range(2, g.nx - 2):
        x = i * g.dx
        gx, gy, gz = C.c_double(), C.c_double(), C.c_double()
        O.lib().o_f3_grad(C.byref(g), O.dptr(u), x, 0.4 * g.ymax, 0.6 * g.zmax, C.byref(gx), C.byref(gy),
                          C.byref(gz))
        assert gx.value == pytest.approx(2 * x, rel=1e-9)
        assert gy.value == 0 and gz.value == 0


def test_poisson3_operator_eigenfunction():
    n = 15
    g, mask = box3(n)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    a, b, c = 2, 3, 1
    f = (np.sin(a * np.pi * i / (n - 1)) * np.sin(b * np.pi * j / (n - 1)) * np.sin(c * np.pi * k / (n - 1)))
    f = f.ravel()
    y = np.zeros_like(f)
    assert O.lib().o_poisson3_apply(C.byref(g), O.dptr(mask), O.dptr(f), O.dptr(y)) == 0
    lam = sum(2 * (np.cos(m * np.pi / (n - 1)) - 1) for m in (a, b, c))
    free = mask == O.FREE
    np.testing.assert_allclose(y[free], lam * f[free], atol=1e-12)
    np.testing.assert_array_equal(y[~free], f[~free])


def test_is_free3_brute_force():
    n = 11
    g, mask = box3(n)
    m3 = mask.reshape(n, n, n)
    m3[4:8, 3:9, 5:9] = O.FIXED
    rng = np.random.default_rng(2)
    for x, y, z in rng.uniform(-0.05 * g.xmax, 1.05 * g.xmax, (2000, 3)):
        i, j, k = int(np.floor(x / g.dx)), int(np.floor(y / g.dy)), int(np.floor(z / g.dz))
        ok = min(x, y, z) >= 0 and max(i, j, k) <= n - 2
        want = ok and (m3[i:i + 2, j:j + 2, k:k + 2] == O.FREE).any()
        assert O.lib().o_is_free3(C.byref(g), O.dptr(mask), x, y, z) == int(want)


def test_advance3d_rotation_handedness_and_angle():
    g, mask = box3()
    u = np.zeros(g.nx * g.ny * g.nz)
    B, dt = 0.5, 1e-9
    f = O.make_fields3(g, mask, u, B=(0.0, 0.0, B))
    s = O.make_species(QE, M_ION, dt)
    p = one(5e-3, 5e-3, 5e-3, 1000.0, 0.0, 0.0)
    O.lib().o_advance3d(C.byref(f), C.byref(s), O.dptr(p), 1, None, None)
    theta = 2 * np.arctan(QE * B * dt / (2 * M_ION))
    # species3d.cpp:38-50 with only tz: vx' = vx - (vy + vx tz) sz, vy' = vy + (vx - vy tz) sz,
    # so (1000, 0, 0) turns towards +vy
    assert p["vx"][0] == pytest.approx(1000.0 * np.cos(theta), rel=1e-12)
    assert p["vy"][0] == pytest.approx(1000.0 * np.sin(theta), rel=1e-12)
    assert p["vz"][0] == 0.0
    assert p["x"][0] == pytest.approx(5e-3 + p["vx"][0] * dt, rel=1e-15)


def test_advance3d_free_acceleration_in_linear_potential():
    g, mask = box3()
    X, Y, Z = nodes(g)
    u = (100.0 * X + 40.0 * Y - 70.0 * Z).ravel()  # E = (-100, -40, 70)
    dt = 1e-9
    f = O.make_fields3(g, mask, u)
    s = O.make_species(QE, M_ION, dt)
    p = one(4e-3, 5e-3, 6e-3, 10.0, -20.0, 30.0)
    O.lib().o_advance3d(C.byref(f), C.byref(s), O.dptr(p), 1, None, None)
    a = QE / M_ION * np.array([-100.0, -40.0, 70.0])
    v1 = np.array([10.0, -20.0, 30.0]) + a * dt
    np.testing.assert_allclose([p["vx"][0], p["vy"][0], p["vz"][0]], v1, rtol=1e-9)
    np.testing.assert_allclose([p["x"][0], p["y"][0], p["z"][0]],
                               np.array([4e-3, 5e-3, 6e-3]) + v1 * dt, rtol=1e-12)


def test_advance3d_init_then_step_keeps_speed_in_pure_B():
    g, mask = box3()
    u = np.zeros(g.nx * g.ny * g.nz)
    f = O.make_fields3(g, mask, u, B=(0.2, -0.3, 0.4))
    s = O.make_species(QE, M_ION, 1e-8)
    p = one(5e-3, 5e-3, 5e-3, 300.0, -200.0, 100.0)
    O.lib().o_advance3d_init(C.byref(f), C.byref(s), O.dptr(p), 1)
    for _ in range(50):
        O.lib().o_advance3d(C.byref(f), C.byref(s), O.dptr(p), 1, None, None)
    sp = np.sqrt(p["vx"][0] ** 2 + p["vy"][0] ** 2 + p["vz"][0] ** 2)
    assert sp == pytest.approx(np.sqrt(300.0 ** 2 + 200.0 ** 2 + 100.0 ** 2), rel=1e-12)


def test_advance3d_boundary_free_periodic_and_deposit():
    n = 13
    g, mask = box3(n)
    m3 = mask.reshape(n, n, n)
    m3[5:8, 5:8, 5:8] = O.FIXED
    u = np.zeros(n ** 3)
    rng = np.random.default_rng(3)
    N = 3000
    L = g.xmax
    for boundary in (0, 1):
        f = O.make_fields3(g, mask, u, boundary=boundary)
        p = O.particles(N)
        for c in ("x", "y", "z"):
            p[c] = rng.uniform(-0.3 * L, 1.3 * L, N)
        x0, y0, z0 = p["x"].copy(), p["y"].copy(), p["z"].copy()
        rho = np.zeros(n ** 3)
        s = O.make_species(-QE, M_ION, 1e-9, rho=rho)
        O.lib().o_advance3d(C.byref(f), C.byref(s), O.dptr(p), N, None, None)
        out = (x0 < 0) | (x0 > L) | (y0 < 0) | (y0 > L) | (z0 < 0) | (z0 > L)

        def mod(a):  # mymath.cpp:62-67
            return np.where((a >= 0) & (a <= L), a, a - L * np.trunc(a / L) + np.where(a < 0, L, 0))
        if boundary == 0:
            xs, ys, zs, cand = x0, y0, z0, ~out
        else:
            xs, ys, zs, cand = mod(x0), mod(y0), mod(z0), np.ones(N, bool)
        free = np.array([O.lib().o_is_free3(C.byref(g), O.dptr(mask), a, b, c) for a, b, c in zip(xs, ys, zs)],
                        bool)
        keep = cand & free
        np.testing.assert_array_equal(p["empty"] == 0, keep)
        np.testing.assert_array_equal(p["x"][keep], xs[keep])
        assert rho.sum() == pytest.approx(-QE * keep.sum(), rel=1e-12)
