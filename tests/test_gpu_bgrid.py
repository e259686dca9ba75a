"""Gridded magnetic field on the device (m2d_fields2d_set_magnetic_field,
m2d_fields2d_B, the B-grid variant of the particle pass) against the oracle's
Fields::B with magnetic_field_const = 0 (fields.hpp:171-176) and
advance_boris with the per-particle field (particles.cpp:946-975).

The device keeps Br, Bz as FP32 tables with FP64 index arithmetic, like the
E gather; tolerances are those of the magnetised push tests."""
import ctypes as C

import numpy as np
import pytest

from m2d_testlib import M_E, QE, Scene, ions_on_disk, species
from mag2d_amd import abi
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L = 2e-2


def bottle(n=41, lo=-1e-3, hi=21e-3, B0=0.1, Lb=8e-3):
    """Bz = B0 (1 + ((z - c)/Lb)^2), Br = -B0 (x - c)(z - c) / Lb^2 on an
    n x n grid from lo to hi on both axes (its own origin and spacing)."""
    h = (hi - lo) / (n - 1)
    r = lo + h * np.arange(n)
    R, Z = np.meshgrid(r, r, indexing="ij")
    c = 0.5 * L
    bz = B0 * (1 + ((Z - c) / Lb) ** 2)
    br = -B0 * (R - c) * (Z - c) / Lb ** 2
    return abi.grid2d(n, n, h, h, lo, lo), O.grid2(n, n, h, h, lo, lo), br.ravel(), bz.ravel()


def caps(g):
    mask, volt = abi.grid_build(g, abi.GEO_EMPTY, L, L, extern_field=0.0)
    abi.square_electrode(g, mask, volt, 2e-3, 1.8e-2, -1, 2e-3, -10.0)
    abi.square_electrode(g, mask, volt, 2e-3, 1.8e-2, 1.8e-2, 1.0, -10.0)
    return mask, volt


def test_B_query_matches_oracle(ctx):
    sc = Scene(ctx, n=64, geometry=abi.GEO_EMPTY, rf=False, solve_oracle=False)
    g, og, br, bz = bottle()
    sc.fields.set_magnetic_field(g, br, bz)
    O.set_b_grid(sc.of, og, br, bz)
    rng = np.random.default_rng(31)
    xs, zs = rng.uniform(0, L, 3000), rng.uniform(0, L, 3000)
    gr, gz, gt = sc.fields.B(xs, zs)
    want = np.zeros((3, len(xs)))
    for k in range(len(xs)):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        assert O.lib().o_fields_B(C.byref(sc.of), xs[k], zs[k], C.byref(a), C.byref(b), C.byref(c)) == 0
        want[:, k] = a.value, b.value, c.value
    scale = np.abs(bz).max()
    assert np.abs(gr - want[0]).max() < 1e-6 * scale
    assert np.abs(gz - want[1]).max() < 1e-6 * scale
    assert np.all(gt == 0.0)
    assert ctx.status() == 0
    # back to the constant field of the params
    sc.fields.set_magnetic_field(None, None, None)
    gr, gz, gt = sc.fields.B(xs[:4], zs[:4])
    assert np.all(gr == 0.0) and np.all(gz == 0.0)


def test_push_in_gridded_B_matches_oracle(ctx):
    """Electrons in a magnetic bottle between end caps at -10 V, 100 steps."""
    sc = Scene(ctx, n=128, custom=caps, rf=False)
    g, og, br, bz = bottle()
    sc.fields.set_magnetic_field(g, br, bz)
    O.set_b_grid(sc.of, og, br, bz)
    rng = np.random.default_rng(32)
    x, z, vx, vy, vz = ions_on_disk(3000, rng, r=4e-3, T_eV=1.0, mass=M_E)
    P = abi.ParticlesSoA.from_host(ctx, x, z, vx, vy, vz)
    op = O.particles(len(x))
    op["x"], op["z"], op["vx"], op["vy"], op["vz"] = x, z, vx, vy, vz
    os_ = O.make_species(-QE, M_E, 1e-11)
    abi.advance_init(ctx, sc.fields, species(-QE, M_E, 1e-11, float("inf"), 0), P)
    O.lib().o_advance_boris_init(C.byref(sc.of), C.byref(os_), O.dptr(op), len(op))
    steps = 100
    for st in range(steps):
        abi.advance(ctx, sc.fields, species(-QE, M_E, 1e-11, float("inf"), st), P)
        os_.niter = st
        O.lib().o_advance_boris(C.byref(sc.of), C.byref(os_), O.dptr(op), len(op), None)
        O.lib().o_advance_boundary(C.byref(sc.of), C.byref(os_), O.dptr(op), len(op), None)
    assert ctx.status() == 0
    alive, oalive = P.alive(), op["empty"] == 0
    assert np.sum(alive != oalive) <= 2 and alive.sum() > 0.5 * len(x)
    both = alive & oalive
    d = P.download()
    tol = steps * float(np.spacing(np.float32(L)))
    assert np.abs(d["x"][both] - op["x"][both]).max() < tol
    assert np.abs(d["z"][both] - op["z"][both]).max() < tol
    vs = np.sqrt(op["vx"] ** 2 + op["vy"] ** 2 + op["vz"] ** 2)[both].max()
    for k in ("vx", "vy", "vz"):
        assert np.abs(d[k][both] - op[k][both]).max() < 1e-4 * vs, k


def test_uniform_B_grid_matches_constant_B(ctx):
    """A grid of one value gives the constant-field pass's trajectories."""
    Bz = 0.1
    sc_c = Scene(ctx, n=64, custom=caps, rf=False, Bz=Bz, solve_oracle=False)
    sc_g = Scene(ctx, n=64, custom=caps, rf=False, solve_oracle=False)
    g = abi.grid2d(5, 5, L / 4, L / 4)
    sc_g.fields.set_magnetic_field(g, np.zeros(25), np.full(25, Bz))
    rng = np.random.default_rng(33)
    x, z, vx, vy, vz = ions_on_disk(2000, rng, r=4e-3, T_eV=1.0, mass=M_E)
    out = []
    for sc in (sc_c, sc_g):
        P = abi.ParticlesSoA.from_host(ctx, x, z, vx, vy, vz)
        for st in range(50):
            abi.advance(ctx, sc.fields, species(-QE, M_E, 1e-11, float("inf"), st), P)
        out.append(P.download())
    for k in ("x", "z"):
        assert np.abs(out[0][k] - out[1][k]).max() < 50 * float(np.spacing(np.float32(L)))
    vs = np.abs(vx).max() * 4
    for k in ("vx", "vy", "vz"):
        assert np.abs(out[0][k] - out[1][k]).max() < 1e-4 * vs


def test_particle_outside_the_B_grid_is_flagged(ctx):
    sc = Scene(ctx, n=64, geometry=abi.GEO_EMPTY, rf=False, solve_oracle=False)
    g = abi.grid2d(5, 5, 2.5e-3, 2.5e-3)  # covers [0, 1 cm]^2 of the 2 cm box
    sc.fields.set_magnetic_field(g, np.zeros(25), np.full(25, 0.05))
    ctx.status(clear=True)
    x = np.array([5e-3, 15e-3], np.float32)
    zero = np.zeros(2, np.float32)
    P = abi.ParticlesSoA.from_host(ctx, x, x, zero + 1e3, zero, zero)
    abi.advance(ctx, sc.fields, species(QE, 40 * 1.66e-27, 1e-9, float("inf"), 0), P)
    assert ctx.status() & abi.FLAG_OUT_OF_RANGE
    with pytest.raises(abi.M2DError, match="garbage loaded"):
        sc.fields.set_magnetic_field(g, np.full(25, np.nan), np.zeros(25))


def test_multicoll_rejects_a_B_grid(ctx):
    sc = Scene(ctx, n=64, geometry=abi.GEO_EMPTY, rf=False, solve_oracle=False)
    sc.fields.set_magnetic_field(abi.grid2d(5, 5, 5e-3, 5e-3), np.zeros(25), np.zeros(25))
    x = np.full(64, 0.01, np.float32)
    P = abi.ParticlesSoA.from_host(ctx, x, x, x * 0, x * 0, x * 0)
    ttd = ctx.to_device(np.ones(64))
    with pytest.raises(abi.M2DError, match="zero B only"):
        abi.advance_multicoll(ctx, sc.fields, species(QE, 40 * 1.66e-27, 1e-9, 1e-9, 0), P, ttd.ptr)
