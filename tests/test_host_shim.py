"""The C++ host layer (mag2d_amd/host/mag2d.hpp): the reference's class API
(BaseSpecies container, Species, Field2D, Fields, Pic) over the C ABI.

CPU: the library loads, and the slot container keeps the reference's
insert/remove/resize semantics and checkpoint format (particles.hpp:223-249,
particles.cpp:31-107).  GPU: a self-consistent trap run through
Pic::advance_init / Pic::advance matches the oracle's restatement of the same
step ordering (pic.cpp:330-384)."""
import ctypes as C
import math
import os
from pathlib import Path

import numpy as np
import pytest

from mag2d_amd import abi
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent
# M2D_HOST_LIB: another build of the same source (the sanitizer build of
# tests/test_sanitizers.py)
HOST_LIB = Path(os.environ.get("M2D_HOST_LIB", ROOT / "mag2d_amd" / "host" / "libmag2d_host.so"))


class TrapCfg(C.Structure):
    _fields_ = [("grid", C.c_int32), ("n_particles", C.c_int32), ("steps", C.c_int32),
                ("selfconsistent", C.c_int32), ("length", C.c_double), ("dt", C.c_double),
                ("Bz", C.c_double), ("end_cap_V", C.c_double), ("temperature_K", C.c_double),
                ("mass", C.c_double), ("charge", C.c_double), ("radius", C.c_double),
                ("density_total", C.c_double), ("seed", C.c_uint64), ("u_smooth", C.c_int32)]


class MccCfg(C.Structure):
    _fields_ = [("grid", C.c_int32), ("n_particles", C.c_int32), ("steps", C.c_int32), ("mover", C.c_int32),
                ("length", C.c_double), ("dt", C.c_double), ("extern_field", C.c_double),
                ("temperature_K", C.c_double), ("mass", C.c_double), ("charge", C.c_double),
                ("radius", C.c_double), ("gas_mass", C.c_double), ("gas_temperature_K", C.c_double),
                ("gas_density", C.c_double), ("rate", C.c_double), ("seed", C.c_uint64)]


ADVANCE_BORIS, ADVANCE_MULTICOLL = 0, 1


def host():
    abi.lib()  # libmag2d_amd first: the host layer links against it
    L = C.CDLL(str(HOST_LIB))
    L.m2dh_list_selftest.restype = C.c_int
    L.m2dh_list_selftest.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    _f = C.POINTER(C.c_float)
    L.m2dh_trap_run.restype = C.c_int
    L.m2dh_trap_run.argtypes = [C.POINTER(TrapCfg), _f, _f, _f, _f, _f, C.c_int64, C.POINTER(C.c_double),
                                C.c_char_p, C.c_int]
    L.m2dh_trap_run_bfile.restype = C.c_int
    L.m2dh_trap_run_bfile.argtypes = [C.POINTER(TrapCfg), C.c_char_p, _f, _f, _f, _f, _f, C.c_int64,
                                      C.POINTER(C.c_double), C.c_char_p, C.c_int]
    L.m2dh_initscript_run.restype = C.c_int
    L.m2dh_initscript_run.argtypes = [C.c_char_p, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                      _f, _f, _f, _f, _f, C.c_int64, C.POINTER(C.c_int64), C.c_char_p, C.c_int]
    L.m2dh_mcc_run.restype = C.c_int
    L.m2dh_mcc_run.argtypes = [C.POINTER(MccCfg), _f, _f, _f, _f, _f, C.POINTER(C.c_double), C.c_int64,
                               C.POINTER(C.c_double), C.c_char_p, C.c_int]
    return L


def test_host_library_loads_and_exports():
    L = host()
    for name in ("m2dh_list_selftest", "m2dh_trap_run", "m2dh_trap_run_bfile", "m2dh_mcc_run",
                 "m2dh_initscript_run"):
        assert hasattr(L, name)


def test_particle_container_semantics_and_checkpoint(tmp_path):
    msg = C.create_string_buffer(256)
    rc = host().m2dh_list_selftest(str(tmp_path / "species.bin").encode(), msg, 256)
    assert rc == 0, msg.value.decode()
    # the checkpoint is the reference's layout: int size, int n_live, 64-byte records
    raw = (tmp_path / "species.bin").read_bytes()
    size, live = np.frombuffer(raw[:8], np.int32)
    assert len(raw) == 8 + 64 * live and live <= size


def oracle_trap(cfg, bgrid=None):
    """Pic<CARTESIAN> advance_init + advance (pic.cpp:330-384) in the oracle;
    bgrid = (Grid2, Br, Bz) for magnetic_field_const = 0."""
    n, L = cfg.grid, cfg.length
    dx = L / (n - 1)
    og = O.grid2(n, n, dx, dx)
    mask = np.zeros(n * n, np.uint8)
    volt = np.zeros(n * n)
    O.lib().o_grid_build(C.byref(og), C.byref(O.Geom(abi.GEO_EMPTY, L, L, 1e-3, 0.0, 0.0)), O.dptr(mask),
                         O.dptr(volt))
    h = 1.5 * dx
    O.lib().o_square_electrode(C.byref(og), O.dptr(mask), O.dptr(volt), 0.0, L, 0.0, h, cfg.end_cap_V, 0)
    O.lib().o_square_electrode(C.byref(og), O.dptr(mask), O.dptr(volt), 0.0, L, L - h, L, cfg.end_cap_V, 0)
    dV = (cfg.n_particles / cfg.density_total) / ((n - 1) ** 2)
    u = np.zeros(n * n)
    of = O.make_fields(og, mask, u, x_max=L, z_max=L, geometry_empty=1, selfconsistent=cfg.selfconsistent,
                       Bz=cfg.Bz)
    u = of._keep[1]  # the array the oracle fields read
    if bgrid is not None:
        O.set_b_grid(of, *bgrid)

    def solve(rho_f):
        b = rho_f.copy()
        O.lib().o_poisson_rhs(C.byref(og), O.dptr(mask), O.dptr(volt), O.dptr(b), O.SOLVE_STATIC, dV)
        out = np.zeros_like(b)
        assert O.lib().o_poisson_solve(C.byref(og), O.dptr(mask), O.dptr(b), O.dptr(out)) == 0
        if cfg.u_smooth:  # Fields::u_smooth after the solve (pic.cpp:336, 372)
            assert O.lib().o_u_smooth(C.byref(og), O.dptr(out), 0, -1.0) == 0
        u[:] = out

    vth = math.sqrt(abi.KB * cfg.temperature_K / cfg.mass)
    p = O.particles(cfg.n_particles)
    O.lib().o_add_particles_on_disk_philox(O.dptr(p), cfg.n_particles, 0.5 * L, 0.5 * L, cfg.radius, vth,
                                           C.byref(of), cfg.seed, 0)
    rho_s = np.zeros(n * n)
    rho_f = np.zeros(n * n)
    s = O.make_species(cfg.charge, cfg.mass, cfg.dt, rho=rho_s)
    N = cfg.n_particles
    if cfg.selfconsistent:
        O.lib().o_accumulate(C.byref(of), C.byref(s), O.dptr(p), N)
        rho_f += rho_s
        solve(rho_f)
        rho_f[:] = 0
        rho_f += rho_s
    O.lib().o_advance_boris_init(C.byref(of), C.byref(s), O.dptr(p), N)
    for k in range(cfg.steps):
        if cfg.selfconsistent:
            solve(rho_f)
            rho_f[:] = 0
            rho_s[:] = 0
        s.niter = k
        O.lib().o_advance_boris(C.byref(of), C.byref(s), O.dptr(p), N, None)
        O.lib().o_advance_boundary(C.byref(of), C.byref(s), O.dptr(p), N, None)
        if cfg.selfconsistent:
            rho_f += rho_s
    return p, rho_f, u


@pytest.mark.gpu
@pytest.mark.parametrize("smooth", [0, 1])
def test_pic_selfconsistent_trap_matches_oracle(smooth):
    """smooth = 1: Param::u_smooth, the potential filtered after every solve."""
    cfg = TrapCfg(grid=64, n_particles=3000, steps=25, selfconsistent=1, length=2e-2, dt=1e-11, Bz=0.1,
                  end_cap_V=-10.0, temperature_K=11604.5, mass=9.11e-31, charge=-abi.QE, radius=5e-3,
                  density_total=1e13, seed=1234, u_smooth=smooth)
    n = cfg.n_particles
    out = {k: np.zeros(n, np.float32) for k in ("x", "z", "vx", "vy", "vz")}
    m = (C.c_double * 6)()
    msg = C.create_string_buffer(512)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    rc = host().m2dh_trap_run(C.byref(cfg), fp(out["x"]), fp(out["z"]), fp(out["vx"]), fp(out["vy"]),
                              fp(out["vz"]), n, m, msg, 512)
    assert rc == 0, msg.value.decode()
    p, rho_f, u = oracle_trap(cfg)
    alive = p["empty"] == 0
    k = int(alive.sum())
    assert int(m[1]) == k
    assert k > 0.5 * n
    steps, L = cfg.steps, cfg.length
    pos_tol = 4 * steps * np.spacing(np.float32(L))
    assert np.abs(out["x"][:k] - p["x"][alive]).max() <= pos_tol
    assert np.abs(out["z"][:k] - p["z"][alive]).max() <= pos_tol
    vscale = np.abs(p["vx"][alive]).max()
    for c in ("vx", "vy", "vz"):
        assert np.abs(out[c][:k] - p[c][alive]).max() <= 1e-4 * vscale, c
    # charge on the grid is exactly the live particles' charge (fixed-point deposit)
    assert math.isclose(m[2], cfg.charge * k, rel_tol=1e-12)
    assert math.isclose(m[2], rho_f.sum(), rel_tol=1e-12)
    # the last solve reached the FP64 tolerance; potential at the centre as the oracle's
    assert m[4] <= 1e-10
    c = (cfg.grid // 2) * cfg.grid + cfg.grid // 2
    assert abs(m[3] - u[c]) <= 1e-6 * max(1.0, abs(u).max())
    e_or = O.lib().o_mean_energy_eV(O.dptr(p), n, cfg.mass)
    assert math.isclose(m[0], e_or, rel_tol=1e-4)


def oracle_mcc(cfg):
    """The same run in the oracle: samples of the Philox disk sampler, every
    time_to_death 0 (the particles come after lifetime_init, as the
    initscript's do), then advance_multicoll + advance_boundary per step with
    niter advancing by two (particles.cpp:858-859 and particles.hpp:346-347)."""
    n, L = cfg.grid, cfg.length
    dx = L / (n - 1)
    og = O.grid2(n, n, dx, dx)
    mask = np.zeros(n * n, np.uint8)
    volt = np.zeros(n * n)
    O.lib().o_grid_build(C.byref(og), C.byref(O.Geom(abi.GEO_EMPTY, L, L, 1e-3, 0.0, cfg.extern_field)),
                         O.dptr(mask), O.dptr(volt))
    of = O.make_fields(og, mask, np.zeros(n * n), x_max=L, z_max=L, geometry_empty=1,
                       extern_field=cfg.extern_field)
    vth = math.sqrt(abi.KB * cfg.temperature_K / cfg.mass)
    p = O.particles(cfg.n_particles)
    O.lib().o_add_particles_on_disk_philox(O.dptr(p), cfg.n_particles, 0.5 * L, 0.5 * L, cfg.radius, vth,
                                           C.byref(of), cfg.seed, 0)
    E_max = cfg.temperature_K * abi.KB / abi.QE * 10.0
    mcc = O.make_mcc(cfg.mass, E_max, [dict(mass=cfg.gas_mass, density=cfg.gas_density,
                                            temperature=cfg.gas_temperature_K,
                                            inters=[O.inter(O.ELASTIC, cfg.mass, cfg.gas_mass, rate=cfg.rate)])])
    s = O.make_species(cfg.charge, cfg.mass, cfg.dt, lifetime=mcc.lifetime, mcc=mcc, seed=cfg.seed, stream=0)
    N = cfg.n_particles
    events = 0
    for k in range(cfg.steps):
        s.niter = 2 * k
        ne = C.c_long(0)
        assert O.lib().o_advance_multicoll(C.byref(of), C.byref(s), O.dptr(p), N, C.byref(ne)) == 0
        O.lib().o_advance_boundary(C.byref(of), C.byref(s), O.dptr(p), N, None)
        events += ne.value
    return p, mcc.lifetime, events


@pytest.mark.gpu
def test_pic_multicoll_run_matches_oracle():
    """Param::mover = ADVANCE_MULTICOLL through Pic::advance: the same events,
    velocities and positions as the oracle's restatement."""
    cfg = MccCfg(grid=64, n_particles=4000, steps=12, mover=ADVANCE_MULTICOLL, length=2e-2, dt=1e-9,
                 extern_field=1000.0, temperature_K=1160.45, mass=3 * 1.66053906660e-27, charge=abi.QE,
                 radius=5e-3, gas_mass=6.69e-27, gas_temperature_K=300.0, gas_density=1e24, rate=1e-15,
                 seed=1234)
    n = cfg.n_particles
    out = {k: np.zeros(n, np.float32) for k in ("x", "z", "vx", "vy", "vz")}
    ttd = np.zeros(n)
    m = (C.c_double * 4)()
    msg = C.create_string_buffer(512)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    rc = host().m2dh_mcc_run(C.byref(cfg), fp(out["x"]), fp(out["z"]), fp(out["vx"]), fp(out["vy"]),
                             fp(out["vz"]), ttd.ctypes.data_as(C.POINTER(C.c_double)), n, m, msg, 512)
    assert rc == 0, msg.value.decode()
    p, lifetime, events = oracle_mcc(cfg)
    assert m[3] == pytest.approx(lifetime, rel=1e-14)
    assert int(m[2]) == 2 * cfg.steps
    # every particle also has an event at t = 0 of its first step (time_to_death 0)
    assert events == pytest.approx(n + cfg.steps * n * cfg.dt / lifetime, rel=0.05)
    alive = p["empty"] == 0
    k = int(alive.sum())
    assert int(m[1]) == k and k > 0.9 * n
    np.testing.assert_allclose(ttd[:k], p["time_to_death"][alive], rtol=1e-12, atol=1e-24)
    vscale = np.abs(p["vx"][alive]).max()
    for c in ("vx", "vy", "vz"):
        assert np.abs(out[c][:k] - p[c][alive]).max() <= 1e-4 * vscale, c
    for c in ("x", "z"):
        assert np.abs(out[c][:k] - p[c][alive]).max() <= 4 * cfg.steps * np.spacing(np.float32(cfg.length)), c


@pytest.mark.gpu
def test_initscript_samplers_match_oracle(tmp_path):
    """add_particles_bessel / add_particles_everywhere lines of an initscript
    (pic.cpp:241-326) through Species: each call draws from its own stream
    (species stream ^ call << 16), as the oracle's Philox samplers do."""
    script = tmp_path / "initscript.txt"
    script.write_text("# samplers\nadd_particles_bessel ION 3000 0.004 0.01 0.005\n"
                      "add_particles_everywhere ION 2000\n")
    n_out = 5000
    out = {k: np.zeros(n_out, np.float32) for k in ("x", "z", "vx", "vy", "vz")}
    live = C.c_int64(0)
    msg = C.create_string_buffer(512)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    mass, T, L, grid = 3 * 1.66053906660e-27, 1160.45, 2e-2, 64
    rc = host().m2dh_initscript_run(str(script).encode(), grid, L, mass, T, 1234, fp(out["x"]), fp(out["z"]),
                                    fp(out["vx"]), fp(out["vy"]), fp(out["vz"]), n_out, C.byref(live), msg, 512)
    assert rc == 0, msg.value.decode()
    dx = L / (grid - 1)
    of = O.make_fields(O.grid2(grid, grid, dx, dx), np.full(grid * grid, O.FREE, np.uint8), np.zeros(grid * grid),
                       x_max=L, z_max=L, geometry_empty=1)
    vth = math.sqrt(abi.KB * T / mass)
    pb = O.particles(3000)
    O.lib().o_add_particles_bessel_philox(O.dptr(pb), 3000, 0.004, 0.01, 0.005, vth, C.byref(of), 1234, 0)
    pe = O.particles(2000)
    O.lib().o_add_particles_everywhere_philox(O.dptr(pe), 2000, vth, float("inf"), C.byref(of), 1234, 1 << 16)
    want = np.concatenate([pb[pb["empty"] == 0], pe])
    assert live.value == len(want) and live.value < 5000
    k = live.value
    for c in ("x", "z"):
        np.testing.assert_allclose(out[c][:k], want[c], rtol=1e-7)
    for c in ("vx", "vy", "vz"):
        np.testing.assert_allclose(out[c][:k], want[c], rtol=1e-6, atol=1e-6 * vth)


@pytest.mark.gpu
def test_pic_trap_with_magnetic_field_file_matches_oracle(tmp_path):
    """magnetic_field_const = 0: Pic loads B from a file of "r z Br Bz" lines
    (Fields::load_magnetic_field, fields.cpp:870-959; rows written with r
    descending, which the reader turns round) and the push takes B per
    particle (Fields::B, fields.hpp:171-176)."""
    cfg = TrapCfg(grid=64, n_particles=3000, steps=25, selfconsistent=1, length=2e-2, dt=1e-11, Bz=0.0,
                  end_cap_V=-10.0, temperature_K=11604.5, mass=9.11e-31, charge=-abi.QE, radius=5e-3,
                  density_total=1e13, seed=1234)
    nb, lo, hi = 26, -1e-3, 2.4e-2
    h = (hi - lo) / (nb - 1)
    r = lo + h * np.arange(nb)
    R, Z = np.meshgrid(r, r, indexing="ij")
    bz = 0.1 * (1 + ((Z - 0.01) / 8e-3) ** 2)
    br = -0.1 * (R - 0.01) * (Z - 0.01) / 8e-3 ** 2
    lines = [f"{R[i, j]:.17g} {Z[i, j]:.17g} {br[i, j]:.17g} {bz[i, j]:.17g}"
             for i in reversed(range(nb)) for j in range(nb)]
    bfile = tmp_path / "B.txt"
    bfile.write_text("# r z Br Bz\n" + "\n".join(lines) + "\n")
    n = cfg.n_particles
    out = {k: np.zeros(n, np.float32) for k in ("x", "z", "vx", "vy", "vz")}
    m = (C.c_double * 6)()
    msg = C.create_string_buffer(512)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    rc = host().m2dh_trap_run_bfile(C.byref(cfg), str(bfile).encode(), fp(out["x"]), fp(out["z"]), fp(out["vx"]),
                                    fp(out["vy"]), fp(out["vz"]), n, m, msg, 512)
    assert rc == 0, msg.value.decode()
    p, rho_f, u = oracle_trap(cfg, (O.grid2(nb, nb, h, h, lo, lo), br.ravel(), bz.ravel()))
    alive = p["empty"] == 0
    k = int(alive.sum())
    assert int(m[1]) == k and k > 0.5 * n
    pos_tol = 4 * cfg.steps * np.spacing(np.float32(cfg.length))
    assert np.abs(out["x"][:k] - p["x"][alive]).max() <= pos_tol
    assert np.abs(out["z"][:k] - p["z"][alive]).max() <= pos_tol
    vscale = np.abs(p["vx"][alive]).max()
    for c in ("vx", "vy", "vz"):
        assert np.abs(out[c][:k] - p[c][alive]).max() <= 1e-4 * vscale, c
