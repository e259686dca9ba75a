"""The on-disk formats of the host layer: Field2D's text format (print and
load, Field2D.cpp:30-130), the particle checkpoint with BaseSpecies::load's
re-draw of positions outside the box (particles.cpp:57-93), and a Pic run
started from files as its constructor reads them (pic.cpp:130-188).

CPU: print is byte-identical to the reference's Field2D::print and load
returns what the reference's Field2D::load returns (golden fixtures written
by the reference's own code, tests/golden/make_golden.py), including the
files it refuses; the checkpoint re-draw takes the oracle's Philox draws.
GPU: a run from loaded u, uRF and particles matches the oracle stepping the
same loaded state."""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

from mag2d_amd import abi
from oracle import oracle as O
from test_host_shim import host

G = Path(__file__).parent / "golden"
_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
PURPOSE_RELOAD = 5


class FileCfg(C.Structure):
    _fields_ = [("grid", C.c_int32), ("steps", C.c_int32), ("geometry", C.c_int32), ("rf", C.c_int32),
                ("length", C.c_double), ("dt", C.c_double), ("rf_amplitude", C.c_double),
                ("rf_omega", C.c_double), ("mass", C.c_double), ("charge", C.c_double), ("seed", C.c_uint64)]


def hostf():
    L = host()
    L.m2dh_field_text_write.restype = C.c_int
    L.m2dh_field_text_write.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                                        _d, C.c_double, C.c_char_p, C.c_char_p, C.c_int]
    L.m2dh_field_text_read.restype = C.c_int
    L.m2dh_field_text_read.argtypes = [C.c_char_p, _d, _d, C.c_int64, C.c_char_p, C.c_int]
    L.m2dh_particles_load.restype = C.c_int
    L.m2dh_particles_load.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                      C.c_uint32, C.c_void_p, C.c_int64, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64), C.c_char_p, C.c_int]
    L.m2dh_param_check.restype = C.c_int
    L.m2dh_param_check.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_char_p, C.c_int]
    L.m2dh_file_run.restype = C.c_int
    L.m2dh_file_run.argtypes = [C.POINTER(FileCfg), C.c_char_p, C.c_char_p, C.c_char_p, _f, _f, _f, _f, _f,
                                C.c_int64, C.POINTER(C.c_int64), C.c_char_p, C.c_int]
    return L


def dptr(a):
    return a.ctypes.data_as(_d)


def text_write(path, data, nx, nz, dx, dz, x0=0.0, z0=0.0, factor=1.0):
    msg = C.create_string_buffer(256)
    data = np.ascontiguousarray(data, np.float64)
    rc = hostf().m2dh_field_text_write(nx, nz, dx, dz, x0, z0, dptr(data), factor, str(path).encode(), msg, 256)
    assert rc == 0, msg.value.decode()


def text_read(path, cap=4096):
    meta, data = np.zeros(6), np.zeros(cap)
    msg = C.create_string_buffer(256)
    rc = hostf().m2dh_field_text_read(str(path).encode(), dptr(meta), dptr(data), cap, msg, 256)
    return rc, meta, data, msg.value.decode()


def test_print_is_byte_identical_to_the_reference(tmp_path):
    d = np.load(G / "field2d_text_ref.npz")
    nx, nz = (int(v) for v in d["print_shape"])
    dx, dz, x0, z0 = d["print_grid"]
    out = tmp_path / "u.txt"
    text_write(out, d["print_data"], nx, nz, dx, dz, x0, z0, float(d["print_factor"]))
    assert out.read_bytes() == (G / "field2d_print_ref.txt").read_bytes()


def test_load_matches_the_reference_and_refuses_what_it_refuses():
    d = np.load(G / "field2d_text_ref.npz")
    for k, name in enumerate("abcd"):
        rc, meta, data, msg = text_read(G / f"field2d_load_{name}.txt")
        assert (rc == 0) == (d["load_rc"][k] == 0), (name, msg)
        if rc == 0:
            # grid and values bit for bit: the same reader arithmetic on the same text
            assert np.array_equal(meta, d["load_meta"][k]), name
            n = int(meta[0] * meta[1])
            assert np.array_equal(data[:n], d["load_data"][k][:n]), name
    # the two orders of the same nodes give the same field
    assert np.array_equal(d["load_data"][0], d["load_data"][1])


def test_print_load_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(7)
    nx, nz, dx, dz = 13, 9, 1e-3 / 3, 2.5e-4
    u = rng.normal(size=nx * nz)
    f = tmp_path / "u.txt"
    text_write(f, u, nx, nz, dx, dz, 1e-3, 0.0)
    rc, meta, data, msg = text_read(f)
    assert rc == 0, msg
    assert meta[0] == nx and meta[1] == nz
    # print keeps 6 significant digits of every number: the spacing, the
    # difference of the first two coordinates, is good to 5e-6 of their sum
    assert abs(meta[2] - dx) <= 5e-6 * (2e-3 + dx) and abs(meta[3] - dz) <= 5e-6 * dz
    assert meta[4] == pytest.approx(1e-3, rel=5e-6) and meta[5] == 0.0
    np.testing.assert_allclose(data[:nx * nz], u, rtol=5e-6, atol=1e-300)
    # a file that is not a grid: the reference throws, so does the host
    (tmp_path / "bad.txt").write_text("0 0 1\n1 0 2\n2.5 0 3\n0 1 4\n1 1 5\n2.5 1 6\n")
    assert text_read(tmp_path / "bad.txt")[0] == -1
    assert text_read(tmp_path / "missing.txt")[0] == -1


def test_param_check():
    msg = C.create_string_buffer(256)
    L = hostf()
    assert L.m2dh_param_check(0, 1, 1, msg, 256) == 0
    assert L.m2dh_param_check(1, 0, 1, msg, 256) == -1
    assert "electric_field_from_file" in msg.value.decode()
    assert L.m2dh_param_check(1, 1, 0, msg, 256) == -1


def write_checkpoint(path, parts, size):
    """The reference's checkpoint (particles.cpp:31-55): int size, int n_live,
    then the live t_particle records."""
    with open(path, "wb") as fw:
        fw.write(np.array([size, len(parts)], np.int32).tobytes())
        fw.write(parts.tobytes())


def oracle_remap(parts, box, seed, stream):
    """particles.cpp:77-83 with the library's uniforms: word 0 of the Philox
    block (seed, stream, slot, step 0, purpose 5) for x, word 1 for z."""
    x0, x1, z0, z1 = box
    out = parts.copy()
    r = O.Rng()
    for i in range(len(out)):
        ox = out["x"][i] < x0 or out["x"][i] > x1
        oz = out["z"][i] < z0 or out["z"][i] > z1
        if not (ox or oz):
            continue
        O.lib().o_rng_philox_begin(C.byref(r), seed, stream, i, 0, PURPOSE_RELOAD)
        ux, uz = O.lib().o_rng_uni(C.byref(r)), O.lib().o_rng_uni(C.byref(r))
        if ox:
            out["x"][i] = (x1 - x0) * ux + x0
        if oz:
            out["z"][i] = (z1 - z0) * uz + z0
    return out


def synthetic_parts(rng, n, L, vscale, frac_out=0.08):
    p = np.zeros(n, O.PARTICLE_DTYPE)
    p["x"] = rng.uniform(0.2 * L, 0.8 * L, n)
    p["z"] = rng.uniform(0.2 * L, 0.8 * L, n)
    p["y"] = rng.normal(size=n)
    for k in ("vx", "vy", "vz"):
        p[k] = rng.normal(size=n) * vscale
    p["time_to_death"] = rng.exponential(1e-6, n)
    out_x = rng.uniform(size=n) < frac_out
    out_z = rng.uniform(size=n) < frac_out
    p["x"][out_x] = np.where(rng.uniform(size=out_x.sum()) < 0.5, -1e-4, L + 2e-4)
    p["z"][out_z] = np.where(rng.uniform(size=out_z.sum()) < 0.5, -3e-4, L + 1e-4)
    return p, out_x | out_z


def test_checkpoint_load_redraws_positions_outside_the_box(tmp_path):
    rng = np.random.default_rng(11)
    L, n, size = 1e-2, 500, 640
    parts, moved = synthetic_parts(rng, n, L, 100.0)
    f = tmp_path / "particles_ION.dat"
    write_checkpoint(f, parts, size)
    out = np.zeros(size, O.PARTICLE_DTYPE)
    ns, nl = C.c_int64(), C.c_int64()
    msg = C.create_string_buffer(256)
    rc = hostf().m2dh_particles_load(str(f).encode(), 0.0, L, 0.0, L, 1234, 3, out.ctypes.data, size,
                                     C.byref(ns), C.byref(nl), msg, 256)
    assert rc == 0, msg.value.decode()
    assert ns.value == size and nl.value == n
    want = oracle_remap(parts, (0.0, L, 0.0, L), 1234, 3)
    assert moved.sum() > 40
    for k in ("x", "y", "z", "vx", "vy", "vz", "time_to_death"):
        assert np.array_equal(out[k][:n], want[k]), k
    assert np.all(out["empty"][:n] == 0) and np.all(out["empty"][n:] == 1)
    inside = (out["x"][:n] >= 0) & (out["x"][:n] <= L) & (out["z"][:n] >= 0) & (out["z"][:n] <= L)
    assert inside.all()
    # a truncated file: the host throws where the reference prints "read error"
    f.write_bytes(f.read_bytes()[:8 + 64 * 10 + 5])
    assert hostf().m2dh_particles_load(str(f).encode(), 0.0, L, 0.0, L, 1, 0, out.ctypes.data, size,
                                       C.byref(ns), C.byref(nl), msg, 256) == -1


@pytest.mark.gpu
def test_pic_run_from_files_matches_oracle(tmp_path):
    """Pic with electric_field_from_file and particle_reload (RF_QUAD
    geometry, rf on, not self-consistent): u and uRF come from Field2D::print
    text, the particles from a checkpoint with 8% of them outside the box;
    no solve runs.  The oracle steps the same loaded state (the same text
    parsed by numpy, the same re-draws) with the same Boris and boundary
    steps."""
    n_grid, Lb = 64, 2e-2
    dx = Lb / (n_grid - 1)
    x = np.arange(n_grid) * dx
    X, Z = np.meshgrid(x, x, indexing="ij")
    c = 0.5 * Lb
    u = 2.0 * ((X - c) ** 2 - (Z - c) ** 2) / c ** 2 + 0.3 * np.sin(X / Lb * 3.0)
    urf = ((X - c) ** 2 - (Z - c) ** 2) / c ** 2
    uf, rf_f = tmp_path / "u.txt", tmp_path / "urf.txt"
    text_write(uf, u.ravel(), n_grid, n_grid, dx, dx)
    text_write(rf_f, urf.ravel(), n_grid, n_grid, dx, dx)
    rng = np.random.default_rng(12)
    n = 4000
    parts, moved = synthetic_parts(rng, n, Lb, 300.0)
    write_checkpoint(tmp_path / "particles_ION.dat", parts, n + 100)
    mass, amp, omega, dt, steps, seed = 40 * 1.66053906660e-27, 30.0, 2 * math.pi * 17e6, 2e-9, 40, 99
    cfg = FileCfg(grid=n_grid, steps=steps, geometry=abi.GEO_RF_QUAD, rf=1, length=Lb, dt=dt, rf_amplitude=amp,
                  rf_omega=omega, mass=mass, charge=abi.QE, seed=seed)
    out = {k: np.zeros(n, np.float32) for k in ("x", "z", "vx", "vy", "vz")}
    nl = C.c_int64()
    msg = C.create_string_buffer(512)
    fp = lambda a: a.ctypes.data_as(_f)  # noqa: E731
    rc = hostf().m2dh_file_run(C.byref(cfg), str(uf).encode(), str(rf_f).encode(), str(tmp_path).encode(),
                               fp(out["x"]), fp(out["z"]), fp(out["vx"]), fp(out["vy"]), fp(out["vz"]), n,
                               C.byref(nl), msg, 512)
    assert rc == 0, msg.value.decode()

    # the oracle: the loaded state
    uo = np.loadtxt(uf)[:, 2].copy()
    urfo = np.loadtxt(rf_f)[:, 2].copy()
    assert np.abs(uo - u.ravel()).max() <= 5e-6 * np.abs(u).max()
    p0 = oracle_remap(parts, (0.0, Lb, 0.0, Lb), seed, 0)
    og = O.grid2(n_grid, n_grid, dx, dx)
    of = O.make_fields(og, np.zeros(n_grid * n_grid, np.uint8), uo, urfo, x_max=Lb, z_max=Lb,
                       electric_field_from_file=1, rf=1, rf_amplitude=amp, rf_omega=omega)
    p = O.particles(n)
    for k in ("x", "y", "z", "vx", "vy", "vz", "time_to_death"):
        p[k] = p0[k]
    p["empty"] = 0
    s = O.make_species(abi.QE, mass, dt)
    # positions are FP32 on the device
    p["x"] = p["x"].astype(np.float32)
    p["z"] = p["z"].astype(np.float32)
    for k in ("vx", "vy", "vz"):
        p[k] = p[k].astype(np.float32)
    O.lib().o_advance_boris_init(C.byref(of), C.byref(s), O.dptr(p), n)
    for k in range(steps):
        s.niter = k
        O.lib().o_advance_boris(C.byref(of), C.byref(s), O.dptr(p), n, None)
        O.lib().o_advance_boundary(C.byref(of), C.byref(s), O.dptr(p), n, None)
    alive = p["empty"] == 0
    k = int(alive.sum())
    assert nl.value == k and k > 0.5 * n, (nl.value, k)
    pos_tol = 4 * steps * np.spacing(np.float32(Lb))
    assert np.abs(out["x"][:k] - p["x"][alive]).max() <= pos_tol
    assert np.abs(out["z"][:k] - p["z"][alive]).max() <= pos_tol
    vscale = np.abs(p["vx"][alive]).max()
    for c_ in ("vx", "vy", "vz"):
        assert np.abs(out[c_][:k] - p[c_][alive]).max() <= 1e-4 * vscale, c_
    # the particles moved: the loaded fields acted
    assert np.abs(out["vx"][:k] - p0["vx"][alive]).max() > 10.0
