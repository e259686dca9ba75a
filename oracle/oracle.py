"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker.  The product library
never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# M2D_ORACLE_LIB: another build of the same source (the sanitizer build of
# tests/test_sanitizers.py)
LIB = Path(os.environ.get("M2D_ORACLE_LIB", HERE / "liboracle.so"))
REF_LIB = HERE / "_ref" / "libm2d_ref.so"
REF_LANGEVIN = HERE / "_ref" / "test_langevin"

EPS0, KB, QE = 8.854187817e-12, 1.380662e-23, 1.602189e-19
FIXED, FIXED_RF, FREE, BOUNDARY = 0, 1, 2, 3
ELASTIC, LANGEVIN, CX, COULOMB, SUPERELASTIC = range(5)
RNG_SHR3, RNG_PHILOX = 0, 1
SOLVE_STATIC, SOLVE_RF, SOLVE_RAW = 0, 1, 2


class Grid2(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("dx", C.c_double), ("dz", C.c_double),
                ("idx", C.c_double), ("idz", C.c_double), ("xmin", C.c_double), ("zmin", C.c_double)]


class Grid3(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)] + [
        (k, C.c_double) for k in ("dx", "dy", "dz", "idx", "idy", "idz", "xmin", "ymin", "zmin",
                                  "xmax", "ymax", "zmax")]


class Fields3(C.Structure):
    _fields_ = [("g", Grid3), ("x_max", C.c_double), ("y_max", C.c_double), ("z_max", C.c_double),
                ("boundary", C.c_int), ("Bx", C.c_double), ("By", C.c_double), ("Bz", C.c_double),
                ("u", C.c_void_p), ("mask", C.c_void_p)]


def grid3(nx, ny, nz, dx, dy, dz):
    return Grid3(nx, ny, nz, dx, dy, dz, 1 / dx, 1 / dy, 1 / dz, 0, 0, 0, (nx - 1) * dx, (ny - 1) * dy,
                 (nz - 1) * dz)


def make_fields3(g, mask, u, *, boundary=0, B=(0.0, 0.0, 0.0)):
    keep = [np.ascontiguousarray(mask, np.uint8), np.ascontiguousarray(u, np.float64)]
    f = Fields3(g, g.xmax, g.ymax, g.zmax, boundary, B[0], B[1], B[2], keep[1].ctypes.data,
                keep[0].ctypes.data)
    f._keep = keep
    return f


class Particle(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("vx", C.c_double),
                ("vy", C.c_double), ("vz", C.c_double), ("time_to_death", C.c_double),
                ("empty", C.c_uint8), ("pad_", C.c_uint8 * 7)]


PARTICLE_DTYPE = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8"), ("vx", "f8"), ("vy", "f8"),
                           ("vz", "f8"), ("time_to_death", "f8"), ("empty", "u1"), ("pad_", "u1", 7)])
assert PARTICLE_DTYPE.itemsize == C.sizeof(Particle) == 64


class Rng(C.Structure):
    _fields_ = [("kind", C.c_int), ("jsr", C.c_uint32), ("key", C.c_uint32 * 2),
                ("ctr", C.c_uint32 * 4), ("buf", C.c_uint32 * 4), ("nbuf", C.c_int)]


class Geom(C.Structure):
    _fields_ = [("geometry", C.c_int), ("x_max", C.c_double), ("z_max", C.c_double),
                ("probe_radius", C.c_double), ("u_probe", C.c_double), ("extern_field", C.c_double)]


class Inter(C.Structure):
    _fields_ = [("type", C.c_int), ("DE", C.c_double), ("rate", C.c_double), ("cutoff", C.c_double),
                ("m1", C.c_double), ("m2", C.c_double), ("t1", C.c_double), ("n1", C.c_double),
                ("q1", C.c_double), ("q2", C.c_double), ("ntab", C.c_int),
                ("tabE", C.POINTER(C.c_double)), ("tabS", C.POINTER(C.c_double))]


class Target(C.Structure):
    _fields_ = [("rate", C.c_double), ("density", C.c_double), ("mass", C.c_double),
                ("vth", C.c_double), ("n_inter", C.c_int), ("inter", C.POINTER(Inter)),
                ("parts", C.c_void_p), ("n_parts", C.c_long)]


class Mcc(C.Structure):
    _fields_ = [("n_targets", C.c_int), ("targets", C.POINTER(Target)), ("lifetime", C.c_double),
                ("mass", C.c_double)]


class Fields(C.Structure):
    _fields_ = [("g", Grid2), ("x_min", C.c_double), ("x_max", C.c_double), ("z_min", C.c_double),
                ("z_max", C.c_double), ("geometry_empty", C.c_int), ("selfconsistent", C.c_int),
                ("electric_field_from_file", C.c_int), ("extern_field", C.c_double), ("rf", C.c_int),
                ("rf_amplitude", C.c_double), ("rf_omega", C.c_double), ("rf_U0", C.c_double),
                ("Br", C.c_double), ("Bz", C.c_double), ("Bt", C.c_double), ("boundary", C.c_int),
                ("u", C.c_void_p), ("uRF", C.c_void_p), ("mask", C.c_void_p),
                ("b_grid", C.c_int), ("gB", Grid2), ("BgR", C.c_void_p), ("BgZ", C.c_void_p)]


class Species(C.Structure):
    _fields_ = [("charge", C.c_double), ("mass", C.c_double), ("dt", C.c_double),
                ("lifetime", C.c_double), ("niter", C.c_ulong), ("rho", C.c_void_p),
                ("mcc", C.POINTER(Mcc)), ("rng_kind", C.c_int), ("seed", C.c_uint64),
                ("stream", C.c_uint32), ("slot0", C.c_long), ("shr3", C.POINTER(Rng))]


_PD = C.POINTER(C.c_double)
_P = C.c_void_p
_SIGS = {
    "o_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "o_rng_shr3_seed": (None, [C.POINTER(Rng), C.c_uint32]),
    "o_rng_philox_begin": (None, [C.POINTER(Rng), C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64,
                                  C.c_uint32]),
    "o_rng_u32": (C.c_uint32, [C.POINTER(Rng)]),
    "o_rng_uni": (C.c_double, [C.POINTER(Rng)]),
    "o_trial_uniform": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64]),
    "o_rng_normal": (C.c_double, [C.POINTER(Rng)]),
    "o_rot_len": (None, [C.POINTER(Rng), C.c_double, _PD, _PD, _PD]),
    "o_rot": (None, [C.POINTER(Rng), _PD, _PD, _PD]),
    "o_deflect": (None, [C.POINTER(Rng), C.c_double, _PD, _PD, _PD]),
    "o_radius": (C.c_double, [C.POINTER(Rng)]),
    "o_mod": (C.c_double, [C.c_double, C.c_double]),
    "o_vec_interpolate": (C.c_double, [_PD, _PD, C.c_int, C.c_double]),
    "o_ellint_k": (C.c_double, [C.c_double]),
    "o_langevin_chi": (C.c_double, [C.c_double]),
    "o_f2_accumulate": (C.c_int, [C.POINTER(Grid2), _P, C.c_double, C.c_double, C.c_double]),
    "o_f2_interpolate": (C.c_int, [C.POINTER(Grid2), _P, C.c_double, C.c_double, _PD]),
    "o_f2_grad": (None, [C.POINTER(Grid2), _P, C.c_double, C.c_double, _PD, _PD]),
    "o_grid_build": (None, [C.POINTER(Grid2), C.POINTER(Geom), _P, _P]),
    "o_square_electrode": (None, [C.POINTER(Grid2), _P, _P, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_int]),
    "o_circle_electrode": (None, [C.POINTER(Grid2), _P, _P, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_int]),
    "o_mark_boundary": (None, [C.POINTER(Grid2), _P]),
    "o_is_free": (C.c_int, [C.POINTER(Grid2), _P, C.c_double, C.c_double]),
    "o_poisson_rhs": (None, [C.POINTER(Grid2), _P, _P, _P, C.c_int, C.c_double]),
    "o_poisson_apply": (C.c_int, [C.POINTER(Grid2), _P, _P, _P]),
    "o_poisson_solve": (C.c_int, [C.POINTER(Grid2), _P, _P, _P]),
    "o_sigma_v": (C.c_double, [C.POINTER(Inter), C.c_double]),
    "o_svmax_find": (C.c_double, [C.POINTER(Inter), C.c_int, C.c_double, C.c_int]),
    "o_scatter": (C.c_int, [C.POINTER(Mcc), _P, C.POINTER(Rng)]),
    "o_fields_E": (None, [C.POINTER(Fields), C.c_double, C.c_double, C.c_double, _PD, _PD]),
    "o_fields_B": (C.c_int, [C.POINTER(Fields), C.c_double, C.c_double, _PD, _PD, _PD]),
    "o_u_smooth": (C.c_int, [C.POINTER(Grid2), _P, C.c_int, C.c_double]),
    "o_advance_boris": (None, [C.POINTER(Fields), C.POINTER(Species), _P, C.c_long,
                               C.POINTER(C.c_long)]),
    "o_advance_boris_init": (None, [C.POINTER(Fields), C.POINTER(Species), _P, C.c_long]),
    "o_rng_rexp": (C.c_double, [C.POINTER(Rng)]),
    "o_ttd_init": (None, [C.POINTER(Species), _P, C.c_long]),
    "o_advance_multicoll": (C.c_int, [C.POINTER(Fields), C.POINTER(Species), _P, C.c_long,
                                      C.POINTER(C.c_long)]),
    "o_advance_boundary": (C.c_long, [C.POINTER(Fields), C.POINTER(Species), _P, C.c_long,
                                      C.POINTER(C.c_long)]),
    "o_accumulate": (C.c_long, [C.POINTER(Fields), C.POINTER(Species), _P, C.c_long]),
    "o_mean_energy_eV": (C.c_double, [_P, C.c_long, C.c_double]),
    "o_add_particles_on_disk": (C.c_long, [_P, C.c_long, C.c_long, C.c_double, C.c_double,
                                           C.c_double, C.c_double, C.POINTER(Fields), C.POINTER(Rng)]),
    "o_add_particles_on_disk_philox": (None, [_P, C.c_long, C.c_double, C.c_double, C.c_double,
                                              C.c_double, C.POINTER(Fields), C.c_uint64, C.c_uint32]),
    "o_add_particles_bessel_philox": (None, [_P, C.c_long, C.c_double, C.c_double, C.c_double,
                                             C.c_double, C.POINTER(Fields), C.c_uint64, C.c_uint32]),
    "o_add_particles_everywhere_philox": (None, [_P, C.c_long, C.c_double, C.c_double, C.POINTER(Fields),
                                                 C.c_uint64, C.c_uint32]),
    "o_f3_accumulate": (C.c_int, [C.POINTER(Grid3), _P, C.c_double, C.c_double, C.c_double,
                                  C.c_double]),
    "o_f3_interpolate": (C.c_int, [C.POINTER(Grid3), _P, C.c_double, C.c_double, C.c_double, _PD]),
    "o_f3_grad": (None, [C.POINTER(Grid3), _P, C.c_double, C.c_double, C.c_double, _PD, _PD, _PD]),
    "o_poisson3_apply": (C.c_int, [C.POINTER(Grid3), _P, _P, _P]),
    "o_poisson3_rhs": (None, [C.POINTER(Grid3), _P, _P, _P, C.c_double]),
    "o_selfconsistent_run": (C.c_double, [C.POINTER(Fields), C.POINTER(Species), _P, C.c_long, _P, C.c_double,
                                          C.c_int]),
    "o_is_free3": (C.c_int, [C.POINTER(Grid3), _P, C.c_double, C.c_double, C.c_double]),
    "o_advance3d": (C.c_long, [C.POINTER(Fields3), C.POINTER(Species), _P, C.c_long, C.POINTER(C.c_long),
                               C.POINTER(C.c_long)]),
    "o_advance3d_init": (None, [C.POINTER(Fields3), C.POINTER(Species), _P, C.c_long]),
    "o_bench_steps": (C.c_double, [C.POINTER(Fields), C.POINTER(Species), _P, C.c_long, C.c_int]),
    "o_bench_steps_omp": (C.c_double, [C.POINTER(Fields), C.POINTER(Species), _P, C.c_long, C.c_int,
                                       C.c_int]),
}

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB))
        for k, (res, args) in _SIGS.items():
            f = getattr(L, k)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def ref_available() -> bool:
    return REF_LIB.exists()


def ref():
    """oracle/_ref: the reference's own Field2D/Field3D/t_random/... code."""
    global _ref
    if _ref is None:
        L = C.CDLL(str(REF_LIB))
        L.ref_rng_new.restype = C.c_void_p
        L.ref_rng_new.argtypes = [C.c_uint]
        L.ref_rng_free.argtypes = [C.c_void_p]
        L.ref_rng_radius.restype = C.c_double
        L.ref_rng_radius.argtypes = [C.c_void_p]
        L.ref_mod.restype = C.c_double
        L.ref_mod.argtypes = [C.c_double, C.c_double]
        L.ref_hist_mean_tot.restype = C.c_double
        _ref = L
    return _ref


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def grid2(nx, nz, dx, dz, xmin=0.0, zmin=0.0) -> Grid2:
    return Grid2(nx, nz, dx, dz, 1.0 / dx, 1.0 / dz, xmin, zmin)


def particles(n) -> np.ndarray:
    p = np.zeros(n, PARTICLE_DTYPE)
    return p


def make_fields(g: Grid2, mask, u, urf=None, *, x_max, z_max, geometry_empty=0, selfconsistent=0,
                electric_field_from_file=0, extern_field=0.0, rf=0, rf_amplitude=0.0, rf_omega=0.0,
                rf_U0=0.0, Br=0.0, Bz=0.0, Bt=0.0, boundary=0):
    keep = [np.ascontiguousarray(mask, np.uint8), np.ascontiguousarray(u, np.float64),
            np.ascontiguousarray(urf if urf is not None else np.zeros_like(u), np.float64)]
    f = Fields(g, g.xmin, x_max, g.zmin, z_max, geometry_empty, selfconsistent,
               electric_field_from_file, extern_field, rf, rf_amplitude, rf_omega, rf_U0, Br, Bz, Bt,
               boundary, keep[1].ctypes.data, keep[2].ctypes.data, keep[0].ctypes.data)
    f._keep = keep
    return f


def set_b_grid(f, gB, Br, Bz):
    """Fields::B with magnetic_field_const = 0: Br, Bz (row-major) on grid gB."""
    keep = [np.ascontiguousarray(Br, np.float64), np.ascontiguousarray(Bz, np.float64)]
    f.b_grid = 1
    f.gB = gB
    f.BgR, f.BgZ = keep[0].ctypes.data, keep[1].ctypes.data
    f._keep_b = keep
    return f


def make_species(charge, mass, dt, lifetime=float("inf"), niter=0, rho=None, mcc=None,
                 rng_kind=RNG_PHILOX, seed=1234, stream=0, shr3=None):
    s = Species(charge, mass, dt, lifetime, niter, rho.ctypes.data if rho is not None else None,
                C.pointer(mcc) if mcc is not None else None, rng_kind, seed, stream, 0,
                C.pointer(shr3) if shr3 is not None else None)
    s._keep = (rho, mcc, shr3)
    return s


def make_mcc(mass, E_max_eV, targets):
    """BaseSpecies::lifetime_init (particles.cpp:142-170) for the primary of
    mass `mass`; targets: list of dict(mass, density, temperature,
    inters=[Inter...], parts=None)."""
    vmax = np.sqrt(E_max_eV * QE / mass * 2.0)
    keep = []
    T = (Target * len(targets))()
    rate_total = 0.0
    for k, t in enumerate(targets):
        inters = (Inter * max(1, len(t["inters"])))(*t["inters"])
        keep.append(inters)
        T[k].density = t["density"]
        T[k].mass = t["mass"]
        T[k].vth = np.sqrt(2.0 * KB * t["temperature"] / t["mass"]) * np.sqrt(0.5)
        T[k].n_inter = len(t["inters"])
        T[k].inter = inters
        T[k].parts = t["parts"].ctypes.data if t.get("parts") is not None else None
        T[k].n_parts = len(t["parts"]) if t.get("parts") is not None else 0
        T[k].rate = lib().o_svmax_find(inters, len(t["inters"]), vmax, 1000) * t["density"]
        rate_total += T[k].rate
    m = Mcc(len(targets), T, 1.0 / rate_total if rate_total > 0 else float("inf"), mass)
    m._keep = (T, keep, targets)
    return m


def inter(type_, m1, m2, rate=0.0, DE=0.0, cutoff=0.0, tabE=None, tabS=None):
    """Interaction record as the reference builds it (DE in J, LANGEVIN rate * cutoff^2)."""
    r = rate * (cutoff ** 2 if type_ == LANGEVIN else 1.0)
    it = Inter(type_, DE * QE, r, cutoff, m1, m2, 0, 0, 0, 0, 0, None, None)
    if tabE is not None:
        tE = np.ascontiguousarray(tabE, np.float64)
        tS = np.ascontiguousarray(tabS, np.float64)
        it.ntab = len(tE)
        it.tabE = tE.ctypes.data_as(_PD)
        it.tabS = tS.ctypes.data_as(_PD)
        it._keep = (tE, tS)
    return it


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().o_philox4x32_10(c, k, o)
    return list(o)
