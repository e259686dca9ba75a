"""ctypes binding of the C ABI in include/mag2d_amd.h.

Thin and explicit: structures mirror the header field for field, every entry
point is declared with its argument types, and every non-zero status raises
:class:`M2DError` with the library's own message.  The library must exist
(``python -m mag2d_amd.build``); there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libmag2d_amd.so"

M2D_OK = 0
ABI_VERSION = 2  # M2D_ABI_VERSION of include/mag2d_amd.h

STATUS = {-1: "BAD_ARG", -2: "OUT_OF_RANGE", -3: "NAN", -4: "HIP_ERROR", -5: "NOT_CONVERGED",
          -6: "UNSUPPORTED", -7: "NO_MEMORY"}
FLAG_OUT_OF_RANGE, FLAG_NAN, FLAG_PARTNER, FLAG_EVENTS = 1, 2, 4, 8
FIXED, FIXED_RF, FREE, BOUNDARY = 0, 1, 2, 3
DEAD_BITS = 0xFFC0DEAD
(GEO_EMPTY, GEO_PROBE, GEO_RF_22PT, GEO_RF_8PT, GEO_RF_HAITRAP, GEO_RF_QUAD, GEO_MAC,
 GEO_PENNING, GEO_PENNING_SIMPLE, GEO_TUBE) = range(10)
BC_FREE, BC_PERIODIC = 0, 1
FIELD_U, FIELD_URF, FIELD_RHO, FIELD_VOLTAGE = 0, 1, 2, 3
SOLVE_STATIC, SOLVE_RF, SOLVE_RAW = 0, 1, 2
ELASTIC, LANGEVIN, CX, COULOMB, SUPERELASTIC = range(5)
STEP_PUSH, STEP_BOUNDARY, STEP_DEPOSIT = 1, 2, 4
PASS_DEPOSIT, PASS_MCC, PASS_TILED, PASS_FAST, PASS_BGRID, PASS_PHT, PASS_HIPOS = 1, 2, 4, 8, 16, 32, 64

EPS0 = 8.854187817e-12
KB = 1.380662e-23
QE = 1.602189e-19


class M2DError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Grid2D(C.Structure):
    _fields_ = [("nx", C.c_int32), ("nz", C.c_int32), ("dx", C.c_double), ("dz", C.c_double),
                ("xmin", C.c_double), ("zmin", C.c_double)]


class Geom2D(C.Structure):
    _fields_ = [("geometry", C.c_int32), ("x_max", C.c_double), ("z_max", C.c_double),
                ("probe_radius", C.c_double), ("u_probe", C.c_double), ("extern_field", C.c_double)]


class Particles2D(C.Structure):
    _fields_ = [("x", C.c_void_p), ("z", C.c_void_p), ("vx", C.c_void_p), ("vy", C.c_void_p),
                ("vz", C.c_void_p), ("n", C.c_int64), ("xl", C.c_void_p), ("zl", C.c_void_p)]


class Species2D(C.Structure):
    _fields_ = [("charge", C.c_double), ("mass", C.c_double), ("dt", C.c_double),
                ("lifetime", C.c_double), ("niter", C.c_uint64), ("stream", C.c_uint32),
                ("reserved", C.c_uint32)]


class FieldsParams(C.Structure):
    _fields_ = [("grid", Grid2D), ("x_max", C.c_double), ("z_max", C.c_double),
                ("geometry_empty", C.c_int32), ("selfconsistent", C.c_int32),
                ("electric_field_from_file", C.c_int32), ("boundary", C.c_int32),
                ("extern_field", C.c_double), ("rf", C.c_int32), ("rf_amplitude", C.c_double),
                ("rf_omega", C.c_double), ("rf_U0", C.c_double), ("Br", C.c_double),
                ("Bz", C.c_double), ("Bt", C.c_double), ("dV", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("outer_iters", C.c_int32), ("inner_iters", C.c_int32), ("rel_residual", C.c_double)]


class Interaction(C.Structure):
    _fields_ = [("type", C.c_int32), ("target", C.c_int32), ("DE", C.c_double), ("rate", C.c_double),
                ("cutoff", C.c_double), ("tab_offset", C.c_int32), ("tab_len", C.c_int32)]


class Target(C.Structure):
    _fields_ = [("mass", C.c_double), ("density", C.c_double), ("temperature", C.c_double),
                ("charge", C.c_double), ("partner", Particles2D)]


class MccDesc(C.Structure):
    _fields_ = [("mass", C.c_double), ("charge", C.c_double), ("temperature", C.c_double),
                ("density", C.c_double), ("E_max", C.c_double), ("n_targets", C.c_int32),
                ("targets", C.POINTER(Target)), ("n_inter", C.c_int32),
                ("inter", C.POINTER(Interaction)), ("n_tab", C.c_int32),
                ("tab_E", C.POINTER(C.c_double)), ("tab_sigma", C.POINTER(C.c_double))]


class StepStats(C.Structure):
    _fields_ = [("removed", C.c_int64), ("collisions", C.c_int64), ("spilled", C.c_int64)]


class AdjointCfg(C.Structure):
    _fields_ = [("steps", C.c_int32), ("checkpoint_every", C.c_int32), ("n_groups", C.c_int32),
                ("group", C.POINTER(C.c_int32)), ("adjoint_tol", C.c_double), ("mcc", C.c_void_p)]


class AdjointInfo(C.Structure):
    _fields_ = [("n_live", C.c_int64), ("forward_ms", C.c_double), ("backward_ms", C.c_double)]


class BatchDesc(C.Structure):
    _fields_ = [("n_runs", C.c_int32), ("offset", C.POINTER(C.c_int64)), ("rf_amplitude", C.POINTER(C.c_double)),
                ("mcc", C.POINTER(C.c_void_p)), ("stream", C.POINTER(C.c_uint32))]


class BatchShape(C.Structure):
    _fields_ = [("block", C.c_int32), ("ppt", C.c_int32), ("seg_steps", C.c_int32)]


class Grid3D(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double)]


class Particles3D(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("x", "y", "z", "vx", "vy", "vz")] + [("n", C.c_int64)]


class Fields3DParams(C.Structure):
    _fields_ = [("grid", Grid3D), ("x_max", C.c_double), ("y_max", C.c_double), ("z_max", C.c_double),
                ("boundary", C.c_int32), ("Bx", C.c_double), ("By", C.c_double), ("Bz", C.c_double),
                ("rhs_scale", C.c_double)]


_P = C.c_void_p
_PD = C.POINTER(C.c_double)
_PU8 = C.POINTER(C.c_uint8)
_SIGS = {
    "m2d_last_error": (C.c_char_p, []),
    "m2d_abi_version": (C.c_int, []),
    "m2d_ctx_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(_P)]),
    "m2d_ctx_destroy": (C.c_int, [_P]),
    "m2d_ctx_set_stream": (C.c_int, [_P, _P]),
    "m2d_ctx_stream": (_P, [_P]),
    "m2d_ctx_seed": (C.c_uint64, [_P]),
    "m2d_ctx_synchronize": (C.c_int, [_P]),
    "m2d_ctx_status": (C.c_int, [_P, C.POINTER(C.c_uint32), C.c_int]),
    "m2d_ctx_profile": (C.c_int, [_P, C.c_int]),
    "m2d_ctx_profile_read": (C.c_int, [_P, _PD, C.POINTER(C.c_int64)]),
    "m2d_ctx_last_pass": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "m2d_malloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "m2d_free": (C.c_int, [_P, _P]),
    "m2d_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "m2d_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "m2d_memcpy_d2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "m2d_memset": (C.c_int, [_P, _P, C.c_int, C.c_size_t]),
    "m2d_field_add": (C.c_int, [_P, _P, _P, C.c_double, C.c_size_t]),
    "m2d_grid_build_2d": (C.c_int, [C.POINTER(Grid2D), C.POINTER(Geom2D), _PU8, _PD]),
    "m2d_grid_square_electrode_2d": (C.c_int, [C.POINTER(Grid2D), _PU8, _PD, C.c_double, C.c_double,
                                               C.c_double, C.c_double, C.c_double, C.c_int]),
    "m2d_grid_circle_electrode_2d": (C.c_int, [C.POINTER(Grid2D), _PU8, _PD, C.c_double, C.c_double,
                                               C.c_double, C.c_double, C.c_int]),
    "m2d_grid_mark_boundary_2d": (C.c_int, [C.POINTER(Grid2D), _PU8]),
    "m2d_grid_is_free_2d": (C.c_int, [C.POINTER(Grid2D), _PU8, C.c_double, C.c_double]),
    "m2d_fields2d_create": (C.c_int, [_P, C.POINTER(FieldsParams), _PU8, _PD, C.POINTER(_P)]),
    "m2d_fields2d_destroy": (C.c_int, [_P]),
    "m2d_fields2d_data": (_P, [_P, C.c_int]),
    "m2d_fields2d_upload": (C.c_int, [_P, C.c_int, _PD]),
    "m2d_fields2d_download": (C.c_int, [_P, C.c_int, _PD]),
    "m2d_fields2d_update": (C.c_int, [_P]),
    "m2d_fields2d_reset": (C.c_int, [_P]),
    "m2d_fields2d_u_smooth": (C.c_int, [_P, C.c_int, C.c_int, C.c_double]),
    "m2d_fields2d_solve": (C.c_int, [_P, C.c_int, C.c_double, C.POINTER(SolveInfo)]),
    "m2d_fields2d_E": (C.c_int, [_P, _P, _P, C.c_int64, C.c_double, _P, _P]),
    "m2d_fields2d_set_magnetic_field": (C.c_int, [_P, C.POINTER(Grid2D), _PD, _PD]),
    "m2d_fields2d_B": (C.c_int, [_P, _P, _P, C.c_int64, _P, _P, _P]),
    "m2d_accumulate_2d": (C.c_int, [_P, C.POINTER(Grid2D), C.POINTER(Particles2D), C.c_double, _P]),
    "m2d_accumulate_2d_noalloc": (C.c_int, [_P, C.POINTER(Grid2D), C.POINTER(Particles2D), C.c_double, _P]),
    "m2d_interpolate_2d": (C.c_int, [_P, C.POINTER(Grid2D), _P, _P, _P, C.c_int64, _P]),
    "m2d_grad_2d": (C.c_int, [_P, C.POINTER(Grid2D), _P, _P, _P, C.c_int64, _P, _P]),
    "m2d_mcc_create": (C.c_int, [_P, C.POINTER(MccDesc), C.POINTER(_P)]),
    "m2d_mcc_destroy": (C.c_int, [_P]),
    "m2d_mcc_lifetime": (C.c_double, [_P]),
    "m2d_mcc_rates": (C.c_int, [_P, _PD]),
    "m2d_mcc_set_partner": (C.c_int, [_P, C.c_int32, C.POINTER(Particles2D)]),
    "m2d_advance_2d": (C.c_int, [_P, _P, C.POINTER(Species2D), C.POINTER(Particles2D), _P, _P,
                                 C.c_uint32, C.POINTER(StepStats)]),
    "m2d_advance_init_2d": (C.c_int, [_P, _P, C.POINTER(Species2D), C.POINTER(Particles2D)]),
    "m2d_advance_multicoll_2d": (C.c_int, [_P, _P, C.POINTER(Species2D), C.POINTER(Particles2D), _P, _P,
                                           C.POINTER(StepStats)]),
    "m2d_time_to_death_init_2d": (C.c_int, [_P, C.POINTER(Species2D), C.POINTER(Particles2D), _P]),
    "m2d_mean_energy_2d": (C.c_int, [_P, C.POINTER(Particles2D), C.c_double, _PD,
                                     C.POINTER(C.c_int64)]),
    "m2d_count_live_2d": (C.c_int, [_P, C.POINTER(Particles2D), C.POINTER(C.c_int64)]),
    "m2d_compact_2d": (C.c_int, [_P, C.POINTER(Particles2D), C.POINTER(Particles2D),
                                 C.POINTER(C.c_int64)]),
    "m2d_sort_2d": (C.c_int, [_P, _P, C.POINTER(Particles2D), C.POINTER(Particles2D),
                              C.POINTER(C.c_int64)]),
    "m2d_sort_forget_2d": (C.c_int, [_P, C.POINTER(Particles2D)]),
    "m2d_add_particles_on_disk_2d": (C.c_int, [_P, _P, C.POINTER(Particles2D), C.c_double,
                                               C.c_double, C.c_double, C.c_double, C.c_uint32]),
    "m2d_add_particles_bessel_2d": (C.c_int, [_P, _P, C.POINTER(Particles2D), C.c_double, C.c_double,
                                              C.c_double, C.c_double, C.c_uint32]),
    "m2d_add_particles_everywhere_2d": (C.c_int, [_P, _P, C.POINTER(Particles2D), C.c_double, _P, C.c_double,
                                                  C.c_uint32]),
    "m2d_adjoint_energy_2d": (C.c_int, [_P, _P, C.POINTER(Species2D), C.POINTER(Particles2D),
                                        C.POINTER(AdjointCfg), _PD, _PD, C.POINTER(AdjointInfo)]),
    "m2d_batch_create": (C.c_int, [_P, _P, C.POINTER(Species2D), C.POINTER(Particles2D), C.POINTER(BatchDesc),
                                   C.POINTER(_P)]),
    "m2d_batch_destroy": (C.c_int, [_P]),
    "m2d_batch_tune": (C.c_int, [_P, C.c_uint64, C.c_int32, C.c_int32, C.POINTER(BatchShape)]),
    "m2d_batch_set_shape": (C.c_int, [_P, C.POINTER(BatchShape)]),
    "m2d_batch_advance": (C.c_int, [_P, C.c_uint64, C.c_int32]),
    "m2d_batch_energy": (C.c_int, [_P, _PD, C.POINTER(C.c_int64)]),
    "m2d_fields3d_create": (C.c_int, [_P, C.POINTER(Fields3DParams), _PU8, _PD, C.POINTER(_P)]),
    "m2d_fields3d_destroy": (C.c_int, [_P]),
    "m2d_fields3d_data": (_P, [_P, C.c_int]),
    "m2d_fields3d_upload": (C.c_int, [_P, C.c_int, _PD]),
    "m2d_fields3d_download": (C.c_int, [_P, C.c_int, _PD]),
    "m2d_fields3d_reset": (C.c_int, [_P]),
    "m2d_fields3d_solve": (C.c_int, [_P, C.c_double, C.POINTER(SolveInfo)]),
    "m2d_accumulate_3d": (C.c_int, [_P, C.POINTER(Grid3D), C.POINTER(Particles3D), C.c_double, _P]),
    "m2d_interpolate_3d": (C.c_int, [_P, C.POINTER(Grid3D), _P, _P, _P, _P, C.c_int64, _P]),
    "m2d_grad_3d": (C.c_int, [_P, C.POINTER(Grid3D), _P, _P, _P, _P, C.c_int64, _P, _P, _P]),
    "m2d_advance_3d": (C.c_int, [_P, _P, C.POINTER(Species2D), C.POINTER(Particles3D), _P, _P,
                                 C.c_uint32, C.POINTER(StepStats)]),
    "m2d_advance_init_3d": (C.c_int, [_P, _P, C.POINTER(Species2D), C.POINTER(Particles3D)]),
    "m2d_sort_3d": (C.c_int, [_P, _P, C.POINTER(Particles3D), C.POINTER(Particles3D),
                              C.POINTER(C.c_int64)]),
    "m2d_count_live_3d": (C.c_int, [_P, C.POINTER(Particles3D), C.POINTER(C.c_int64)]),
    "m2d_sort_forget_3d": (C.c_int, [_P, C.POINTER(Particles3D)]),
    "m2d_mean_energy_3d": (C.c_int, [_P, C.POINTER(Particles3D), C.c_double, _PD,
                                     C.POINTER(C.c_int64)]),
}

_lib = None


def lib():
    """Load libmag2d_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise ImportError(f"{_LIB_PATH} is missing: build it with `python -m mag2d_amd.build`")
        L = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        # the structures below follow this layout version (m2d_step_stats,
        # m2d_adjoint_cfg): a library of another one would read them wrongly
        if L.m2d_abi_version() != ABI_VERSION:
            raise ImportError(f"{_LIB_PATH} has ABI version {L.m2d_abi_version()}, this binding "
                              f"{ABI_VERSION}: rebuild it with `python -m mag2d_amd.build`")
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(code: int):
    if code != M2D_OK:
        raise M2DError(code, lib().m2d_last_error().decode())


def _ptr(a: np.ndarray, dtype):
    assert a.dtype == dtype and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(np.ctypeslib.as_ctypes_type(dtype)))


# --------------------------------------------------------------------- objects

class Context:
    """m2d_ctx: device, stream and RNG seed."""

    def __init__(self, device: int = 0, seed: int = 1234):
        h = C.c_void_p()
        check(lib().m2d_ctx_create(device, seed, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().m2d_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().m2d_ctx_synchronize(self.h))

    def status(self, clear=True) -> int:
        f = C.c_uint32()
        check(lib().m2d_ctx_status(self.h, C.byref(f), int(clear)))
        return f.value

    def set_stream(self, stream_ptr: int | None):
        check(lib().m2d_ctx_set_stream(self.h, C.c_void_p(stream_ptr or 0)))

    def profile(self, enable: bool):
        check(lib().m2d_ctx_profile(self.h, int(enable)))

    def profile_read(self):
        ms, n = C.c_double(), C.c_int64()
        check(lib().m2d_ctx_profile_read(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def last_pass(self) -> int:
        """PASS_* bits of the particle pass the last advance launched"""
        v = C.c_uint32()
        check(lib().m2d_ctx_last_pass(self.h, C.byref(v)))
        return v.value

    # device memory -------------------------------------------------------
    def malloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        check(lib().m2d_malloc(self.h, nbytes, C.byref(p)))
        return p.value

    def free(self, p: int):
        check(lib().m2d_free(self.h, C.c_void_p(p)))

    def h2d(self, dst: int, a: np.ndarray):
        a = np.ascontiguousarray(a)
        check(lib().m2d_memcpy_h2d(self.h, C.c_void_p(dst), a.ctypes.data_as(C.c_void_p), a.nbytes))

    def d2h(self, src: int, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        check(lib().m2d_memcpy_d2h(self.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(src), out.nbytes))
        return out

    def to_device(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        self.h2d(d.ptr, a)
        return d


class DeviceArray:
    def __init__(self, ctx: Context, shape, dtype):
        self.ctx, self.shape, self.dtype = ctx, tuple(np.atleast_1d(shape)), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctx.malloc(max(self.nbytes, 1))

    def get(self) -> np.ndarray:
        return self.ctx.d2h(self.ptr, self.shape, self.dtype)

    def set(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        self.ctx.h2d(self.ptr, a)

    def zero(self):
        check(lib().m2d_memset(self.ctx.h, C.c_void_p(self.ptr), 0, self.nbytes))

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.ctx.free(self.ptr)
        except Exception:
            pass
        self.ptr = None


class ParticlesSoA:
    """Device SoA x, z, vx, vy, vz (FP32) of n slots.  With ``hipos`` the
    store also keeps the FP32 remainders xl, zl of the positions (long-time
    positions: x + xl is the position in FP64)."""

    NAMES = ("x", "z", "vx", "vy", "vz")

    def __init__(self, ctx: Context, n: int, hipos: bool = False):
        self.ctx, self.n, self.hipos = ctx, int(n), bool(hipos)
        names = self.NAMES + (("xl", "zl") if self.hipos else ())
        self.arr = {k: DeviceArray(ctx, (max(self.n, 1),), np.float32) for k in names}

    @classmethod
    def from_host(cls, ctx, x, z, vx, vy, vz, hipos: bool = False):
        p = cls(ctx, len(x), hipos=hipos)
        p.upload(x, z, vx, vy, vz)
        return p

    def upload(self, x, z, vx, vy, vz):
        """Positions may be FP64: a store with remainders keeps them to
        about 2^-48 (the FP32 value and its rounded remainder)."""
        x64, z64 = np.asarray(x, np.float64), np.asarray(z, np.float64)
        cols = {"x": x64.astype(np.float32), "z": z64.astype(np.float32), "vx": vx, "vy": vy, "vz": vz}
        if self.hipos:
            cols["xl"] = (x64 - cols["x"].astype(np.float64)).astype(np.float32)
            cols["zl"] = (z64 - cols["z"].astype(np.float64)).astype(np.float32)
        for k, a in cols.items():
            a = np.asarray(a, dtype=np.float32)
            assert a.shape == (self.n,)
            if self.n:
                self.arr[k].set(a)
        sort_forget(self.ctx, self)

    def download(self):
        """The arrays; with remainders also x64, z64, the FP64 positions."""
        d = {k: self.arr[k].get()[: self.n] for k in self.arr}
        if self.hipos:
            d["x64"] = d["x"].astype(np.float64) + d["xl"].astype(np.float64)
            d["z64"] = d["z"].astype(np.float64) + d["zl"].astype(np.float64)
        return d

    def struct(self, n: int | None = None) -> Particles2D:
        xl = self.arr["xl"].ptr if self.hipos else None
        zl = self.arr["zl"].ptr if self.hipos else None
        return Particles2D(*(self.arr[k].ptr for k in self.NAMES), self.n if n is None else n, xl, zl)

    def alive(self):
        x = self.arr["x"].get()[: self.n]
        return x.view(np.uint32) != DEAD_BITS


def grid2d(nx, nz, dx, dz, xmin=0.0, zmin=0.0) -> Grid2D:
    return Grid2D(nx, nz, dx, dz, xmin, zmin)


def grid_build(g: Grid2D, geometry, x_max, z_max, probe_radius=1e-4, u_probe=-10.0,
               extern_field=100.0):
    mask = np.zeros(g.nx * g.nz, np.uint8)
    volt = np.zeros(g.nx * g.nz, np.float64)
    geo = Geom2D(geometry, x_max, z_max, probe_radius, u_probe, extern_field)
    check(lib().m2d_grid_build_2d(C.byref(g), C.byref(geo), _ptr(mask, np.uint8), _ptr(volt, np.float64)))
    return mask, volt


def square_electrode(g, mask, volt, rmin, rmax, zmin, zmax, v, mtype=FIXED):
    check(lib().m2d_grid_square_electrode_2d(C.byref(g), _ptr(mask, np.uint8), _ptr(volt, np.float64),
                                             rmin, rmax, zmin, zmax, v, mtype))


def circle_electrode(g, mask, volt, xc, zc, r, v, mtype=FIXED):
    check(lib().m2d_grid_circle_electrode_2d(C.byref(g), _ptr(mask, np.uint8), _ptr(volt, np.float64),
                                             xc, zc, r, v, mtype))


def mark_boundary(g, mask):
    check(lib().m2d_grid_mark_boundary_2d(C.byref(g), _ptr(mask, np.uint8)))


class Fields2D:
    """m2d_fields2d: Fields (u, uRF, rho, mask, voltage) on the device."""

    def __init__(self, ctx: Context, params: FieldsParams, mask: np.ndarray, voltage: np.ndarray):
        self.ctx, self.params = ctx, params
        self.nx, self.nz = params.grid.nx, params.grid.nz
        mask = np.ascontiguousarray(mask, np.uint8).ravel()
        voltage = np.ascontiguousarray(voltage, np.float64).ravel()
        h = C.c_void_p()
        check(lib().m2d_fields2d_create(ctx.h, C.byref(params), _ptr(mask, np.uint8),
                                        _ptr(voltage, np.float64), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().m2d_fields2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ptr(self, which) -> int:
        return lib().m2d_fields2d_data(self.h, which)

    def upload(self, which, a):
        a = np.ascontiguousarray(a, np.float64).ravel()
        assert a.size == self.nx * self.nz
        check(lib().m2d_fields2d_upload(self.h, which, _ptr(a, np.float64)))

    def download(self, which) -> np.ndarray:
        out = np.empty(self.nx * self.nz, np.float64)
        check(lib().m2d_fields2d_download(self.h, which, _ptr(out, np.float64)))
        return out

    def u_smooth(self, which=FIELD_U, symmetry=False, radius=-1.0):
        check(lib().m2d_fields2d_u_smooth(self.h, which, int(symmetry), radius))

    def reset(self):
        check(lib().m2d_fields2d_reset(self.h))

    def update(self):
        check(lib().m2d_fields2d_update(self.h))

    def solve(self, mode, tol=1e-10) -> SolveInfo:
        info = SolveInfo()
        check(lib().m2d_fields2d_solve(self.h, mode, tol, C.byref(info)))
        return info

    def set_magnetic_field(self, g, Br, Bz):
        """Fields::load_magnetic_field as data: Br, Bz (row-major) on grid g; None clears."""
        if Br is None:
            check(lib().m2d_fields2d_set_magnetic_field(self.h, None, None, None))
            return
        br = np.ascontiguousarray(Br, np.float64)
        bz = np.ascontiguousarray(Bz, np.float64)
        check(lib().m2d_fields2d_set_magnetic_field(self.h, C.byref(g), _ptr(br, np.float64), _ptr(bz, np.float64)))

    def B(self, x, z):
        x = np.ascontiguousarray(x, np.float64)
        z = np.ascontiguousarray(z, np.float64)
        n = len(x)
        dx, dz = self.ctx.to_device(x), self.ctx.to_device(z)
        out = [DeviceArray(self.ctx, (n,), np.float64) for _ in range(3)]
        check(lib().m2d_fields2d_B(self.h, C.c_void_p(dx.ptr), C.c_void_p(dz.ptr), n, C.c_void_p(out[0].ptr),
                                   C.c_void_p(out[1].ptr), C.c_void_p(out[2].ptr)))
        return tuple(o.get() for o in out)

    def E(self, x, z, t=0.0):
        x = np.ascontiguousarray(x, np.float64)
        z = np.ascontiguousarray(z, np.float64)
        n = len(x)
        dx, dz = self.ctx.to_device(x), self.ctx.to_device(z)
        ex, ez = DeviceArray(self.ctx, (n,), np.float64), DeviceArray(self.ctx, (n,), np.float64)
        check(lib().m2d_fields2d_E(self.h, C.c_void_p(dx.ptr), C.c_void_p(dz.ptr), n, t,
                                   C.c_void_p(ex.ptr), C.c_void_p(ez.ptr)))
        return ex.get(), ez.get()


class Mcc:
    """m2d_mcc built from species-conf style records."""

    def __init__(self, ctx: Context, mass, targets, interactions, E_max, tab_E=(), tab_sigma=(),
                 charge=0.0, temperature=0.0, density=0.0):
        self.ctx = ctx
        T = (Target * len(targets))(*targets)
        I = (Interaction * max(1, len(interactions)))(*interactions)
        tE = np.ascontiguousarray(tab_E, np.float64)
        tS = np.ascontiguousarray(tab_sigma, np.float64)
        self._keep = (T, I, tE, tS)
        d = MccDesc(mass, charge, temperature, density, E_max, len(targets), T, len(interactions), I,
                    len(tE), tE.ctypes.data_as(_PD) if len(tE) else None,
                    tS.ctypes.data_as(_PD) if len(tS) else None)
        h = C.c_void_p()
        check(lib().m2d_mcc_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        self.n_targets = len(targets)

    @property
    def lifetime(self) -> float:
        return lib().m2d_mcc_lifetime(self.h)

    def rates(self):
        out = np.zeros(self.n_targets)
        check(lib().m2d_mcc_rates(self.h, _ptr(out, np.float64)))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().m2d_mcc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def advance(ctx, fields, species: Species2D, parts: ParticlesSoA, mcc=None, rho_ptr=None,
            flags=STEP_PUSH | STEP_BOUNDARY, want_stats=False):
    st = StepStats()
    s = parts.struct()
    check(lib().m2d_advance_2d(ctx.h, fields.h, C.byref(species), C.byref(s),
                               mcc.h if mcc is not None else None,
                               C.c_void_p(rho_ptr) if rho_ptr else None, flags,
                               C.byref(st) if want_stats else None))
    return st


def advance_init(ctx, fields, species, parts):
    s = parts.struct()
    check(lib().m2d_advance_init_2d(ctx.h, fields.h, C.byref(species), C.byref(s)))


def advance_multicoll(ctx, fields, species: Species2D, parts: ParticlesSoA, ttd_ptr: int, mcc=None,
                      want_stats=False):
    """Species<CARTESIAN>::advance_multicoll; ttd_ptr: device FP64 time_to_death per slot."""
    st = StepStats()
    s = parts.struct()
    check(lib().m2d_advance_multicoll_2d(ctx.h, fields.h, C.byref(species), C.byref(s), C.c_void_p(ttd_ptr),
                                         mcc.h if mcc is not None else None,
                                         C.byref(st) if want_stats else None))
    return st


def time_to_death_init(ctx, species: Species2D, parts: ParticlesSoA, ttd_ptr: int):
    s = parts.struct()
    check(lib().m2d_time_to_death_init_2d(ctx.h, C.byref(species), C.byref(s), C.c_void_p(ttd_ptr)))


def accumulate(ctx, g: Grid2D, parts: ParticlesSoA, charge, rho_ptr):
    s = parts.struct()
    check(lib().m2d_accumulate_2d(ctx.h, C.byref(g), C.byref(s), charge, C.c_void_p(rho_ptr)))


def mean_energy(ctx, parts, mass):
    m, n = C.c_double(), C.c_int64()
    s = parts.struct()
    check(lib().m2d_mean_energy_2d(ctx.h, C.byref(s), mass, C.byref(m), C.byref(n)))
    return m.value, n.value


def count_live(ctx, parts):
    n = C.c_int64()
    s = parts.struct()
    check(lib().m2d_count_live_2d(ctx.h, C.byref(s), C.byref(n)))
    return n.value


def compact(ctx, src: ParticlesSoA, dst: ParticlesSoA):
    n = C.c_int64()
    a, b = src.struct(), dst.struct()
    check(lib().m2d_compact_2d(ctx.h, C.byref(a), C.byref(b), C.byref(n)))
    return n.value


def sort(ctx, fields, src: ParticlesSoA, dst: ParticlesSoA, want_count=True):
    """Cell-sort src into dst; returns the live count (None if not requested)."""
    n = C.c_int64()
    a, b = src.struct(), dst.struct()
    check(lib().m2d_sort_2d(ctx.h, fields.h, C.byref(a), C.byref(b), C.byref(n) if want_count else None))
    return n.value if want_count else None


def cell_keys(x, z, nx, nz, dx, dz, xmin=0.0, zmin=0.0):
    """Host copy of the sort key (tile << 6 | cell in tile, 8 x 8 tiles) for tests;
    the sort orders by the tile bits (key >> 6) only."""
    X = (np.asarray(x, np.float64) - xmin) * (1.0 / dx)
    Z = (np.asarray(z, np.float64) - zmin) * (1.0 / dz)
    i = np.where(X >= 0, np.minimum(X, nx - 2), 0).astype(np.int64)
    j = np.where(Z >= 0, np.minimum(Z, nz - 2), 0).astype(np.int64)
    ntz = (nz - 2) // 8 + 1
    return (((i // 8) * ntz + (j // 8)) << 6) | ((i % 8) << 3) | (j % 8)


def sort_forget(ctx, parts):
    s = parts.struct()
    check(lib().m2d_sort_forget_2d(ctx.h, C.byref(s)))


def add_particles_on_disk(ctx, fields, parts, cx, cz, radius, vth, stream=0):
    s = parts.struct()
    check(lib().m2d_add_particles_on_disk_2d(ctx.h, fields.h, C.byref(s), cx, cz, radius, vth, stream))


def add_particles_bessel(ctx, fields, parts, cx, cz, radius, vth, stream=0):
    s = parts.struct()
    check(lib().m2d_add_particles_bessel_2d(ctx.h, fields.h, C.byref(s), cx, cz, radius, vth, stream))


def add_particles_everywhere(ctx, fields, parts, vth, ttd_ptr=None, lifetime=float("inf"), stream=0):
    s = parts.struct()
    check(lib().m2d_add_particles_everywhere_2d(ctx.h, fields.h, C.byref(s), vth,
                                                C.c_void_p(ttd_ptr) if ttd_ptr else None, lifetime, stream))


def interpolate(ctx, g, data_ptr, x, z):
    n = len(x)
    dx, dz = ctx.to_device(np.asarray(x, np.float64)), ctx.to_device(np.asarray(z, np.float64))
    out = DeviceArray(ctx, (n,), np.float64)
    check(lib().m2d_interpolate_2d(ctx.h, C.byref(g), C.c_void_p(data_ptr), C.c_void_p(dx.ptr),
                                   C.c_void_p(dz.ptr), n, C.c_void_p(out.ptr)))
    return out.get()


def grad(ctx, g, data_ptr, x, z, init=0.0):
    n = len(x)
    dx, dz = ctx.to_device(np.asarray(x, np.float64)), ctx.to_device(np.asarray(z, np.float64))
    gx = ctx.to_device(np.full(n, init))
    gz = ctx.to_device(np.full(n, init))
    check(lib().m2d_grad_2d(ctx.h, C.byref(g), C.c_void_p(data_ptr), C.c_void_p(dx.ptr),
                            C.c_void_p(dz.ptr), n, C.c_void_p(gx.ptr), C.c_void_p(gz.ptr)))
    return gx.get(), gz.get()


def rf_phase(niter, dt, rf_omega, rf_amplitude, rf_U0):
    """Phase factor of Fields::E (fields.hpp:140-144)."""
    t = niter * dt
    ph = rf_omega * t
    y = 10000000 * math.pi
    if not (0.0 <= ph <= y):
        ph = ph - y * int(ph / y) + (y if ph < 0 else 0)
    return rf_amplitude * math.cos(ph) + rf_U0


# --------------------------------------------------------------------- 3D

def grid3d(nx, ny, nz, dx, dy, dz) -> Grid3D:
    return Grid3D(nx, ny, nz, dx, dy, dz)


class ParticlesSoA3D:
    """Device SoA x, y, z, vx, vy, vz (FP32) of n slots."""

    NAMES = ("x", "y", "z", "vx", "vy", "vz")

    def __init__(self, ctx: Context, n: int):
        self.ctx, self.n = ctx, int(n)
        self.arr = {k: DeviceArray(ctx, (max(self.n, 1),), np.float32) for k in self.NAMES}

    @classmethod
    def from_host(cls, ctx, *arrays):
        p = cls(ctx, len(arrays[0]))
        p.upload(*arrays)
        return p

    def upload(self, *arrays):
        for k, a in zip(self.NAMES, arrays):
            a = np.asarray(a, dtype=np.float32)
            assert a.shape == (self.n,)
            if self.n:
                self.arr[k].set(a)
        s = self.struct()
        check(lib().m2d_sort_forget_3d(self.ctx.h, C.byref(s)))

    def download(self):
        return {k: self.arr[k].get()[: self.n] for k in self.NAMES}

    def struct(self) -> Particles3D:
        return Particles3D(*(self.arr[k].ptr for k in self.NAMES), self.n)

    def alive(self):
        return self.arr["x"].get()[: self.n].view(np.uint32) != DEAD_BITS


class Fields3D:
    """m2d_fields3d: ElMag3D / Solver (u, rho, mask, voltage) on the device."""

    def __init__(self, ctx: Context, params: Fields3DParams, mask: np.ndarray, voltage: np.ndarray):
        self.ctx, self.params = ctx, params
        g = params.grid
        self.shape = (g.nx, g.ny, g.nz)
        self.n = g.nx * g.ny * g.nz
        mask = np.ascontiguousarray(mask, np.uint8).ravel()
        voltage = np.ascontiguousarray(voltage, np.float64).ravel()
        h = C.c_void_p()
        check(lib().m2d_fields3d_create(ctx.h, C.byref(params), _ptr(mask, np.uint8),
                                        _ptr(voltage, np.float64), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().m2d_fields3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ptr(self, which) -> int:
        return lib().m2d_fields3d_data(self.h, which)

    def upload(self, which, a):
        a = np.ascontiguousarray(a, np.float64).ravel()
        assert a.size == self.n
        check(lib().m2d_fields3d_upload(self.h, which, _ptr(a, np.float64)))

    def download(self, which) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        check(lib().m2d_fields3d_download(self.h, which, _ptr(out, np.float64)))
        return out

    def reset(self):
        check(lib().m2d_fields3d_reset(self.h))

    def solve(self, tol=1e-10) -> SolveInfo:
        info = SolveInfo()
        check(lib().m2d_fields3d_solve(self.h, tol, C.byref(info)))
        return info


def advance_3d(ctx, fields: Fields3D, species: Species2D, parts: ParticlesSoA3D, mcc=None, rho_ptr=None,
               flags=STEP_PUSH | STEP_BOUNDARY, want_stats=False):
    st = StepStats()
    s = parts.struct()
    check(lib().m2d_advance_3d(ctx.h, fields.h, C.byref(species), C.byref(s),
                               mcc.h if mcc is not None else None,
                               C.c_void_p(rho_ptr) if rho_ptr else None, flags,
                               C.byref(st) if want_stats else None))
    return st


def advance_init_3d(ctx, fields, species, parts):
    s = parts.struct()
    check(lib().m2d_advance_init_3d(ctx.h, fields.h, C.byref(species), C.byref(s)))


def accumulate_3d(ctx, g: Grid3D, parts: ParticlesSoA3D, charge, rho_ptr):
    s = parts.struct()
    check(lib().m2d_accumulate_3d(ctx.h, C.byref(g), C.byref(s), charge, C.c_void_p(rho_ptr)))


def sort_3d(ctx, fields, src: ParticlesSoA3D, dst: ParticlesSoA3D):
    n = C.c_int64()
    a, b = src.struct(), dst.struct()
    check(lib().m2d_sort_3d(ctx.h, fields.h, C.byref(a), C.byref(b), C.byref(n)))
    return n.value


def count_live_3d(ctx, parts):
    n = C.c_int64()
    s = parts.struct()
    check(lib().m2d_count_live_3d(ctx.h, C.byref(s), C.byref(n)))
    return n.value


def mean_energy_3d(ctx, parts, mass):
    m, n = C.c_double(), C.c_int64()
    s = parts.struct()
    check(lib().m2d_mean_energy_3d(ctx.h, C.byref(s), mass, C.byref(m), C.byref(n)))
    return m.value, n.value


def interpolate_3d(ctx, g, data_ptr, x, y, z):
    n = len(x)
    d = [ctx.to_device(np.asarray(a, np.float64)) for a in (x, y, z)]
    out = DeviceArray(ctx, (n,), np.float64)
    check(lib().m2d_interpolate_3d(ctx.h, C.byref(g), C.c_void_p(data_ptr), *(C.c_void_p(a.ptr) for a in d),
                                   n, C.c_void_p(out.ptr)))
    return out.get()


def grad_3d(ctx, g, data_ptr, x, y, z):
    n = len(x)
    d = [ctx.to_device(np.asarray(a, np.float64)) for a in (x, y, z)]
    out = [DeviceArray(ctx, (n,), np.float64) for _ in range(3)]
    check(lib().m2d_grad_3d(ctx.h, C.byref(g), C.c_void_p(data_ptr), *(C.c_void_p(a.ptr) for a in d), n,
                            *(C.c_void_p(o.ptr) for o in out)))
    return tuple(o.get() for o in out)


# --------------------------------------------------------------------- adjoint

def adjoint_energy(ctx, fields: Fields2D, species: Species2D, parts: ParticlesSoA, steps, group=None,
                   n_groups=0, checkpoint_every=100, adjoint_tol=0.0, mcc=None):
    """J (mean kinetic energy in eV after `steps` self-consistent steps) and,
    when `group` (nx*nz int array, -1 for none) is given, dJ/dV per group;
    adjoint_tol: relative residual of the adjoint solves (0: 1e-10); mcc: the
    species' collisions (its lifetime is the species')."""
    g = None if group is None else np.ascontiguousarray(group, np.int32)
    cfg = AdjointCfg(steps, checkpoint_every, n_groups,
                     g.ctypes.data_as(C.POINTER(C.c_int32)) if g is not None else None, adjoint_tol,
                     mcc.h.value if mcc is not None else None)
    J = C.c_double()
    grad = np.zeros(max(n_groups, 1))
    info = AdjointInfo()
    s = parts.struct()
    check(lib().m2d_adjoint_energy_2d(ctx.h, fields.h, C.byref(species), C.byref(s), C.byref(cfg), C.byref(J),
                                      _ptr(grad, np.float64) if g is not None else None, C.byref(info)))
    return J.value, (grad[:n_groups] if g is not None else None), info


# --------------------------------------------------------------------- batches

class Batch:
    """m2d_batch: independent runs in one store, advanced by the persistent pass."""

    def __init__(self, ctx, fields, species: Species2D, parts: ParticlesSoA, offsets, rf_amplitude, mccs=None,
                 streams=None):
        self.ctx, self.fields, self.parts = ctx, fields, parts
        self.offsets = np.ascontiguousarray(offsets, np.int64)
        self.amp = np.ascontiguousarray(rf_amplitude, np.float64)
        R = len(self.offsets) - 1
        self.mcc_ptrs = (C.c_void_p * R)(*[(m.h.value if m is not None else None) for m in mccs]) if mccs else None
        self.streams = np.ascontiguousarray(streams, np.uint32) if streams is not None else None
        self._mccs = mccs
        d = BatchDesc(R, self.offsets.ctypes.data_as(C.POINTER(C.c_int64)),
                      self.amp.ctypes.data_as(C.POINTER(C.c_double)),
                      C.cast(self.mcc_ptrs, C.POINTER(C.c_void_p)) if self.mcc_ptrs is not None else None,
                      self.streams.ctypes.data_as(C.POINTER(C.c_uint32)) if self.streams is not None else None)
        self._s = parts.struct()
        h = C.c_void_p()
        check(lib().m2d_batch_create(ctx.h, fields.h, C.byref(species), C.byref(self._s), C.byref(d), C.byref(h)))
        self.h = h
        self.n_runs = R

    def close(self):
        if getattr(self, "h", None):
            lib().m2d_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, niter0, probe_steps, seg_steps):
        sh = BatchShape()
        check(lib().m2d_batch_tune(self.h, niter0, probe_steps, seg_steps, C.byref(sh)))
        return sh.block, sh.ppt, sh.seg_steps

    def set_shape(self, block, ppt, seg_steps):
        sh = BatchShape(block, ppt, seg_steps)
        check(lib().m2d_batch_set_shape(self.h, C.byref(sh)))

    def advance(self, niter0, steps):
        check(lib().m2d_batch_advance(self.h, niter0, steps))

    def energy(self):
        e = np.zeros(self.n_runs)
        n = np.zeros(self.n_runs, np.int64)
        check(lib().m2d_batch_energy(self.h, _ptr(e, np.float64), n.ctypes.data_as(C.POINTER(C.c_int64))))
        return e, n
