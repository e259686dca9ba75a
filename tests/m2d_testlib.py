"""Shared set-up for the GPU parity tests: the same scene on the device
(through the C ABI) and in the oracle."""
import ctypes as C
import math

import numpy as np

from mag2d_amd import abi
from oracle import oracle as O

AMU = 1.66053906660e-27
M_ION = 3 * AMU
M_HE = 6.69e-27
M_E = 9.11e-31
QE = abi.QE
KB = abi.KB


class Scene:
    """A 2D trap on an n x n grid: mask/voltage, potentials from the oracle's
    FP64 direct solve, device Fields and matching oracle Fields."""

    def __init__(self, ctx, n=128, L=2e-2, geometry=abi.GEO_RF_22PT, rf=True, Bz=0.0, Bt=0.0,
                 rf_amplitude=30.0, rf_freq=17e6, boundary=abi.BC_FREE, custom=None,
                 selfconsistent=False, dV=1e-12, solve_oracle=True):
        self.n, self.L = n, L
        dx = L / (n - 1)
        self.g = abi.grid2d(n, n, dx, dx)
        self.og = O.grid2(n, n, dx, dx)
        if custom is not None:
            self.mask, self.volt = custom(self.g)
        else:
            self.mask, self.volt = abi.grid_build(self.g, geometry, L, L)
        self.dV = dV
        self.u = np.zeros(n * n)
        self.urf = np.zeros(n * n)
        if solve_oracle:
            self.u = self.oracle_solve(np.zeros(n * n), O.SOLVE_STATIC)
            if rf:
                self.urf = self.oracle_solve(np.zeros(n * n), O.SOLVE_RF)
        omega = 2 * math.pi * rf_freq
        self.params = abi.FieldsParams(
            self.g, L, L, int(geometry == abi.GEO_EMPTY and custom is None), int(selfconsistent), 0,
            boundary, 0.0, int(rf), rf_amplitude, omega, 0.0, 0.0, Bz, Bt, dV)
        self.fields = abi.Fields2D(ctx, self.params, self.mask, self.volt)
        self.fields.upload(abi.FIELD_U, self.u)
        self.fields.upload(abi.FIELD_URF, self.urf)
        self.of = O.make_fields(self.og, self.mask, self.u, self.urf, x_max=L, z_max=L,
                                geometry_empty=int(geometry == abi.GEO_EMPTY and custom is None),
                                selfconsistent=int(selfconsistent), rf=int(rf),
                                rf_amplitude=rf_amplitude, rf_omega=omega, Bz=Bz, Bt=Bt,
                                boundary=boundary)

    def oracle_solve(self, rho, mode):
        b = np.array(rho, dtype=np.float64)
        O.lib().o_poisson_rhs(C.byref(self.og), O.dptr(self.mask), O.dptr(self.volt), O.dptr(b), mode,
                              self.dV)
        u = np.zeros_like(b)
        assert O.lib().o_poisson_solve(C.byref(self.og), O.dptr(self.mask), O.dptr(b), O.dptr(u)) == 0
        return u


def ions_on_disk(n, rng, cx=1e-2, cz=1e-2, r=5e-3, T_eV=0.1, mass=M_ION):
    ang = rng.uniform(0, 2 * np.pi, n)
    rad = r * np.sqrt(rng.uniform(0, 1, n))
    x = (cx + rad * np.cos(ang)).astype(np.float32)
    z = (cz + rad * np.sin(ang)).astype(np.float32)
    vth = math.sqrt(T_eV * QE / mass)
    v = (rng.normal(size=(3, n)) * vth).astype(np.float32)
    return x, z, v[0], v[1], v[2]


def to_oracle(x, z, vx, vy, vz, alive=None):
    p = O.particles(len(x))
    p["x"], p["z"], p["vx"], p["vy"], p["vz"] = x, z, vx, vy, vz
    if alive is not None:
        p["empty"] = (~alive).astype(np.uint8)
    return p


def species(charge, mass, dt, lifetime=float("inf"), niter=0, stream=0):
    return abi.Species2D(charge, mass, dt, lifetime, niter, stream, 0)
