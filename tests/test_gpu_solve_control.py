"""The solves' convergence control on the device.

The 2D and 3D solves queue their FP64 refinement passes ahead; one-thread
control kernels stop them on the device, and the host reads the control
block once a solve (SolveCtl and queue_refinement, m2d_internal.hpp).  These
cases drive the paths a host-side loop would take by itself:
- a residual that is not finite: M2D_ERR_NAN, with nothing added to u;
- a tolerance that 12 refinement passes cannot reach: M2D_ERR_NOT_CONVERGED
  after exactly 12 passes, with u refined all the same;
- solves whose needs jump past what the previous solve on the same fields
  queued: a pass that runs out of queued iterations is suspended on the
  device and resumed, so a cold solve after a warm one (queued for no pass)
  repeats the first cold solve bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from mag2d_amd import abi
from test_gpu_poisson3d_shapes import rows7, scene3
from test_gpu_poisson_shapes import free_residual, scene

pytestmark = pytest.mark.gpu

ERR_NAN, ERR_NOT_CONVERGED = -3, -5
NX, NZ = 130, 67


def fields2d(ctx, seed=5):
    L = 2e-2
    dx = L / (max(NX, NZ) - 1)
    dV = 1e-12
    mask, volt, rho = scene(NX, NZ, seed)
    g = abi.grid2d(NX, NZ, dx, dx)
    params = abi.FieldsParams(g, dx * (NX - 1), dx * (NZ - 1), 0, 1, 0, abi.BC_FREE, 0.0, 0, 0.0, 0.0, 0.0, 0.0,
                              0.0, 0.0, dV)
    f = abi.Fields2D(ctx, params, mask, volt)
    f.upload(abi.FIELD_U, np.zeros(NX * NZ))
    f.upload(abi.FIELD_RHO, rho)
    scale = -(dx * dx) / abi.EPS0 / dV  # fields.cpp:307
    return f, mask, volt, rho, scale


def solve2d(f, rho, tol):
    """(status, info) of m2d_fields2d_solve on rho, without raising (the
    solve overwrites the charge array with its right-hand side, so rho goes
    up each time)"""
    f.upload(abi.FIELD_RHO, rho)
    info = abi.SolveInfo()
    code = abi.lib().m2d_fields2d_solve(f.h, abi.SOLVE_STATIC, tol, C.byref(info))
    return code, info


def rel2d(f, mask, volt, rho, scale):
    rn, bn = free_residual(f.download(abi.FIELD_U), mask, volt, rho, NX, NZ, scale)
    return rn / bn


def test_2d_residual_not_finite_adds_nothing(ctx):
    f, mask, volt, rho, scale = fields2d(ctx)
    assert solve2d(f, rho, 1e-11)[0] == 0
    u1 = f.download(abi.FIELD_U)
    bad = rho.copy()
    bad[np.flatnonzero(mask == abi.FREE)[100]] = np.nan
    code, info = solve2d(f, bad, 1e-11)
    assert code == ERR_NAN and info.outer_iters == 0 and info.inner_iters == 0
    assert np.array_equal(f.download(abi.FIELD_U), u1)
    # the next solve starts from the same u, which already meets the tolerance
    code, info = solve2d(f, rho, 1e-11)
    assert code == 0 and info.outer_iters == 0 and info.inner_iters == 0
    assert np.array_equal(f.download(abi.FIELD_U), u1)


def test_2d_unreachable_tolerance_stops_after_12_passes(ctx):
    f, mask, volt, rho, scale = fields2d(ctx)
    code, info = solve2d(f, rho, 1e-30)
    assert code == ERR_NOT_CONVERGED
    assert info.outer_iters == 12 and info.inner_iters >= 12
    assert info.rel_residual <= 1e-12
    assert rel2d(f, mask, volt, rho, scale) <= 1e-12


def test_2d_solves_whose_needs_jump_converge(ctx):
    """A warm solve that needs no pass, then a cold one, then one that needs
    longer passes than any before it: each meets the tolerance, and the cold
    solve repeats the first one's result bit for bit."""
    f, mask, volt, rho, scale = fields2d(ctx)
    tol = 1e-11
    code, first = solve2d(f, rho, tol)
    u_first = f.download(abi.FIELD_U)
    assert code == 0 and first.outer_iters >= 1 and rel2d(f, mask, volt, rho, scale) <= 1.01 * tol
    code, warm = solve2d(f, rho, tol)
    assert code == 0 and warm.outer_iters == 0
    # cold again, queued from the warm solve's estimate of no passes
    f.upload(abi.FIELD_U, np.zeros(NX * NZ))
    code, cold = solve2d(f, rho, tol)
    assert code == 0 and (cold.outer_iters, cold.inner_iters) == (first.outer_iters, first.inner_iters)
    assert np.array_equal(f.download(abi.FIELD_U), u_first)
    # a tighter tolerance from zero: longer passes than any solve queued before
    f.upload(abi.FIELD_U, np.zeros(NX * NZ))
    code, tight = solve2d(f, rho, 1e-13)
    assert code == 0 and tight.inner_iters > first.inner_iters
    assert rel2d(f, mask, volt, rho, scale) <= 1.01e-13


def fields3d(ctx, shape=(33, 17, 65), seed=11):
    nx, ny, nz = shape
    h, scale = 1e-4, 2.9e-2
    mask, volt, rho = scene3(nx, ny, nz, seed)
    g = abi.grid3d(nx, ny, nz, h, h, h)
    params = abi.Fields3DParams(g, h * (nx - 1), h * (ny - 1), h * (nz - 1), abi.BC_FREE, 0.0, 0.0, 0.0, scale)
    f = abi.Fields3D(ctx, params, mask, volt)
    return f, mask, volt, rho, scale


def solve3d(f, rho, tol):
    """(status, info) of m2d_fields3d_solve on rho (the solve overwrites the
    charge array with its right-hand side, so rho goes up each time)"""
    f.upload(abi.FIELD_RHO, rho)
    info = abi.SolveInfo()
    code = abi.lib().m2d_fields3d_solve(f.h, tol, C.byref(info))
    return code, info


def rel3d(f, mask, volt, rho, scale, shape=(33, 17, 65)):
    u = f.download(abi.FIELD_U)
    free = mask == abi.FREE
    b = np.where(free, -rho * scale, volt)  # Solver::solve, fields3d.cpp:86-93
    A = rows7(mask, *shape)
    ud = np.where(free, 0.0, b)
    return np.linalg.norm((b - A @ u)[free]) / np.linalg.norm((b - A @ ud)[free])


def test_3d_residual_not_finite_adds_nothing(ctx):
    f, mask, volt, rho, scale = fields3d(ctx)
    code, _ = solve3d(f, rho, 1e-10)
    assert code == 0
    u1 = f.download(abi.FIELD_U)
    bad = rho.copy()
    bad[np.flatnonzero(mask == abi.FREE)[50]] = np.inf
    code, info = solve3d(f, bad, 1e-10)
    assert code == ERR_NAN and info.outer_iters == 0 and info.inner_iters == 0
    assert np.array_equal(f.download(abi.FIELD_U), u1)
    code, info = solve3d(f, rho, 1e-10)
    assert code == 0 and info.outer_iters == 0
    assert np.array_equal(f.download(abi.FIELD_U), u1)


def test_3d_unreachable_tolerance_stops_after_12_passes(ctx):
    f, mask, volt, rho, scale = fields3d(ctx)
    code, info = solve3d(f, rho, 1e-30)
    assert code == ERR_NOT_CONVERGED
    assert info.outer_iters == 12 and info.rel_residual <= 1e-12
    assert rel3d(f, mask, volt, rho, scale) <= 1e-12


def test_3d_solves_whose_needs_jump_converge(ctx):
    f, mask, volt, rho, scale = fields3d(ctx)
    code, first = solve3d(f, rho, 1e-10)
    assert code == 0 and first.outer_iters >= 1
    u_first = f.download(abi.FIELD_U)
    code, warm = solve3d(f, rho, 1e-10)
    assert code == 0 and warm.outer_iters == 0
    f.upload(abi.FIELD_U, np.zeros(mask.size))
    code, cold = solve3d(f, rho, 1e-10)
    assert code == 0 and (cold.outer_iters, cold.inner_iters) == (first.outer_iters, first.inner_iters)
    assert np.array_equal(f.download(abi.FIELD_U), u_first)
    f.upload(abi.FIELD_U, np.zeros(mask.size))
    code, tight = solve3d(f, rho, 1e-13)
    assert code == 0 and tight.inner_iters > first.inner_iters
    assert rel3d(f, mask, volt, rho, scale) <= 1.01e-13
