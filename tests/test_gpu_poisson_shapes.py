"""Poisson solve on ragged grid shapes.

The solve runs its multigrid levels as LDS tiles with halos, a one-workgroup
coarse V-cycle and in-kernel reductions; the tile edges, the level count and
the kernel chosen depend on the grid's shape.  These cases cover odd and
non-square grids, thin grids whose coarsening stops early, grids above and
below the small-tile threshold (65536 nodes), and interior electrodes that
cut through the tiles.  Each is checked against the FP64 rows of the operator
(Fields::Fields, fields.cpp:229-237) and, where the oracle's direct solve is
quick, against that solution.
"""
import ctypes as C

import numpy as np
import pytest

from mag2d_amd import abi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def scene(nx, nz, seed):
    """Grounded frame, a few interior electrodes at +-5 V, random space charge."""
    rng = np.random.default_rng(seed)
    mask = np.full((nx, nz), abi.FREE, np.uint8)
    volt = np.zeros((nx, nz))
    mask[0, :] = mask[-1, :] = mask[:, 0] = mask[:, -1] = abi.FIXED
    for _ in range(4):
        if nx < 8 or nz < 8:
            break
        i0, j0 = rng.integers(2, nx - 5), rng.integers(2, nz - 5)
        di, dj = rng.integers(1, max(2, nx // 6)), rng.integers(1, max(2, nz // 6))
        sl = (slice(i0, min(nx - 2, i0 + di)), slice(j0, min(nz - 2, j0 + dj)))
        mask[sl] = abi.FIXED
        volt[sl] = rng.choice([-5.0, 5.0])
    rho = np.where(mask == abi.FREE, np.abs(rng.normal(size=(nx, nz))) * 1e-17, 0.0)
    return mask.ravel(), volt.ravel(), rho.ravel()


def free_residual(u, mask, volt, rho, nx, nz, scale):
    """|b - A u| and |b_eff| over the free rows, in FP64 numpy."""
    U = u.reshape(nx, nz)
    free = (mask == abi.FREE).reshape(nx, nz)
    pinned = ~free
    Au = np.zeros_like(U)
    Au[1:-1, 1:-1] = U[:-2, 1:-1] + U[2:, 1:-1] + U[1:-1, :-2] + U[1:-1, 2:] - 4 * U[1:-1, 1:-1]
    b = rho.reshape(nx, nz) * scale
    r = np.where(free, b - Au, 0.0)
    beff = b.copy()
    Up = np.where(pinned, U, 0.0)
    beff[1:-1, 1:-1] -= Up[:-2, 1:-1] + Up[2:, 1:-1] + Up[1:-1, :-2] + Up[1:-1, 2:]
    beff = np.where(free, beff, 0.0)
    return np.linalg.norm(r), np.linalg.norm(beff)


SHAPES = [(37, 53), (130, 67), (5, 200), (200, 6), (257, 129), (300, 260), (513, 97)]


@pytest.mark.parametrize("nx,nz", SHAPES)
def test_poisson_ragged_shapes(ctx, nx, nz):
    L = 2e-2
    dx = L / (max(nx, nz) - 1)
    dV = 1e-12
    mask, volt, rho = scene(nx, nz, seed=nx * 1000 + nz)
    g = abi.grid2d(nx, nz, dx, dx)
    params = abi.FieldsParams(g, dx * (nx - 1), dx * (nz - 1), 0, 1, 0, abi.BC_FREE, 0.0, 0, 0.0, 0.0, 0.0, 0.0,
                              0.0, 0.0, dV)
    f = abi.Fields2D(ctx, params, mask, volt)
    f.upload(abi.FIELD_U, np.zeros(nx * nz))
    f.upload(abi.FIELD_RHO, rho)
    tol = 1e-11
    info = f.solve(abi.SOLVE_STATIC, tol)
    u = f.download(abi.FIELD_U)
    assert info.rel_residual <= tol
    pinned = mask != abi.FREE
    assert np.array_equal(u[pinned], volt[pinned])
    scale = -(dx * dx) / abi.EPS0 / dV  # fields.cpp:307
    rn, bn = free_residual(u, mask, volt, rho, nx, nz, scale)
    assert rn <= 1.01 * tol * bn, (rn, bn)
    if nx * nz <= 20000:
        og = O.grid2(nx, nz, dx, dx)
        b = rho.copy()
        O.lib().o_poisson_rhs(C.byref(og), O.dptr(mask), O.dptr(volt), O.dptr(b), O.SOLVE_STATIC, dV)
        want = np.zeros_like(b)
        assert O.lib().o_poisson_solve(C.byref(og), O.dptr(mask), O.dptr(b), O.dptr(want)) == 0
        assert np.max(np.abs(u - want)) <= 1e-9 * np.max(np.abs(want))


def test_poisson_repeated_solves_are_bit_identical(ctx):
    """Same input, same bits: the in-kernel reductions add in a fixed order."""
    nx, nz = 300, 260
    L = 2e-2
    dx = L / (nx - 1)
    mask, volt, rho = scene(nx, nz, seed=5)
    g = abi.grid2d(nx, nz, dx, dx)
    params = abi.FieldsParams(g, dx * (nx - 1), dx * (nz - 1), 0, 1, 0, abi.BC_FREE, 0.0, 0, 0.0, 0.0, 0.0, 0.0,
                              0.0, 0.0, 1e-12)
    f = abi.Fields2D(ctx, params, mask, volt)
    out = []
    for _ in range(2):
        f.upload(abi.FIELD_U, np.zeros(nx * nz))
        f.upload(abi.FIELD_RHO, rho)
        f.solve(abi.SOLVE_STATIC, 1e-11)
        out.append(f.download(abi.FIELD_U))
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))


def test_poisson_large_grid_is_deterministic_and_converges_fast(ctx):
    """A grid whose V-cycle tiles outnumber the blocks the GPU holds at once
    (1537 x 1537: 9409 fine tiles, 37 249 on the first Galerkin level).  The
    ascent kernels read each tile's halo from the descent's array and write
    another, so blocks that start after their neighbours have finished still
    read the descent's values: the preconditioner stays the same linear map,
    the CG needs the same few iterations as on small grids, and two solves
    from the same start give the same bits."""
    nx = nz = 1537
    L = 2e-2
    dx = L / (nx - 1)
    dV = 1e-12
    mask, volt, rho = scene(nx, nz, seed=77)
    g = abi.grid2d(nx, nz, dx, dx)
    params = abi.FieldsParams(g, L, L, 0, 1, 0, abi.BC_FREE, 0.0, 0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, dV)
    f = abi.Fields2D(ctx, params, mask, volt)
    us, infos = [], []
    for _ in range(2):
        f.upload(abi.FIELD_U, np.zeros(nx * nz))
        f.upload(abi.FIELD_RHO, rho)
        infos.append(f.solve(abi.SOLVE_STATIC, 1e-10))
        us.append(f.download(abi.FIELD_U))
    assert np.array_equal(us[0], us[1])
    assert infos[0].rel_residual <= 1e-10
    # a V-cycle that mixed descent and ascent values took 144 iterations at 512^2
    assert infos[0].inner_iters <= 20, infos[0].inner_iters
    scale = -(dx * dx) / abi.EPS0 / dV
    rn, bn = free_residual(us[0], mask, volt, rho, nx, nz, scale)
    assert rn <= 1.01e-10 * bn


def test_grids_past_32bit_table_offsets_are_rejected(ctx):
    """The 2D particle gathers address the FP32 field tables by 32-bit byte
    offsets, so a 2D grid needs (nx+1) (nz+1) < 2^30: creation refuses a
    larger grid before it reads mask or voltage (tiny arrays stand in for
    them here)."""
    mask, volt = np.zeros(16, np.uint8), np.zeros(16)
    h = 1e-6
    g = abi.grid2d(32767, 32767, h, h)  # (nx+1) (nz+1) = 2^30
    params = abi.FieldsParams(g, h * 32766, h * 32766, 0, 0, 0, abi.BC_FREE, 0.0, 0, 0.0, 0.0, 0.0, 0.0, 0.0,
                              0.0, 1e-12)
    with pytest.raises(abi.M2DError, match="too large"):
        abi.Fields2D(ctx, params, mask, volt)
