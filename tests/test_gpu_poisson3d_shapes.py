"""3D Poisson solve on ragged grid shapes.

The 3D V-cycle runs colour-compact sweeps on the fine levels and the small
levels in one workgroup (in LDS when the hierarchy has the common shape);
which path runs, and how many levels there are, depends on the grid.  These
cases cover non-cubic grids, a thin grid whose coarsening stops early, and the
LDS path, with interior electrodes, against a direct sparse solve of the
reference's 7-point rows (fields3d.cpp:35-76) and the FP64 residual.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from mag2d_amd import abi

pytestmark = pytest.mark.gpu


def scene3(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    mask = np.full((nx, ny, nz), abi.FREE, np.uint8)
    volt = np.zeros((nx, ny, nz))
    mask[0], mask[-1] = abi.FIXED, abi.FIXED
    mask[:, 0], mask[:, -1] = abi.FIXED, abi.FIXED
    mask[:, :, 0], mask[:, :, -1] = abi.FIXED, abi.FIXED
    for _ in range(3):
        if min(nx, ny, nz) < 8:
            break
        o = [rng.integers(2, s - 5) for s in (nx, ny, nz)]
        e = [rng.integers(1, max(2, s // 5)) for s in (nx, ny, nz)]
        sl = tuple(slice(o[q], min((nx, ny, nz)[q] - 2, o[q] + e[q])) for q in range(3))
        mask[sl] = abi.FIXED
        volt[sl] = rng.choice([-3.0, 3.0])
    rho = np.where(mask == abi.FREE, rng.normal(size=(nx, ny, nz)) * 1e-3, 0.0)
    return mask.ravel(), volt.ravel(), rho.ravel()


def rows7(mask, nx, ny, nz):
    """Identity on Dirichlet nodes, [1 x6, -6] on free ones, row-major (i, j, k)."""
    N = nx * ny * nz
    idx = np.arange(N)
    d = mask != abi.FREE
    rows, cols, vals = [idx[d]], [idx[d]], [np.ones(d.sum())]
    f = idx[~d]
    rows.append(f); cols.append(f); vals.append(np.full(f.size, -6.0))
    for off in (1, nz, ny * nz):
        for sgn in (-1, 1):
            rows.append(f); cols.append(f + sgn * off); vals.append(np.ones(f.size))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


SHAPES = [(33, 17, 65), (48, 40, 24), (9, 9, 129), (41, 41, 41), (20, 70, 12)]


@pytest.mark.parametrize("nx,ny,nz", SHAPES)
def test_poisson3d_ragged_shapes(ctx, nx, ny, nz):
    h = 1e-4
    scale = 2.9e-2
    mask, volt, rho = scene3(nx, ny, nz, seed=nx * 10000 + ny * 100 + nz)
    g = abi.grid3d(nx, ny, nz, h, h, h)
    params = abi.Fields3DParams(g, h * (nx - 1), h * (ny - 1), h * (nz - 1), abi.BC_FREE, 0.0, 0.0, 0.0, scale)
    f = abi.Fields3D(ctx, params, mask, volt)
    f.upload(abi.FIELD_RHO, rho)
    tol = 1e-10
    info = f.solve(tol)
    u = f.download(abi.FIELD_U)
    assert info.rel_residual <= tol
    free = mask == abi.FREE
    assert np.array_equal(u[~free], volt[~free])
    # right-hand side of the rows: Solver::solve, fields3d.cpp:86-93
    b = np.where(free, -rho * scale, volt)
    A = rows7(mask, nx, ny, nz)
    r = (b - A @ u)[free]
    ud = np.where(free, 0.0, b)
    beff = (b - A @ ud)[free]
    assert np.linalg.norm(r) <= 1.01 * tol * np.linalg.norm(beff)
    if nx * ny * nz <= 40000:  # the direct solve stays quick
        uref = spla.spsolve(A.tocsc(), b)
        assert np.abs(u - uref).max() <= 1e-8 * np.abs(uref).max()
