"""Fields::u_smooth (fields.cpp:28-113) on the device (m2d_fields2d_u_smooth)
against the oracle: the filter multiplies only by powers of two, so the FP64
results are bit-identical; the gather tables follow the smoothed field."""
import ctypes as C

import numpy as np
import pytest

from m2d_testlib import Scene
from mag2d_amd import abi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def fields(ctx, nx, nz, h=3e-4):
    g = abi.grid2d(nx, nz, h, h)
    mask, volt = abi.grid_build(g, abi.GEO_EMPTY, h * (nx - 1), h * (nz - 1))
    params = abi.FieldsParams(g, h * (nx - 1), h * (nz - 1), 0, 1, 0, abi.BC_FREE, 0.0, 1, 10.0, 1e8, 0.0, 0.0,
                              0.0, 0.0, 1e-12)
    return abi.Fields2D(ctx, params, mask, volt), O.grid2(nx, nz, h, h)


@pytest.mark.parametrize("nx,nz,symmetry,radius,which", [
    (37, 53, 0, -1.0, abi.FIELD_U),
    (65, 65, 1, -1.0, abi.FIELD_U),
    (65, 65, 1, 6e-3, abi.FIELD_URF),
    (200, 130, 0, 1.5e-2, abi.FIELD_U),
])
def test_u_smooth_is_bit_identical_to_oracle(ctx, nx, nz, symmetry, radius, which):
    f, og = fields(ctx, nx, nz)
    rng = np.random.default_rng(nx * nz + symmetry)
    u = rng.normal(size=nx * nz) * 3.0
    f.upload(which, u)
    f.u_smooth(which, symmetry=bool(symmetry), radius=radius)
    got = f.download(which)
    want = u.copy()
    assert O.lib().o_u_smooth(C.byref(og), O.dptr(want), symmetry, radius) == 0
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    assert not np.array_equal(got, u)


def test_u_smooth_rejects_symmetry_on_a_non_square_grid(ctx):
    f, _ = fields(ctx, 33, 40)
    with pytest.raises(abi.M2DError, match="non-square"):
        f.u_smooth(abi.FIELD_U, symmetry=True)


def test_gather_tables_follow_the_smoothed_field(ctx):
    """Fields::E after u_smooth is the gradient of the smoothed potential."""
    sc = Scene(ctx, n=96)
    sc.fields.u_smooth(abi.FIELD_U, symmetry=True)
    u_s = sc.u.copy()
    assert O.lib().o_u_smooth(C.byref(sc.og), O.dptr(u_s), 1, -1.0) == 0
    sc.of._keep[1][:] = u_s  # the array the oracle fields read
    rng = np.random.default_rng(7)
    xs, zs = rng.uniform(0.003, 0.017, 2000), rng.uniform(0.003, 0.017, 2000)
    ex, ez = sc.fields.E(xs, zs, 0.0)
    rx, rz = np.zeros(len(xs)), np.zeros(len(xs))
    for k in range(len(xs)):
        a, b = C.c_double(), C.c_double()
        O.lib().o_fields_E(C.byref(sc.of), xs[k], zs[k], 0.0, C.byref(a), C.byref(b))
        rx[k], rz[k] = a.value, b.value
    scale = max(np.abs(rx).max(), np.abs(rz).max())
    assert np.max(np.abs(ex - rx)) < 1e-5 * scale
    assert np.max(np.abs(ez - rz)) < 1e-5 * scale
