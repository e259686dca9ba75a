"""m2d_accumulate_2d_noalloc (the entry point that allocates nothing: its
fixed-point grid is a static per-GPU device array) against
m2d_accumulate_2d, bit for bit, on two calls in a row, the second with fewer
particles on a smaller grid, called through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from mag2d_amd import abi

pytestmark = pytest.mark.gpu


def _call(fn, ctx, g, P, q, n_nodes):
    rho = ctx.to_device(np.full(n_nodes, 0.25))  # the deposit adds to what rho holds
    s = P.struct()
    abi.check(fn(ctx.h, C.byref(g), C.byref(s), q, C.c_void_p(rho.ptr)))
    return rho.get()


@pytest.mark.parametrize("seed", [0])
def test_noalloc_accumulate_matches_accumulate(ctx, seed):
    rng = np.random.default_rng(seed)
    L = abi.lib()
    for n_grid, N, q in ((512, 400_000, 1.602189e-19), (96, 30_001, -1.602189e-19)):
        dx = 2e-2 / (n_grid - 1)
        g = abi.grid2d(n_grid, n_grid, dx, dx)
        x = rng.uniform(0, 2e-2, N).astype(np.float32)
        z = rng.uniform(0, 2e-2, N).astype(np.float32)
        dead = rng.uniform(size=N) < 0.1
        x[dead] = np.array([abi.DEAD_BITS], np.uint32).view(np.float32)[0]
        zero = np.zeros(N, np.float32)
        P = abi.ParticlesSoA.from_host(ctx, x, z, zero, zero, zero)
        print(f"noalloc: grid {n_grid}^2, {N} particles", flush=True)
        got = _call(L.m2d_accumulate_2d_noalloc, ctx, g, P, q, n_grid * n_grid)
        want = _call(L.m2d_accumulate_2d, ctx, g, P, q, n_grid * n_grid)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
        assert (got - 0.25).sum() == pytest.approx((~dead).sum() * q, rel=1e-9)
    assert ctx.status(clear=True) == 0


def test_noalloc_rejects_grids_beyond_its_scratch(ctx):
    g = abi.grid2d(1025, 1024, 1e-5, 1e-5)
    P = abi.ParticlesSoA(ctx, 4)
    rho = ctx.to_device(np.zeros(1))
    s = P.struct()
    code = abi.lib().m2d_accumulate_2d_noalloc(ctx.h, C.byref(g), C.byref(s), 1.0, C.c_void_p(rho.ptr))
    assert code == -1
    assert b"static scratch" in abi.lib().m2d_last_error()
