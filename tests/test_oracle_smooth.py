"""Pin the oracle's Fields::u_smooth (fields.cpp:28-113) by what its
stencil fixes.  The filter (c + edges/2 + diagonals/4) / 4 has weights
summing to 1 and zero first moments: it keeps linear fields exactly, adds
h^2/2 to x^2 (second moment (1 + 1) / 4 in cells^2), and removes the
checkerboard (1 - 4/2 + 4/4 = 0).  The symmetry pass makes the field
invariant under the square's 8 symmetries, and the filter keeps that."""
import ctypes as C

import numpy as np

from oracle import oracle as O


def smooth(u, nx, nz, h=1e-3, symmetry=0, radius=-1.0):
    g = O.grid2(nx, nz, h, h)
    out = np.array(u, np.float64).ravel().copy()
    rc = O.lib().o_u_smooth(C.byref(g), O.dptr(out), symmetry, radius)
    return rc, out.reshape(nx, nz)


def test_filter_keeps_linear_adds_half_h2_to_quadratic_removes_checkerboard():
    nx, nz, h = 23, 31, 1e-3
    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    lin = 3.0 + 2.0 * i - 5.0 * j
    rc, s = smooth(lin, nx, nz, h)
    assert rc == 0 and np.array_equal(s, lin)
    x = i * h
    rc, s = smooth(x * x, nx, nz, h)
    np.testing.assert_allclose(s[1:-1, 1:-1], (x * x + 0.5 * h * h)[1:-1, 1:-1], rtol=1e-12, atol=1e-20)
    chk = (-1.0) ** (i + j)
    rc, s = smooth(chk, nx, nz, h)
    assert np.all(s[1:-1, 1:-1] == 0.0)
    # the frame is left as it is
    assert np.array_equal(s[0], chk[0]) and np.array_equal(s[-1], chk[-1])
    assert np.array_equal(s[:, 0], chk[:, 0]) and np.array_equal(s[:, -1], chk[:, -1])


def test_radius_limits_the_filter():
    nx = nz = 41
    h = 1e-3
    rng = np.random.default_rng(1)
    u = rng.normal(size=(nx, nz))
    rc, s = smooth(u, nx, nz, h, radius=8.5e-3)
    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    far = (i - 20.0) ** 2 + (j - 20.0) ** 2 > 8.5 ** 2
    assert np.array_equal(s[far], u[far])
    inner = ~far
    inner[0], inner[-1], inner[:, 0], inner[:, -1] = False, False, False, False
    assert np.all(s[inner] != u[inner])


def test_symmetry_pass_gives_a_symmetric_field():
    n = 33
    rng = np.random.default_rng(2)
    u = rng.normal(size=(n, n))
    rc, s = smooth(u, n, n, symmetry=1)
    assert rc == 0
    for t in (s[::-1, :], s[:, ::-1], s.T, s[::-1, ::-1].T):
        np.testing.assert_allclose(t, s, rtol=1e-14, atol=1e-15)
    # a node away from the centre takes the mean of its 8 mirror images (a
    # radius below one cell leaves only the centre node to the filter)
    rc, sym_only = smooth(u, n, n, symmetry=1, radius=1e-9)
    c = n - 1
    orbit = [(5, 2), (c - 5, 2), (5, c - 2), (c - 5, c - 2), (c - 2, c - 5), (2, 5), (c - 2, 5), (2, c - 5)]
    want = sum(u[a, b] for a, b in orbit) / 8.0
    for a, b in orbit:
        assert sym_only[a, b] == want
    assert smooth(u[:, :-1], n, n - 1, symmetry=1)[0] == -1
