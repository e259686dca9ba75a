"""Pins for the oracle's t_grid builders (o_grid_build, after
fields.cpp:417-868), from what each geometry fixes rather than from the
builder's own loops: connected components of the electrode nodes, their
areas, centroids, signs and voltages against the electrode list of
fields.cpp (typed here as a table of shapes), the grid symmetries of the
symmetric geometries, the angular harmonic of the solved RF basis of the
multipoles, and the BOUNDARY ring as the free 4-neighbours of FIXED nodes
inside [2, M-2).  The device builders equal the oracle's node for node
(tests/test_abi_cpu.py), so these pins cover both."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O

GEO_EMPTY, GEO_PROBE, GEO_RF_22PT, GEO_RF_8PT, GEO_RF_HAITRAP, GEO_RF_QUAD, GEO_MAC, GEO_PENNING, \
    GEO_PENNING_SIMPLE, GEO_TUBE = range(10)


def build(geo, nx, nz, x_max, z_max, probe_radius=2e-3, u_probe=-7.0, extern_field=250.0):
    g = O.grid2(nx, nz, x_max / (nx - 1), z_max / (nz - 1))
    mask = np.zeros(nx * nz, np.uint8)
    volt = np.zeros(nx * nz)
    O.lib().o_grid_build(C.byref(g), C.byref(O.Geom(geo, x_max, z_max, probe_radius, u_probe, extern_field)),
                         O.dptr(mask), O.dptr(volt))
    return g, mask.reshape(nx, nz), volt.reshape(nx, nz)


def components(sel):
    lab, n = ndimage.label(sel)  # 4-connectivity
    return lab, n


def check_boundary_ring(m):
    """BOUNDARY = nodes of [2, M-2) x [2, N-2) that are not FIXED and have a
    FIXED 4-neighbour; no BOUNDARY elsewhere (fields.cpp:507-518 and each
    builder's copy of that loop)."""
    M, N = m.shape
    fixed = m == O.FIXED
    nb = np.zeros_like(fixed)
    nb[1:, :] |= fixed[:-1, :]
    nb[:-1, :] |= fixed[1:, :]
    nb[:, 1:] |= fixed[:, :-1]
    nb[:, :-1] |= fixed[:, 1:]
    inner = np.zeros_like(fixed)
    inner[2:M - 2, 2:N - 2] = True
    want = nb & ~fixed & inner
    np.testing.assert_array_equal(m == O.BOUNDARY, want)


def check_frame(g, m, v, sides, linear, extern_field=250.0):
    """Outer frame: FIXED on i = M-1, j = 0, j = N-1 and, for 4 sides, i = 0;
    voltage 0 or the uniform field -E dz (j - N/2) (fields.cpp:421-433)."""
    M, N = m.shape
    frame = np.zeros((M, N), bool)
    frame[M - 1, :] = frame[:, 0] = frame[:, N - 1] = True
    if sides == 4:
        frame[0, :] = True
    assert (m[frame] == O.FIXED).all()
    j = np.broadcast_to(np.arange(N), (M, N))
    want = -extern_field * g.dz * (j - N // 2) if linear else np.zeros((M, N))
    np.testing.assert_allclose(v[frame], want[frame], rtol=0, atol=1e-12 * max(1.0, abs(want).max()))
    if sides == 3:
        # the axis i = 0 is open: no electrode touches it in these geometries
        inner_axis = m[0, 1:N - 1]
        assert ((inner_axis == O.FREE) | (inner_axis == O.BOUNDARY)).any()
    return frame


# ------------------------------------------------------------------ multipoles

# fields.cpp:494-505, 543-552, 584-593, 683-686: centre (1 cm, 1 cm); rod k at
# angle 2 pi (k + 1/32) / npoles measured from +z towards +x, sign -1 for even k
MULTIPOLES = {
    GEO_RF_22PT: dict(npoles=22, R=0.75e-2, r=0.05e-2, linear=False),
    GEO_RF_8PT: dict(npoles=8, R=0.4e-2, r=0.1e-2, linear=True),
    GEO_RF_HAITRAP: dict(npoles=8, R=0.31e-2, r=0.01e-2, linear=True),
}


@pytest.mark.parametrize("geo", sorted(MULTIPOLES))
def test_multipole_rods(geo):
    spec = MULTIPOLES[geo]
    n = 801 if geo == GEO_RF_HAITRAP else 401  # the 0.1 mm rods need dx < 0.05 mm
    L = 2e-2
    g, m, v = build(geo, n, n, L, L)
    frame = check_frame(g, m, v, 4, spec["linear"])
    check_boundary_ring(m)
    rods = m == O.FIXED_RF
    lab, k = components(rods)
    assert k == spec["npoles"]
    assert not (rods & frame).any()
    dx = g.dx
    area = math.pi * spec["r"] ** 2 / dx ** 2
    per = 2 * math.pi * spec["r"] / dx
    found = []
    for c in range(1, k + 1):
        ii, jj = np.nonzero(lab == c)
        # a closed disk of radius r on the node lattice
        assert abs(len(ii) - area) <= per + 4, (len(ii), area)
        xc, zc = ii.mean() * dx, jj.mean() * g.dz
        d = math.hypot(xc - 1e-2, zc - 1e-2)
        assert abs(d - spec["R"]) < dx, (d, spec["R"])
        r_max = np.sqrt((ii * dx - xc) ** 2 + (jj * g.dz - zc) ** 2).max()
        assert r_max <= spec["r"] + dx
        vals = np.unique(v[lab == c])
        assert len(vals) == 1 and abs(vals[0]) == 1.0
        found.append((math.atan2(xc - 1e-2, zc - 1e-2) % (2 * math.pi), vals[0]))
    found.sort()
    npo = spec["npoles"]
    for kk, (ang, sign) in enumerate(found):
        want = 2 * math.pi * (kk + 1.0 / 32) / npo
        assert abs(ang - want) < 2 * dx / spec["R"], (kk, ang, want)
        assert sign == (-1.0 if kk % 2 == 0 else 1.0)


def test_quadrupole_rods_and_symmetry():
    # fields.cpp:683-686: rods of r = 2 mm at (5, 10) and (15, 10) mm at +1,
    # at (10, 5) and (10, 15) mm at -1.  No node lies on a rod's rim (where
    # rounding would decide), so a quarter turn maps the node sets exactly.
    n = 190
    L = 2e-2
    g, m, v = build(GEO_RF_QUAD, n, n, L, L)
    check_frame(g, m, v, 4, False)
    check_boundary_ring(m)
    lab, k = components(m == O.FIXED_RF)
    assert k == 4
    want = {(5e-3, 10e-3): 1.0, (15e-3, 10e-3): 1.0, (10e-3, 5e-3): -1.0, (10e-3, 15e-3): -1.0}
    for c in range(1, 5):
        ii, jj = np.nonzero(lab == c)
        xc, zc = ii.mean() * g.dx, jj.mean() * g.dz
        key = min(want, key=lambda p: math.hypot(p[0] - xc, p[1] - zc))
        assert math.hypot(key[0] - xc, key[1] - zc) < 0.5 * g.dx
        assert abs(len(ii) - math.pi * 2e-3 ** 2 / g.dx ** 2) <= 2 * math.pi * 2e-3 / g.dx + 4
        assert (v[lab == c] == want[key]).all()
    # a quarter turn about the centre maps the rods onto each other with the sign flipped
    rot = np.rot90(m)
    np.testing.assert_array_equal(rot == O.FIXED_RF, m == O.FIXED_RF)
    rods = m == O.FIXED_RF
    np.testing.assert_array_equal(np.rot90(v)[rods], -v[rods])


def _solve_rf_basis(geo, n):
    L = 2e-2
    g, m, v = build(geo, n, n, L, L)
    b = np.zeros(n * n)
    mk, vk = m.ravel().copy(), v.ravel().copy()
    O.lib().o_poisson_rhs(C.byref(g), O.dptr(mk), O.dptr(vk), O.dptr(b), O.SOLVE_RF, 1e-12)
    u = np.zeros(n * n)
    assert O.lib().o_poisson_solve(C.byref(g), O.dptr(mk), O.dptr(b), O.dptr(u)) == 0
    return g, u.reshape(n, n)


@pytest.mark.parametrize("geo,n,harmonic,radius", [(GEO_RF_22PT, 257, 11, 0.55e-2), (GEO_RF_8PT, 161, 4, 0.25e-2),
                                                    (GEO_RF_QUAD, 161, 2, 0.25e-2)])
def test_rf_basis_is_the_multipole_field(geo, n, harmonic, radius):
    """The solved basis uRF (boundary_solve_rf, fields.cpp:314-348) on a circle
    inside the rods: alternating rods make uRF change sign under a turn by
    one rod pitch, so only odd multiples of npoles/2 appear in its angular
    spectrum, and npoles/2 dominates."""
    g, u = _solve_rf_basis(geo, n)
    th = np.linspace(0, 2 * np.pi, 512, endpoint=False)
    x = 1e-2 + radius * np.sin(th)
    z = 1e-2 + radius * np.cos(th)
    vals = np.array([_interp(g, u, a, b) for a, b in zip(x, z)])
    spec = np.abs(np.fft.rfft(vals)) / len(vals)
    k = int(np.argmax(spec[1:]) + 1)
    assert k == harmonic, (k, spec[:30])
    even = [spec[j] for j in range(1, 60) if j % harmonic == 0 and (j // harmonic) % 2 == 0]
    assert max(even) < 0.05 * spec[harmonic]
    assert spec[0] < 0.05 * spec[harmonic]


def _interp(g, u, x, z):
    X, Z = x / g.dx, z / g.dz
    i, j = int(X), int(Z)
    a, b = X - i, Z - j
    return (1 - a) * (1 - b) * u[i, j] + a * (1 - b) * u[i + 1, j] + (1 - a) * b * u[i, j + 1] + a * b * u[i + 1, j + 1]


# ------------------------------------------------------------------ probe, tube, empty

def test_probe_disk_at_the_centre():
    n = 161
    L = 2e-2
    g, m, v = build(GEO_PROBE, n, n, L, L, probe_radius=2.5e-3, u_probe=-7.0)
    frame = check_frame(g, m, v, 4, True)
    check_boundary_ring(m)
    lab, k = components((m == O.FIXED) & ~frame)
    assert k == 1
    ii, jj = np.nonzero(lab == 1)
    assert abs(ii.mean() - (n - 1) / 2) < 1e-9 and abs(jj.mean() - (n - 1) / 2) < 1e-9
    assert abs(len(ii) - math.pi * 2.5e-3 ** 2 / g.dx ** 2) <= 2 * math.pi * 2.5e-3 / g.dx + 4
    assert (v[lab == 1] == -7.0).all()
    # the probe and its ring are symmetric under the square's 8 symmetries
    core = (m == O.FIXED) & ~frame
    for t in (np.transpose, np.flipud, np.fliplr, lambda a: np.rot90(a, 2)):
        np.testing.assert_array_equal(t(core), core)
        np.testing.assert_array_equal(t(m == O.BOUNDARY), m == O.BOUNDARY)


def test_tube_is_the_complement_of_an_open_disk():
    n = 161
    L = 2e-2
    R = 7e-3
    g, m, v = build(GEO_TUBE, n, n, L, L, probe_radius=R)
    check_frame(g, m, v, 4, False)
    check_boundary_ring(m)
    fixed = m == O.FIXED
    assert (v[fixed] == 0.0).all()
    lab, k = components(~fixed)
    assert k == 1
    ii, jj = np.nonzero(~fixed)
    assert abs(ii.mean() - (n - 1) / 2) < 1e-9 and abs(jj.mean() - (n - 1) / 2) < 1e-9
    # the open disk r < R: every inner node strictly inside, none outside
    r = np.hypot(ii * g.dx - L / 2, jj * g.dz - L / 2)
    assert r.max() < R
    assert abs(len(ii) - math.pi * R ** 2 / g.dx ** 2) <= 2 * math.pi * R / g.dx + 4
    for t in (np.transpose, np.flipud, np.fliplr):
        np.testing.assert_array_equal(t(fixed), fixed)


def test_empty_is_the_frame_alone():
    n = 64
    g, m, v = build(GEO_EMPTY, n, n, 2e-2, 2e-2)
    frame = check_frame(g, m, v, 4, True)
    assert (m[~frame] == O.FREE).all()
    assert (m == O.BOUNDARY).sum() == 0


# ------------------------------------------------------------------ rectangle electrodes
# square_electrode(rmin, rmax, zmin, zmax, V) fills the nodes strictly inside
# the rectangle; later electrodes overwrite earlier ones.  Tables in mm, from
# fields.cpp:716-733 (MAC, threshold = u_probe), 759-771 (PENNING) and
# 806-818 (PENNING_SIMPLE, trap_voltage = -1).
TH = -7.0
MAC = [(5, 45, 10, 15, 0.0), (5, 7, 20, 80, 0.0), (5, 45, 85, 90, 0.0),
       (30, 33, 140, 170, 0.8 * TH), (45, 48, 180, 280, TH), (30, 33, 290, 320, TH), (25, 28, 320, 335, TH),
       (15, 45, 350, 353, 0.0), (0, 45, 395, 400, 3e3)]
PENNING = [(7.85, 8.35, 1, 25, -5.0), (0, 7.3, 15, 16, 10.0), (0, 3.5, 12, 19, 10.0), (0, 1.9, 1, 12, 10.0),
           (4, 7, 52, 53, -5.0), (2.5, 7, 46, 47, 0.0)]
PENNING_SIMPLE = [(10, 11, 0, 10, -0.5), (10, 11, 11, 20, 0.0), (10, 11, 21, 60, -1.0), (10, 11, 61, 75, -10.0)]


@pytest.mark.parametrize("geo,table,nx,nz,x_max,z_max", [
    (GEO_MAC, MAC, 200, 1601, 5.25e-2, 0.42),
    (GEO_PENNING, PENNING, 173, 1187, 9e-3, 6e-2),
    (GEO_PENNING_SIMPLE, PENNING_SIMPLE, 233, 1554, 1.2e-2, 8e-2),
])
def test_rectangle_electrodes(geo, table, nx, nz, x_max, z_max):
    g, m, v = build(geo, nx, nz, x_max, z_max, u_probe=TH)
    frame = check_frame(g, m, v, 3, False)
    check_boundary_ring(m)
    r = np.arange(nx)[:, None] * g.dx * 1e3
    z = np.arange(nz)[None, :] * g.dz * 1e3
    # the grids are chosen so that no node lies within 1e-4 mm of an edge
    # (edges at 0 exclude the axis node exactly)
    owner = np.full((nx, nz), -1)
    for k, (r0, r1, z0, z1, V) in enumerate(table):
        for a, e in ((r, r0), (r, r1), (z, z0), (z, z1)):
            assert e == 0 or np.abs(a - e).min() > 1e-4
        inside = (r > r0) & (r < r1) & (z > z0) & (z < z1)
        assert inside.any(), k
        owner[inside] = k
    elec = owner >= 0
    # every electrode node, and no other node off the frame, is FIXED
    assert (m[elec] == O.FIXED).all()
    assert (m[~elec & ~frame] != O.FIXED).all()
    want = np.array([t[4] for t in table])
    np.testing.assert_array_equal(v[elec], want[owner[elec]])
    # each connected piece of electrode nodes is a union of table rectangles
    lab, k = components(elec)
    assert k >= 1
    for c in range(1, k + 1):
        ks = np.unique(owner[lab == c])
        assert ((lab == c) == np.isin(owner, ks)).all()


def test_rectangle_electrode_bounds_on_a_coarse_grid():
    """The strict inequalities of square_electrode (fields.cpp:835-850): a
    rectangle whose edges fall between nodes covers exactly the nodes between
    them, counted independently of the builder."""
    nx, nz = 21, 31
    g = O.grid2(nx, nz, 1e-3, 1e-3)
    mask = np.full(nx * nz, O.FREE, np.uint8)
    volt = np.zeros(nx * nz)
    O.lib().o_square_electrode(C.byref(g), O.dptr(mask), O.dptr(volt), 2.5e-3, 7.5e-3, 10.2e-3, 20.7e-3, 3.0,
                               O.FIXED)
    m = mask.reshape(nx, nz)
    ii, jj = np.nonzero(m == O.FIXED)
    assert (ii.min(), ii.max(), jj.min(), jj.max()) == (3, 7, 11, 20)
    assert len(ii) == 5 * 10
    # disks on the node lattice: radius just above and just below 2 nodes
    for r, count in ((2.0001e-3, 13), (1.9999e-3, 9)):
        mask[:] = O.FREE
        O.lib().o_circle_electrode(C.byref(g), O.dptr(mask), O.dptr(volt), 10e-3, 15e-3, r, 1.0, O.FIXED_RF)
        m = mask.reshape(nx, nz)
        ii, jj = np.nonzero(m == O.FIXED_RF)
        assert len(ii) == count
        assert ii.mean() == 10 and jj.mean() == 15
