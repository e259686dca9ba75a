"""GPU parity of the 3D path (Field3D, Species<CARTESIAN3D>, the 3D solve)
against the oracle (tests/test_oracle_3d.py pins it) and scipy."""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from mag2d_amd import abi
from oracle import oracle as O

pytestmark = pytest.mark.gpu

QE = abi.QE
AMU = 1.66053906660e-27
M40 = 40 * AMU


def quad_trap(n=33, L=1e-2, V=2.0, r0=0.3, a=0.08):
    """Grounded box with four rods along z at radius r0*L around the axis
    (+V, -V, +V, -V): a linear quadrupole."""
    h = L / (n - 1)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    X, Y = i * h - L / 2, j * h - L / 2
    mask = np.full((n, n, n), abi.FREE, np.uint8)
    volt = np.zeros((n, n, n))
    face = (i == 0) | (j == 0) | (k == 0) | (i == n - 1) | (j == n - 1) | (k == n - 1)
    mask[face] = abi.FIXED
    for q, (cx, cy) in enumerate([(r0, 0), (0, r0), (-r0, 0), (0, -r0)]):
        rod = (X - cx * L) ** 2 + (Y - cy * L) ** 2 <= (a * L) ** 2
        mask[rod] = abi.FIXED
        volt[rod] = V if q % 2 == 0 else -V
    return mask.ravel(), volt.ravel(), h


class Scene3:
    def __init__(self, ctx, n=33, L=1e-2, B=(0.0, 0.0, 0.5), boundary=abi.BC_FREE, rhs_scale=1e-3):
        self.n, self.L = n, L
        self.mask, self.volt, self.h = quad_trap(n, L)
        self.g = abi.grid3d(n, n, n, self.h, self.h, self.h)
        self.params = abi.Fields3DParams(self.g, L, L, L, boundary, B[0], B[1], B[2], rhs_scale)
        self.fields = abi.Fields3D(ctx, self.params, self.mask, self.volt)
        self.og = O.grid3(n, n, n, self.h, self.h, self.h)
        self.B, self.boundary = B, boundary

    def oracle_fields(self, u):
        return O.make_fields3(self.og, self.mask, u, boundary=self.boundary, B=self.B)


def laplace_matrix(mask, n):
    """The reference's 7-point rows (fields3d.cpp:35-76): identity on
    Dirichlet nodes, [1 x6, -6] elsewhere."""
    N = n ** 3
    idx = np.arange(N)
    d = (mask == abi.FIXED) | (mask == abi.FIXED_RF)
    rows, cols, vals = [idx[d]], [idx[d]], [np.ones(d.sum())]
    f = idx[~d]
    rows.append(f); cols.append(f); vals.append(np.full(f.size, -6.0))
    for off in (1, n, n * n):
        for sgn in (-1, 1):
            rows.append(f); cols.append(f + sgn * off); vals.append(np.ones(f.size))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


def test_field3d_point_kernels_match_oracle(ctx):
    n, h = 17, 1e-3
    g = abi.grid3d(n, n, n, h, h, h)
    og = O.grid3(n, n, n, h, h, h)
    rng = np.random.default_rng(0)
    data = rng.normal(size=n ** 3)
    d = ctx.to_device(data)
    L = (n - 1) * h
    pts = rng.uniform(0, L, (400, 3))
    pts[:4] = [[0, 0, 0], [L * 0.999999, L * 0.5, 0], [L * 0.5, 1e-12, L * 0.999], [h, 2 * h, 3 * h]]
    x, y, z = pts.T.copy()
    gi = abi.interpolate_3d(ctx, g, d.ptr, x, y, z)
    gx, gy, gz = abi.grad_3d(ctx, g, d.ptr, x, y, z)
    for k in range(len(x)):
        o = C.c_double()
        assert O.lib().o_f3_interpolate(C.byref(og), O.dptr(data), x[k], y[k], z[k], C.byref(o)) == 0
        assert gi[k] == pytest.approx(o.value, rel=1e-12, abs=1e-12)
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        O.lib().o_f3_grad(C.byref(og), O.dptr(data), x[k], y[k], z[k], C.byref(a), C.byref(b), C.byref(c))
        assert (gx[k], gy[k], gz[k]) == pytest.approx((a.value, b.value, c.value), rel=1e-11, abs=1e-6)


def test_accumulate_3d_matches_oracle_and_is_order_independent(ctx):
    n, h = 17, 1e-3
    g = abi.grid3d(n, n, n, h, h, h)
    og = O.grid3(n, n, n, h, h, h)
    rng = np.random.default_rng(1)
    N = 50000
    pos = rng.uniform(0, (n - 1) * h * 0.9999, (3, N)).astype(np.float32)
    zeros = np.zeros(N, np.float32)
    ref = np.zeros(n ** 3)
    for k in range(N):
        O.lib().o_f3_accumulate(C.byref(og), O.dptr(ref), -QE, float(pos[0, k]), float(pos[1, k]), float(pos[2, k]))
    outs = []
    for perm in (np.arange(N), rng.permutation(N)):
        P = abi.ParticlesSoA3D.from_host(ctx, pos[0][perm], pos[1][perm], pos[2][perm], zeros, zeros, zeros)
        rho = abi.DeviceArray(ctx, (n ** 3,), np.float64)
        rho.zero()
        abi.accumulate_3d(ctx, g, P, -QE, rho.ptr)
        outs.append(rho.get())
    assert np.array_equal(outs[0], outs[1])  # integer sums: bit for bit in any order
    np.testing.assert_allclose(outs[0], ref, rtol=1e-6, atol=1e-6 * QE)
    assert outs[0].sum() == pytest.approx(-QE * N, rel=1e-12)


def test_poisson3d_matches_direct_solve_and_fp64_residual(ctx):
    sc = Scene3(ctx, n=25, rhs_scale=3.7e-2)
    rng = np.random.default_rng(2)
    rho = rng.normal(size=sc.n ** 3) * 1e-3
    sc.fields.upload(abi.FIELD_RHO, rho)
    info = sc.fields.solve(1e-10)
    u = sc.fields.download(abi.FIELD_U)
    assert info.rel_residual <= 1e-10
    b = rho.copy()
    O.lib().o_poisson3_rhs(C.byref(sc.og), O.dptr(sc.mask), O.dptr(sc.volt), O.dptr(b), 3.7e-2)
    A = laplace_matrix(sc.mask, sc.n)
    # the reference solves with the transposed CSC (UMFPACK_At): the row form here
    uref = spla.spsolve(A.tocsc(), b)
    assert np.abs(u - uref).max() <= 1e-9 * np.abs(uref).max()
    Au = np.zeros_like(u)
    assert O.lib().o_poisson3_apply(C.byref(sc.og), O.dptr(sc.mask), O.dptr(u), O.dptr(Au)) == 0
    free = sc.mask == abi.FREE
    # b_eff: the free rows' rhs with the Dirichlet neighbours moved over (what the solver normalises by)
    ud = np.where(free, 0.0, b)
    Aud = np.zeros_like(ud)
    O.lib().o_poisson3_apply(C.byref(sc.og), O.dptr(sc.mask), O.dptr(ud), O.dptr(Aud))
    beff = (b - Aud)[free]
    assert np.linalg.norm((b - Au)[free]) <= 1.01e-10 * np.linalg.norm(beff)


def _ions(n, rng, L, T_eV=0.05, mass=M40, spread=0.12):
    pos = (L / 2 + rng.normal(size=(3, n)) * spread * L).astype(np.float32)
    vth = math.sqrt(T_eV * QE / mass)
    vel = (rng.normal(size=(3, n)) * vth).astype(np.float32)
    return pos, vel


def _oracle_parts(pos, vel):
    p = O.particles(pos.shape[1])
    p["x"], p["y"], p["z"] = pos
    p["vx"], p["vy"], p["vz"] = vel
    return p


def test_advance3d_magnetised_quadrupole_100_steps_matches_oracle(ctx):
    sc = Scene3(ctx)
    sc.fields.reset()
    sc.fields.solve(1e-10)
    u = sc.fields.download(abi.FIELD_U)
    of = sc.oracle_fields(u)
    rng = np.random.default_rng(3)
    N = 4000
    pos, vel = _ions(N, rng, sc.L)
    P = abi.ParticlesSoA3D.from_host(ctx, *pos, *vel)
    op = _oracle_parts(pos, vel)
    dt = 5e-9
    sp_ = abi.Species2D(QE, M40, dt, float("inf"), 0, 0, 0)
    os_ = O.make_species(QE, M40, dt)
    abi.advance_init_3d(ctx, sc.fields, sp_, P)
    O.lib().o_advance3d_init(C.byref(of), C.byref(os_), O.dptr(op), N)
    steps = 100
    for k in range(steps):
        sp_.niter = k
        os_.niter = k
        abi.advance_3d(ctx, sc.fields, sp_, P, flags=abi.STEP_PUSH | abi.STEP_BOUNDARY)
        O.lib().o_advance3d(C.byref(of), C.byref(os_), O.dptr(op), N, None, None)
    d = P.download()
    alive = P.alive()
    np.testing.assert_array_equal(alive, op["empty"] == 0)
    assert alive.sum() > 0.5 * N
    tol = steps * np.spacing(np.float32(sc.L))
    for c in ("x", "y", "z"):
        assert np.abs(d[c][alive] - op[c][alive]).max() <= tol, c
    vs = np.abs(op["vx"][alive]).max()
    for c in ("vx", "vy", "vz"):
        assert np.abs(d[c][alive] - op[c][alive]).max() <= 1e-4 * vs, c


@pytest.mark.parametrize("itype", [abi.ELASTIC, abi.CX])
def test_advance3d_mcc_same_draws_as_oracle(ctx, itype):
    sc = Scene3(ctx, B=(0.0, 0.0, 0.0))
    u = sc.fields.download(abi.FIELD_U)
    of = sc.oracle_fields(u)
    rng = np.random.default_rng(4)
    N = 3000
    pos, vel = _ions(N, rng, sc.L, spread=0.05)
    M_HE = 6.69e-27
    rate = 5e-14
    mcc = abi.Mcc(ctx, M40, [abi.Target(M_HE, 1e20, 300.0, 0.0, abi.Particles2D())],
                  [abi.Interaction(itype, 0, 0.0, rate, 0.0, 0, 0)], 1.0)
    omcc = O.make_mcc(M40, 1.0, [dict(mass=M_HE, density=1e20, temperature=300.0,
                                      inters=[O.inter(itype, M40, M_HE, rate=rate)])])
    assert mcc.lifetime == pytest.approx(omcc.lifetime, rel=1e-12)
    dt = 1e-8
    P = abi.ParticlesSoA3D.from_host(ctx, *pos, *vel)
    op = _oracle_parts(pos, vel)
    sp_ = abi.Species2D(QE, M40, dt, mcc.lifetime, 0, 7, 0)
    os_ = O.make_species(QE, M40, dt, lifetime=omcc.lifetime, mcc=omcc, seed=1234, stream=7)
    coll = 0
    for k in range(20):
        sp_.niter = k
        os_.niter = k
        st = abi.advance_3d(ctx, sc.fields, sp_, P, mcc=mcc, flags=abi.STEP_PUSH | abi.STEP_BOUNDARY,
                            want_stats=True)
        nc = C.c_long()
        O.lib().o_advance3d(C.byref(of), C.byref(os_), O.dptr(op), N, C.byref(nc), None)
        assert st.collisions == nc.value
        coll += nc.value
    assert coll > 50
    d = P.download()
    alive = P.alive()
    np.testing.assert_array_equal(alive, op["empty"] == 0)
    vs = np.abs(op["vx"][alive]).max()
    for c in ("vx", "vy", "vz"):
        assert np.abs(d[c][alive] - op[c][alive]).max() <= 1e-4 * vs, c


def test_advance3d_periodic_wrap_matches_oracle(ctx):
    n, L = 17, 1e-2
    h = L / (n - 1)
    mask = np.full((n, n, n), abi.FREE, np.uint8)
    mask[[0, -1], :, :] = abi.FIXED
    mask[:, [0, -1], :] = abi.FIXED
    mask[:, :, [0, -1]] = abi.FIXED
    mask = mask.ravel()
    volt = np.zeros(n ** 3)
    g = abi.grid3d(n, n, n, h, h, h)
    fields = abi.Fields3D(ctx, abi.Fields3DParams(g, L, L, L, abi.BC_PERIODIC, 0, 0, 0, 1.0), mask, volt)
    of = O.make_fields3(O.grid3(n, n, n, h, h, h), mask, np.zeros(n ** 3), boundary=1)
    rng = np.random.default_rng(5)
    N = 5000
    pos = rng.uniform(0.01 * L, 0.99 * L, (3, N)).astype(np.float32)
    vel = (rng.normal(size=(3, N)) * 3e4).astype(np.float32)
    P = abi.ParticlesSoA3D.from_host(ctx, *pos, *vel)
    op = _oracle_parts(pos, vel)
    dt = 2e-8
    sp_ = abi.Species2D(QE, M40, dt, float("inf"), 0, 0, 0)
    os_ = O.make_species(QE, M40, dt)
    for k in range(10):
        abi.advance_3d(ctx, fields, sp_, P, flags=abi.STEP_PUSH | abi.STEP_BOUNDARY)
        O.lib().o_advance3d(C.byref(of), C.byref(os_), O.dptr(op), N, None, None)
    d = P.download()
    alive = P.alive()
    np.testing.assert_array_equal(alive, op["empty"] == 0)
    for c in ("x", "y", "z"):
        assert np.abs(d[c][alive] - op[c][alive]).max() <= 10 * 2 * np.spacing(np.float32(L)), c


def test_fused_deposit_3d_matches_oracle(ctx):
    sc = Scene3(ctx, B=(0.0, 0.0, 0.3))
    sc.fields.solve(1e-10)
    u = sc.fields.download(abi.FIELD_U)
    of = sc.oracle_fields(u)
    rng = np.random.default_rng(6)
    N = 20000
    pos, vel = _ions(N, rng, sc.L)
    P = abi.ParticlesSoA3D.from_host(ctx, *pos, *vel)
    op = _oracle_parts(pos, vel)
    rho = abi.DeviceArray(ctx, (sc.n ** 3,), np.float64)
    rho.zero()
    orho = np.zeros(sc.n ** 3)
    sp_ = abi.Species2D(-QE, M40, 5e-9, float("inf"), 0, 0, 0)
    os_ = O.make_species(-QE, M40, 5e-9, rho=orho)
    abi.advance_3d(ctx, sc.fields, sp_, P, rho_ptr=rho.ptr,
                   flags=abi.STEP_PUSH | abi.STEP_BOUNDARY | abi.STEP_DEPOSIT)
    O.lib().o_advance3d(C.byref(of), C.byref(os_), O.dptr(op), N, None, None)
    r = rho.get()
    alive = P.alive()
    np.testing.assert_array_equal(alive, op["empty"] == 0)
    # one deposition on the same (post-step FP32) positions: within 1e-6 of the FP64 CIC
    d = P.download()
    ref = np.zeros(sc.n ** 3)
    for x, y, z in zip(d["x"][alive], d["y"][alive], d["z"][alive]):
        assert O.lib().o_f3_accumulate(C.byref(sc.og), O.dptr(ref), -QE, float(x), float(y), float(z)) == 0
    np.testing.assert_allclose(r, ref, rtol=1e-6, atol=1e-6 * QE)
    # and close to the oracle's own step (positions differ by FP32 rounding only)
    np.testing.assert_allclose(r, orho, rtol=1e-4, atol=1e-5 * QE)
    assert r.sum() == pytest.approx(-QE * alive.sum(), rel=1e-12)


def test_sort_3d_orders_by_cell_stably_and_drops_empty(ctx):
    sc = Scene3(ctx)
    rng = np.random.default_rng(7)
    N = 30000
    pos, vel = _ions(N, rng, sc.L, spread=0.2)
    pos = np.clip(pos, 0.0, sc.L * 0.999).astype(np.float32)
    pos[0, ::9] = np.frombuffer(np.uint32(abi.DEAD_BITS).tobytes() * len(pos[0, ::9]), np.float32)
    P = abi.ParticlesSoA3D.from_host(ctx, *pos, *vel)
    Q = abi.ParticlesSoA3D(ctx, N)
    live = abi.sort_3d(ctx, sc.fields, P, Q)
    alive = pos[0].view(np.uint32) != abi.DEAD_BITS
    assert live == alive.sum()
    i, j, k = (np.minimum((pos[c][alive].astype(np.float64) * (1.0 / sc.h)).astype(np.int64), sc.n - 2)
               for c in range(3))
    nt = (sc.n - 2) // 4 + 1  # 4 x 4 x 4-cell tiles, tile-major
    key = ((((i // 4) * nt + j // 4) * nt + k // 4) << 6) | ((i % 4) * 16 + (j % 4) * 4 + k % 4)
    order = np.argsort(key, kind="stable")  # tile-major cell order
    d = Q.download()
    np.testing.assert_array_equal(d["x"][:live], pos[0][alive][order])
    np.testing.assert_array_equal(d["vy"][:live], vel[1][alive][order])
    assert (d["x"][live:].view(np.uint32) == abi.DEAD_BITS).all()
    assert abi.count_live_3d(ctx, Q) == live


def test_tiled_and_untiled_3d_passes_agree_bit_for_bit(ctx):
    sc = Scene3(ctx, B=(0.0, 0.0, 0.5))
    sc.fields.solve(1e-10)
    rng = np.random.default_rng(8)
    N = 60000
    pos, vel = _ions(N, rng, sc.L, T_eV=2.0, spread=0.1)
    A = abi.ParticlesSoA3D.from_host(ctx, *pos, *vel)
    S = abi.ParticlesSoA3D(ctx, N)
    live = abi.sort_3d(ctx, sc.fields, A, S)  # S: sorted, with a tile schedule
    d = S.download()
    U = abi.ParticlesSoA3D.from_host(ctx, *(d[k] for k in abi.ParticlesSoA3D.NAMES))  # same data, no schedule
    rs, ru = (abi.DeviceArray(ctx, (sc.n ** 3,), np.float64) for _ in range(2))
    sp_ = abi.Species2D(QE, M40, 2e-8, float("inf"), 0, 0, 0)
    fl = abi.STEP_PUSH | abi.STEP_BOUNDARY | abi.STEP_DEPOSIT
    # each slot's sort tile (4 x 4 x 4 cells) and its deposit window: the
    # tile's cells plus 2 on every side
    idx = 1.0 / sc.h
    tile = [np.clip(np.floor(d[c][:live].astype(np.float64) * idx), 0, sc.n - 2).astype(np.int64) // 4
            for c in ("x", "y", "z")]
    spilled = []
    for k in range(10):
        sp_.niter = k
        rs.zero()
        ru.zero()
        st = abi.advance_3d(ctx, sc.fields, sp_, S, rho_ptr=rs.ptr, flags=fl, want_stats=True)
        su = abi.advance_3d(ctx, sc.fields, sp_, U, rho_ptr=ru.ptr, flags=fl, want_stats=True)
        assert np.array_equal(rs.get(), ru.get()), k
        assert su.spilled == 0 and st.removed == su.removed, k  # the untiled pass has no windows
        # the spill count is the number of deposits outside their window
        e = S.download()
        alive = e["x"][:live].view(np.uint32) != abi.DEAD_BITS
        out = np.zeros(live, bool)
        for c, t in zip(("x", "y", "z"), tile):
            cell = (np.where(alive, e[c][:live], 0.0).astype(np.float64) * idx).astype(np.int64)
            out |= (cell < 4 * t - 2) | (cell > 4 * t + 5)
        assert st.spilled == int((out & alive).sum()), k
        spilled.append(st.spilled)
    assert spilled[0] == 0 and spilled[-1] > 0  # ions drift out of their windows
    a, b = S.download(), U.download()
    for c in abi.ParticlesSoA3D.NAMES:
        assert np.array_equal(a[c].view(np.uint32), b[c].view(np.uint32)), c
    assert abi.count_live_3d(ctx, S) < live  # ions reached the rods or walls meanwhile
