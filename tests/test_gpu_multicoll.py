"""The multicoll mover on the device (m2d_advance_multicoll_2d, the time_to_death
draw m2d_time_to_death_init_2d) against the oracle's restatement of
Species<CARTESIAN>::advance_multicoll (particles.cpp:813-861) and
BaseSpecies::lifetime_init (particles.cpp:163-169).

Both sides draw from the same Philox streams keyed by (slot, step), and the
event times are FP64 on both, so the same events fire in the same steps: the
event counts agree exactly, time_to_death to FP64 rounding, velocities and
positions to the FP32 rounding of the device store.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

from m2d_testlib import M_HE, M_ION, QE, to_oracle
from mag2d_amd import abi
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N_GRID, L = 64, 2e-2
DT = 1e-9


def empty_fields(ctx, E, Bz=0.0, selfconsistent=0):
    dx = L / (N_GRID - 1)
    g = abi.grid2d(N_GRID, N_GRID, dx, dx)
    mask, volt = abi.grid_build(g, abi.GEO_EMPTY, L, L)
    params = abi.FieldsParams(g, L, L, 1, selfconsistent, 0, abi.BC_FREE, E, 0, 0.0, 0.0, 0.0, 0.0, Bz, 0.0,
                              1e-12)
    f = abi.Fields2D(ctx, params, mask, volt)
    of = O.make_fields(O.grid2(N_GRID, N_GRID, dx, dx), mask, np.zeros(N_GRID * N_GRID), x_max=L, z_max=L,
                       geometry_empty=1, selfconsistent=selfconsistent, extern_field=E, Bz=Bz)
    return f, of


def elastic_pair(ctx, density, T_gas=300.0, rate=1e-15):
    tgt = abi.Target(M_HE, density, T_gas, 0.0, abi.Particles2D())
    it = abi.Interaction(abi.ELASTIC, 0, 0.0, rate, 0.0, 0, 0)
    E_max = 10 * T_gas * abi.KB / QE
    dev = abi.Mcc(ctx, M_ION, [tgt], [it], E_max)
    orc = O.make_mcc(M_ION, E_max, [dict(mass=M_HE, density=density, temperature=T_gas,
                                          inters=[O.inter(abi.ELASTIC, M_ION, M_HE, rate=rate)])])
    return dev, orc


def ions(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.006, 0.014, n).astype(np.float32)
    z = rng.uniform(0.006, 0.014, n).astype(np.float32)
    v = (rng.normal(size=(3, n)) * 500.0).astype(np.float32)
    return x, z, v


def test_multicoll_same_events_as_oracle(ctx):
    E = 1000.0
    f, of = empty_fields(ctx, E)
    dev, orc = elastic_pair(ctx, density=1e24)  # lifetime 1e-9 s: about one event per step
    assert dev.lifetime == pytest.approx(orc.lifetime, rel=1e-14)
    n = 20000
    x, z, v = ions(n, 21)
    P = abi.ParticlesSoA.from_host(ctx, x, z, v[0], v[1], v[2])
    ttd = ctx.to_device(np.zeros(n))
    sp = abi.Species2D(QE, M_ION, DT, dev.lifetime, 0, 7, 0)
    abi.time_to_death_init(ctx, sp, P, ttd.ptr)
    os_ = O.make_species(QE, M_ION, DT, lifetime=orc.lifetime, mcc=orc, seed=1234, stream=7)
    op = to_oracle(x, z, v[0], v[1], v[2])
    O.lib().o_ttd_init(C.byref(os_), O.dptr(op), n)
    np.testing.assert_allclose(ttd.get(), op["time_to_death"], rtol=1e-13)
    total = 0
    for step in range(6):
        sp.niter = os_.niter = step
        st = abi.advance_multicoll(ctx, f, sp, P, ttd.ptr, mcc=dev, want_stats=True)
        ne = C.c_long(0)
        assert O.lib().o_advance_multicoll(C.byref(of), C.byref(os_), O.dptr(op), n, C.byref(ne)) == 0
        assert st.collisions == ne.value, step
        total += ne.value
    assert total == pytest.approx(6 * n * DT / orc.lifetime, rel=0.05)
    assert ctx.status() == 0
    d = P.download()
    np.testing.assert_allclose(ttd.get(), op["time_to_death"], rtol=1e-12, atol=1e-24)
    for k in ("vx", "vy", "vz"):
        assert np.abs(d[k] - op[k]).max() < 2e-5 * 500.0, k
    for k in ("x", "z"):
        assert np.abs(d[k] - op[k]).max() < 1e-8, k


def test_multicoll_event_statistics(ctx):
    """Exp(lifetime) waiting times: Poisson events per step, and the
    time_to_death left over is again Exp(lifetime)."""
    f, _ = empty_fields(ctx, 0.0)
    n = 200_000
    lam = 2.5
    x, z, v = ions(n, 22)
    P = abi.ParticlesSoA.from_host(ctx, x, z, v[0], v[1], v[2])
    ttd = ctx.to_device(np.zeros(n))
    sp = abi.Species2D(QE, M_ION, DT, DT / lam, 3, 9, 0)
    abi.time_to_death_init(ctx, sp, P, ttd.ptr)
    t0 = ttd.get() / (DT / lam)
    assert stats.kstest(t0, "expon").pvalue > 1e-3
    st = abi.advance_multicoll(ctx, f, sp, P, ttd.ptr, mcc=None, want_stats=True)
    assert abs(st.collisions - n * lam) < 5 * np.sqrt(n * lam)
    rest = ttd.get() / (DT / lam)
    assert rest.min() > 0
    assert stats.kstest(rest, "expon").pvalue > 1e-3
    # without collisions (mcc NULL: null events) and E = 0 nothing but x, z moves
    d = P.download()
    assert np.array_equal(d["vx"], v[0]) and np.array_equal(d["vz"], v[2])
    np.testing.assert_allclose(d["x"], x.astype(np.float64) + v[0].astype(np.float64) * DT, rtol=0, atol=2e-9)


def test_multicoll_free_flight_matches_the_closed_form(ctx):
    """No events: n steps of the reference's update give
    z = z0 + v0 t + a t^2/2 + a t dt; each step rounds the state to FP32."""
    E = -2000.0
    f, _ = empty_fields(ctx, E)
    n = 1024
    x, z, v = ions(n, 23)
    P = abi.ParticlesSoA.from_host(ctx, x, z, v[0], v[1], v[2])
    ttd = ctx.to_device(np.full(n, np.inf))
    sp = abi.Species2D(QE, M_ION, DT, float("inf"), 0, 0, 0)
    steps = 20
    for k in range(steps):
        sp.niter = k
        st = abi.advance_multicoll(ctx, f, sp, P, ttd.ptr, want_stats=True)
        assert st.collisions == 0
    a = E * QE / M_ION
    t = steps * DT
    d = P.download()
    np.testing.assert_allclose(d["vz"], v[2].astype(np.float64) + a * t, rtol=0, atol=5e-3)
    # the a t dt term (1.3e-6 m here) is far above the FP32 rounding of 20 steps
    want = z.astype(np.float64) + v[2].astype(np.float64) * t + 0.5 * a * t * t + a * t * DT
    np.testing.assert_allclose(d["z"], want, rtol=0, atol=5e-8)


def test_multicoll_event_budget_flag(ctx):
    """A time_to_death of -inf never reaches dt: the particle stops at the
    per-step event budget and the step raises M2D_FLAG_EVENTS."""
    f, _ = empty_fields(ctx, 0.0)
    n = 256
    x, z, v = ions(n, 24)
    P = abi.ParticlesSoA.from_host(ctx, x, z, v[0], v[1], v[2])
    t = np.full(n, 1.0)
    t[5] = -np.inf
    ttd = ctx.to_device(t)
    ctx.status(clear=True)
    sp = abi.Species2D(QE, M_ION, DT, DT, 0, 0, 0)
    st = abi.advance_multicoll(ctx, f, sp, P, ttd.ptr, want_stats=True)
    assert st.collisions == 4096
    assert ctx.status() & abi.FLAG_EVENTS


def test_multicoll_rejects_what_the_reference_rejects(ctx):
    n = 64
    x, z, v = ions(n, 25)
    P = abi.ParticlesSoA.from_host(ctx, x, z, v[0], v[1], v[2])
    ttd = ctx.to_device(np.ones(n))
    sp = abi.Species2D(QE, M_ION, DT, DT, 0, 0, 0)
    f, _ = empty_fields(ctx, 0.0, Bz=0.1)
    with pytest.raises(abi.M2DError, match="zero B only"):
        abi.advance_multicoll(ctx, f, sp, P, ttd.ptr)
    f, _ = empty_fields(ctx, 0.0, selfconsistent=1)
    with pytest.raises(abi.M2DError, match="const extern field only"):
        abi.advance_multicoll(ctx, f, sp, P, ttd.ptr)
    f, _ = empty_fields(ctx, 0.0)
    sp.lifetime = DT / 1000.0
    with pytest.raises(abi.M2DError, match="at most 256"):
        abi.advance_multicoll(ctx, f, sp, P, ttd.ptr)
