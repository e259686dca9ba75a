"""Pin the oracle's multicoll mover, Species<CARTESIAN>::advance_multicoll
(particles.cpp:813-861), and the time_to_death draw of lifetime_init
(particles.cpp:163-169).  The reference's tests hold no vectors for this
mover, so each check comes from what the mathematics fixes:

- free flights: the reference's update x += (v + a t/2) t after the kick sums
  over n steps to x0 + v0 t + a t^2/2 + a t dt (not the exact a t^2/2);
- one event at a known time splits the step into two such flights;
- with Exp(lifetime) waiting times the events of one step are Poisson with
  mean dt/lifetime, and the time_to_death left afterwards is again
  Exp(lifetime) (memorylessness);
- elastic events against a heavy target at rest keep the speed and
  randomise the direction.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O

QE = O.QE
M_ION = 3 * 1.66053906660e-27


def fields(extern_field, **kw):
    n = 16
    g = O.grid2(n, n, 1e-3, 1e-3)
    mask = np.full(n * n, O.FREE, np.uint8)
    return O.make_fields(g, mask, np.zeros(n * n), x_max=15e-3, z_max=15e-3, geometry_empty=1,
                         extern_field=extern_field, **kw)


def store(n, rng, ttd):
    p = O.particles(n)
    p["x"] = rng.uniform(4e-3, 11e-3, n)
    p["z"] = rng.uniform(4e-3, 11e-3, n)
    for k in ("vx", "vy", "vz"):
        p[k] = rng.normal(size=n) * 400.0
    p["time_to_death"] = ttd
    return p


def step(f, s, p):
    ne = C.c_long(0)
    rc = O.lib().o_advance_multicoll(C.byref(f), C.byref(s), O.dptr(p), len(p), C.byref(ne))
    assert rc == 0
    s.niter += 1
    return ne.value


def test_free_flights_sum_to_the_closed_form():
    E, q, m, dt, n_steps = 1000.0, QE, M_ION, 1e-9, 37
    f = fields(E)
    s = O.make_species(q, m, dt, lifetime=float("inf"))
    p = store(5, np.random.default_rng(1), np.inf)
    p0 = p.copy()
    for _ in range(n_steps):
        assert step(f, s, p) == 0
    a = E * q / m  # Fields::E(0, 0) is (0, +extern_field), un-negated (fields.hpp:126-131)
    t = n_steps * dt
    np.testing.assert_allclose(p["vz"], p0["vz"] + a * t, rtol=1e-12)
    np.testing.assert_allclose(p["z"], p0["z"] + p0["vz"] * t + 0.5 * a * t * t + a * t * dt, rtol=1e-12)
    np.testing.assert_allclose(p["x"], p0["x"] + p0["vx"] * t, rtol=1e-12)
    assert np.array_equal(p["vx"], p0["vx"]) and np.array_equal(p["vy"], p0["vy"])
    assert np.all(np.isinf(p["time_to_death"]))


@pytest.mark.parametrize("frac", [0.3, 0.0])
def test_one_null_event_splits_the_flight(frac):
    """A null collision at tau = frac dt (mcc NULL) leaves v alone but splits
    the step: z = z0 + v0 dt + a (1.5 tau^2 + tau r + 1.5 r^2), r = dt - tau."""
    E, q, m, dt = -750.0, QE, M_ION, 1e-9
    f = fields(E)
    s = O.make_species(q, m, dt, lifetime=1e300)  # the redrawn time_to_death never comes
    tau = frac * dt
    p = store(4, np.random.default_rng(2), tau + dt)  # no event in the first step
    assert step(f, s, p) == 0
    np.testing.assert_allclose(p["time_to_death"], tau, rtol=1e-9, atol=1e-24)
    p["time_to_death"] = tau  # exactly, so the closed form below uses the same tau
    p0 = p.copy()
    assert step(f, s, p) == len(p)
    a = E * q / m
    r = dt - tau
    np.testing.assert_allclose(p["vz"], p0["vz"] + a * dt, rtol=1e-12)
    want = p0["z"] + p0["vz"] * dt + a * (1.5 * tau * tau + tau * r + 1.5 * r * r)
    np.testing.assert_allclose(p["z"] - p0["z"], want - p0["z"], rtol=1e-9)
    assert np.all(p["time_to_death"] > 1e290)


def test_event_counts_are_poisson_and_the_rest_is_exponential():
    dt, lam = 1e-9, 1.7
    L = dt / lam
    f = fields(0.0)
    n = 20000
    s = O.make_species(QE, M_ION, dt, lifetime=L)
    p = store(n, np.random.default_rng(3), 0.0)
    O.lib().o_ttd_init(C.byref(s), O.dptr(p), n)
    assert stats.kstest(p["time_to_death"] / L, "expon").pvalue > 1e-3
    counts = np.zeros(n, np.int64)
    for k in range(n):  # one slot per call: the per-particle event counts
        s.slot0 = k
        counts[k] = step(f, s, p[k:k + 1])
        s.niter -= 1
    kmax = 7
    obs = np.bincount(np.minimum(counts, kmax), minlength=kmax + 1)
    pmf = stats.poisson.pmf(np.arange(kmax), lam)
    exp_ = n * np.append(pmf, 1 - pmf.sum())
    assert stats.chisquare(obs, exp_).pvalue > 1e-3
    rest = p["time_to_death"]
    assert np.all(rest > 0)
    assert stats.kstest(rest / L, "expon").pvalue > 1e-3


def test_elastic_events_against_a_heavy_cold_target_keep_the_speed():
    dt = 1e-9
    m2 = 1e9 * M_ION
    mcc = O.make_mcc(M_ION, 1.0, [dict(mass=m2, density=3e24, temperature=0.0,
                                       inters=[O.inter(O.ELASTIC, M_ION, m2, rate=1e-15)])])
    L = mcc.lifetime
    assert L == pytest.approx(1.0 / (1e-15 * 3e24), rel=1e-12)
    f = fields(0.0)
    n = 4000
    s = O.make_species(QE, M_ION, dt, lifetime=L, mcc=mcc)
    p = store(n, np.random.default_rng(4), 0.0)
    O.lib().o_ttd_init(C.byref(s), O.dptr(p), n)
    v0 = np.stack([p["vx"], p["vy"], p["vz"]])
    ev = sum(step(f, s, p) for _ in range(4))
    assert ev == pytest.approx(4 * n * dt / L, rel=0.05)
    v1 = np.stack([p["vx"], p["vy"], p["vz"]])
    sp0, sp1 = np.linalg.norm(v0, axis=0), np.linalg.norm(v1, axis=0)
    # a target at rest only takes energy: at most 2 m/m2 of the speed per event
    loss = (sp0 - sp1) / sp0
    assert loss.min() > -1e-14 and loss.max() < 1e-7
    cosang = (v0 * v1).sum(0) / (sp0 * sp1)
    assert abs(cosang.mean()) < 0.05  # about 12 events: the direction is forgotten


def test_rejects_what_the_reference_rejects():
    p = store(2, np.random.default_rng(5), 0.0)
    s = O.make_species(QE, M_ION, 1e-9, lifetime=1e-9)
    f = fields(0.0, Bz=0.1)
    assert O.lib().o_advance_multicoll(C.byref(f), C.byref(s), O.dptr(p), 2, None) == -1
    f = fields(0.0, selfconsistent=1)
    assert O.lib().o_advance_multicoll(C.byref(f), C.byref(s), O.dptr(p), 2, None) == -2
