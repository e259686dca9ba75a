"""The MCC physics beyond the constant-rate elastic buffer gas, against the
oracle's restatement of BaseSpecies::scatter (particles.cpp:208-365) on the
same Philox draws: SUPERELASTIC (particles.cpp:268-288), tabulated cross
sections (particles.hpp:61-70 on the e-He table of species_conf_MCC.txt),
partners sampled from a live ensemble (particles.cpp:228-238) and COULOMB
with its partner update (particles.cpp:19-26, 309-314)."""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

from mag2d_amd import abi
from oracle import oracle as O
from m2d_testlib import M_E, M_HE, M_ION, QE, Scene, species, to_oracle

pytestmark = pytest.mark.gpu

AMU = 1.66053906660e-27
G = Path(__file__).parent / "golden"


def _no_field_scene(ctx):
    # EMPTY geometry, not self-consistent: Fields::E is the constant extern
    # field (0 here), fields.hpp:126-131
    return Scene(ctx, n=64, geometry=abi.GEO_EMPTY, rf=False, solve_oracle=False)


def _step_both(ctx, sc, P, op, q, m, dt, steps, dev, orc, partner_dev=None, partner_or=None):
    coll = 0
    os_ = O.make_species(q, m, dt, lifetime=orc.lifetime, mcc=orc, seed=1234, stream=0)
    for st in range(steps):
        r = abi.advance(ctx, sc.fields, species(q, m, dt, dev.lifetime, st), P, mcc=dev, want_stats=True)
        coll += r.collisions
        os_.niter = st
        O.lib().o_advance_boris(C.byref(sc.of), C.byref(os_), O.dptr(op), len(op), None)
        O.lib().o_advance_boundary(C.byref(sc.of), C.byref(os_), O.dptr(op), len(op), None)
    return coll


def _store(ctx, rng, n, vscale, x0=0.01):
    x = rng.uniform(x0 - 0.004, x0 + 0.004, n).astype(np.float32)
    z = rng.uniform(0.006, 0.014, n).astype(np.float32)
    v = (rng.normal(size=(3, n)) * vscale).astype(np.float32)
    return x, z, v[0], v[1], v[2]


def _compare(P, op, vscale, what):
    d = P.download()
    alive = P.alive()
    assert np.array_equal(alive, op["empty"] == 0), what
    err = max(np.abs(d[k][alive] - op[k][alive]).max() for k in ("vx", "vy", "vz"))
    # FP32 state against the FP64 oracle on the same draws: 1e-5 of the velocity scale
    assert err < 1e-5 * vscale, (what, err / vscale)
    return err / vscale


def test_superelastic_same_draws_as_oracle(ctx):
    """SUPERELASTIC with DE = 0.3 eV (released into the relative motion)."""
    sc = _no_field_scene(ctx)
    DE = 0.3
    dev = abi.Mcc(ctx, M_ION, [abi.Target(M_HE, 1e23, 300.0, 0.0, abi.Particles2D())],
                  [abi.Interaction(abi.SUPERELASTIC, 0, DE, 1e-15, 0.0, 0, 0)], 1.0)
    orc = O.make_mcc(M_ION, 1.0, [dict(mass=M_HE, density=1e23, temperature=300.0,
                                       inters=[O.inter(O.SUPERELASTIC, M_ION, M_HE, rate=1e-15, DE=DE)])])
    assert dev.lifetime == pytest.approx(orc.lifetime, rel=1e-14)
    rng = np.random.default_rng(51)
    n, vs = 6000, 500.0
    st = _store(ctx, rng, n, vs)
    P = abi.ParticlesSoA.from_host(ctx, *st)
    op = to_oracle(*st)
    coll = _step_both(ctx, sc, P, op, QE, M_ION, 1e-9, 8, dev, orc)
    assert coll > 0.5 * n * 8 * (1 - math.exp(-1e-9 / orc.lifetime))
    # each collision releases DE into the relative motion: the ions end far
    # above their initial energy, on both sides alike
    d = P.download()
    e_dev = 0.5 * M_ION * sum(d[k].astype(np.float64) ** 2 for k in ("vx", "vy", "vz")).mean() / QE
    e0 = 0.5 * M_ION * sum(a.astype(np.float64) ** 2 for a in st[2:]).mean() / QE
    assert e_dev > e0 + 0.05
    _compare(P, op, 4000.0, "superelastic")


def test_tabulated_cross_section_same_draws_as_oracle(ctx):
    """Electrons in He with the tabulated elastic cross section of
    species_conf_MCC.txt (the table of the vec_interpolate fixture):
    svmax over the table sets the lifetime, sigma(E) v the null-collision
    test of every trial."""
    sc = _no_field_scene(ctx)
    d = np.load(G / "vecinterp_ref.npz")
    E, S = np.ascontiguousarray(d["E"]), np.ascontiguousarray(d["S"])
    dens, E_max = 3e22, 10.0
    dev = abi.Mcc(ctx, M_E, [abi.Target(M_HE, dens, 300.0, 0.0, abi.Particles2D())],
                  [abi.Interaction(abi.ELASTIC, 0, 0.0, 0.0, 0.0, 0, len(E))], E_max, tab_E=E, tab_sigma=S)
    orc = O.make_mcc(M_E, E_max, [dict(mass=M_HE, density=dens, temperature=300.0,
                                       inters=[O.inter(O.ELASTIC, M_E, M_HE, tabE=E, tabS=S)])])
    assert dev.lifetime == pytest.approx(orc.lifetime, rel=1e-13)
    rng = np.random.default_rng(52)
    n = 8000
    vs = math.sqrt(2.0 * QE / M_E)  # 1 eV electrons
    st = _store(ctx, rng, n, vs / math.sqrt(3))
    P = abi.ParticlesSoA.from_host(ctx, *st)
    op = to_oracle(*st)
    dt = 2e-12
    coll = _step_both(ctx, sc, P, op, -QE, M_E, dt, 10, dev, orc)
    prob = 1 - math.exp(-dt / orc.lifetime)
    assert coll > 0.5 * n * 10 * prob
    _compare(P, op, vs, "tabulated")


def _ensemble(ctx, rng, n, mass, T):
    vth = math.sqrt(abi.KB * T / mass)
    x, z, vx, vy, vz = _store(ctx, rng, n, vth)
    dead = rng.uniform(size=n) < 0.15  # holes in the partner store
    x[dead] = np.array([abi.DEAD_BITS], np.uint32).view(np.float32)[0]
    return (x, z, vx, vy, vz), dead


@pytest.mark.parametrize("itype", [abi.ELASTIC, abi.COULOMB])
def test_live_partner_ensemble_same_draws_as_oracle(ctx, itype):
    """The target is a species with a live particle store (2000 slots, 15%
    of them empty): every collision draws a random live partner
    (random_particle, particles.hpp:201-211).  With COULOMB the partner turns
    too (particles.cpp:309-314); the device applies the queue in slot order,
    as the reference's loop does, so several hits on one partner within a
    step compose in the same order.  Primary and partner velocities are
    compared with the oracle after each of 6 steps."""
    sc = _no_field_scene(ctx)
    rng = np.random.default_rng(53 + itype)
    M_AR = 40 * AMU
    pst, pdead = _ensemble(ctx, rng, 2000, M_AR, 3000.0)
    PB = abi.ParticlesSoA.from_host(ctx, *pst)
    opb = to_oracle(*pst, alive=~pdead)
    # the primary's Debye length (T1, n1) is 0.2 um, so sigma v peaks near the
    # ions' relative speeds and most trials end in a real collision
    T1, n1, dens = 1000.0, 1e20, 1e22
    if itype == abi.COULOMB:
        it_dev = abi.Interaction(abi.COULOMB, 0, 0.0, 0.0, 0.0, 0, 0)
        it_or = O.inter(O.COULOMB, M_ION, M_AR)
        it_or.t1, it_or.n1, it_or.q1, it_or.q2 = T1, n1, QE, QE
        E_max = 0.1
    else:
        it_dev = abi.Interaction(abi.ELASTIC, 0, 0.0, 2e-15, 0.0, 0, 0)
        it_or = O.inter(O.ELASTIC, M_ION, M_AR, rate=2e-15)
        E_max = 1.0
    dev = abi.Mcc(ctx, M_ION, [abi.Target(M_AR, dens, 3000.0, QE, PB.struct())], [it_dev], E_max,
                  charge=QE, temperature=T1, density=n1)
    orc = O.make_mcc(M_ION, E_max, [dict(mass=M_AR, density=dens, temperature=3000.0, inters=[it_or], parts=opb)])
    assert dev.lifetime == pytest.approx(orc.lifetime, rel=1e-12)
    n, vs = 3000, 600.0
    st = _store(ctx, rng, n, vs)
    P = abi.ParticlesSoA.from_host(ctx, *st)
    op = to_oracle(*st)
    dt = 0.2 * orc.lifetime  # about 18% of the ions collide each step
    coll = 0
    os_ = O.make_species(QE, M_ION, dt, lifetime=orc.lifetime, mcc=orc, seed=1234, stream=0)
    worst = 0.0
    for s in range(6):
        r = abi.advance(ctx, sc.fields, species(QE, M_ION, dt, dev.lifetime, s), P, mcc=dev, want_stats=True)
        coll += r.collisions
        os_.niter = s
        O.lib().o_advance_boris(C.byref(sc.of), C.byref(os_), O.dptr(op), n, None)
        O.lib().o_advance_boundary(C.byref(sc.of), C.byref(os_), O.dptr(op), n, None)
        worst = max(worst, _compare(P, op, 2 * vs, f"primary step {s}"))
        db = PB.download()
        for k in ("vx", "vy", "vz"):
            err = np.abs(db[k][~pdead] - opb[k][~pdead]).max()
            assert err < 1e-5 * 2 * vs, (k, s, err)
    assert coll > 0.1 * n * 6
    moved = np.abs(PB.download()["vx"][~pdead] - pst[2][~pdead]) > 0
    if itype == abi.COULOMB:
        assert moved.sum() > 100  # partners turned
    else:
        assert moved.sum() == 0   # ELASTIC leaves the partner as it is
    assert ctx.status(clear=True) == 0


def test_coulomb_is_rejected_where_hits_run_in_parallel(ctx):
    """The batched, 3D and multicoll passes scatter their hits in parallel,
    so an interaction that turns its partner is refused there."""
    sc = _no_field_scene(ctx)
    PB = abi.ParticlesSoA.from_host(ctx, *(np.full(4, 0.01, np.float32),) * 2, *(np.zeros(4, np.float32),) * 3)
    dev = abi.Mcc(ctx, M_ION, [abi.Target(M_ION, 1e20, 300.0, QE, PB.struct())],
                  [abi.Interaction(abi.COULOMB, 0, 0.0, 0.0, 0.0, 0, 0)], 0.1, charge=QE, temperature=300.0,
                  density=1e16)
    P = abi.ParticlesSoA.from_host(ctx, *(np.full(8, 0.01, np.float32),) * 2, *(np.zeros(8, np.float32),) * 3)
    with pytest.raises(abi.M2DError) as e:
        abi.Batch(ctx, sc.fields, species(QE, M_ION, 1e-9), P, np.array([0, 8], np.int64), [0.0], [dev])
    assert e.value.code == -6
    ttd = ctx.to_device(np.ones(8))
    with pytest.raises(abi.M2DError) as e:
        abi.advance_multicoll(ctx, sc.fields, species(QE, M_ION, 1e-9, dev.lifetime), P, ttd.ptr, mcc=dev)
    assert e.value.code == -6
    with pytest.raises(abi.M2DError):
        abi.Mcc(ctx, M_ION, [abi.Target(M_ION, 1e20, 300.0, QE, PB.struct())],
                [abi.Interaction(abi.COULOMB, 0, 0.0, 0.0, 0.0, 0, 0)], 0.1)  # no primary temperature, density, charge
    # partners drawn from the colliding store itself: those passes refuse it too
    own = abi.Mcc(ctx, M_ION, [abi.Target(M_ION, 1e20, 300.0, QE, P.struct())],
                  [abi.Interaction(abi.ELASTIC, 0, 0.0, 1e-15, 0.0, 0, 0)], 1.0)
    with pytest.raises(abi.M2DError) as e:
        abi.Batch(ctx, sc.fields, species(QE, M_ION, 1e-9), P, np.array([0, 8], np.int64), [0.0], [own])
    assert e.value.code == -6
    with pytest.raises(abi.M2DError) as e:
        abi.advance_multicoll(ctx, sc.fields, species(QE, M_ION, 1e-9, own.lifetime), P, ttd.ptr, mcc=own)
    assert e.value.code == -6


def test_self_target_partners_are_read_as_the_push_left_them(ctx):
    """A species whose collision partners come from its own store (ELASTIC,
    random_particle over the same ensemble).  The reference scatters inside
    its particle loop, so a partner later in slot order still has the
    previous step's velocity there (particles.cpp:924-995); the device pushes
    every particle first, and each scatter of the step reads its partner's
    velocity as the push left it (the new velocities land after all of them,
    so the result does not depend on thread timing).  With no field the push
    leaves velocities as they are, so the expected velocity of every slot
    whose trial fired is the oracle's scatter on the same Philox draws against
    the step's starting velocities; two runs give the same bits."""
    sc = _no_field_scene(ctx)
    rng = np.random.default_rng(61)
    n, vs = 3000, 800.0
    st = _store(ctx, rng, n, vs)
    dens, k_el, dt = 1e22, 2e-15, 1e-8
    runs = []
    for rep in range(2):
        P = abi.ParticlesSoA.from_host(ctx, *st)
        dev = abi.Mcc(ctx, M_ION, [abi.Target(M_ION, dens, 300.0, QE, P.struct())],
                      [abi.Interaction(abi.ELASTIC, 0, 0.0, k_el, 0.0, 0, 0)], 1.0)
        prob = 1 - math.exp(-dt / dev.lifetime)
        assert 0.1 < prob < 0.3
        for s in range(3):
            before = P.download()
            snap = to_oracle(before["x"], before["z"], before["vx"], before["vy"], before["vz"])
            orc = O.make_mcc(M_ION, 1.0, [dict(mass=M_ION, density=dens, temperature=300.0,
                                               inters=[O.inter(O.ELASTIC, M_ION, M_ION, rate=k_el)], parts=snap)])
            abi.advance(ctx, sc.fields, species(QE, M_ION, dt, dev.lifetime, s), P, mcc=dev)
            after = P.download()
            assert P.alive().all()
            hits = 0
            for p in range(n):
                u = O.lib().o_trial_uniform(1234, 0, p, s)
                if u >= prob:
                    for k in ("vx", "vy", "vz"):
                        assert after[k][p] == before[k][p], (s, p, k)
                    continue
                hits += 1
                one = snap[p:p + 1].copy()
                r = O.Rng()
                O.lib().o_rng_philox_begin(C.byref(r), 1234, 0, p, s, 0)
                O.lib().o_scatter(C.byref(orc), O.dptr(one), C.byref(r))
                for k in ("vx", "vy", "vz"):
                    assert abs(after[k][p] - one[k][0]) <= 2e-6 * vs, (s, p, k, after[k][p], one[k][0])
            assert hits > 0.5 * prob * n
        runs.append(P.download())
    for k in ("x", "z", "vx", "vy", "vz"):
        assert np.array_equal(runs[0][k].view(np.uint32), runs[1][k].view(np.uint32)), k
    assert ctx.status(clear=True) == 0
