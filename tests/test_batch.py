"""Batched ensembles: independent RF-trap runs with ragged particle counts,
their own RF amplitudes and buffer-gas pressures, advanced by the persistent
pass as one hipGraph, against each run stepped alone with m2d_advance_2d."""
import math

import numpy as np
import pytest

from mag2d_amd import abi
from m2d_testlib import M_HE, M_ION, QE, Scene, ions_on_disk

pytestmark = pytest.mark.gpu


def _runs(ctx, sc, counts, amps, densities, rng):
    stores, mccs = [], []
    for k, n in enumerate(counts):
        x, z, vx, vy, vz = ions_on_disk(n, rng, r=6e-3, T_eV=0.3) if n else ([], [], [], [], [])
        stores.append((np.asarray(x, np.float32), np.asarray(z, np.float32), np.asarray(vx, np.float32),
                       np.asarray(vy, np.float32), np.asarray(vz, np.float32)))
        if densities[k] is None:
            mccs.append(None)
        else:
            mccs.append(abi.Mcc(ctx, M_ION, [abi.Target(M_HE, densities[k], 300.0, 0.0, abi.Particles2D())],
                                [abi.Interaction(abi.ELASTIC, 0, 0.0, 1e-15, 0.0, 0, 0)], 1.0))
    return stores, mccs


@pytest.mark.parametrize("tuned", [False, True])
def test_batch_matches_runs_stepped_alone_bit_for_bit(ctx, tuned):
    sc = Scene(ctx, n=64)
    rng = np.random.default_rng(11)
    counts = [1000, 37, 0, 2048, 513, 3]
    amps = [10.0, 50.0, 30.0, 22.5, 41.0, 17.0]
    dens = [1e21, 3e21, None, 1e22, None, 5e20]
    stores, mccs = _runs(ctx, sc, counts, amps, dens, rng)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    cat = [np.concatenate([s[c] for s in stores]).astype(np.float32) for c in range(5)]
    B = abi.ParticlesSoA.from_host(ctx, *cat)
    dt, steps = 1e-9, 120
    sp = abi.Species2D(QE, M_ION, dt, float("inf"), 0, 0, 0)
    batch = abi.Batch(ctx, sc.fields, sp, B, offsets, amps, mccs)
    if tuned:
        block, ppt, seg = batch.tune(0, 20, 50)
        assert block in (64, 128, 256) and ppt in (1, 2, 4) and seg == 50
        assert np.array_equal(B.download()["x"].view(np.uint32), cat[0].view(np.uint32))  # tuning left it alone
    else:
        batch.set_shape(128, 2, 50)
    batch.advance(0, steps)
    got = B.download()
    alive_b = B.alive()
    mean_b, live_b = batch.energy()
    omega = sc.params.rf_omega
    for r, n in enumerate(counts):
        lo, hi = offsets[r], offsets[r + 1]
        if n == 0:
            assert live_b[r] == 0
            continue
        P = abi.ParticlesSoA.from_host(ctx, *stores[r])
        # the same run alone: its own RF amplitude through a Fields copy, its own collisions and stream
        params = abi.FieldsParams(sc.g, sc.L, sc.L, 0, 0, 0, abi.BC_FREE, 0.0, 1, amps[r], omega, 0.0, 0.0, 0.0,
                                  0.0, sc.dV)
        fr = abi.Fields2D(ctx, params, sc.mask, sc.volt)
        fr.upload(abi.FIELD_U, sc.u)
        fr.upload(abi.FIELD_URF, sc.urf)
        life = mccs[r].lifetime if mccs[r] is not None else float("inf")
        s1 = abi.Species2D(QE, M_ION, dt, life, 0, r, 0)
        for k in range(steps):
            s1.niter = k
            abi.advance(ctx, fr, s1, P, mcc=mccs[r], flags=abi.STEP_PUSH | abi.STEP_BOUNDARY)
        d = P.download()
        alive = P.alive()
        np.testing.assert_array_equal(alive, alive_b[lo:hi])
        for c in ("x", "z", "vx", "vy", "vz"):
            assert np.array_equal(d[c][alive].view(np.uint32), got[c][lo:hi][alive].view(np.uint32)), (r, c)
        m, nl = abi.mean_energy(ctx, P, M_ION)
        assert nl == live_b[r]
        assert mean_b[r] == pytest.approx(m, rel=1e-12)
    assert ctx.status(clear=True) == 0


def _oracle_run(sc, store, amp, density, stream, steps, dt):
    """One run stepped alone by the oracle (advance_boris + advance_boundary,
    particles.cpp:924-995, particles.hpp:370-411) on the run's Philox
    stream, slots counted from the run's first slot."""
    import ctypes as C
    from oracle import oracle as O
    from m2d_testlib import to_oracle
    of = O.make_fields(sc.og, sc.mask, sc.u, sc.urf, x_max=sc.L, z_max=sc.L, rf=1, rf_amplitude=amp,
                       rf_omega=sc.params.rf_omega)
    omcc = None
    if density is not None:
        omcc = O.make_mcc(M_ION, 1.0, [dict(mass=M_HE, density=density, temperature=300.0,
                                             inters=[O.inter(O.ELASTIC, M_ION, M_HE, rate=1e-15)])])
    op = to_oracle(*store)
    os_ = O.make_species(QE, M_ION, dt, lifetime=omcc.lifetime if omcc else float("inf"), mcc=omcc,
                         seed=1234, stream=stream)
    for k in range(steps):
        os_.niter = k
        O.lib().o_advance_boris(C.byref(of), C.byref(os_), O.dptr(op), len(op), None)
        O.lib().o_advance_boundary(C.byref(of), C.byref(os_), O.dptr(op), len(op), None)
    return op


def test_ragged_batch_matches_oracle_run_by_run(ctx):
    """Each run of a ragged batch (an empty run, a one-ion run, runs with and
    without MCC, different amplitudes and densities) against the oracle
    stepping that run alone on the run's Philox stream and slot numbering.
    Bounds as for single runs: positions within steps * ulp(L), velocities
    within 1e-4 of the run's velocity scale; removal sets may differ only for
    ions within rounding of a wall."""
    sc = Scene(ctx, n=64)
    rng = np.random.default_rng(31)
    counts = [0, 1, 1500, 37, 777, 2049, 256, 5]
    amps = [30.0, 12.0, 50.0, 10.0, 33.3, 21.0, 45.0, 27.0]
    dens = [None, 1e22, None, 5e21, 2e22, 1e21, None, 3e22]
    stores, mccs = _runs(ctx, sc, counts, amps, dens, rng)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    cat = [np.concatenate([s[c] for s in stores]).astype(np.float32) for c in range(5)]
    B = abi.ParticlesSoA.from_host(ctx, *cat)
    dt, steps = 1e-9, 100
    sp = abi.Species2D(QE, M_ION, dt, float("inf"), 0, 0, 0)
    batch = abi.Batch(ctx, sc.fields, sp, B, offsets, amps, mccs)
    batch.set_shape(128, 2, 25)
    batch.advance(0, steps)
    got = B.download()
    alive_b = B.alive()
    mean_b, live_b = batch.energy()
    tol = steps * float(np.spacing(np.float32(sc.L)))
    for r, n in enumerate(counts):
        lo, hi = offsets[r], offsets[r + 1]
        op = _oracle_run(sc, stores[r], amps[r], dens[r], r, steps, dt)
        oalive = op["empty"] == 0
        if n == 0:
            assert live_b[r] == 0
            continue
        alive = alive_b[lo:hi]
        assert np.sum(alive != oalive) <= max(1, n // 500), r
        both = alive & oalive
        if not both.any():
            continue
        for c in ("x", "z"):
            assert np.abs(got[c][lo:hi][both] - op[c][both]).max() < tol, (r, c)
        vs = np.sqrt(op["vx"] ** 2 + op["vy"] ** 2 + op["vz"] ** 2)[both].max()
        for c in ("vx", "vy", "vz"):
            assert np.abs(got[c][lo:hi][both] - op[c][both]).max() < 1e-4 * vs, (r, c)
        if live_b[r] == both.sum() == oalive.sum():
            from oracle import oracle as O
            assert mean_b[r] == pytest.approx(O.lib().o_mean_energy_eV(O.dptr(op), len(op), M_ION), rel=1e-4)
    assert ctx.status(clear=True) == 0


def test_ensemble_at_full_shape_properties(ctx):
    """The ensemble of the batched workload at its full shape: 256 runs on
    128 x 128, counts lognormal in [1e3, 1e5], buffer-gas pressure
    log-uniform in [1e-4, 1e-1] Pa, RF amplitude uniform in [10, 50] V, 5000
    steps.  Properties that hold whatever the trajectories: live counts never
    grow, no device flag (NaN, off-grid gather), every run's mean energy
    finite, positive and below the trap's depth scale; the three smallest
    runs, stepped by the oracle on the same streams, end with the same
    energy distribution (KS, p > 0.01) and live fraction (within 5 %)."""
    from scipy import stats
    sc = Scene(ctx, n=128)
    rng = np.random.default_rng(2024)
    R = 256
    counts = np.clip(np.exp(rng.normal(math.log(1e4), 1.0, R)), 1e3, 1e5).astype(np.int64)
    pressure = 10 ** rng.uniform(-4, -1, R)
    dens = list(pressure / (abi.KB * 300.0))
    amps = list(rng.uniform(10, 50, R))
    stores, mccs = _runs(ctx, sc, counts, amps, dens, rng)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    cat = [np.concatenate([s[c] for s in stores]).astype(np.float32) for c in range(5)]
    B = abi.ParticlesSoA.from_host(ctx, *cat)
    dt, steps = 1e-9, 5000
    sp = abi.Species2D(QE, M_ION, dt, float("inf"), 0, 0, 0)
    batch = abi.Batch(ctx, sc.fields, sp, B, offsets, amps, mccs)
    batch.tune(0, 20, 250)
    batch.advance(0, steps)
    mean_b, live_b = batch.energy()
    assert ctx.status(clear=True) == 0
    assert np.all(live_b <= counts) and np.all(live_b > 0)
    assert np.all(np.isfinite(mean_b)) and np.all(mean_b > 0) and np.all(mean_b < 60.0)
    got = B.download()
    alive_b = B.alive()
    for r in np.argsort(counts)[:3]:
        lo, hi = offsets[r], offsets[r + 1]
        op = _oracle_run(sc, stores[r], amps[r], dens[r], int(r), steps, dt)
        oalive = op["empty"] == 0
        a = alive_b[lo:hi]
        assert abs(a.sum() - oalive.sum()) <= 0.05 * counts[r], (r, a.sum(), oalive.sum())
        e_dev = sum(got[c][lo:hi][a].astype(np.float64) ** 2 for c in ("vx", "vy", "vz"))
        e_or = op["vx"][oalive] ** 2 + op["vy"][oalive] ** 2 + op["vz"][oalive] ** 2
        assert stats.ks_2samp(e_dev, e_or).pvalue > 0.01, r
