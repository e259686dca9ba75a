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
