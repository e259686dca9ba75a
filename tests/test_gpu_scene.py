"""The RF 22-pole scene with He buffer-gas MCC run the way the headline
benchmark runs it -- cell-sorted store, compile-time specialised tiled pass
with the MCC trial, the fused deposit and the phase-applied gather tables --
checked step by step against the oracle's Philox step (particles.cpp:924-995
advance_boris with the collision trial and scatter; particles.hpp:370-411
advance_boundary; Field2D.hpp:45-62 for the deposit).

What one run covers: the phase-table builds (every step after the first),
particles leaving their tile's deposit window (the global-atomic fallback of
the tiled deposit), and a re-sort in the middle of the run."""
import ctypes as C
import math

import numpy as np
import pytest

from mag2d_amd import abi
from oracle import oracle as O
from m2d_testlib import M_HE, M_ION, QE, Scene, ions_on_disk, species, to_oracle

pytestmark = pytest.mark.gpu

BENCH_PASS = abi.PASS_DEPOSIT | abi.PASS_MCC | abi.PASS_TILED | abi.PASS_FAST | abi.PASS_PHT


def _apply_sort(op, x, z, alive, sc):
    """The device sort's permutation (stable by the tile of the device
    positions, empty slots dropped to the end) applied to the oracle store."""
    idx = np.flatnonzero(alive)
    keys = abi.cell_keys(x[idx], z[idx], sc.n, sc.n, sc.g.dx, sc.g.dz)
    perm = np.argsort(keys >> 6, kind="stable")
    out = O.particles(len(op))
    out[: len(idx)] = op[idx[perm]]
    out["empty"][len(idx):] = 1
    return out, keys[perm]


def _oracle_rho(sc, d, alive, q):
    """Oracle CIC (Field2D::accumulate) of the device's own surviving positions."""
    p = to_oracle(d["x"], d["z"], d["vx"], d["vy"], d["vz"], alive)
    rho = np.zeros(sc.n * sc.n)
    os_ = O.make_species(q, M_ION, 1e-9, rho=rho)
    O.lib().o_accumulate(C.byref(sc.of), C.byref(os_), O.dptr(p), len(p))
    return rho


def test_rf22pt_scene_on_the_bench_pass_matches_oracle_step_by_step(ctx):
    sc = Scene(ctx, n=128)
    density = 3e21  # p_coll = 1 - exp(-dt n k) = 3.0e-3 per step (the headline scene's 1e20 gives 1e-4)
    tgt = abi.Target(M_HE, density, 300.0, 0.0, abi.Particles2D())
    it = abi.Interaction(abi.ELASTIC, 0, 0.0, 1e-15, 0.0, 0, 0)
    E_max = 1.0
    mcc = abi.Mcc(ctx, M_ION, [tgt], [it], E_max)
    omcc = O.make_mcc(M_ION, E_max, [dict(mass=M_HE, density=density, temperature=300.0,
                                          inters=[O.inter(O.ELASTIC, M_ION, M_HE, rate=1e-15)])])
    assert mcc.lifetime == pytest.approx(omcc.lifetime, rel=1e-14)
    rng = np.random.default_rng(21)
    # 90% cold ions of the headline scene, 10% at 20 eV: those cross the
    # 8-cell halo of their deposit window within the 40 steps between sorts
    n_cold, n_hot = 54_000, 6_000
    cold = ions_on_disk(n_cold, rng, T_eV=0.1)
    hot = ions_on_disk(n_hot, rng, r=3e-3, T_eV=20.0)
    x, z, vx, vy, vz = (np.concatenate([a, b]).astype(np.float32) for a, b in zip(cold, hot))
    n = len(x)
    dt, steps, resort_at = 1e-9, 80, 40
    P = abi.ParticlesSoA.from_host(ctx, x, z, vx, vy, vz)
    Q = abi.ParticlesSoA(ctx, n)
    op = to_oracle(x, z, vx, vy, vz)
    # the oracle's time_to_death field is not used by this mover: it carries
    # each particle's index through the sorts
    op["time_to_death"] = np.arange(n)
    os_ = O.make_species(QE, M_ION, dt, lifetime=omcc.lifetime, mcc=omcc, seed=1234, stream=0)
    abi.advance_init(ctx, sc.fields, species(QE, M_ION, dt, mcc.lifetime, 0), P)
    O.lib().o_advance_boris_init(C.byref(sc.of), C.byref(os_), O.dptr(op), n)
    bufs, cur = [P, Q], 0
    rho = abi.DeviceArray(ctx, (sc.n * sc.n,), np.float64)
    flags = abi.STEP_PUSH | abi.STEP_BOUNDARY | abi.STEP_DEPOSIT
    coll_dev = coll_or = 0
    outside_window = 0
    tile_of_slot = None
    worst_rho = 0.0
    passes = []
    for st in range(steps):
        if st in (0, resort_at):
            d = bufs[cur].download()
            alive = bufs[cur].alive()
            live = abi.sort(ctx, sc.fields, bufs[cur], bufs[1 - cur])
            cur = 1 - cur
            op, skeys = _apply_sort(op, d["x"], d["z"], alive, sc)
            assert live == alive.sum()
            tile_of_slot = skeys >> 6  # sort tile of every live slot at this sort
        rho.zero()
        r = abi.advance(ctx, sc.fields, species(QE, M_ION, dt, mcc.lifetime, st), bufs[cur], mcc=mcc,
                        rho_ptr=rho.ptr, flags=flags, want_stats=True)
        passes.append(ctx.last_pass())
        coll_dev += r.collisions
        nc = C.c_long()
        os_.niter = st
        O.lib().o_advance_boris(C.byref(sc.of), C.byref(os_), O.dptr(op), n, C.byref(nc))
        O.lib().o_advance_boundary(C.byref(sc.of), C.byref(os_), O.dptr(op), n, None)
        coll_or += nc.value
        d = bufs[cur].download()
        alive = bufs[cur].alive()
        # each step's grid: the oracle's CIC of the same surviving positions,
        # to 1e-6 relative on every node above 1e-3 of the peak
        ref = _oracle_rho(sc, d, alive, QE)
        got = rho.get()
        big = np.abs(ref) > 1e-3 * np.abs(ref).max()
        worst_rho = max(worst_rho, float(np.max(np.abs(got[big] - ref[big]) / np.abs(ref[big]))))
        assert got.sum() == pytest.approx(alive.sum() * QE, rel=1e-12)
        # particles outside the deposit window of the tile they were sorted into
        k = len(tile_of_slot)
        a = alive[:k]
        ci = np.minimum((d["x"][:k][a].astype(np.float64) / sc.g.dx).astype(np.int64), sc.n - 2)
        cj = np.minimum((d["z"][:k][a].astype(np.float64) / sc.g.dz).astype(np.int64), sc.n - 2)
        ntz = (sc.n - 2) // 8 + 1
        ti, tj = tile_of_slot[a] // ntz, tile_of_slot[a] % ntz
        out = (ci < 8 * ti - 8) | (ci + 1 > 8 * ti + 16) | (cj < 8 * tj - 8) | (cj + 1 > 8 * tj + 16)
        outside_window = max(outside_window, int(out.sum()))
    d = bufs[cur].download()
    alive = bufs[cur].alive()
    oalive = op["empty"] == 0
    mismatch = int(np.sum(alive != oalive))
    print(f"scene: worst rho rel {worst_rho:.3g}, passes {sorted(set(hex(p) for p in passes))}, "
          f"outside window {outside_window}, removal mismatch {mismatch}, collisions {coll_dev} / {coll_or}",
          flush=True)
    assert worst_rho < 1e-6, worst_rho
    # the form of the pass that the benchmark times ran at every step but the
    # first, whose phase-applied tables the flush of a previous step had not built
    assert passes[0] == BENCH_PASS & ~abi.PASS_PHT
    assert all(p == BENCH_PASS for p in passes[1:]), [hex(p) for p in passes]
    assert outside_window > 0  # the window-overflow path of the deposit ran
    assert mismatch <= 3, mismatch
    # the same trials fire on both sides (same Philox words, same slots)
    assert abs(coll_dev - coll_or) <= mismatch, (coll_dev, coll_or)
    prob = 1 - math.exp(-dt / omcc.lifetime)
    assert coll_dev > 0.5 * prob * n * steps
    both = alive & oalive
    # stated bounds: FP32 positions within steps * ulp(L) of the FP64 oracle,
    # velocities within 1e-4 of the velocity scale of each population
    L = sc.L
    for k in ("x", "z"):
        assert np.abs(d[k][both] - op[k][both]).max() < steps * float(np.spacing(np.float32(L))), k
    hot_ion = op["time_to_death"] >= n_cold
    for pop in (both & ~hot_ion, both & hot_ion):
        assert pop.sum() > 0
        vs = np.sqrt(op["vx"] ** 2 + op["vy"] ** 2 + op["vz"] ** 2)[pop].max()
        for k in ("vx", "vy", "vz"):
            assert np.abs(d[k][pop] - op[k][pop]).max() < 1e-4 * vs, k
    assert ctx.status(clear=True) == 0
