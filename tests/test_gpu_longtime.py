"""Long-time positions (m2d_particles2d::xl, zl): positions kept as an FP32
value plus its FP32 remainder, read and written as their FP64 sum, against
the FP64 oracle over the run lengths of the RF-trap workloads (2000 steps
single run, 5000 steps batched), beside the same run with FP32 positions.
The reference drifts in FP64 (particles.cpp:986-987); FP32 positions round
every drift x += v dt to half an ulp of x, and over thousands of steps of an
ion crossing a field-free region those roundings add up; the pair carries the
position to ~2^-48 of its size, so what remains is the FP32 velocity's
rounding.  The errors of both are printed and bounded here."""
import ctypes as C
import math

import numpy as np
import pytest

from mag2d_amd import abi
from oracle import oracle as O
from m2d_testlib import M_HE, M_ION, QE, Scene, ions_on_disk, species, to_oracle

pytestmark = pytest.mark.gpu


def _rf_scene(ctx, n):
    """The RF 22-pole scene on n x n with uRF from the device solve (FP64
    residual 1e-12), shared by the device and the oracle."""
    sc = Scene(ctx, n=n, solve_oracle=False)
    info = sc.fields.solve(abi.SOLVE_RF, 1e-12)
    assert info.rel_residual <= 1e-12
    sc.urf = sc.fields.download(abi.FIELD_URF)
    sc.of = O.make_fields(sc.og, sc.mask, sc.u, sc.urf, x_max=sc.L, z_max=sc.L, rf=1, rf_amplitude=30.0,
                          rf_omega=sc.params.rf_omega)
    return sc


def _mcc(ctx, density):
    dev = abi.Mcc(ctx, M_ION, [abi.Target(M_HE, density, 300.0, 0.0, abi.Particles2D())],
                  [abi.Interaction(abi.ELASTIC, 0, 0.0, 1e-15, 0.0, 0, 0)], 1.0)
    orc = O.make_mcc(M_ION, 1.0, [dict(mass=M_HE, density=density, temperature=300.0,
                                       inters=[O.inter(O.ELASTIC, M_ION, M_HE, rate=1e-15)])])
    return dev, orc


def _errors(d64x, d64z, ox, oz, alive):
    e = np.maximum(np.abs(d64x - ox), np.abs(d64z - oz))[alive]
    return np.percentile(e, 50), np.percentile(e, 99), e.max()


def test_field_free_drift_is_fp64_exact(ctx):
    """With E = 0 and B = 0 (the RF 22-pole scene at zero amplitude: u = 0,
    uRF weighted by 0) velocities stay what they are and only the drift acts:
    over 5000 steps the long-time positions stay within 5e-13 m of the FP64
    drift (the pair rounds each drift to ~2^-48 of the position, 4e-17 m at
    1 cm, so 5000 steps stay below 2e-13 m; measured max 1.2e-13 m), while
    FP32 positions drift away by their per-step rounding (measured 2.3e-6
    m; bound: above 1e-8 m)."""
    sc = Scene(ctx, n=128, rf_amplitude=0.0)
    assert np.abs(sc.u).max() == 0.0
    rng = np.random.default_rng(47)
    n = 4000
    # about 250 m/s: over 5000 steps the ions move ~1 mm and stay off the rods
    x, z, vx, vy, vz = ions_on_disk(n, rng, r=3e-3, T_eV=0.002)
    A = abi.ParticlesSoA.from_host(ctx, x, z, vx, vy, vz, hipos=True)
    B = abi.ParticlesSoA.from_host(ctx, x, z, vx, vy, vz)
    op = to_oracle(x, z, vx, vy, vz)
    dt, steps = 1e-9, 5000
    os_ = O.make_species(QE, M_ION, dt)
    passes = set()
    for st in range(steps):
        s = species(QE, M_ION, dt, niter=st)
        abi.advance(ctx, sc.fields, s, A)
        passes.add(ctx.last_pass())
        abi.advance(ctx, sc.fields, s, B)
        os_.niter = st
        O.lib().o_advance_boris(C.byref(sc.of), C.byref(os_), O.dptr(op), n, None)
        O.lib().o_advance_boundary(C.byref(sc.of), C.byref(os_), O.dptr(op), n, None)
    assert passes == {abi.PASS_FAST | abi.PASS_HIPOS}, passes
    a, b = A.download(), B.download()
    alive = A.alive() & B.alive() & (op["empty"] == 0)
    assert alive.sum() > 0.9 * n
    for k in ("vx", "vy", "vz"):
        assert np.array_equal(a[k][alive].astype(np.float64), op[k][alive])  # no field: velocities untouched
    hi = _errors(a["x64"], a["z64"], op["x"], op["z"], alive)
    lo = _errors(b["x"].astype(np.float64), b["z"].astype(np.float64), op["x"], op["z"], alive)
    print(f"field-free 5000 steps: long-time median/99%/max {hi[0]:.2e} {hi[1]:.2e} {hi[2]:.2e} m; "
          f"FP32 {lo[0]:.2e} {lo[1]:.2e} {lo[2]:.2e} m", flush=True)
    assert hi[2] < 5e-13
    assert lo[2] > 1e-8


def test_rf22pt_2000_steps_long_time_positions_track_fp64(ctx):
    """The headline scene (512 x 512, He at 1e20 m^-3, deposit on, sorted
    store: the tiled specialised pass with phase-applied tables) for 2000
    steps.  What remains of the error of long-time positions is the FP32
    velocities' (the push stays FP32 for long-time positions): measured
    median / 99% / max 2.8e-9 / 1.6e-8 / 3.4e-8 m against 1.7e-8 / 1.3e-7 /
    3.4e-7 m for FP32 positions.  Bounds with margin: median within 1e-8 m
    and at least 3 times below the FP32 positions', maximum within 1e-7 m
    (5e-6 of the box); velocities within 1e-4 of their scale; every 250 steps
    the grid equals the oracle's CIC of the device positions to 1e-6."""
    sc = _rf_scene(ctx, 512)
    dev, orc = _mcc(ctx, 1e20)
    rng = np.random.default_rng(41)
    n = 6000
    x, z, vx, vy, vz = ions_on_disk(n, rng, T_eV=0.1)
    A = abi.ParticlesSoA.from_host(ctx, x, z, vx, vy, vz, hipos=True)
    B = abi.ParticlesSoA.from_host(ctx, x, z, vx, vy, vz)
    As, Bs = abi.ParticlesSoA(ctx, n, hipos=True), abi.ParticlesSoA(ctx, n)
    dt, steps = 1e-9, 2000
    for P in (A, B):
        abi.advance_init(ctx, sc.fields, species(QE, M_ION, dt, dev.lifetime, 0), P)
    op = to_oracle(x, z, vx, vy, vz)
    os_ = O.make_species(QE, M_ION, dt, lifetime=orc.lifetime, mcc=orc, seed=1234, stream=0)
    O.lib().o_advance_boris_init(C.byref(sc.of), C.byref(os_), O.dptr(op), n)
    # one sort of both stores (same cells: the positions are still the FP32 inputs)
    assert abi.sort(ctx, sc.fields, A, As) == n and abi.sort(ctx, sc.fields, B, Bs) == n
    order = np.argsort(abi.cell_keys(x, z, sc.n, sc.n, sc.g.dx, sc.g.dz) >> 6, kind="stable")
    op = op[order].copy()
    rho = abi.DeviceArray(ctx, (sc.n * sc.n,), np.float64)
    flags = abi.STEP_PUSH | abi.STEP_BOUNDARY | abi.STEP_DEPOSIT
    worst_rho = 0.0
    passes = set()
    for st in range(steps):
        s = species(QE, M_ION, dt, dev.lifetime, st)
        rho.zero()
        abi.advance(ctx, sc.fields, s, As, mcc=dev, rho_ptr=rho.ptr, flags=flags)
        passes.add(ctx.last_pass())
        if st % 250 == 249:
            d = As.download()
            alive = As.alive()
            p = to_oracle(d["x64"], d["z64"], d["vx"], d["vy"], d["vz"], alive)
            ref = np.zeros(sc.n * sc.n)
            o2 = O.make_species(QE, M_ION, dt, rho=ref)
            O.lib().o_accumulate(C.byref(sc.of), C.byref(o2), O.dptr(p), n)
            got = rho.get()
            big = np.abs(ref) > 1e-3 * np.abs(ref).max()
            worst_rho = max(worst_rho, float(np.max(np.abs(got[big] - ref[big]) / np.abs(ref[big]))))
        abi.advance(ctx, sc.fields, s, Bs, mcc=dev, flags=abi.STEP_PUSH | abi.STEP_BOUNDARY)
        os_.niter = st
        O.lib().o_advance_boris(C.byref(sc.of), C.byref(os_), O.dptr(op), n, None)
        O.lib().o_advance_boundary(C.byref(sc.of), C.byref(os_), O.dptr(op), n, None)
    want = abi.PASS_DEPOSIT | abi.PASS_MCC | abi.PASS_TILED | abi.PASS_FAST | abi.PASS_PHT | abi.PASS_HIPOS
    assert want in passes, sorted(hex(p) for p in passes)
    a, b = As.download(), Bs.download()
    alive = As.alive() & Bs.alive() & (op["empty"] == 0)
    assert alive.sum() > 0.95 * n
    hi = _errors(a["x64"], a["z64"], op["x"], op["z"], alive)
    lo = _errors(b["x"].astype(np.float64), b["z"].astype(np.float64), op["x"], op["z"], alive)
    vs = np.sqrt(op["vx"] ** 2 + op["vy"] ** 2 + op["vz"] ** 2)[alive].max()
    err_v = max(np.abs(a[k] - op[k])[alive].max() for k in ("vx", "vy", "vz"))
    print(f"2000 steps: |x - x_fp64| long-time median/99%/max {hi[0]:.2e} {hi[1]:.2e} {hi[2]:.2e} m; "
          f"FP32 {lo[0]:.2e} {lo[1]:.2e} {lo[2]:.2e} m; velocity {err_v / vs:.2e} of scale; rho {worst_rho:.2e}",
          flush=True)
    assert hi[0] < 1e-8 and hi[0] * 3 < lo[0], (hi, lo)
    assert hi[2] < 1e-7, hi
    assert err_v < 1e-4 * vs
    assert worst_rho < 1e-6
    assert ctx.status(clear=True) == 0


def test_batched_5000_steps_long_time_positions_track_fp64(ctx):
    """Three runs of the batched ensemble (RF amplitudes 12, 30 and 47 V,
    He from 1e18 to 1e21 m^-3, 128 x 128) for 5000 steps with long-time
    positions, run by run against the oracle on the run's stream.  Measured
    for the 12 V run: median / 99% / max 1.5e-8 / 7.1e-7 / 2.4e-6 m against
    5.2e-8 / 2.0e-6 / 4.8e-6 m for FP32 positions; the largest errors belong
    to ions that meet the rods' RF field, where the FP32 push's velocity
    error grows fastest.  Bounds with margin: median within 5e-8 m and at
    least 2 times below the FP32 positions', maximum within 1e-5 m (5e-4 of
    the box)."""
    sc = _rf_scene(ctx, 128)
    rng = np.random.default_rng(43)
    counts = [700, 1500, 999]
    amps = [12.0, 30.0, 47.0]
    dens = [1e18, 2e20, 1e21]
    mccs = [_mcc(ctx, d) for d in dens]
    stores = [ions_on_disk(c, rng, r=5e-3, T_eV=0.2) for c in counts]
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    cat = [np.concatenate([s[c] for s in stores]).astype(np.float32) for c in range(5)]
    dt, steps = 1e-9, 5000
    sp = abi.Species2D(QE, M_ION, dt, float("inf"), 0, 0, 0)
    res = {}
    for hipos in (True, False):
        P = abi.ParticlesSoA.from_host(ctx, *cat, hipos=hipos)
        batch = abi.Batch(ctx, sc.fields, sp, P, offsets, amps, [m[0] for m in mccs])
        batch.set_shape(128, 2, 500)
        batch.advance(0, steps)
        res[hipos] = (P.download(), P.alive())
    assert ctx.status(clear=True) == 0
    omega = sc.params.rf_omega
    found = []
    for r, n in enumerate(counts):
        lo, hi = offsets[r], offsets[r + 1]
        of = O.make_fields(sc.og, sc.mask, sc.u, sc.urf, x_max=sc.L, z_max=sc.L, rf=1, rf_amplitude=amps[r],
                           rf_omega=omega)
        op = to_oracle(*stores[r])
        orc = mccs[r][1]
        os_ = O.make_species(QE, M_ION, dt, lifetime=orc.lifetime, mcc=orc, seed=1234, stream=r)
        for k in range(steps):
            os_.niter = k
            O.lib().o_advance_boris(C.byref(of), C.byref(os_), O.dptr(op), n, None)
            O.lib().o_advance_boundary(C.byref(of), C.byref(os_), O.dptr(op), n, None)
        a, aa = res[True]
        b, ba = res[False]
        alive = aa[lo:hi] & ba[lo:hi] & (op["empty"] == 0)
        assert alive.sum() > 0.5 * n
        e_hi = _errors(a["x64"][lo:hi], a["z64"][lo:hi], op["x"], op["z"], alive)
        e_32 = _errors(b["x"][lo:hi].astype(np.float64), b["z"][lo:hi].astype(np.float64), op["x"], op["z"], alive)
        print(f"run {r}: 5000 steps, |x - x_fp64| long-time median/99%/max {e_hi[0]:.2e} {e_hi[1]:.2e} "
              f"{e_hi[2]:.2e} m; FP32 {e_32[0]:.2e} {e_32[1]:.2e} {e_32[2]:.2e} m", flush=True)
        found.append((r, e_hi, e_32))
    for r, e_hi, e_32 in found:
        assert e_hi[0] < 5e-8 and e_hi[0] * 2 < e_32[0], (r, e_hi, e_32)
        assert e_hi[2] < 1e-5, (r, e_hi)


def test_long_time_store_through_sort_compact_and_samplers(ctx):
    """The remainders travel with their slots through the cell sort and the
    compaction, and the samplers split their FP64 positions into value and
    remainder."""
    sc = Scene(ctx, n=64, geometry=abi.GEO_EMPTY, rf=False, solve_oracle=False)
    n = 5000
    P = abi.ParticlesSoA(ctx, n, hipos=True)
    vth = math.sqrt(0.1 * QE / M_ION)
    abi.add_particles_on_disk(ctx, sc.fields, P, 0.01, 0.01, 0.005, vth, stream=3)
    d = P.download()
    op = O.particles(n)
    O.lib().o_add_particles_on_disk_philox(O.dptr(op), n, 0.01, 0.01, 0.005, vth, C.byref(sc.of), 1234, 3)
    alive = P.alive()
    assert np.array_equal(alive, op["empty"] == 0)
    # the pair holds the sampled FP64 position to ~2^-48 of it
    np.testing.assert_allclose(d["x64"][alive], op["x"][alive], rtol=4e-15, atol=0)
    np.testing.assert_allclose(d["z64"][alive], op["z"][alive], rtol=4e-15, atol=0)
    Q = abi.ParticlesSoA(ctx, n, hipos=True)
    live = abi.sort(ctx, sc.fields, P, Q)
    keys = abi.cell_keys(d["x64"][alive], d["z64"][alive], sc.n, sc.n, sc.g.dx, sc.g.dz)
    order = np.argsort(keys >> 6, kind="stable")  # tile order
    q = Q.download()
    assert live == alive.sum()
    for k in ("x", "xl", "z", "zl", "vx"):
        assert np.array_equal(q[k][:live], d[k][alive][order]), k
    R = abi.ParticlesSoA(ctx, n, hipos=True)
    assert abi.compact(ctx, P, R) == live
    r = R.download()
    for k in ("x", "xl", "zl", "vz"):
        assert np.array_equal(r[k][:live], d[k][alive]), k
    with pytest.raises(abi.M2DError):
        abi.sort(ctx, sc.fields, P, abi.ParticlesSoA(ctx, n))  # remainders on one side only
