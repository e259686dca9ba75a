#!/usr/bin/env python3
"""Headline benchmark: particle-steps/s of the fused PIC-MCC particle pass on
one MI355X, for the RF 22-pole ion trap with buffer-gas MCC:

  4 194 304 singly charged ions of 3 u on a 512 x 512 grid (2 cm x 2 cm),
  uRF = Laplace basis of the 22 alternating +-1 rods (solved on the GPU),
  E = -grad(u + V_rf cos(2 pi 17 MHz t) uRF), V_rf = 30 V, B = 0,
  He buffer at 300 K, 1e20 m^-3, constant k = 1e-15 m^3/s elastic,
  Maxwellian 0.1 eV on a 5 mm disk, dt = 1e-9 s.

One step = push + E gather + MCC trial/scatter + boundary + CIC deposit
(64-bit fixed point) of every live ion, inputs resident in HBM.  Prints one
JSON line.  The CPU baseline is the oracle's serial FP64 loop (AoS, the
reference's own loop order and SHR3 stream) on a bounded sample of the same
workload, timed on this host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from mag2d_amd import abi  # noqa: E402

METRIC = "particle-steps/s on 1 GPU (4M ions, 512² grid) vs CPU; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
ADJ_TOL = 1e-8  # FP64 relative residual of the adjoint solves in the adjoint2d workload
M_ION = 3 * 1.66053906660e-27
M_HE = 6.69e-27


def emit(out):
    """Print the result line; a run whose kernels raised a device flag (NaN
    positions, gathers or deposits off the grid) is marked invalid and exits
    non-zero."""
    flags = out.get("device_flags", out.get("config", {}).get("device_flags", 0))
    if flags:
        out["invalid"] = f"device flags 0x{flags:x} raised during the run"
    print(json.dumps(out), flush=True)
    if flags:
        sys.exit(4)


def build_scene(ctx, n_grid, n_particles):
    L = 2e-2
    dx = L / (n_grid - 1)
    g = abi.grid2d(n_grid, n_grid, dx, dx)
    mask, volt = abi.grid_build(g, abi.GEO_RF_22PT, L, L)
    dV = (n_particles / 1e11) / ((n_grid - 1) ** 2)  # Param::dV with density_total 1e11
    params = abi.FieldsParams(g, L, L, 0, 0, 0, abi.BC_FREE, 0.0, 1, 30.0, 2 * math.pi * 17e6, 0.0,
                              0.0, 0.0, 0.0, dV)
    fields = abi.Fields2D(ctx, params, mask, volt)
    # the first solve of a process loads the solver's kernels; the timed one
    # starts again from uRF = 0
    fields.solve(abi.SOLVE_RF, 1e-10)
    fields.upload(abi.FIELD_URF, np.zeros(n_grid * n_grid))
    t0 = time.perf_counter()
    info_rf = fields.solve(abi.SOLVE_RF, 1e-10)
    t1 = time.perf_counter()
    fields.reset()
    info_st = fields.solve(abi.SOLVE_STATIC, 1e-10)
    fields.reset()
    poisson = {"grid": [n_grid, n_grid], "rf_basis_solve_ms": round((t1 - t0) * 1e3, 3),
               "outer_iters": info_rf.outer_iters, "inner_pcg_iters": info_rf.inner_iters,
               "rel_residual_fp64": info_rf.rel_residual, "static_rel_residual": info_st.rel_residual}
    tgt = abi.Target(M_HE, 1e20, 300.0, 0.0, abi.Particles2D())
    it = abi.Interaction(abi.ELASTIC, 0, 0.0, 1e-15, 0.0, 0, 0)
    T_ion_eV = 0.1
    E_max = 10 * T_ion_eV  # BaseSpecies default: 10 k T of the species (particles.hpp:160)
    mcc = abi.Mcc(ctx, M_ION, [tgt], [it], E_max)
    return g, mask, volt, params, fields, mcc, poisson


def cpu_baseline(fields, params, mask, n_grid, budget_s, n):
    """Oracle serial FP64 step (boris + MCC with the SHR3 stream, boundary +
    deposit) on the same scene's full store of n ions for as many steps as
    fit the budget; returns particle-steps/s."""
    from oracle import oracle as O
    import ctypes as C
    L = 2e-2
    dx = L / (n_grid - 1)
    og = O.grid2(n_grid, n_grid, dx, dx)
    u = fields.download(abi.FIELD_U)
    urf = fields.download(abi.FIELD_URF)
    of = O.make_fields(og, mask, u, urf, x_max=L, z_max=L, selfconsistent=1, rf=1, rf_amplitude=30.0,
                       rf_omega=2 * math.pi * 17e6)
    mcc = O.make_mcc(M_ION, 1.0, [dict(mass=M_HE, density=1e20, temperature=300.0,
                                        inters=[O.inter(O.ELASTIC, M_ION, M_HE, rate=1e-15)])])
    r = O.Rng()
    O.lib().o_rng_shr3_seed(C.byref(r), 1234)

    def sample(n):
        rng = np.random.default_rng(1)
        p = O.particles(n)
        ang = rng.uniform(0, 2 * np.pi, n)
        rad = 5e-3 * np.sqrt(rng.uniform(0, 1, n))
        p["x"] = 1e-2 + rad * np.cos(ang)
        p["z"] = 1e-2 + rad * np.sin(ang)
        vth = math.sqrt(0.1 * abi.QE / M_ION)
        for k in ("vx", "vy", "vz"):
            p[k] = rng.normal(size=n) * vth
        return p

    rho = np.zeros(n_grid * n_grid)
    s = O.make_species(abi.QE, M_ION, 1e-9, lifetime=mcc.lifetime, rho=rho, mcc=mcc,
                       rng_kind=O.RNG_SHR3, shr3=r)
    n_cal = 200_000
    p = sample(n_cal)
    t_cal = O.lib().o_bench_steps(C.byref(of), C.byref(s), O.dptr(p), n_cal, 2)
    per_ps = t_cal / (2 * n_cal)
    steps = max(2, int(budget_s / (per_ps * n)))
    p = sample(n)
    t = O.lib().o_bench_steps(C.byref(of), C.byref(s), O.dptr(p), n, steps)
    out = {"value": n * steps / t, "unit": "particle-steps/s", "cores": 1, "kind": "port",
           "sample": f"{n} ions x {steps} steps of the same RF 22-pole scene (uRF from the GPU solve), "
                     f"FP64 AoS, serial reference loop order, SHR3 stream, deposit on; {t:.1f} s"}
    # the OpenMP form of the same loop (tests/test_particles_omp.cpp's idiom):
    # particles split over the host's threads, a Philox stream per slot and a
    # private rho per thread summed after each step
    threads = min(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1), 64)
    steps_omp = max(2, int(0.3 * budget_s * threads / (per_ps * n)))
    p = sample(n)
    t = O.lib().o_bench_steps_omp(C.byref(of), C.byref(s), O.dptr(p), n, steps_omp, threads)
    out["omp"] = {"value": n * steps_omp / t, "unit": "particle-steps/s", "threads": threads,
                  "sample": f"{n} ions x {steps_omp} steps, same scene and loop, {threads} OpenMP threads; "
                            f"{t:.1f} s"}
    return out


TRAFFIC_FILE = ROOT / "profiles" / "k_advance2d_rf22pt_512x512_4M_traffic.md"
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def copy_calibration(ctx, mib, reps=5):
    """Plain device-to-device copies of a known size (the runtime's blit
    kernel): each moves `mib` MiB in and `mib` MiB out, which calibrates
    the FETCH_SIZE / WRITE_SIZE counters of a --pmc run of this program."""
    nbytes = mib << 20
    a = abi.DeviceArray(ctx, (nbytes,), np.uint8)
    b = abi.DeviceArray(ctx, (nbytes,), np.uint8)
    a.zero()
    for _ in range(reps):
        abi.check(abi.lib().m2d_memcpy_d2d(ctx.h, abi.C.c_void_p(b.ptr), abi.C.c_void_p(a.ptr), nbytes))
    ctx.sync()


def _counter_means(csv_path):
    """Mean counter value per (kernel, counter) of a rocprofv3 csv counter file."""
    import csv
    acc = {}
    with open(csv_path) as fh:
        for row in csv.DictReader(fh):
            key = (row["Kernel_Name"], row["Counter_Name"])
            acc.setdefault(key, []).append(float(row["Counter_Value"]))
    return acc


def summarize_pmc(fetch_dir, write_dir, copy_mib, kernel="k_advance2d"):
    """Markdown summary of a FETCH_SIZE run and a WRITE_SIZE run: every
    counter is in KiB per dispatch; each is scaled by the factor that maps
    the copy kernel's reading onto the bytes it is known to move."""
    import glob
    fetch = _counter_means(glob.glob(os.path.join(fetch_dir, "*counter_collection.csv"))[0])
    write = _counter_means(glob.glob(os.path.join(write_dir, "*counter_collection.csv"))[0])
    copy_kib = copy_mib * 1024.0

    def copy_reading(table, counter):
        vals = [v for (k, c), vs in table.items() if c == counter and COPY_KERNEL in k for v in vs]
        big = [v for v in vals if v > 0.5 * copy_kib]
        return sum(big) / len(big) if big else None

    f_copy, w_copy = copy_reading(fetch, "FETCH_SIZE"), copy_reading(write, "WRITE_SIZE")
    f_scale = copy_kib / f_copy if f_copy else None
    w_scale = copy_kib / w_copy if w_copy else None
    lines = ["| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | calibrated MB read | calibrated MB written |",
             "|---|---|---|---|---|---|"]
    names = sorted({k for (k, c) in fetch} | {k for (k, c) in write})
    target = None
    for k in names:
        fv = fetch.get((k, "FETCH_SIZE"), [])
        wv = write.get((k, "WRITE_SIZE"), [])
        fm = sum(fv) / len(fv) if fv else float("nan")
        wm = sum(wv) / len(wv) if wv else float("nan")
        rd = fm * 1024 * f_scale / 1e6 if f_scale else float("nan")
        wr = wm * 1024 * w_scale / 1e6 if w_scale else float("nan")
        short = k.split("(")[0][:60]
        lines.append(f"| `{short}` | {max(len(fv), len(wv))} | {fm:.1f} | {wm:.1f} | {rd:.2f} | {wr:.2f} |")
        if kernel in k and "<" in k and fv and wv:
            target = (rd + wr) * 1e6
    out = ["copy calibration: %d MiB copies read %s KiB (FETCH_SIZE) and wrote %s KiB (WRITE_SIZE)"
           % (copy_mib, f"{f_copy:.1f}" if f_copy else "n/a", f"{w_copy:.1f}" if w_copy else "n/a"),
           f"fetch_scale: {f_scale}", f"write_scale: {w_scale}", ""] + lines + [""]
    if target is not None:
        out.append(f"traffic_bytes_per_launch: {target:.0f}")
    return "\n".join(out)


def summarize_counters(dirs, particles, kernel="k_advance2d"):
    """Markdown table of the SQ counters of `kernel` (mean per dispatch) from
    rocprofv3 --pmc runs, with the instruction counts per particle: a
    wavefront instruction runs once for 64 lanes, and a lane carries one
    particle per pass slot, so per particle = count * 64 / particles."""
    import glob
    vals = {}
    for d in dirs:
        for (k, c), vs in _counter_means(glob.glob(os.path.join(d, "*counter_collection.csv"))[0]).items():
            if kernel in k and "<" in k:
                vals[c] = sum(vs) / len(vs)
    lines = ["| counter | mean per dispatch | per particle |", "|---|---|---|"]
    for c in sorted(vals):
        per = vals[c] * 64 / particles if c.startswith("SQ_INSTS") else None
        lines.append(f"| {c} | {vals[c]:.4g} | {'' if per is None else f'{per:.1f}'} |")
    if "SQ_ACTIVE_INST_VALU" in vals and "SQ_WAVE_CYCLES" in vals:
        lines += ["", f"SQ_ACTIVE_INST_VALU / SQ_WAVE_CYCLES: {vals['SQ_ACTIVE_INST_VALU'] / vals['SQ_WAVE_CYCLES']:.3f}"]
    if "SQ_ACTIVE_INST_ANY" in vals and "SQ_WAVE_CYCLES" in vals:
        lines += [f"SQ_ACTIVE_INST_ANY / SQ_WAVE_CYCLES: {vals['SQ_ACTIVE_INST_ANY'] / vals['SQ_WAVE_CYCLES']:.3f}"]
    return "\n".join(lines)


def summarize_trace(trace_dir):
    """Markdown table of a rocprofv3 --kernel-trace --stats csv run."""
    import csv
    import glob
    path = glob.glob(os.path.join(trace_dir, "*kernel_stats.csv"))[0]
    lines = ["| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    with open(path) as fh:
        for row in csv.DictReader(fh):
            lines.append("| `%s` | %s | %.3f | %.2f | %.2f | %.2f | %.1f |" % (
                row["Name"].split("(")[0][:70], row["Calls"], float(row["TotalDurationNs"]) / 1e6,
                float(row["AverageNs"]) / 1e3, float(row["MinNs"]) / 1e3, float(row["MaxNs"]) / 1e3,
                float(row["Percentage"])))
    return "\n".join(lines)


def committed_traffic(particles, grid):
    """HBM bytes per launch of the particle pass from the committed counter
    summary, when it was measured on this same workload."""
    try:
        text = TRAFFIC_FILE.read_text()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch: (\d+)", text, re.M)
    w = re.search(r"^workload: particles=(\d+) grid=(\d+)", text, re.M)
    if not m or not w or int(w.group(1)) != particles or int(w.group(2)) != grid:
        return None
    return int(m.group(1))


M_E = 9.11e-31
M_40 = 40 * 1.66053906660e-27


def _timed(ctx, one_step, steps, warmup, on_timed=None):
    """warmup untimed steps, then `steps` timed ones bracketed by device
    synchronisation; returns (elapsed s, particle-pass ms, launches)."""
    for k in range(warmup):
        one_step(k)
    ctx.sync()
    if on_timed:
        on_timed()
    ctx.profile(True)
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        one_step(k)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    kern_ms, launches = ctx.profile_read()
    ctx.profile(False)
    return elapsed, kern_ms, launches


def run_penning2d(args):
    """Self-consistent 2D Penning trap: 100 000 electrons, 128 x 128 grid over a
    3 cm box, grounded frame with end caps at -10 V, Bz = 0.1 T, Maxwellian
    1 eV in a centred 2 cm disk, dt = 1e-11 s; every step solves Poisson on
    the deposited charge (Pic::advance, pic.cpp:330-358)."""
    n, N, L = args.grid or 128, args.particles or 100_000, 3e-2
    steps, warmup = args.steps or 1000, args.warmup if args.warmup is not None else 20
    ctx = abi.Context(0, seed=1234)
    dx = L / (n - 1)
    g = abi.grid2d(n, n, dx, dx)
    mask, volt = abi.grid_build(g, abi.GEO_EMPTY, L, L, extern_field=0.0)
    h = 1.5 * dx
    abi.square_electrode(g, mask, volt, 0.0, L, 0.0, h, -10.0)
    abi.square_electrode(g, mask, volt, 0.0, L, L - h, L, -10.0)
    density_total = 1e12
    dV = (N / density_total) / ((n - 1) ** 2)  # Param::dV (param.cpp:132-134)
    params = abi.FieldsParams(g, L, L, 1, 1, 0, abi.BC_FREE, 0.0, 0, 0.0, 0.0, 0.0, 0.0, 0.1, 0.0, dV)
    fields = abi.Fields2D(ctx, params, mask, volt)
    P = abi.ParticlesSoA(ctx, N)
    T = 11604.5
    vth = math.sqrt(abi.KB * T / M_E)
    abi.add_particles_on_disk(ctx, fields, P, L / 2, L / 2, 1e-2, vth, stream=0)
    sp = abi.Species2D(-abi.QE, M_E, 1e-11, float("inf"), 0, 0, 0)
    rho = fields.ptr(abi.FIELD_RHO)
    # Pic::advance_init (pic.cpp:359-384): deposit, solve, rebuild rho, leapfrog start
    fields.reset()
    abi.accumulate(ctx, g, P, -abi.QE, rho)
    fields.solve(abi.SOLVE_STATIC, 1e-10)
    fields.reset()
    abi.accumulate(ctx, g, P, -abi.QE, rho)
    abi.advance_init(ctx, fields, sp, P)
    bufs = [P, abi.ParticlesSoA(ctx, N)]
    cur = [0]
    solve_info = []

    def one_step(k):
        if args.sort_every and k % args.sort_every == 0:
            abi.sort(ctx, fields, bufs[cur[0]], bufs[1 - cur[0]], want_count=False)
            cur[0] = 1 - cur[0]
        solve_info.append(fields.solve(abi.SOLVE_STATIC, 1e-10))
        fields.reset()
        sp.niter = k
        abi.advance(ctx, fields, sp, bufs[cur[0]], rho_ptr=rho,
                    flags=abi.STEP_PUSH | abi.STEP_BOUNDARY | abi.STEP_DEPOSIT)

    elapsed, kern_ms, launches = _timed(ctx, one_step, steps, warmup)
    live = abi.count_live(ctx, bufs[cur[0]])
    inner = [s.inner_iters for s in solve_info[warmup:]]
    pass_us = kern_ms / max(launches, 1) * 1e3
    out = {
        "metric": "particle-steps/s, self-consistent 2D Penning trap (solve every step)",
        "value": live * steps / elapsed, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 push/deposit, f64 Poisson residual", "data": "synthetic",
        "config": {"workload": "2D Penning trap, 100k electrons, 128x128, Bz 0.1 T, end caps -10 V, "
                               "Poisson every step", "particles": N, "live_particles_end": live,
                   "grid": [n, n], "dt": 1e-11, "sort_every": args.sort_every},
        "particle_pass_us": pass_us,
        "poisson": {"tol": 1e-10, "mean_inner_pcg_iters": float(np.mean(inner)),
                    "max_rel_residual": max(s.rel_residual for s in solve_info[warmup:])},
        "device_flags": ctx.status(clear=True),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_penning(fields, mask, volt, n, L, dV, N, vth, args.cpu_seconds)
    emit(out)


def cpu_baseline_penning(fields, mask, volt, n, L, dV, N, vth, budget_s):
    """The oracle's Pic ordering on the same trap: FP64 direct Poisson solve
    every step, Boris, boundary and deposit, serial."""
    from oracle import oracle as O
    import ctypes as C
    dx = L / (n - 1)
    og = O.grid2(n, n, dx, dx)
    u = np.zeros(n * n)
    of = O.make_fields(og, mask, u, x_max=L, z_max=L, geometry_empty=1, selfconsistent=1, Bz=0.1)
    u = of._keep[1]
    p = O.particles(N)
    O.lib().o_add_particles_on_disk_philox(O.dptr(p), N, L / 2, L / 2, 1e-2, vth, C.byref(of), 1234, 0)
    rho = np.zeros(n * n)
    s = O.make_species(-abi.QE, M_E, 1e-11, rho=rho)
    O.lib().o_accumulate(C.byref(of), C.byref(s), O.dptr(p), N)

    def step():
        b = rho.copy()
        O.lib().o_poisson_rhs(C.byref(og), O.dptr(mask), O.dptr(volt), O.dptr(b), O.SOLVE_STATIC, dV)
        out = np.zeros_like(b)
        O.lib().o_poisson_solve(C.byref(og), O.dptr(mask), O.dptr(b), O.dptr(out))
        u[:] = out
        rho[:] = 0
        O.lib().o_advance_boris(C.byref(of), C.byref(s), O.dptr(p), N, None)
        O.lib().o_advance_boundary(C.byref(of), C.byref(s), O.dptr(p), N, None)
        s.niter += 1

    t0 = time.perf_counter()
    step()
    per = time.perf_counter() - t0
    k = max(1, int(budget_s / per) - 1)
    t0 = time.perf_counter()
    for _ in range(k):
        step()
    t = time.perf_counter() - t0
    live = int((p["empty"] == 0).sum())
    return {"value": live * k / t, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{N} electrons x {k} steps of the same trap, FP64 banded direct solve every step, "
                      f"serial; {t:.1f} s"}


def run_adjoint2d(args):
    """Adjoint sensitivity: the mean ion kinetic energy after 2000 self-consistent
    steps of a 2D trap (256 x 256 over 2 cm, Bz = 0.1 T, 500 000 ions of 3 u at
    0.1 eV on a 5 mm disk, dt = 1e-9 s), differentiated with respect to 8
    segmented wall electrodes (voltages U[-5, 5] V, seed 1234) with
    checkpoints every 100 steps.  Reports the gradient's wall time over one
    plain forward run of the same steps."""
    n, N, L = args.grid or 256, args.particles or 500_000, 2e-2
    steps = args.steps or 2000
    K = 100
    ctx = abi.Context(0, seed=1234)
    dx = L / (n - 1)
    g = abi.grid2d(n, n, dx, dx)
    mask, volt = abi.grid_build(g, abi.GEO_EMPTY, L, L, extern_field=0.0)
    group = np.full(n * n, -1, np.int32)
    G = group.reshape(n, n)
    h = n // 2
    G[0, :h], G[0, h:], G[-1, :h], G[-1, h:] = 0, 1, 2, 3
    G[1:-1, 0][: h - 1], G[1:-1, 0][h - 1:] = 4, 5
    G[1:-1, -1][: h - 1], G[1:-1, -1][h - 1:] = 6, 7
    V = np.random.default_rng(1234).uniform(-5, 5, 8)
    volt = volt + np.where(group >= 0, V[np.maximum(group, 0)], 0.0)
    dV = (N / 1e12) / ((n - 1) ** 2)
    params = abi.FieldsParams(g, L, L, 1, 1, 0, abi.BC_FREE, 0.0, 0, 0.0, 0.0, 0.0, 0.0, 0.1, 0.0, dV)
    fields = abi.Fields2D(ctx, params, mask, volt)
    vth = math.sqrt(0.1 * abi.QE / M_ION)
    sp = abi.Species2D(abi.QE, M_ION, 1e-9, float("inf"), 0, 0, 0)

    def fresh():
        # the disk sample, tile-sorted as a production run keeps its store
        Q = abi.ParticlesSoA(ctx, N)
        abi.add_particles_on_disk(ctx, fields, Q, L / 2, L / 2, 5e-3, vth, stream=0)
        P = abi.ParticlesSoA(ctx, N)
        abi.sort(ctx, fields, Q, P, want_count=False)
        fields.upload(abi.FIELD_U, np.zeros(n * n))
        ctx.sync()
        return P

    P = fresh()
    t0 = time.perf_counter()
    J0, _, info0 = abi.adjoint_energy(ctx, fields, sp, P, steps)
    t_fwd = time.perf_counter() - t0
    P = fresh()
    t0 = time.perf_counter()
    # the adjoint solves stop at an FP64 relative residual of 1e-8: their error then sits four
    # orders below the 1e-4 finite-difference check (tests/test_adjoint.py runs the same setting)
    J1, grad, info1 = abi.adjoint_energy(ctx, fields, sp, P, steps, group=group, n_groups=8, checkpoint_every=K,
                                         adjoint_tol=ADJ_TOL)
    t_grad = time.perf_counter() - t0
    out = {
        "metric": "adjoint/forward time ratio (gradient of mean ion energy w.r.t. 8 electrode voltages)",
        "value": t_grad / t_fwd, "unit": "x forward", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": t_grad / steps * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 push/deposit, f64 adjoint and Poisson residual", "data": "synthetic",
        "config": {"workload": "2D trap, 500k ions of 3 u, 256x256, 8 segmented electrodes, Poisson every "
                               "step, checkpoint every 100", "particles": N, "live_particles_end": info1.n_live,
                   "grid": [n, n], "dt": 1e-9, "checkpoint_every": K, "forward_tol": 1e-10,
                   "adjoint_tol": ADJ_TOL},
        "forward_s": t_fwd, "gradient_s": t_grad,
        "gradient_split_ms": {"forward": info1.forward_ms, "replay_and_reverse": info1.backward_ms},
        "J_eV": J1, "J_forward_only_eV": J0, "grad_eV_per_V": list(grad),
        "device_flags": ctx.status(clear=True),
    }
    emit(out)


def run_batch2d(args):
    """Batched ensemble: 256 independent RF 22-pole runs on a 128 x 128 grid
    (2 cm), ions of 3 u at 0.1 eV on a 5 mm disk, dt = 1e-9 s; per run
    (seed 1234) N ~ lognormal clipped to [1e3, 1e5], He pressure log-uniform
    in [1e-4, 1e-1] Pa at 300 K (constant k = 1e-15 m^3/s elastic) and V_rf
    uniform in [10, 50] V; 5000 steps in one hipGraph of persistent launches
    with the tuned shape.  Compared with the same runs stepped one after
    another with the plain pass (a sample of runs and steps, as a rate)."""
    n, L = args.grid or 128, 2e-2
    steps = args.steps or 5000
    R = 256
    ctx = abi.Context(0, seed=1234)
    dx = L / (n - 1)
    g = abi.grid2d(n, n, dx, dx)
    mask, volt = abi.grid_build(g, abi.GEO_RF_22PT, L, L)
    omega = 2 * math.pi * 17e6

    def fields_for(amp):
        prm = abi.FieldsParams(g, L, L, 0, 0, 0, abi.BC_FREE, 0.0, 1, amp, omega, 0.0, 0.0, 0.0, 0.0, 1e-12)
        return abi.Fields2D(ctx, prm, mask, volt)

    fields = fields_for(30.0)
    fields.solve(abi.SOLVE_RF, 1e-10)
    urf = fields.download(abi.FIELD_URF)
    rng = np.random.default_rng(1234)
    counts = np.clip(np.exp(rng.normal(math.log(1e4), 1.0, R)), 1e3, 1e5).astype(np.int64)
    pressure = 10 ** rng.uniform(-4, -1, R)
    amps = rng.uniform(10, 50, R)
    dens = pressure / (abi.KB * 300.0)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    N = int(offsets[-1])
    P = abi.ParticlesSoA(ctx, N)
    vth = math.sqrt(0.1 * abi.QE / M_ION)
    abi.add_particles_on_disk(ctx, fields, P, L / 2, L / 2, 5e-3, vth, stream=0)
    mccs = [abi.Mcc(ctx, M_ION, [abi.Target(M_HE, d, 300.0, 0.0, abi.Particles2D())],
                    [abi.Interaction(abi.ELASTIC, 0, 0.0, 1e-15, 0.0, 0, 0)], 1.0) for d in dens]
    sp = abi.Species2D(abi.QE, M_ION, 1e-9, float("inf"), 0, 0, 0)
    batch = abi.Batch(ctx, fields, sp, P, offsets, amps, mccs)
    t0 = time.perf_counter()
    shape = batch.tune(0, 50, 1000)
    t_tune = time.perf_counter() - t0
    ctx.sync()
    t0 = time.perf_counter()
    batch.advance(0, steps)
    ctx.sync()
    t_batch = time.perf_counter() - t0
    mean, live = batch.energy()
    rate_batch = N * steps / t_batch
    # sequential plain runs: every 32nd run alone, its own Fields (amplitude) and collisions
    seq_steps = min(200, steps)
    sample = list(range(0, R, 32))
    t_seq, ps_seq = 0.0, 0
    for r in sample:
        fr = fields_for(float(amps[r]))
        fr.upload(abi.FIELD_URF, urf)
        Q = abi.ParticlesSoA(ctx, int(counts[r]))
        abi.add_particles_on_disk(ctx, fr, Q, L / 2, L / 2, 5e-3, vth, stream=r)
        s1 = abi.Species2D(abi.QE, M_ION, 1e-9, mccs[r].lifetime, 0, r, 0)
        ctx.sync()
        t0 = time.perf_counter()
        for k in range(seq_steps):
            s1.niter = k
            abi.advance(ctx, fr, s1, Q, mcc=mccs[r], flags=abi.STEP_PUSH | abi.STEP_BOUNDARY)
        ctx.sync()
        t_seq += time.perf_counter() - t0
        ps_seq += int(counts[r]) * seq_steps
    rate_seq = ps_seq / t_seq
    out = {
        "metric": "batched ensemble speed-up over sequential plain runs (particle-steps/s ratio)",
        "value": rate_batch / rate_seq, "unit": "x sequential", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": t_batch / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "256 RF 22-pole runs, 128x128, ragged N lognormal [1e3,1e5], pressure log-uniform "
                               "[1e-4,1e-1] Pa, V_rf U[10,50] V", "runs": R, "particles": N,
                   "grid": [n, n], "dt": 1e-9, "shape": {"block": shape[0], "ppt": shape[1], "seg_steps": shape[2]}},
        "batched_particle_steps_per_s": rate_batch, "batched_s": t_batch, "tune_s": t_tune,
        "sequential_particle_steps_per_s": rate_seq,
        "sequential_sample": f"{len(sample)} runs x {seq_steps} steps stepped alone with the plain pass",
        "live_particles_end": int(live.sum()), "mean_energy_eV_range": [float(mean.min()), float(mean.max())],
        "device_flags": ctx.status(clear=True),
    }
    emit(out)


def quad_trap_3d(n, L, V_rod=2.0, V_end=5.0, r0=0.3, a=0.08):
    """Linear quadrupole in a box: four rods along z (+V, -V, +V, -V) at
    radius r0 L, end caps (the z faces) at V_end, the other faces grounded."""
    h = L / (n - 1)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    X, Y = i * h - L / 2, j * h - L / 2
    mask = np.full((n, n, n), abi.FREE, np.uint8)
    volt = np.zeros((n, n, n))
    face = (i == 0) | (j == 0) | (k == 0) | (i == n - 1) | (j == n - 1) | (k == n - 1)
    mask[face] = abi.FIXED
    volt[(k == 0) | (k == n - 1)] = V_end
    for q, (cx, cy) in enumerate([(r0, 0), (0, r0), (-r0, 0), (0, -r0)]):
        rod = (X - cx * L) ** 2 + (Y - cy * L) ** 2 <= (a * L) ** 2
        mask[rod & ~face] = abi.FIXED
        volt[rod & ~face] = V_rod if q % 2 == 0 else -V_rod
    return mask.ravel(), volt.ravel(), h


def run_trap3d(args):
    """3D trap: 128^3 grid over a 1 cm box, linear quadrupole rods with end
    caps, Bz = 0.5 T, 16 777 216 ions of 40 u (Gaussian cloud, 0.05 eV),
    space-charge Poisson (7-point multigrid PCG, FP64 refinement to 1e-10)
    every step, fixed-point CIC, tile sort every 20 steps."""
    n, N, L = args.grid or 128, args.particles or 16 * 1024 * 1024, 1e-2
    steps, warmup = args.steps or 100, args.warmup if args.warmup is not None else 5
    ctx = abi.Context(0, seed=1234)
    mask, volt, h = quad_trap_3d(n, L)
    g = abi.grid3d(n, n, n, h, h, h)
    params = abi.Fields3DParams(g, L, L, L, abi.BC_FREE, 0.0, 0.0, 0.5, 1e13)
    fields = abi.Fields3D(ctx, params, mask, volt)
    rng = np.random.default_rng(1234)
    pos = np.clip(L / 2 + rng.normal(size=(3, N)).astype(np.float32) * np.float32(0.1 * L),
                  np.float32(0.2 * L), np.float32(0.8 * L)).astype(np.float32)
    vth = math.sqrt(0.05 * abi.QE / M_40)
    vel = (rng.normal(size=(3, N)).astype(np.float32) * np.float32(vth)).astype(np.float32)
    P = abi.ParticlesSoA3D.from_host(ctx, *pos, *vel)
    del pos, vel
    sp = abi.Species2D(abi.QE, M_40, 1e-8, float("inf"), 0, 0, 0)
    rho = fields.ptr(abi.FIELD_RHO)
    fields.reset()
    abi.accumulate_3d(ctx, g, P, abi.QE, rho)
    t0 = time.perf_counter()
    info0 = fields.solve(1e-10)
    ctx.sync()
    solve0_ms = (time.perf_counter() - t0) * 1e3
    fields.reset()
    abi.advance_init_3d(ctx, fields, sp, P)
    bufs = [P, abi.ParticlesSoA3D(ctx, N)]
    cur = [0]
    infos = []
    # Re-sort policy.  A deposit outside its tile's LDS window costs a
    # contended global atomic per node: about 0.43 ns a particle against
    # 0.8 ms for a sort of the 16.8M store, so once the spills since the last
    # sort reach about a tenth of the store they have cost one sort (the
    # rent-or-buy rule).  It needs the step statistics every step, whose read
    # stalls the queue: here it measured 2% slower than sorting every 10 steps.
    adaptive = args.resort_spill is not None
    spilled = [0]
    sorts = [0]

    def one_step(k):
        if (k == 0 or spilled[0] >= args.resort_spill * N) if adaptive else \
                (args.sort_every and k % args.sort_every == 0):
            abi.sort_3d(ctx, fields, bufs[cur[0]], bufs[1 - cur[0]])
            cur[0] = 1 - cur[0]
            spilled[0] = 0
            sorts[0] += 1
        if k > 0:
            infos.append(fields.solve(1e-10))
            fields.reset()
        sp.niter = k
        st = abi.advance_3d(ctx, fields, sp, bufs[cur[0]], rho_ptr=rho,
                            flags=abi.STEP_PUSH | abi.STEP_BOUNDARY | abi.STEP_DEPOSIT, want_stats=adaptive)
        spilled[0] += st.spilled

    sorts_warm = []
    elapsed, kern_ms, launches = _timed(ctx, one_step, steps, warmup,
                                        on_timed=lambda: sorts_warm.append(sorts[0]))
    live = abi.count_live_3d(ctx, bufs[cur[0]])
    kern_us = kern_ms / max(launches, 1) * 1e3
    bytes_launch = 48.0 * live + 12.0 * n ** 3  # 24 B state read+written, FP32 gradient tables once
    out = {
        "metric": "particle-steps/s, 3D trap with space-charge Poisson every step",
        "value": live * steps / elapsed, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 push/deposit, f64 Poisson residual", "data": "synthetic",
        "config": {"workload": "3D linear quadrupole + end caps, 128^3, Bz 0.5 T, 16M ions of 40 u, "
                               "Poisson every step", "particles": N, "live_particles_end": live,
                   "grid": [n, n, n], "dt": 1e-8,
                   "sort": (f"re-sort once the spilled deposits since the last sort reach {args.resort_spill} "
                            "of the store" if adaptive else f"every {args.sort_every} steps"),
                   "sorts_timed": sorts[0] - sorts_warm[0]},
        "roofline": {"kernel": "k_advance3d", "bound": "hbm", "achieved": bytes_launch / (kern_us * 1e-6) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bytes_launch / (kern_us * 1e-6) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel_avg_us": kern_us, "bytes_per_launch": bytes_launch},
        "poisson": {"tol": 1e-10, "first_solve_ms": solve0_ms, "first_outer": info0.outer_iters,
                    "first_inner": info0.inner_iters,
                    "mean_inner_pcg_iters": float(np.mean([s.inner_iters for s in infos])) if infos else None,
                    "max_rel_residual": max(s.rel_residual for s in infos) if infos else None},
        "device_flags": ctx.status(clear=True),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_3d(fields, mask, n, L, h, args.cpu_seconds)
    emit(out)


def cpu_baseline_3d(fields, mask, n, L, h, budget_s):
    """The oracle's Species<CARTESIAN3D> step (push, boundary, deposit) on a
    sample of the same trap; the 3D solve has no CPU restatement here."""
    from oracle import oracle as O
    import ctypes as C
    u = fields.download(abi.FIELD_U)
    og = O.grid3(n, n, n, h, h, h)
    of = O.make_fields3(og, mask, u, B=(0.0, 0.0, 0.5))
    rng = np.random.default_rng(1)
    N = 200_000
    p = O.particles(N)
    for c in ("x", "y", "z"):
        p[c] = np.clip(L / 2 + rng.normal(size=N) * 0.1 * L, 0.2 * L, 0.8 * L)
    vth = math.sqrt(0.05 * abi.QE / M_40)
    for c in ("vx", "vy", "vz"):
        p[c] = rng.normal(size=N) * vth
    rho = np.zeros(n ** 3)
    s = O.make_species(abi.QE, M_40, 1e-8, rho=rho)
    t0 = time.perf_counter()
    O.lib().o_advance3d(C.byref(of), C.byref(s), O.dptr(p), N, None, None)
    per = time.perf_counter() - t0
    k = max(2, int(budget_s / per))
    t0 = time.perf_counter()
    for _ in range(k):
        O.lib().o_advance3d(C.byref(of), C.byref(s), O.dptr(p), N, None, None)
    t = time.perf_counter() - t0
    return {"value": N * k / t, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{N} ions x {k} steps of the same 3D trap (push, boundary, deposit; no solve), "
                      f"serial FP64; {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("rf22pt", "penning2d", "trap3d", "adjoint2d", "batch2d"),
                    default="rf22pt",
                    help="rf22pt: the headline RF trap (default); penning2d: self-consistent 2D Penning "
                         "trap; trap3d: 3D quadrupole with space charge")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--grid", type=int, default=None)
    ap.add_argument("--sort-every", type=int, default=None,
                    help="tile-sort the particle store every K steps (inside the timed region); default "
                         "60 for rf22pt (the 2D deposit windows reach 8 cells past their tile), 40 for "
                         "the other 2D workloads, 10 for trap3d, whose ions leave the 2-cell 3D halo "
                         "sooner (5 and 20 measured slower)")
    ap.add_argument("--resort-spill", type=float, default=None,
                    help="trap3d: instead of a fixed period, re-sort once the deposits that fell outside "
                         "their tile's LDS window since the last sort add up to this fraction of the store "
                         "(read from the step statistics: one more synchronisation a step)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the timed steps, run plain copies of this many MiB (counter calibration)")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="print a markdown table of a rocprofv3 --kernel-trace --stats csv output and exit")
    ap.add_argument("--summarize-counters", nargs="+", metavar="DIR",
                    help="print the SQ counters of a kernel (--counter-kernel) from rocprofv3 --pmc run "
                         "directories")
    ap.add_argument("--counter-kernel", default="k_advance2d",
                    help="kernel whose counters --summarize-counters prints (default k_advance2d)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="print the calibrated FETCH_SIZE/WRITE_SIZE summary of two rocprofv3 --pmc csv "
                         "outputs (with --copy-calibration-mib) and exit")
    args = ap.parse_args()
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace))
        return
    if args.summarize_pmc:
        print(summarize_pmc(*args.summarize_pmc, args.copy_calibration_mib))
        return
    if args.summarize_counters:
        print(summarize_counters(args.summarize_counters, args.particles or 4194304, args.counter_kernel))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}): the "
              "particle pass would shard by particles with an int64 all-reduce of the deposit grid, "
              "which this build does not implement", file=sys.stderr)
        sys.exit(2)
    if args.sort_every is None:
        args.sort_every = {"trap3d": 10, "rf22pt": 60}.get(args.workload, 40)
    if args.workload == "penning2d":
        return run_penning2d(args)
    if args.workload == "trap3d":
        return run_trap3d(args)
    if args.workload == "adjoint2d":
        return run_adjoint2d(args)
    if args.workload == "batch2d":
        return run_batch2d(args)
    args.steps = args.steps if args.steps is not None else 200
    args.warmup = args.warmup if args.warmup is not None else 20
    args.particles = args.particles or 4 * 1024 * 1024
    args.grid = args.grid or 512

    ctx = abi.Context(0, seed=1234)
    g, mask, volt, params, fields, mcc, poisson = build_scene(ctx, args.grid, args.particles)
    N = args.particles
    P = abi.ParticlesSoA(ctx, N)
    vth = math.sqrt(0.1 * abi.QE / M_ION)
    abi.add_particles_on_disk(ctx, fields, P, 1e-2, 1e-2, 5e-3, vth, stream=0)
    sp = abi.Species2D(abi.QE, M_ION, 1e-9, mcc.lifetime, 0, 0, 0)
    abi.advance_init(ctx, fields, sp, P)
    rho = abi.DeviceArray(ctx, (args.grid * args.grid,), np.float64)
    rho.zero()
    flags = abi.STEP_PUSH | abi.STEP_BOUNDARY | abi.STEP_DEPOSIT

    bufs = [P, abi.ParticlesSoA(ctx, N)]
    cur = 0

    def one_step(step):
        nonlocal cur
        if args.sort_every and step % args.sort_every == 0:
            abi.sort(ctx, fields, bufs[cur], bufs[1 - cur], want_count=False)
            cur = 1 - cur
        sp.niter = step
        abi.advance(ctx, fields, sp, bufs[cur], mcc=mcc, rho_ptr=rho.ptr, flags=flags)

    step = 0
    for _ in range(args.warmup):
        one_step(step)
        step += 1
    ctx.sync()
    live0 = abi.count_live(ctx, bufs[cur])
    ctx.profile(True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_step(step)
        step += 1
    ctx.sync()
    elapsed = time.perf_counter() - t0
    kern_ms, launches = ctx.profile_read()
    ctx.profile(False)
    sorts_in_window = sum(1 for k in range(args.warmup, args.warmup + args.steps)
                          if args.sort_every and k % args.sort_every == 0)
    live1 = abi.count_live(ctx, bufs[cur])
    flags_dev = ctx.status(clear=True)
    live_mean = 0.5 * (live0 + live1)
    # the sort's cost, timed on the live store after the window: the
    # steady-state step carries 1/sort_every of it whether or not the
    # window happened to hold a sort
    sort_ms = 0.0
    if args.sort_every:
        n_sorts = 3
        ctx.sync()
        t1 = time.perf_counter()
        for _ in range(n_sorts):
            abi.sort(ctx, fields, bufs[cur], bufs[1 - cur], want_count=False)
            cur = 1 - cur
        ctx.sync()
        sort_ms = (time.perf_counter() - t1) / n_sorts * 1e3
    window_ms = elapsed / args.steps * 1e3
    step_ms = (elapsed * 1e3 - sorts_in_window * sort_ms) / args.steps + (
        sort_ms / args.sort_every if args.sort_every else 0.0)
    value = live_mean / (step_ms * 1e-3)
    Nc = args.grid * args.grid
    kern_avg_s = kern_ms / max(launches, 1) * 1e-3
    # algorithmic bytes: the particle pass reads and writes the 20 B SoA state
    # of every live ion and reads the two FP32 potentials (u, uRF) once; the
    # deposit grid's zero + flush (16 B per node) belongs to the tile reduction
    bytes_per_launch = 40.0 * live_mean + 8.0 * Nc
    bytes_per_step = 40.0 * live_mean + 24.0 * Nc
    achieved = bytes_per_launch / kern_avg_s / 1e9
    if args.copy_calibration_mib:
        copy_calibration(ctx, args.copy_calibration_mib)
    traffic = committed_traffic(N, args.grid)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "RF 22-pole ion trap with He buffer-gas MCC (push+gather+MCC+boundary+"
                               "CIC deposit per step)",
                   "particles": N, "live_particles_mean": live_mean, "grid": [args.grid, args.grid],
                   "dt": 1e-9,
                   "layout": "SoA FP32 x,z,vx,vy,vz (20 B per ion; the long-time form adds FP32 "
                             "position remainders xl,zl, 28 B, not used here); int64 2^-31 fixed-point "
                             "deposit grid",
                   "sort_every": args.sort_every,
                   "sort": "stable tile order (8x8-cell tiles), timed separately: ms_per_step = "
                           "(window - sorts in window) / steps + sort_ms / sort_every",
                   "sort_ms": sort_ms, "sorts_in_window": sorts_in_window,
                   "window_ms_per_step": window_ms,
                   "device_flags": flags_dev},
        "roofline": {"kernel": "k_advance2d", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel_avg_us": kern_avg_s * 1e6, "bytes_per_launch": bytes_per_launch,
                     "launches_timed": launches,
                     "step_bytes": bytes_per_step,
                     "step_achieved_GBs": bytes_per_step / (step_ms * 1e-3) / 1e9,
                     "step_frac": bytes_per_step / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "poisson": poisson,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(fields, params, mask, args.grid, args.cpu_seconds, N)
    else:
        out["cpu_baseline"] = None
    emit(out)


if __name__ == "__main__":
    main()
