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
M_ION = 3 * 1.66053906660e-27
M_HE = 6.69e-27


def build_scene(ctx, n_grid, n_particles):
    L = 2e-2
    dx = L / (n_grid - 1)
    g = abi.grid2d(n_grid, n_grid, dx, dx)
    mask, volt = abi.grid_build(g, abi.GEO_RF_22PT, L, L)
    dV = (n_particles / 1e11) / ((n_grid - 1) ** 2)  # Param::dV with density_total 1e11
    params = abi.FieldsParams(g, L, L, 0, 0, 0, abi.BC_FREE, 0.0, 1, 30.0, 2 * math.pi * 17e6, 0.0,
                              0.0, 0.0, 0.0, dV)
    fields = abi.Fields2D(ctx, params, mask, volt)
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


def cpu_baseline(fields, params, mask, n_grid, budget_s):
    """Oracle serial FP64 step (boris + MCC with the SHR3 stream, boundary +
    deposit) on a sample of the same scene; returns particle-steps/s."""
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
    n = 1_000_000
    steps = max(2, int(budget_s / (per_ps * n)))
    p = sample(n)
    t = O.lib().o_bench_steps(C.byref(of), C.byref(s), O.dptr(p), n, steps)
    return {"value": n * steps / t, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} ions x {steps} steps of the same RF 22-pole scene (uRF from the GPU solve), "
                      f"FP64 AoS, serial reference loop order, SHR3 stream, deposit on; {t:.1f} s"}


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


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--particles", type=int, default=4 * 1024 * 1024)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--sort-every", type=int, default=20,
                    help="cell-sort the particle store every K steps (inside the timed region)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the timed steps, run plain copies of this many MiB (counter calibration)")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="print a markdown table of a rocprofv3 --kernel-trace --stats csv output and exit")
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
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}): the "
              "particle pass would shard by particles with an int64 all-reduce of the deposit grid, "
              "which this build does not implement", file=sys.stderr)
        sys.exit(2)

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
    live1 = abi.count_live(ctx, bufs[cur])
    flags_dev = ctx.status(clear=True)
    live_mean = 0.5 * (live0 + live1)
    value = live_mean * args.steps / elapsed
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
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "RF 22-pole ion trap with He buffer-gas MCC (push+gather+MCC+boundary+"
                               "CIC deposit per step)",
                   "particles": N, "live_particles_mean": live_mean, "grid": [args.grid, args.grid],
                   "dt": 1e-9, "layout": "SoA FP32 x,z,vx,vy,vz (20 B); int64 fixed-point deposit",
                   "sort_every": args.sort_every,
                   "device_flags": flags_dev},
        "roofline": {"kernel": "k_advance2d", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel_avg_us": kern_avg_s * 1e6, "bytes_per_launch": bytes_per_launch,
                     "launches_timed": launches,
                     "step_bytes": bytes_per_step,
                     "step_achieved_GBs": bytes_per_step / (elapsed / args.steps) / 1e9},
        "poisson": poisson,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(fields, params, mask, args.grid, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
