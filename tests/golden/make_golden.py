"""Write the golden fixtures of tests/golden/ from the reference's own code.

Run in a container that has the reference sources (default /root/reference)
after `make -C oracle ref`, which compiles oracle/_ref/ from those sources:

    python tests/golden/make_golden.py [/root/reference]

Fixtures (data only: inputs and the reference's outputs):
  langevin_chi.txt   stdout of the reference's tests/test_langevin.cpp
  field2d_ref.npz    Field2D::accumulate / interpolate / grad (Field2D.hpp:45-168)
  field3d_ref.npz    Field3D::accumulate / interpolate (Field3D.hpp:40-96)
  rng_ref.npz        t_random(1234): iuni, uni, rot, deflect, radius (random.cpp)
  vecinterp_ref.npz  vec_interpolate (tabulate.cpp:89-141) on the e-He elastic
                     cross section of species_conf_MCC.txt
  misc_ref.npz       mod (mymath.cpp:62-67), Histogram::mean_tot (histogram.cpp)
  field2d_print_ref.txt, field2d_load_[abcd].txt, field2d_text_ref.npz
                     Field2D::print of field2d_text_ref's print_* data, and
                     Field2D::load of the four files (Field2D.cpp:30-130)
"""
from __future__ import annotations

import ctypes as C
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402

REF = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")


def dp(a):
    return a.ctypes.data_as(C.c_void_p)


def field2d_points(rng, nx, nz, dx, dz, n):
    xs = rng.uniform(0, (nx - 1) * dx, n)
    zs = rng.uniform(0, (nz - 1) * dz, n)
    # nodes, cell faces and the last row/column exactly
    ex = np.array([0.0, dx, (nx - 1) * dx, (nx - 2) * dx, 0.5 * dx, (nx - 1.5) * dx, 3 * dx])
    ez = np.array([0.0, (nz - 1) * dz, dz, (nz - 2) * dz, (nz - 1.5) * dz, 0.5 * dz, 2 * dz])
    xs = np.concatenate([xs, ex, rng.uniform(0, (nx - 1) * dx, 7)])
    zs = np.concatenate([zs, rng.uniform(0, (nz - 1) * dz, 7), ez])
    return xs, zs


def main():
    L = O.ref()
    rng = np.random.default_rng(20241202)

    # Langevin table: the reference's test program, as is
    out = subprocess.run([str(O.REF_LANGEVIN)], capture_output=True, text=True, check=True).stdout
    (HERE / "langevin_chi.txt").write_text(out)

    # Field2D
    nx, nz, dx, dz = 9, 7, 1.25e-3, 1.25e-3
    data = rng.normal(size=nx * nz)
    xs, zs = field2d_points(rng, nx, nz, dx, dz, 200)
    acc = np.zeros(nx * nz)
    xs_in = xs[xs < (nx - 1) * dx - 1e-12]
    zs_in = zs[xs < (nx - 1) * dx - 1e-12]
    keep = zs_in < (nz - 1) * dz - 1e-12
    xs_in, zs_in = xs_in[keep], zs_in[keep]
    bad = L.ref_f2_accumulate_many(nx, nz, C.c_double(dx), C.c_double(dz), C.c_double(0), C.c_double(0),
                                   C.c_double(1.5), C.c_long(len(xs_in)), dp(xs_in), dp(zs_in), dp(acc))
    assert bad == 0
    interp = np.zeros(len(xs_in))
    L.ref_f2_interpolate_many(nx, nz, C.c_double(dx), C.c_double(dz), C.c_double(0), C.c_double(0),
                              dp(data), C.c_long(len(xs_in)), dp(xs_in), dp(zs_in), dp(interp))
    gx = np.zeros(len(xs))
    gz = np.zeros(len(xs))
    L.ref_f2_grad_many(nx, nz, C.c_double(dx), C.c_double(dz), C.c_double(0), C.c_double(0), dp(data),
                       C.c_long(len(xs)), dp(xs), dp(zs), dp(gx), dp(gz))
    np.savez(HERE / "field2d_ref.npz", nx=nx, nz=nz, dx=dx, dz=dz, data=data, xs=xs, zs=zs, gx=gx, gz=gz,
             xs_in=xs_in, zs_in=zs_in, q=1.5, acc=acc, interp=interp)

    # Field3D (accumulate / interpolate; grad reads members the reference never sets)
    n3 = (6, 5, 7)
    d3 = (1e-3, 1.1e-3, 0.9e-3)
    data3 = rng.normal(size=int(np.prod(n3)))
    p3 = [rng.uniform(0, (n3[a] - 1) * d3[a] * 0.999, 150) for a in range(3)]
    acc3 = np.zeros(int(np.prod(n3)))
    bad = L.ref_f3_accumulate_many(*n3, *map(C.c_double, d3), C.c_double(0.7), C.c_long(150),
                                   dp(p3[0]), dp(p3[1]), dp(p3[2]), dp(acc3))
    assert bad == 0
    int3 = np.zeros(150)
    L.ref_f3_interpolate_many(*n3, *map(C.c_double, d3), dp(data3), C.c_long(150), dp(p3[0]), dp(p3[1]),
                              dp(p3[2]), dp(int3))
    np.savez(HERE / "field3d_ref.npz", n=np.array(n3), d=np.array(d3), data=data3, x=p3[0], y=p3[1],
             z=p3[2], q=0.7, acc=acc3, interp=int3)

    # t_random(1234)
    L.ref_rng_new.restype = C.c_void_p
    r = C.c_void_p(L.ref_rng_new(1234))
    u32 = np.zeros(64, np.uint32)
    L.ref_rng_u32_many(r, C.c_long(64), u32.ctypes.data_as(C.c_void_p))
    uni = np.zeros(64)
    L.ref_rng_uni_many(r, C.c_long(64), dp(uni))
    rot = np.zeros((16, 3))
    for k in range(16):
        x, y, z = C.c_double(), C.c_double(), C.c_double()
        L.ref_rng_rot_len(r, C.c_double(2.5), C.byref(x), C.byref(y), C.byref(z))
        rot[k] = (x.value, y.value, z.value)
    vin = rng.normal(size=(16, 3))
    angles = rng.uniform(0.01, 3.0, 16)
    defl = np.zeros((16, 3))
    for k in range(16):
        x, y, z = (C.c_double(v) for v in vin[k])
        L.ref_rng_deflect(r, C.c_double(angles[k]), C.byref(x), C.byref(y), C.byref(z))
        defl[k] = (x.value, y.value, z.value)
    rad = np.array([L.ref_rng_radius(r) for _ in range(16)])
    L.ref_rng_free(r)
    np.savez(HERE / "rng_ref.npz", seed=1234, u32=u32, uni=uni, rot_len=2.5, rot=rot, deflect_in=vin,
             deflect_angle=angles, deflect=defl, radius=rad)

    # vec_interpolate on the e-He elastic table of species_conf_MCC.txt
    E, S = [], []
    lines = (REF / "species_conf_MCC.txt").read_text().splitlines()
    on = False
    for ln in lines:
        t = ln.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "CROSS_SECTION":
            on = True
            continue
        if t[0] == "END_CROSS_SECTION":
            break
        if on:
            E.append(float(t[0]))
            S.append(float(t[1]))
    E, S = np.array(E), np.array(S)
    q = np.concatenate([rng.uniform(-1, E[-1] * 1.2, 300), E[:5], [E[-1], E[0]]])
    out = np.zeros(len(q))
    L.ref_vec_interpolate_many(C.c_int(len(E)), dp(E), dp(S), C.c_long(len(q)), dp(q), dp(out))
    np.savez(HERE / "vecinterp_ref.npz", E=E, S=S, q=q, out=out)

    # mod and Histogram::mean_tot
    mx = np.concatenate([rng.uniform(-1e8, 1e8, 50), [0.0, 3.0, -3.0]])
    my = np.full(len(mx), 10000000 * np.pi)
    mo = np.array([L.ref_mod(C.c_double(a), C.c_double(b)) for a, b in zip(mx, my)])
    vals = rng.exponential(0.3, 500)
    L.ref_hist_mean_tot.argtypes = [C.c_int, C.c_double, C.c_double, C.c_long, C.c_void_p]
    mt = L.ref_hist_mean_tot(200, 0.0, 1.0, len(vals), dp(vals))
    np.savez(HERE / "misc_ref.npz", mod_x=mx, mod_y=my, mod=mo, hist_vals=vals, hist_mean_tot=mt)

    field_text(L)
    print("golden fixtures written to", HERE)


def field_text(L):
    """Field2D::print and Field2D::load (Field2D.cpp:30-130)."""
    rng = np.random.default_rng(20241203)
    # print: values over many decades, a negative zero and integers, a
    # non-zero origin and a factor
    nx, nz, dx, dz, x0, z0, fac = 5, 4, 1.25e-3, 7.5e-4, -2.5e-3, 1e-3, 2.5
    data = rng.normal(size=nx * nz) * 10.0 ** rng.integers(-9, 10, nx * nz)
    data[3], data[7], data[11] = -0.0, 4.0, 123456789.0
    out = HERE / "field2d_print_ref.txt"
    L.ref_f2_print(nx, nz, C.c_double(dx), C.c_double(dz), C.c_double(x0), C.c_double(z0), dp(data),
                   C.c_double(fac), str(out).encode())
    # load: a 6 x 5 grid written x-descending (a), z-major (b), with a node
    # missing (c) and with one node twice and another missing (d)
    mx, mz, hx, hz, ox, oz = 6, 5, 2e-4, 3e-4, 1e-3, -6e-4
    vals = rng.normal(size=(mx, mz))
    rows = {(i, j): f"{ox + i * hx:.6g}\t{oz + j * hz:.6g}\t{vals[i, j]:.17g}" for i in range(mx) for j in range(mz)}
    a = [rows[i, j] for i in reversed(range(mx)) for j in range(mz)]
    b = [rows[i, j] for j in range(mz) for i in range(mx)]
    c = a[:7] + a[8:]
    d = a[:7] + [a[9]] + a[8:]
    meta = np.zeros((4, 6))
    got = np.zeros((4, mx * mz))
    rc = np.zeros(4, np.int32)
    L.ref_f2_load.restype = C.c_int
    for k, (name, lines) in enumerate(zip("abcd", (a, b, c, d))):
        f = HERE / f"field2d_load_{name}.txt"
        f.write_text("# x z value\n" + "\n".join(lines) + "\n\n")
        m = np.zeros(6)
        o = np.zeros(mx * mz)
        rc[k] = L.ref_f2_load(str(f).encode(), dp(m), dp(o), C.c_long(mx * mz))
        meta[k], got[k] = m, o
    np.savez(HERE / "field2d_text_ref.npz", print_shape=np.array([nx, nz]),
             print_grid=np.array([dx, dz, x0, z0]), print_factor=fac, print_data=data, load_rc=rc,
             load_meta=meta, load_data=got)


if __name__ == "__main__":
    main()
