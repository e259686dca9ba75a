"""Pin the CPU oracle against the reference's own code.

The fixtures in tests/golden/ were written by tests/golden/make_golden.py
from oracle/_ref (the reference's Field2D.hpp, Field3D.hpp, random.cpp,
tabulate.cpp, mymath.cpp, histogram.cpp compiled where they lie) and from the
reference's tests/test_langevin.cpp.  Where oracle/_ref is built, the same
checks also run live on fresh random inputs.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest
from scipy import special

from oracle import oracle as O

G = Path(__file__).resolve().parent / "golden"


def _g2(d):
    return O.grid2(int(d["nx"]), int(d["nz"]), float(d["dx"]), float(d["dz"]))


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32-10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_field2d_accumulate_matches_reference():
    d = np.load(G / "field2d_ref.npz")
    g = _g2(d)
    acc = np.zeros(g.nx * g.nz)
    for x, z in zip(d["xs_in"], d["zs_in"]):
        assert O.lib().o_f2_accumulate(C.byref(g), O.dptr(acc), float(d["q"]), x, z) == 0
    np.testing.assert_allclose(acc, d["acc"], rtol=1e-14, atol=1e-15)
    assert abs(acc.sum() - d["q"] * len(d["xs_in"])) < 1e-11


def test_field2d_interpolate_and_grad_match_reference():
    d = np.load(G / "field2d_ref.npz")
    g = _g2(d)
    data = np.ascontiguousarray(d["data"])
    for x, z, ri in zip(d["xs_in"], d["zs_in"], d["interp"]):
        out = C.c_double()
        assert O.lib().o_f2_interpolate(C.byref(g), O.dptr(data), x, z, C.byref(out)) == 0
        assert out.value == pytest.approx(ri, rel=1e-13, abs=1e-14)
    for x, z, rgx, rgz in zip(d["xs"], d["zs"], d["gx"], d["gz"]):
        gx, gz = C.c_double(0), C.c_double(0)
        O.lib().o_f2_grad(C.byref(g), O.dptr(data), x, z, C.byref(gx), C.byref(gz))
        assert gx.value == pytest.approx(rgx, rel=1e-12, abs=1e-9)
        assert gz.value == pytest.approx(rgz, rel=1e-12, abs=1e-9)


def test_field3d_matches_reference():
    d = np.load(G / "field3d_ref.npz")
    n, h = d["n"], d["d"]
    g = O.Grid3(int(n[0]), int(n[1]), int(n[2]), h[0], h[1], h[2], 1 / h[0], 1 / h[1], 1 / h[2], 0, 0, 0,
                (n[0] - 1) * h[0], (n[1] - 1) * h[1], (n[2] - 1) * h[2])
    acc = np.zeros(int(np.prod(n)))
    data = np.ascontiguousarray(d["data"])
    for x, y, z, ri in zip(d["x"], d["y"], d["z"], d["interp"]):
        assert O.lib().o_f3_accumulate(C.byref(g), O.dptr(acc), float(d["q"]), x, y, z) == 0
        out = C.c_double()
        assert O.lib().o_f3_interpolate(C.byref(g), O.dptr(data), x, y, z, C.byref(out)) == 0
        assert out.value == pytest.approx(ri, rel=1e-13, abs=1e-14)
    np.testing.assert_allclose(acc, d["acc"], rtol=1e-13, atol=1e-15)


def test_shr3_stream_and_samplers_match_reference():
    d = np.load(G / "rng_ref.npz")
    r = O.Rng()
    O.lib().o_rng_shr3_seed(C.byref(r), int(d["seed"]))
    words = [O.lib().o_rng_u32(C.byref(r)) for _ in range(64)]
    assert words == list(d["u32"])
    uni = [O.lib().o_rng_uni(C.byref(r)) for _ in range(64)]
    assert uni == list(d["uni"])
    for k in range(16):
        x, y, z = C.c_double(), C.c_double(), C.c_double()
        O.lib().o_rot_len(C.byref(r), float(d["rot_len"]), C.byref(x), C.byref(y), C.byref(z))
        np.testing.assert_allclose([x.value, y.value, z.value], d["rot"][k], rtol=1e-14, atol=1e-14)
    for k in range(16):
        x, y, z = (C.c_double(v) for v in d["deflect_in"][k])
        O.lib().o_deflect(C.byref(r), float(d["deflect_angle"][k]), C.byref(x), C.byref(y), C.byref(z))
        np.testing.assert_allclose([x.value, y.value, z.value], d["deflect"][k], rtol=1e-13, atol=1e-13)
    rad = [O.lib().o_radius(C.byref(r)) for _ in range(16)]
    np.testing.assert_allclose(rad, d["radius"], rtol=1e-15)


def test_vec_interpolate_matches_reference():
    d = np.load(G / "vecinterp_ref.npz")
    E, S = np.ascontiguousarray(d["E"]), np.ascontiguousarray(d["S"])
    pe = E.ctypes.data_as(C.POINTER(C.c_double))
    ps = S.ctypes.data_as(C.POINTER(C.c_double))
    got = [O.lib().o_vec_interpolate(pe, ps, len(E), q) for q in d["q"]]
    np.testing.assert_allclose(got, d["out"], rtol=1e-15, atol=0)


def test_mod_and_mean_tot_match_reference():
    d = np.load(G / "misc_ref.npz")
    got = [O.lib().o_mod(a, b) for a, b in zip(d["mod_x"], d["mod_y"])]
    np.testing.assert_array_equal(got, d["mod"])
    # Histogram::mean_tot is the plain mean of every value added
    assert np.mean(d["hist_vals"]) == pytest.approx(float(d["hist_mean_tot"]), rel=1e-14)


def test_langevin_chi_matches_reference_table():
    rows = [list(map(float, ln.split())) for ln in (G / "langevin_chi.txt").read_text().splitlines()]
    for beta, chi, chi0 in rows:
        mine = O.lib().o_langevin_chi(beta)
        if np.isnan(chi):
            assert np.isnan(mine) or beta == 1.0
            continue
        # the table is printed with 6 significant digits
        assert mine == pytest.approx(chi, rel=6e-6, abs=1e-6)
        assert -3 * np.pi / (16 * beta ** 4) == pytest.approx(chi0, rel=6e-6, abs=1e-6)


def test_ellint_k_against_scipy():
    for k in [0.0, 0.1, 0.5, 0.9, 0.99, 0.999999]:
        # tr1::comp_ellint_1 takes the modulus k; scipy's ellipk takes m = k^2
        assert O.lib().o_ellint_k(k) == pytest.approx(special.ellipk(k * k), rel=1e-14)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_field2d_grad_live_against_reference_random_grids():
    rng = np.random.default_rng(7)
    L = O.ref()
    for nx, nz in [(5, 4), (16, 23), (64, 64)]:
        dx, dz = 1e-3, 1e-3
        g = O.grid2(nx, nz, dx, dz)
        data = rng.normal(size=nx * nz)
        xs = rng.uniform(0, (nx - 1) * dx, 500)
        zs = rng.uniform(0, (nz - 1) * dz, 500)
        gx = np.zeros(500)
        gz = np.zeros(500)
        L.ref_f2_grad_many(nx, nz, C.c_double(dx), C.c_double(dz), C.c_double(0), C.c_double(0),
                           O.dptr(data), C.c_long(500), O.dptr(xs), O.dptr(zs), O.dptr(gx), O.dptr(gz))
        for k in range(500):
            a, b = C.c_double(0), C.c_double(0)
            O.lib().o_f2_grad(C.byref(g), O.dptr(data), xs[k], zs[k], C.byref(a), C.byref(b))
            assert a.value == gx[k] and b.value == gz[k]


def _sv_interp(E, S, v, m1, m2, mul=1.0):
    """sigma(E_rel[eV]) * v with the table read by linear interpolation,
    clamped at both ends (numpy's interp), E_rel = mu v^2 / 2 in eV
    (particles.hpp:61-70,415-420)."""
    mu = m1 * m2 / (m1 + m2)
    e = 0.5 * mu * v * v / O.QE
    return mul * np.interp(e, E, S) * v


@pytest.mark.parametrize("scale", [1.0, 0.37])
def test_svmax_with_tables_brackets_the_true_maximum(scale):
    """BaseSpecies::svmax_find (particles.cpp:190-206) over two tabulated
    interactions (the e-He elastic table of species_conf_MCC.txt and a second,
    scaled copy on a shifted energy axis): the result equals the largest of
    the 1000 sampled sums (the accumulated v += dv samples, v < vmax), and it
    lies within the sampling error below the maximum of the piecewise-linear
    sigma(E) v found by brute force on a grid 1000 times finer."""
    d = np.load(G / "vecinterp_ref.npz")
    E, S = np.ascontiguousarray(d["E"]), np.ascontiguousarray(d["S"])
    E2 = np.ascontiguousarray(E * 3.0 + 0.5)
    S2 = np.ascontiguousarray(S * scale)
    m1, m2 = 9.11e-31, 6.69e-27
    its = (O.Inter * 2)(O.inter(O.ELASTIC, m1, m2, tabE=E, tabS=S), O.inter(O.ELASTIC, m1, m2, tabE=E2, tabS=S2))
    for E_max in (0.3, 5.0, 40.0):
        vmax = np.sqrt(2.0 * E_max * O.QE / m1)
        got = O.lib().o_svmax_find(its, 2, vmax, 1000)
        dv = vmax / 1000
        v = np.add.accumulate(np.concatenate([[0.0], np.full(1100, dv)]))
        v = v[v < vmax]
        sv = _sv_interp(E, S, v, m1, m2) + _sv_interp(E2, S2, v, m1, m2)
        assert got == pytest.approx(sv.max(), rel=1e-13)
        fine = np.linspace(0, vmax, 1_000_001)[:-1]
        f = _sv_interp(E, S, fine, m1, m2) + _sv_interp(E2, S2, fine, m1, m2)
        true = f.max()
        # between grid points the sum changes by at most its slope times the spacing
        slope = np.abs(np.diff(f)).max() / fine[1]
        assert got <= true + slope * fine[1]
        assert true - got <= slope * dv


def test_lifetime_with_tables_and_constant_rate():
    """lifetime = 1 / sum over targets of density * svmax (particles.cpp:142-170)
    with a tabulated and a constant-rate interaction on one target and a
    second target of another density."""
    d = np.load(G / "vecinterp_ref.npz")
    E, S = np.ascontiguousarray(d["E"]), np.ascontiguousarray(d["S"])
    m1, m2, m3 = 9.11e-31, 6.69e-27, 3.3e-26
    E_max = 8.0
    mcc = O.make_mcc(m1, E_max, [
        dict(mass=m2, density=2e20, temperature=300.0,
             inters=[O.inter(O.ELASTIC, m1, m2, tabE=E, tabS=S), O.inter(O.CX, m1, m2, rate=3e-16)]),
        dict(mass=m3, density=5e19, temperature=300.0, inters=[O.inter(O.ELASTIC, m1, m3, rate=1e-15)]),
    ])
    vmax = np.sqrt(2.0 * E_max * O.QE / m1)
    v = np.add.accumulate(np.concatenate([[0.0], np.full(1100, vmax / 1000)]))
    v = v[v < vmax]
    r1 = 2e20 * (_sv_interp(E, S, v, m1, m2) + 3e-16).max()
    r2 = 5e19 * 1e-15
    assert mcc.lifetime == pytest.approx(1.0 / (r1 + r2), rel=1e-13)
    assert mcc.targets[0].rate == pytest.approx(r1, rel=1e-13)
