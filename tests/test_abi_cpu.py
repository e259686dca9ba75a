"""The C-ABI library without a GPU: it loads, exports every entry point the
header declares, and its host-side geometry code (t_grid) matches the oracle
bit for bit."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from mag2d_amd import abi
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent


def header_functions():
    text = (ROOT / "include" / "mag2d_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(m2d_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = abi.lib()
    names = header_functions()
    assert len(names) >= 40
    for n in names:
        assert hasattr(lib, n), n
    # the binding declares exactly the header's entry points
    assert sorted(abi.exported_symbols()) == names


def test_abi_version():
    assert abi.lib().m2d_abi_version() == abi.ABI_VERSION == 2


@pytest.mark.parametrize("geo", range(10))
def test_grid_builders_match_oracle(geo):
    n = 96
    L = 2e-2
    g = abi.grid2d(n, n, L / (n - 1), L / (n - 1))
    mask, volt = abi.grid_build(g, geo, L, L, probe_radius=3e-3, u_probe=-7.0, extern_field=250.0)
    og = O.grid2(n, n, L / (n - 1), L / (n - 1))
    om = np.zeros(n * n, np.uint8)
    ov = np.zeros(n * n)
    O.lib().o_grid_build(C.byref(og), C.byref(O.Geom(geo, L, L, 3e-3, -7.0, 250.0)), O.dptr(om), O.dptr(ov))
    np.testing.assert_array_equal(mask, om)
    np.testing.assert_array_equal(volt, ov)


def test_is_free_matches_oracle():
    n = 64
    L = 2e-2
    g = abi.grid2d(n, n, L / (n - 1), L / (n - 1))
    mask, _ = abi.grid_build(g, abi.GEO_RF_22PT, L, L)
    og = O.grid2(n, n, L / (n - 1), L / (n - 1))
    rng = np.random.default_rng(0)
    pts = np.concatenate([rng.uniform(0, L, (3000, 2)), [[L, L], [0, 0], [L, 0.5 * L], [-1e-9, 0.01]]])
    for x, z in pts:
        a = abi.lib().m2d_grid_is_free_2d(C.byref(g), mask.ctypes.data_as(C.POINTER(C.c_uint8)), x, z)
        b = O.lib().o_is_free(C.byref(og), O.dptr(mask), x, z)
        assert a == b, (x, z)


def test_errors_are_status_codes_not_crashes():
    lib = abi.lib()
    assert lib.m2d_ctx_create(0, 1, None) == -1
    assert "out is NULL" in lib.m2d_last_error().decode()
    g = abi.grid2d(2, 2, 1.0, 1.0)
    assert lib.m2d_grid_build_2d(C.byref(g), None, None, None) == -1
