"""C-ABI library checks that need no GPU: it loads, exports every entry point
declared in include/blgpu.h, and its host-side set-up helpers agree with the
oracle's restatement of the reference initialisation."""
import ctypes
import os
import re

import numpy as np
import pytest

from . import cases, orc
from .blgpu_import import ROOT, blgpu


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "blgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bl_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = blgpu.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(blgpu.SIGNATURES), set(syms) ^ set(blgpu.SIGNATURES)


def test_config_struct_matches_header():
    txt = open(os.path.join(ROOT, "include", "blgpu.h")).read()
    body = txt[txt.index("typedef struct bl_config {"):txt.index("} bl_config;")]
    names = re.findall(r"\*?\s*([A-Za-z_][A-Za-z_0-9]*)\s*[,;]", body)
    assert [f[0] for f in blgpu.BLConfig._fields_] == names


def test_blasius_profile_reproduces_reference_fixture():
    eta0, f0, df0 = cases.load_blasius()
    eta, f, df = blgpu.blasius_profile(n=1000, eta_max=20.0, fpp0=0.46960, sub=64)
    assert np.allclose(eta, eta0, rtol=0, atol=1e-13)
    assert np.abs(df - df0).max() < 1e-9
    assert np.abs(f - f0).max() < 1e-8


@pytest.mark.parametrize("grid", ["tanh", "sinh"])
def test_blasius_tables_and_field_match_oracle(grid):
    c = cases.laminar_small(nx=30, ny=20, nz=8, grid=grid)
    t = blgpu.blasius_tables(c.x, c.y, c.nu, c.eta, c.f, c.df)
    for k, v in t.items():
        assert np.array_equal(v, c.tables[k]), k
    U, V, W = blgpu.blasius_field(c.x, c.y, c.z, c.nu, c.eta, c.f, c.df)
    if grid == "sinh":
        x, y, z, U0, V0, W0 = orc.init_flow(30, 20, 8, 0.15, 0.015, 0.006, 2.0, c.nu, c.eta, c.f, c.df)
    else:
        U0, V0, W0 = c.U0, c.V0, c.W0
    assert np.allclose(U, U0, rtol=0, atol=1e-14)
    assert np.allclose(V, V0, rtol=0, atol=1e-16)
    assert np.all(W == 0)


def test_noise_is_deterministic_and_interior_only():
    shp = [(10, 9, 7), (11, 8, 7), (11, 9, 6)]
    a = [np.zeros(s, order="F") for s in shp]
    b = [np.zeros(s, order="F") for s in shp]
    blgpu.add_noise(*a, 0.01, 0)
    blgpu.add_noise(*b, 0.01, 0)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
        assert np.all(x[0] == 0) and np.all(x[-1] == 0) and np.all(x[:, 0] == 0) and np.all(x[:, :, -1] == 0)
        inner = x[1:-1, 1:-1, 1:-1]
        assert np.abs(inner).max() <= 0.01 and abs(inner.mean()) < 0.002 and inner.std() > 0.004
    c = [np.zeros(s, order="F") for s in shp]
    blgpu.add_noise(*c, 0.01, 1)
    assert not np.array_equal(a[0], c[0])


def test_tanh_grid():
    x, y, z = blgpu.grid_tanh(17, 12, 9, 0.6, 0.06, 0.05, 2.0, 1.0)
    x0, y0, z0 = cases.tanh_grid(17, 12, 9, 0.6, 0.06, 0.05, 2.0)
    assert np.allclose(x, x0, rtol=1e-15) and np.allclose(y, y0, rtol=1e-14, atol=1e-18)
    assert np.allclose(z, z0, rtol=1e-15) and y[0] == 0.0 and abs(y[-1] - 0.06) < 1e-16


def test_setup_helpers_reject_bad_arguments():
    L = blgpu.lib()
    assert L.bl_grid_tanh(1, 1, 1, 1.0, 1.0, 1.0, 2.0, 1.0, None, None, None) != 0
    assert L.bl_blasius_profile(1, 20.0, 0.4696, 1, None, None, None) != 0
