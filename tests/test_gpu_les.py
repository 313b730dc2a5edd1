"""LES path on the device: dynamic Smagorinsky eddy viscosity (LES = 2)
(subgrid.f90:261-414) and the variable-viscosity RHS (equations.f90:97-132,
231-270, 367-404), compared with the FP64 oracle through the C ABI.

Tolerances: nu_t to 1e-11 of max nu_t (the model divides z averages of
products of differences, which loses a few digits to cancellation), the RHS
to a relative L2 of 1e-12 for a given nu_t, and full LES steps to 1e-10 after
ten steps.
"""
import sys

import numpy as np
import pytest

from . import cases
from .blgpu_import import solver_for
from .test_oracle_dsm import random_case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


LES_CASES = {
    "tanh_rk3": dict(nx=24, ny=16, nz=10, noise=0.05, seed=11, scheme=3, les=2),
    "sinh_rk2": dict(nx=26, ny=14, nz=9, grid="sinh", alpha=2.5, noise=0.05, seed=12, scheme=2, les=2),
    "modes_rk3": dict(nx=24, ny=20, nz=10, Ly=0.02, Lz=0.1083, modes=True, noise=0.02, seed=13, scheme=3,
                      les=2),
}


@pytest.fixture(params=sorted(LES_CASES))
def les_case(request):
    say("les case", request.param, LES_CASES[request.param])
    return cases.laminar_small(**LES_CASES[request.param])


@pytest.mark.parametrize("nut_bc", [0, 1])
def test_dsm_parity(les_case, nut_bc):
    o = les_case.oracle()
    o.set_nut_bc(nut_bc)
    s, _ = solver_for(les_case, o, nut_bc=nut_bc)
    ref = o.eddy_viscosity()
    s.eddy_viscosity()
    got = s.download(4)
    scale = np.abs(ref).max()
    assert scale > 0
    err = np.abs(got - ref).max() / scale
    say("nu_t max", scale, "err", err)
    assert err < 1e-11


def test_dsm_parity_arbitrary_ghosts():
    """Ghost planes that are not periodic copies exercise every index map."""
    x, y, z, U, V, W = random_case(3)
    rng = np.random.default_rng(9)
    U[:, :, 0] += 0.3 * rng.standard_normal(U[:, :, 0].shape)
    V[:, :, -1] += 0.3 * rng.standard_normal(V[:, :, -1].shape)
    W[:, :, 0] += 0.3 * rng.standard_normal(W[:, :, 0].shape)
    from . import orc
    from .blgpu_import import blgpu

    o = orc.Oracle(x, y, z, 1e-5, les=2)
    o.set_fields(U, V, W)
    ref = o.eddy_viscosity()
    eta, f, df = cases.load_blasius()
    tables = orc.blasius_tables(x, y, 1e-5, eta, f, df)
    s = blgpu.Solver(x, y, z, 1e-5, tables, les=2)
    s.set_fields(U, V, W)
    s.eddy_viscosity()
    got = s.download(4)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say("arbitrary ghosts err", err)
    assert err < 1e-11


def test_les_rhs_parity(les_case):
    o = les_case.oracle()
    s, _ = solver_for(les_case, o)
    nut = o.eddy_viscosity()
    s.upload(4, nut)
    ro = o.rhs()
    rg = s.rhs()
    for c in range(3):
        e = rel_l2(rg[c], ro[c])
        say("rhs", c, e)
        assert e < 1e-12, c


@pytest.mark.parametrize("nut_bc", [0, 1])
def test_les_step_parity(les_case, nut_bc):
    o = les_case.oracle()
    o.set_nut_bc(nut_bc)
    s, _ = solver_for(les_case, o, nut_bc=nut_bc)
    for n in range(1, 11):
        o.compute_dt()
        s.compute_dt()
        o.step()
        s.step()
        if n in (1, 10):
            tol = 1e-12 if n == 1 else 1e-10
            for name, a, b in zip("UVW", s.fields(), o.fields()):
                e = rel_l2(a, b) if name != "W" else np.linalg.norm(a - b) / np.linalg.norm(o.fields()[0])
                say("step", n, name, e)
                assert e < tol, (n, name, e)
            e = np.abs(s.download(4) - o.get(4)).max() / max(np.abs(o.get(4)).max(), 1e-300)
            say("step", n, "nu_t", e)
            assert e < 1e-9
    cf_o, cf_g = o.wall_stats(), s.wall_stats()
    assert rel_l2(cf_g, cf_o) < 1e-10
    mo, mg = o.monitor(), s.monitor()
    np.testing.assert_allclose(mg, mo, rtol=1e-10, atol=1e-14)


# ---- LES = 1 (constant-coefficient Smagorinsky, subgrid.f90:609-662) and
# ---- LES = 3 (dynamic Smagorinsky per cell, subgrid.f90:436-583)
OTHER_MODELS = {
    "smag_tanh": dict(nx=24, ny=16, nz=10, noise=0.05, seed=21, scheme=3, les=1, dPdx=4e-4),
    "noavg_tanh": dict(nx=24, ny=16, nz=10, noise=0.05, seed=22, scheme=3, les=3),
    "noavg_sinh_rk2": dict(nx=26, ny=14, nz=9, grid="sinh", alpha=2.5, noise=0.05, seed=23, scheme=2, les=3),
}


@pytest.mark.parametrize("name", sorted(OTHER_MODELS))
@pytest.mark.parametrize("nut_bc", [0, 1])
def test_other_models_nu_t_parity(name, nut_bc):
    say("model case", name, OTHER_MODELS[name], "nut_bc", nut_bc)
    c = cases.laminar_small(**OTHER_MODELS[name])
    o = c.oracle()
    o.set_nut_bc(nut_bc)
    s, _ = solver_for(c, o, nut_bc=nut_bc)
    # LES = 1 leaves the ghost entries alone: both start from the same values
    seed = np.random.default_rng(5).random(o.shapes[4])
    o.set(4, seed)
    s.upload(4, seed)
    ref = o.eddy_viscosity()
    s.eddy_viscosity()
    got = s.download(4)
    scale = np.abs(ref[1:-1, 1:-1, 1:-1]).max()
    assert scale > 0
    err = np.abs(got - ref).max() / scale
    say("nu_t max", scale, "err", err)
    assert err < 1e-11


@pytest.mark.parametrize("name", sorted(OTHER_MODELS))
def test_other_models_step_parity(name):
    say("model case", name, OTHER_MODELS[name])
    c = cases.laminar_small(**OTHER_MODELS[name])
    o = c.oracle()
    o.set_nut_bc(1)
    s, _ = solver_for(c, o, nut_bc=1)
    for n in range(1, 11):
        o.compute_dt()
        s.compute_dt()
        o.step()
        s.step()
        if n in (1, 10):
            tol = 1e-12 if n == 1 else 1e-10
            U = o.fields()[0]
            for fname, a, b in zip("UVW", s.fields(), o.fields()):
                e = np.linalg.norm(a - b) / np.linalg.norm(U)
                say("step", n, fname, e)
                assert e < tol, (n, fname, e)
            e = np.abs(s.download(4) - o.get(4)).max() / max(np.abs(o.get(4)).max(), 1e-300)
            say("step", n, "nu_t", e)
            assert e < 1e-9
    assert rel_l2(s.wall_stats(), o.wall_stats()) < 1e-10
