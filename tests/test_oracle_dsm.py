"""Pins for the oracle's dynamic Smagorinsky model (oracle/orc_dsm.c,
subgrid.f90:261-414).

- closed form: on a uniform grid the plane strain u = a x, v = -a y gives, away
  from the x and y ends, L_11 = a^2 hx^2/3, L_22 = a^2 hy^2/3 (the second
  moment of the Simpson test filter), M_ij = (fil_size^2 - 1)|S| S_ij with
  fil_size = 2, hence nu_t = a (hy^2 - hx^2)/36 exactly;
- isotropic spacing gives nu_t = 0, and compression (a < 0) is clipped to 0;
- the index-map numpy statement (tests/dsm_ref.py) agrees with the
  loop-by-loop C restatement on random fields, for both nutBC choices;
- nu_t >= 0 on the interior rows, ghost rows follow nutBC, and the x-ghost
  columns and z-ghost planes stay 0 as on one rank.
"""
import numpy as np
import pytest

from tests import orc
from tests.dsm_ref import dsm as dsm_numpy


def uniform_grid(nx, ny, nz, hx, hy, hz):
    return 1.0 + hx * np.arange(nx), hy * np.arange(ny), hz * np.arange(nz)


def oracle_dsm(x, y, z, U, V, W, nut_bc=0):
    o = orc.Oracle(x, y, z, 1e-5, les=2)
    o.set_nut_bc(nut_bc)
    o.set_fields(U, V, W)
    nut = o.eddy_viscosity()
    o.close()
    return nut


def plane_strain(x, y, z, a):
    nx, ny, nz = x.size, y.size, z.size
    U = np.zeros((nx, ny + 1, nz + 1), order="F")
    V = np.zeros((nx + 1, ny, nz + 1), order="F")
    W = np.zeros((nx + 1, ny + 1, nz), order="F")
    U[:] = a * x[:, None, None]
    V[:] = -a * y[None, :, None]
    return U, V, W


@pytest.mark.parametrize("hx,hy,a", [(0.01, 0.02, 1.0), (0.02, 0.05, 0.5), (0.03, 0.01, -2.0)])
def test_dsm_plane_strain_closed_form(hx, hy, a):
    x, y, z = uniform_grid(14, 12, 7, hx, hy, 0.01)
    nut = oracle_dsm(x, y, z, *plane_strain(x, y, z, a))
    nxg, nyg = x.size + 1, y.size + 1
    core = nut[2:nxg - 2, 2:nyg - 2, 1:-1]  # 0-based cells whose stencils miss the x and y ends
    expect = max(a * (hy ** 2 - hx ** 2) / 36.0, 0.0)
    np.testing.assert_allclose(core, expect, rtol=1e-9, atol=1e-12 * abs(a) * hy ** 2)


def test_dsm_isotropic_plane_strain_is_zero():
    x, y, z = uniform_grid(12, 10, 7, 0.02, 0.02, 0.02)
    nut = oracle_dsm(x, y, z, *plane_strain(x, y, z, 1.5))
    assert np.abs(nut[2:-2, 2:-2, 1:-1]).max() < 1e-15


def random_case(seed, nx=11, ny=9, nz=8):
    rng = np.random.default_rng(seed)
    x = 1.0 + np.cumsum(np.r_[0, 0.01 + 0.005 * rng.random(nx - 1)])
    y = np.cumsum(np.r_[0, 0.002 + 0.004 * rng.random(ny - 1)])
    z = 0.013 * np.arange(nz)
    U = np.asfortranarray(1.0 + 0.2 * rng.standard_normal((nx, ny + 1, nz + 1)))
    V = np.asfortranarray(0.1 * rng.standard_normal((nx + 1, ny, nz + 1)))
    W = np.asfortranarray(0.1 * rng.standard_normal((nx + 1, ny + 1, nz)))
    # periodic ghosts in z as after the boundary conditions
    for A in (U, V):
        A[:, :, 0] = A[:, :, -3]
        A[:, :, -2] = A[:, :, 1]
        A[:, :, -1] = A[:, :, 2]
    W[:, :, 0] = W[:, :, -2]
    W[:, :, -1] = W[:, :, 1]
    return x, y, z, U, V, W


@pytest.mark.parametrize("seed,nut_bc", [(0, 0), (1, 1), (2, 0)])
def test_dsm_matches_index_map_statement(seed, nut_bc):
    x, y, z, U, V, W = random_case(seed)
    a = oracle_dsm(x, y, z, U, V, W, nut_bc)
    b = dsm_numpy(x, y, z, U, V, W, nut_bc)
    assert np.abs(b).max() > 0
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())


def test_dsm_non_periodic_input_matches_index_map_statement():
    """Ghost planes that are not periodic copies: the maps must reproduce
    which array each ghost value is read from."""
    x, y, z, U, V, W = random_case(3)
    rng = np.random.default_rng(9)
    U[:, :, 0] += 0.3 * rng.standard_normal(U[:, :, 0].shape)
    V[:, :, -1] += 0.3 * rng.standard_normal(V[:, :, -1].shape)
    W[:, :, 0] += 0.3 * rng.standard_normal(W[:, :, 0].shape)
    a = oracle_dsm(x, y, z, U, V, W, 0)
    b = dsm_numpy(x, y, z, U, V, W, 0)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())


@pytest.mark.parametrize("nut_bc", [0, 1])
def test_dsm_sign_and_ghosts(nut_bc):
    x, y, z, U, V, W = random_case(4)
    nut = oracle_dsm(x, y, z, U, V, W, nut_bc)
    assert (nut[:, 1:-1] >= 0).all()
    assert (nut[1:-1, 1:-1, 1:-1] > 0).any()
    sg = -1.0 if nut_bc == 1 else 1.0
    np.testing.assert_array_equal(nut[:, 0], sg * nut[:, 1])
    np.testing.assert_array_equal(nut[:, -1], sg * nut[:, -2])
    # one rank: the x-ghost columns and z-ghost planes are never filled
    assert not nut[0].any() and not nut[-1].any()
    assert not nut[:, :, 0].any() and not nut[:, :, -1].any()
