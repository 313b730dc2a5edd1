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

The per-cell model (LES = 3, subgrid.f90:436-583) has the same closed form
away from the x ends, agrees with the index-map statement with its periodic
x maps, and its x and z ghosts are periodic copies.  The constant-coefficient
model (LES = 1, :609-662) gives (Cs Delta_j f_j)^2 |b| for the uniform shear
u = b y, with Delta_j^3 = dx dz dy_j and the van Driest factor f of the
distance to the nearer wall in units nu / sqrt(dPdx), and leaves the ghost
entries of nu_t as they were.
"""
import numpy as np
import pytest

from tests import orc
from tests.dsm_ref import dsm as dsm_numpy


def uniform_grid(nx, ny, nz, hx, hy, hz):
    return 1.0 + hx * np.arange(nx), hy * np.arange(ny), hz * np.arange(nz)


def oracle_dsm(x, y, z, U, V, W, nut_bc=0, les=2):
    o = orc.Oracle(x, y, z, 1e-5, les=les)
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


# ---------------------------------------------------------------- LES = 3 ----
@pytest.mark.parametrize("hx,hy,a", [(0.01, 0.02, 1.0), (0.03, 0.01, -2.0)])
def test_dsm_noavg_plane_strain_closed_form(hx, hy, a):
    x, y, z = uniform_grid(14, 12, 7, hx, hy, 0.01)
    nut = oracle_dsm(x, y, z, *plane_strain(x, y, z, a), les=3)
    nxg, nyg = x.size + 1, y.size + 1
    core = nut[2:nxg - 3, 2:nyg - 2, 1:-1]
    expect = max(a * (hy ** 2 - hx ** 2) / 36.0, 0.0)
    np.testing.assert_allclose(core, expect, rtol=1e-9, atol=1e-12 * abs(a) * hy ** 2)


@pytest.mark.parametrize("seed,nut_bc", [(0, 0), (1, 1), (5, 0)])
def test_dsm_noavg_matches_index_map_statement(seed, nut_bc):
    x, y, z, U, V, W = random_case(seed)
    a = oracle_dsm(x, y, z, U, V, W, nut_bc, les=3)
    b = dsm_numpy(x, y, z, U, V, W, nut_bc, noavg=True)
    assert np.abs(b).max() > 0
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())


def test_dsm_noavg_periodic_ghosts():
    x, y, z, U, V, W = random_case(6)
    nut = oracle_dsm(x, y, z, U, V, W, 0, les=3)
    n1, n3 = nut.shape[0], nut.shape[2]
    np.testing.assert_array_equal(nut[0], nut[n1 - 3])
    np.testing.assert_array_equal(nut[n1 - 2], nut[1])
    np.testing.assert_array_equal(nut[n1 - 1], nut[2])
    np.testing.assert_array_equal(nut[:, :, 0], nut[:, :, n3 - 3])
    np.testing.assert_array_equal(nut[:, :, n3 - 1], nut[:, :, 2])
    assert (nut >= 0).all() and (nut[1:-1, 1:-1, 1:-1] > 0).any()


# ---------------------------------------------------------------- LES = 1 ----
@pytest.mark.parametrize("nut_bc", [0, 1])
def test_smagorinsky_uniform_shear(nut_bc):
    nu, dPdx, b = 1e-5, 4e-4, 3.0
    rng = np.random.default_rng(2)
    x = 1.0 + 0.02 * np.arange(12)
    y = np.cumsum(np.r_[0, 0.002 + 0.004 * rng.random(11)])
    z = 0.013 * np.arange(8)
    _, yg = orc.centers(y)
    U = np.zeros((x.size, y.size + 1, z.size + 1), order="F")
    U[:] = b * yg[None, :, None]
    V = np.zeros((x.size + 1, y.size, z.size + 1), order="F")
    W = np.zeros((x.size + 1, y.size + 1, z.size), order="F")
    o = orc.Oracle(x, y, z, nu, dPdx, 0.0, 3, 1)
    o.set_nut_bc(nut_bc)
    o.set_fields(U, V, W)
    sentinel = np.full(o.shapes[4], 7.5, order="F")
    o.set(4, sentinel)
    nut = o.eddy_viscosity()
    _, xg = orc.centers(x)
    _, zg = orc.centers(z)
    dx, dz = np.diff(xg).min(), np.diff(zg).min()
    nyg = y.size + 1
    for j in range(1, nyg - 1):  # 0-based interior rows
        delta = (dx * dz * (y[j] - y[j - 1])) ** (1.0 / 3.0)
        # distance to the bottom wall for the lower rows (j1 < nyg/2 in 1-based
        # rows), to the top wall otherwise, in viscous units nu / utau
        dist = yg[j] - y[0] if (j + 1) < nyg // 2 else y[-1] - yg[j]
        f = 1.0 - np.exp(-dist / 25.0 / (nu / np.sqrt(dPdx))) if nut_bc == 1 else 1.0
        np.testing.assert_allclose(nut[1:-1, j, 1:-1], (0.11 * delta * f) ** 2 * abs(b), rtol=1e-12)
    # ghost entries keep what they held
    for g in (nut[0], nut[-1], nut[:, 0], nut[:, -1], nut[:, :, 0], nut[:, :, -1]):
        assert np.all(g == 7.5)


def test_unknown_les_model_is_rejected():
    x, y, z = uniform_grid(8, 6, 6, 0.01, 0.01, 0.01)
    with pytest.raises(ValueError):
        orc.Oracle(x, y, z, 1e-5, les=4)

