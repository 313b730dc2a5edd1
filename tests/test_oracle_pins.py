"""Closed-form pins for the oracle's RHS terms and Cf.

The remaining convective cross terms (-d(vv)/dy, -d(uw)/dx, -d(ww)/dz), the
variable-viscosity (LES) viscous terms of all three momentum equations
(equations.f90:97-132, 231-270, 367-404) and the SGS term of Cf
(statistics.f90:137-160) are checked against what the continuous operators
give for fields on which the discrete operators are exact: linear or
quadratic velocity profiles with a linear eddy viscosity on uniform grids,
and linear profiles on the stretched tanh grid, where the y weights
weight_y_0/1 (initialization.f90:482-484) interpolate linear fields exactly.
A face weight applied to the wrong neighbour, an index shifted by one cell
or a dropped term changes these answers.  Second-order convergence to the
exact operator on a smooth field then covers every term of rhs_v and rhs_w
and, with a smooth eddy viscosity, the LES branch of all three.
"""
import numpy as np
import pytest

from . import cases, orc

NU = 0.013
NT0, C1, C2, C3 = 0.007, 0.02, 0.3, 0.5  # nu_t = NT0 + C1 x + C2 y + C3 z


def _grid(stretched, nx=10, ny=9, nz=8, Lx=0.15, Ly=0.015, Lz=0.006):
    x, y, z = cases.tanh_grid(nx, ny, nz, Lx, Ly, Lz)
    if not stretched:
        y = Ly * np.arange(ny) / (ny - 1)
    _, xg = orc.centers(x)
    _, yg = orc.centers(y)
    _, zg = orc.centers(z)
    return x, y, z, xg, yg, zg


def _points(comp, x, y, z, xg, yg, zg):
    """coordinates of the staggered points of u (0), v (1), w (2)"""
    axes = [(x, yg, zg), (xg, y, zg), (xg, yg, z)][comp]
    return np.meshgrid(*axes, indexing="ij")


def _viscous(x, y, z, nu, nut, fields):
    """viscous part of the LES-branch RHS: rhs(nu, nu_t) - rhs(0, 0)"""
    o1 = orc.Oracle(x, y, z, nu, 0.0, 0.0, 3, 2, -1e-4)
    o1.set_fields(*fields)
    o1.set(4, nut)
    r1 = o1.rhs()
    o0 = orc.Oracle(x, y, z, 0.0, 0.0, 0.0, 3, 2, -1e-4)
    o0.set_fields(*fields)
    o0.set(4, np.zeros_like(nut))
    r0 = o0.rhs()
    return [a - b for a, b in zip(r1, r0)]


def _nut(xg, yg, zg):
    X, Y, Z = np.meshgrid(xg, yg, zg, indexing="ij")
    return np.asfortranarray(NT0 + C1 * X + C2 * Y + C3 * Z)


INTERIOR = (slice(1, -2), slice(1, -1), slice(1, -1))


def _zero_fields(x, y, z, xg, yg, zg):
    return [np.zeros((x.size, yg.size, zg.size), order="F"), np.zeros((xg.size, y.size, zg.size), order="F"),
            np.zeros((xg.size, yg.size, z.size), order="F")]


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_les_viscous_quadratic_profile(comp, d):
    """velocity component comp = s_d^2 along direction d, nu_t linear in x,
    y and z, uniform grids: the viscous term is d/ds((nu + nu_t) dq/ds) =
    2 (nu + nu_t(P)) + 2 c_d s_d(P) at every interior point P"""
    x, y, z, xg, yg, zg = _grid(stretched=False)
    P = _points(comp, x, y, z, xg, yg, zg)
    f = _zero_fields(x, y, z, xg, yg, zg)
    f[comp] = np.asfortranarray(P[d] ** 2)
    visc = _viscous(x, y, z, NU, _nut(xg, yg, zg), f)
    nut_p = NT0 + C1 * P[0] + C2 * P[1] + C3 * P[2]
    exact = 2 * (NU + nut_p) + 2 * (C1, C2, C3)[d] * P[d]
    assert np.allclose(visc[comp][INTERIOR], exact[INTERIOR], rtol=1e-9, atol=0), (comp, d)
    for other in range(3):
        if other != comp:
            assert np.all(visc[other][INTERIOR] == 0.0)


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_les_viscous_linear_profile_stretched(comp, d):
    """velocity component comp = a s_d, nu_t linear, stretched y: the
    viscous term is a c_d (the y-interpolated face viscosities are exact for
    linear nu_t on the stretched grid)"""
    a = 1.7
    x, y, z, xg, yg, zg = _grid(stretched=True)
    P = _points(comp, x, y, z, xg, yg, zg)
    f = _zero_fields(x, y, z, xg, yg, zg)
    f[comp] = np.asfortranarray(a * P[d])
    visc = _viscous(x, y, z, NU, _nut(xg, yg, zg), f)
    exact = a * (C1, C2, C3)[d]
    assert np.allclose(visc[comp][INTERIOR], exact, rtol=1e-8, atol=0), (comp, d)


def _rhs(x, y, z, fields, nu=0.0, les=0):
    o = orc.Oracle(x, y, z, nu, 0.0, 0.0, 3, les, -1e-4)
    o.set_fields(*fields)
    return o.rhs()


def test_convective_vv_uw_ww():
    """-d(vv)/dy = -2 a^2 y for v = a y and -d(ww)/dz = -2 a^2 z for w = a z
    (uniform grids), -d(uw)/dx = -u0 a for w = a x, u = u0 (stretched y)"""
    a, u0 = 1.3, 0.8
    x, y, z, xg, yg, zg = _grid(stretched=False)
    f = _zero_fields(x, y, z, xg, yg, zg)
    Pv = _points(1, x, y, z, xg, yg, zg)
    f[1] = np.asfortranarray(a * Pv[1])
    rv = _rhs(x, y, z, f)[1]
    assert np.allclose(rv[INTERIOR], (-2 * a * a * Pv[1])[INTERIOR], rtol=1e-10, atol=1e-14)
    f = _zero_fields(x, y, z, xg, yg, zg)
    Pw = _points(2, x, y, z, xg, yg, zg)
    f[2] = np.asfortranarray(a * Pw[2])
    rw = _rhs(x, y, z, f)[2]
    assert np.allclose(rw[INTERIOR], (-2 * a * a * Pw[2])[INTERIOR], rtol=1e-10, atol=1e-14)
    x, y, z, xg, yg, zg = _grid(stretched=True)
    f = _zero_fields(x, y, z, xg, yg, zg)
    f[0][:] = u0
    Pw = _points(2, x, y, z, xg, yg, zg)
    f[2] = np.asfortranarray(a * Pw[0])
    rw = _rhs(x, y, z, f)[2]
    assert np.allclose(rw[INTERIOR], -u0 * a, rtol=1e-10, atol=1e-12)


def test_cf_with_eddy_viscosity():
    """Cf = 2 (-u_wall v_wall + (nu + nu_t at the wall face) dU/dy): U =
    u_w + b yg (so its interpolation to the wall is u_w and dU/dy = b),
    V = v_w on the wall row, nu_t = NT0 + C1 xg + C3 zg uniform in y; the
    face value averages x neighbours and the z mean runs over planes
    k = 2..nzg-1 (statistics.f90:44-52, 137-160).  At the last station the
    x neighbour is column nxg of nu_t_mean, which compute_statistics never
    fills (statistics.f90:77 runs ii = 1..nx) and so holds its initial 0
    (initialization.f90:501): half the face value there."""
    u_w, b, v_w = 0.02, 40.0, 0.003
    x, y, z, xg, yg, zg = _grid(stretched=True)
    f = _zero_fields(x, y, z, xg, yg, zg)
    f[0] = np.asfortranarray(np.broadcast_to((u_w + b * yg)[None, :, None], f[0].shape))
    f[1][:, 0, :] = v_w
    X, Y, Z = np.meshgrid(xg, yg, zg, indexing="ij")
    nut = np.asfortranarray(NT0 + C1 * X + C3 * Z)
    o = orc.Oracle(x, y, z, NU, 0.0, 0.0, 3, 2, -1e-4)
    o.set_fields(*f)
    o.set(4, nut)
    cf = o.wall_stats()
    zbar = zg[1:-1].mean()
    nut_face = NT0 + C1 * x + C3 * zbar
    nut_face[-1] = 0.5 * (NT0 + C1 * xg[-2] + C3 * zbar)
    exact = 2 * (-u_w * v_w + (NU + nut_face) * b)
    assert np.allclose(cf, exact, rtol=1e-11, atol=0)


def _manufactured(n, nut_fn=None, nu=0.05):
    x = 1.0 + np.arange(n) / (n - 1)
    y = np.arange(n) / (n - 1)
    z = np.arange(n) / (n - 2)
    _, xg = orc.centers(x)
    _, yg = orc.centers(y)
    _, zg = orc.centers(z)
    u = lambda X, Y, Z: np.sin(X) * np.cos(Y) + 0.3 * Z
    v = lambda X, Y, Z: 0.5 * np.cos(X + Y) + 0.1 * np.sin(Z)
    w = lambda X, Y, Z: 0.2 * X * Y + 0.1 * np.cos(Z)
    Pu = np.meshgrid(x, yg, zg, indexing="ij")
    Pv = np.meshgrid(xg, y, zg, indexing="ij")
    Pw = np.meshgrid(xg, yg, z, indexing="ij")
    F = [np.asfortranarray(u(*Pu)), np.asfortranarray(v(*Pv)), np.asfortranarray(w(*Pw))]
    o = orc.Oracle(x, y, z, nu, 0.0, 0.0, 3, 0 if nut_fn is None else 2, -1e-4)
    o.set_fields(*F)
    if nut_fn is not None:
        o.set(4, np.asfortranarray(nut_fn(*np.meshgrid(xg, yg, zg, indexing="ij"))))
    return o.rhs(), (Pu, Pv, Pw)


def _exact(P, comp, nu, nut=None):
    """-div(u u_comp) + div((nu + nu_t) grad u_comp) of the manufactured field
    at the points P of component comp"""
    X, Y, Z = P
    u = np.sin(X) * np.cos(Y) + 0.3 * Z
    ux, uy, uz = np.cos(X) * np.cos(Y), -np.sin(X) * np.sin(Y), 0.3
    v = 0.5 * np.cos(X + Y) + 0.1 * np.sin(Z)
    vx, vy, vz = -0.5 * np.sin(X + Y), -0.5 * np.sin(X + Y), 0.1 * np.cos(Z)
    w = 0.2 * X * Y + 0.1 * np.cos(Z)
    wx, wy, wz = 0.2 * Y, 0.2 * X, -0.1 * np.sin(Z)
    q, qx, qy, qz = [(u, ux, uy, uz), (v, vx, vy, vz), (w, wx, wy, wz)][comp]
    lap = [-2 * np.sin(X) * np.cos(Y), -np.cos(X + Y) - 0.1 * np.sin(Z), -0.1 * np.cos(Z)][comp]
    conv = -((ux * q + u * qx) + (vy * q + v * qy) + (wz * q + w * qz))
    if nut is None:
        return conv + nu * lap
    n, nx_, ny_, nz_ = nut(X, Y, Z)
    return conv + (nu + n) * lap + nx_ * qx + ny_ * qy + nz_ * qz


@pytest.mark.parametrize("les", [False, True])
def test_rhs_uvw_second_order(les):
    """every term of rhs_u, rhs_v, rhs_w (DNS, and LES with a smooth nu_t)
    converges at second order to the continuous operator"""
    nu = 0.05
    nut_fn = (lambda X, Y, Z: 0.02 * (1.0 + 0.5 * np.sin(2 * X) * np.cos(Z) + 0.3 * Y)) if les else None
    nut_d = (lambda X, Y, Z: (0.02 * (1.0 + 0.5 * np.sin(2 * X) * np.cos(Z) + 0.3 * Y),
                              0.02 * np.cos(2 * X) * np.cos(Z), 0.006 + 0 * X,
                              -0.01 * np.sin(2 * X) * np.sin(Z))) if les else None
    errs = np.zeros((3, 3))
    for a, n in enumerate((16, 32, 64)):
        r, P = _manufactured(n, nut_fn, nu)
        for comp in range(3):
            ex = _exact(P[comp], comp, nu, nut_d)
            sl = (slice(2, -3), slice(2, -3), slice(2, -3))
            errs[comp, a] = np.abs(r[comp][sl] - ex[sl]).max()
    rates = np.log2(errs[:, :-1] / errs[:, 1:])
    assert np.all(rates > 1.7), (errs, rates)
