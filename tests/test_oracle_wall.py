"""Pins of the oracle's wall models (compute_wall_model wallmodel.f90:25-144)
and of the bottom-wall conditions that use them (boundary_conditions.f90:
73-108), by what each model and condition is built to satisfy:

- Robin walls (every model but 0, 11, 15): the ghost row makes the wall value
  (U1 + U2)/2 equal alpha times the wall gradient (U2 - U1)/(yg2 - yg1), for U
  with alpha_x and W with alpha_z;
- pseudo-Robin V: V at the wall is alpha_y dVo/dy of the step-start field
  less its wall mean, so the wall flux sums to 0 over the interior wall, and
  the global mass balance (inflow - outflow + wall - top) stays 0;
- model 5: alpha_mean (1 + alpha_std sin(omega t)) at a t where the sine is 1;
- models 6, 8 (log layer, utau = dPdx^0.5): alpha is the slip length that
  gives the log-law velocity at y(2), U+ = ln(Delta+)/kappa + B, for the wall
  gradient utau^2/(nu + nu_t) (6), and kappa Delta U+ (8);
- istress_model 1: utau_model solves the log law at yg(3) for the z-mean U of
  row 3: a field built from chosen utau(x) returns them;
- models 11 and 15 (Neumann U): the wall gradient of U is utau^2/nu (11, with
  nu_t = 0) and the Cf correlation's tau_w/nu (15);
- model 4 (channel momentum balance): alpha^2 <dU/dy dV/dy> = nu <dU/dy> -
  dPdx Ly/2 over both walls (nu_t = 0);
- models 9, 12 (Cf-fitted slip): alpha_x is the slip length whose viscous
  stress nu <U>_wall/alpha_x is frac_vis tau_w of White's correlation with
  U_inf = U(2, nyg, 2), and with the dynamic fraction (WMnutflag 2) it is
  y(3) <U>_wall/<U>_rows3:4, in which nu and tau_w cancel; alpha_y is the
  two-column mean of the slip length that closes the wall stress balance
  nu <dU/dy> - tau_w - <tau_uv> (+ beta_y <U>_wall for 9) = alpha_x alpha_y
  <dU/dy dV/dy> (12) or alpha_y <U dVo/dy> (9), 12 clipped to 2e-3;
- a step with a Robin wall keeps the divergence at round-off.
"""
import math

import numpy as np
import pytest

from . import cases, orc

KAPPA, B = 0.38, 5.2


def _case(**kw):
    args = dict(nx=26, ny=16, nz=8, noise=0.02, seed=4, cfl=-1e-4)
    args.update(kw)
    return cases.laminar_small(**args)


def _geo(c):
    xm, xg = orc.centers(c.x)
    ym, yg = orc.centers(c.y)
    return xg, yg


def test_robin_walls_and_pseudo_robin_v():
    c = _case()
    o = c.oracle()
    U0, V0, W0 = o.fields()
    Vo = V0 * 0.97 + 0.001 * np.sin(np.arange(V0.size).reshape(V0.shape, order="F"))
    o.set_old(U0, Vo, W0)
    ax, ay, az = 3e-4, 2e-4, 5e-4
    o.set_wall_model(5, alpha_mean_x=ax, alpha_mean_y=ay, alpha_mean_z=az)
    o.wall_model()
    o.apply_bc(1, 1e-4)
    U, V, W = o.fields()
    _, yg = _geo(c)
    dyg = yg[1] - yg[0]
    # U without the outflow columns nx-1, nx, which the mass correction
    # shifts after the wall condition (boundary_conditions.f90:720-725)
    for F, a in ((U[:-2], ax), (W, az)):
        wall = 0.5 * (F[:, 0, :] + F[:, 1, :])
        grad = (F[:, 1, :] - F[:, 0, :]) / dyg
        assert np.abs(wall - a * grad).max() < 1e-12 * np.abs(F[:, 1, :]).max()
    # V: alpha_y dVo/dy less one constant, whose removal zeroes the wall mean
    raw = ay * (Vo[:, 1, :] - Vo[:, 0, :]) / (c.y[1] - c.y[0])
    off = V[:, 0, :] - raw
    assert np.ptp(off) < 1e-15 * max(np.abs(raw).max(), 1e-300) + 1e-18
    inner = V[1:-2, 0, 1:-2]
    assert abs(inner.mean()) < 1e-14 * np.abs(raw).max()
    assert abs(off.mean()) > 0  # the removed mean is not trivially 0
    # global mass balance after the outflow correction
    x, y = c.x, c.y
    nzg = U.shape[2]
    qx = sum(((U[0, j, 1:nzg - 2] - U[-2, j, 1:nzg - 2]) * (y[j] - y[j - 1])).sum() for j in range(1, U.shape[1] - 1))
    qy = sum(((V[i, 0, 1:nzg - 2] - V[i, -1, 1:nzg - 2]) * (x[i] - x[i - 1])).sum() for i in range(1, V.shape[0] - 2))
    assert abs(qx + qy) < 1e-15
    o.close()


def test_oscillating_alpha_model5():
    dPdx = 4e-4
    c = _case(dPdx=dPdx)
    o = c.oracle()
    f = 0.3
    omega = 2 * math.pi * f / (c.y[1] - c.y[0]) * math.sqrt(dPdx)
    o.t = 0.5 * math.pi / omega  # sin(omega t) = 1
    o.set_wall_model(5, alpha_mean_x=2e-4, alpha_mean_y=1e-4, alpha_mean_z=3e-4, alpha_std=0.25, freq_mult=f)
    o.wall_model()
    ax, ay, az, _ = o.get_wall_model()
    for a, m in ((ax, 2e-4), (ay, 1e-4), (az, 3e-4)):
        assert np.allclose(a, 1.25 * m, rtol=1e-14, atol=0)
    o.close()


@pytest.mark.parametrize("les,nut_bc", [(0, 1), (1, 1), (1, 0)])
def test_log_layer_models(les, nut_bc):
    dPdx = 2.5e-3
    c = _case(dPdx=dPdx, les=les)
    o = c.oracle()
    o.set_nut_bc(nut_bc)
    utau = math.sqrt(dPdx)
    delta = c.y[1] - c.y[0]
    uplus = math.log(delta * utau / c.nu) / KAPPA + B
    o.eddy_viscosity()
    # the wall value of the wall-parallel mean eddy viscosity (LES 1:
    # interior mean of row 2, clipped, its ghost antisymmetric under nutBC 1)
    nt = o.get(4) if les else None
    if les:
        m = max(nt[1:-1, 1, 1:-1].mean(), 0.0)
        g = -m if nut_bc == 1 else m
        _, yg = _geo(c)
        w0 = (yg[1] - c.y[0]) / (yg[1] - yg[0])
        nut_wall = w0 * g + (1 - w0) * m
    else:
        nut_wall = 0.0
    o.set_wall_model(6)
    o.wall_model()
    a6 = o.get_wall_model()[0]
    # slip length = U(y2)/(dU/dy at the wall) with dU/dy = utau^2/(nu + nu_t)
    want6 = utau * uplus / (utau ** 2 / (c.nu + nut_wall))
    assert np.allclose(a6, want6, rtol=1e-13)
    o.set_wall_model(8)
    o.wall_model()
    assert np.allclose(o.get_wall_model()[0], KAPPA * delta * uplus, rtol=1e-13)
    o.close()


def _law_field(c, o, utau_x):
    """U row 3 (1-based) at the log-law velocity of utau(x) at yg(3)."""
    U, V, W = o.fields()
    _, yg = _geo(c)
    u3 = utau_x / KAPPA * np.log(yg[2] * utau_x / c.nu) + B * utau_x
    U[:, 2, :] = u3[:, None]
    o.set_fields(U, V, W)
    return U


def test_law_of_the_wall_newton():
    c = _case()
    o = c.oracle()
    nx = c.x.size
    utau_x = 0.04 + 0.01 * np.sin(np.linspace(0, 3, nx))
    _law_field(c, o, utau_x)
    o.set_wall_model(0, istress_model=1)
    o.wall_model()
    ut = o.get_wall_model()[3]
    assert np.allclose(ut[1:-1], utau_x[1:-1], rtol=1e-9)
    assert ut[0] == ut[1] and ut[-1] == ut[-2]
    o.close()


def test_neumann_models_11_and_15():
    c = _case()
    o = c.oracle()
    nx = c.x.size
    utau_x = 0.05 + 0.005 * np.cos(np.linspace(0, 2, nx))
    _law_field(c, o, utau_x)
    _, yg = _geo(c)
    dyg = yg[1] - yg[0]
    o.set_wall_model(11, istress_model=1)
    o.wall_model()
    o.apply_bc(1, 1e-4)
    U, V, W = o.fields()
    grad = (U[:-2, 1, :] - U[:-2, 0, :]) / dyg
    ut = o.get_wall_model()[3]
    assert np.allclose(ut[1:-1], utau_x[1:-1], rtol=1e-9)
    want = np.maximum(ut ** 2 / c.nu, 0.0)
    assert np.allclose(grad, want[:-2, None], rtol=1e-10)
    assert np.all(V[:, 0, :] == 0.0) and np.allclose(W[:, 0, :], -W[:, 1, :], rtol=0, atol=0)
    o.set_wall_model(15)
    o.wall_model()
    o.apply_bc(1, 1e-4)
    U = o.fields()[0]
    tau_w = 0.5 * 0.027 * (c.x / c.nu) ** (-1.0 / 7.0)  # White's Cf = 0.027 Re_x^(-1/7), U_inf = 1
    grad = (U[:-2, 1, :] - U[:-2, 0, :]) / dyg
    assert np.allclose(grad, (tau_w / c.nu)[:-2, None], rtol=1e-10)
    o.close()


def test_momentum_balance_model4():
    dPdx = 0.1  # large enough that the balance, and alpha^2, are positive
    c = _case(dPdx=dPdx, noise=0.05)
    o = c.oracle()
    o.set_wall_model(4)
    o.wall_model()
    a = o.get_wall_model()[0]
    U, V, W = o.fields()
    x, y = c.x, c.y
    _, yg = _geo(c)
    ny = y.size
    s = np.s_[1:-1]
    Vx = 0.5 * (V[:-1] + V[1:])  # V on the cell edges x(i)
    dub = (U[s, 1, s] - U[s, 0, s]) / (yg[1] - yg[0])
    dut = (U[s, -1, s] - U[s, -2, s]) / (yg[-1] - yg[-2])
    dvb = (Vx[s, 1, s] - Vx[s, 0, s]) / (y[1] - y[0])
    dvt = (Vx[s, ny - 1, s] - Vx[s, ny - 2, s]) / (y[-1] - y[-2])
    duv = 0.5 * ((dub * dvb).mean() - (dut * dvt).mean())
    du = 0.5 * (dub.mean() - dut.mean())
    rhs = c.nu * du - dPdx * (y[-1] - y[0]) / 2
    assert rhs / duv > 0
    assert np.allclose(a[0] ** 2 * duv, rhs, rtol=1e-12)
    assert np.all(a == a[0])
    o.close()


def test_step_with_robin_wall_keeps_divergence():
    c = _case()
    o = c.oracle()
    o.set_wall_model(5, alpha_mean_x=2e-4, alpha_mean_y=1e-4, alpha_mean_z=2e-4)
    for _ in range(3):
        o.compute_dt()
        o.step()
    V = o.fields()[1]
    assert np.abs(V[:, 0, :]).max() > 0
    assert o.max_divergence() < 1e-11
    o.close()


def _cf_field(c, o):
    """U's first row made positive and its wall ghost row 0.3 of it, so the
    wall mean of U is positive; U_inf = U(2, nyg, 2) set to 1.02"""
    U, V, W = o.fields()
    U[:, 1, :] = np.abs(U[:, 1, :]) + 0.002
    U[:, 0, :] = 0.3 * U[:, 1, :]
    U[1, -1, 1] = 1.02
    o.set_fields(U, V, W)
    return U, V, W


def _white_tau(c, uinf):
    return 0.5 * 0.027 * (uinf * c.x / c.nu) ** (-1.0 / 7.0) * uinf ** 2


@pytest.mark.parametrize("model", [9, 12])
def test_cf_fitted_slip_models(model):
    c = _case(noise=0.05)
    o = c.oracle()
    U, V, W = _cf_field(c, o)
    Vo = 0.9 * V + 1e-4
    o.set_old(U, Vo, W)
    _, yg = _geo(c)
    y = c.y
    nzg = U.shape[2]
    ks = slice(1, nzg - 1)
    um = 0.5 * (U[:, 0, ks] + U[:, 1, ks]).mean(axis=1)
    tw = _white_tau(c, 1.02)
    frac = 0.05
    o.set_wall_model(model, iwall_model_nut=0, frac_vis=frac)
    beta = 0.0
    if model == 9:
        # model 9 reads beta_y of the last pseudo-Robin wall condition: one
        # model + boundary call sets it, then the field is put back
        o.wall_model()
        o.apply_bc(1, 1e-4)
        o.set_fields(U, V, W)
        beta = o.wall_beta()
        assert beta != 0.0
    o.wall_model()
    ax, ay, az, _ = o.get_wall_model()
    # viscous stress of the slip velocity = frac tau_w
    assert np.all(um > 0)
    assert np.allclose(c.nu * um / ax, frac * tw, rtol=1e-13)
    # the stress balance per column, averaged over the two columns of a face
    g = ((U[:, 1, ks] - U[:, 0, ks]) / (yg[1] - yg[0])).mean(axis=1)
    VV = V if model == 12 else Vo
    vx = 0.5 * (VV[:-1] + VV[1:])
    dv = (vx[:, 1, ks] - vx[:, 0, ks]) / (y[1] - y[0])
    if model == 12:
        prod = (((U[:, 1, ks] - U[:, 0, ks]) / (yg[1] - yg[0])) * dv).mean(axis=1)
        h = (c.nu * g - tw) / (ax * prod)
    else:
        prod = (0.5 * (U[:, 1, ks] + U[:, 0, ks]) * dv).mean(axis=1)
        h = (c.nu * g - tw + um * beta) / prod
    want = np.empty(ay.size)
    want[1:-1] = 0.5 * (h[:-1] + h[1:])
    want[0], want[-1] = want[1], want[-2]
    if model == 12:
        want = np.where(np.abs(want) > 2e-3, 2e-3 * np.sign(want), want)
    assert np.allclose(ay, want, rtol=1e-11, atol=1e-18)
    assert np.array_equal(ay, az)
    o.close()


def test_cf_model12_dynamic_fraction_and_clip():
    c = _case(noise=0.05)
    o = c.oracle()
    U, V, W = _cf_field(c, o)
    o.set_wall_model(12, iwall_model_nut=2)
    o.wall_model()
    ax, ay, _, _ = o.get_wall_model()
    nzg = U.shape[2]
    ks = slice(1, nzg - 1)
    um = 0.5 * (U[:, 0, ks] + U[:, 1, ks]).mean(axis=1)
    ur = 0.5 * (U[:, 2, ks] + U[:, 3, ks]).mean(axis=1)
    assert np.allclose(ax, c.y[2] * um / ur, rtol=1e-13)
    assert np.abs(ay).max() <= 2e-3
    assert np.any(np.abs(ay) == 2e-3)  # the clip is reached in this case
    o.close()


def test_unsupported_models_rejected():
    c = _case()
    o = c.oracle()
    for m in (1, 2, 3, 7, 13, 14, 16):
        with pytest.raises(ValueError):
            o.set_wall_model(m)
    o.close()
