"""Pins of the oracle's general-dispatcher boundary flags
(boundary_conditions.f90:22-121) by what the formulas fix, not by their
retyping:

- blowing/suction (top_flag 1, Coleman 2018, :657-661; top_flag 2, Abe 2017,
  :663-666): the profile's extremum sits at (x_bs - x)/sigma = 1/sqrt(2),
  where it is Vmax + phi (Coleman) or Vmax exp(phi - 1/2) (Abe); a grid
  point is put there;
- zero_wz_top (:264-283): the discrete spanwise vorticity dV/dx - dU/dy at
  the top vanishes column by column (a wrong sign, index or spacing fails),
  dW/dy = 0 there, columns 1..i_rescale keep U_top, W_top;
- Falkner-Skan (:979-1007): log U(x) is linear in log(x + 1) with slope
  beta/(2 - beta), U matched to 1 at x(1), V = 0 at the top;
- global mass conservation (:693-727) with the new top V: the net flux
  through the x and y boundaries after the correction is zero;
- random inflow (inflow_flag 4, :412-451): every perturbation lies within
  +-Amp (1 - U_inlet)/2, their normalised values pass a Kolmogorov-Smirnov
  test for uniformity on [0, 1), they are drawn once per step (a second
  boundary call of the same step leaves them) and change from step to step.
"""
import numpy as np
import pytest

from . import cases, orc


def _case(**kw):
    c = cases.laminar_small(nx=30, ny=14, nz=8, noise=0.02, seed=4, **kw)
    o = c.oracle()
    return c, o


def _xg(c):
    return orc.centers(c.x)[1]


def _yg(c):
    return orc.centers(c.y)[1]


def _net_flux(c, o):
    """Qx + Qy of apply_global_mass_conservation over the same ranges."""
    U, V, _ = o.fields()
    nx, nyg, nzg = U.shape
    nxg = nx + 1
    x, y = c.x, c.y
    qx = sum(((U[0, j, 1:nzg - 2] - U[nx - 2, j, 1:nzg - 2]) * (y[j] - y[j - 1])).sum() for j in range(1, nyg - 1))
    qy = sum(((V[i, 0, 1:nzg - 2] - V[i, -1, 1:nzg - 2]) * (x[i] - x[i - 1])).sum() for i in range(1, nxg - 2))
    return qx + qy, abs(U[0]).max() * (y[-1] - y[0]) * nzg


@pytest.mark.parametrize("flag", [1, 2])
def test_blowing_suction_extremum(flag):
    c, o = _case()
    xg = _xg(c)
    i0 = 17
    sigma, vmax, phi = 0.01, 0.3, 0.004
    x_bs = xg[i0] + sigma / np.sqrt(2.0)
    o.set_boundary(top_flag=flag, vbs_max=vmax, x_bs=x_bs, sigma_bs=sigma, phi_bs=phi, i_rescale=0)
    o.apply_bc(1, 1e-4)
    V = o.fields()[1]
    top = V[:, -1, 3]
    want = vmax + phi if flag == 1 else vmax * np.exp(phi - 0.5)
    assert abs(top[i0] - want) < 1e-14
    # it is the extremum over the columns (the profile is unimodal on each side)
    assert np.argmax(top) == i0
    assert np.all(V[:, -1, :] == top[:, None])
    o.close()


def test_zero_vorticity_top():
    c, o = _case()
    xg, yg = _xg(c), _yg(c)
    irs = 4
    o.set_boundary(top_flag=1, vbs_max=0.2, x_bs=xg[12], sigma_bs=0.02, phi_bs=0.001, i_rescale=irs)
    o.apply_bc(2, 1e-4)
    U, V, W = o.fields()
    dx = np.diff(xg).min()  # the reference's dx (initialization.f90:254)
    dy = yg[-1] - yg[-2]
    nx = U.shape[0]
    for k in range(U.shape[2]):
        wz = (V[1:nx + 1, -1, k] - V[:nx, -1, k]) / dx - (U[:, -1, k] - U[:, -2, k]) / dy
        # the last column is the outflow's copy of column nx - 1 (:722-723)
        assert np.abs(wz[irs:nx - 1]).max() < 1e-9 * np.abs(V[:, -1, k]).max() / dx
    top_u, top_v, top_w = (c.tables[k] for k in ("U_top", "V_top", "W_top"))
    assert np.all(U[:irs, -1, :] == top_u[:irs, None])
    assert np.all(W[:irs, -1, :] == top_w[:irs, None])
    assert np.all(W[irs:, -1, :] == W[irs:, -2, :])
    # V keeps the blowing/suction value in those columns too (see the oracle)
    assert not np.allclose(V[:irs, -1, 3], top_v[:irs])
    q, scale = _net_flux(c, o)
    assert abs(q) < 1e-12 * scale
    o.close()


def test_falkner_skan_top():
    c, o = _case()
    xg = _xg(c)
    beta = 0.2
    m = beta / (2.0 - beta)
    o.set_boundary(top_flag=3, beta_hartree=beta)
    o.apply_bc(1, 1e-4)
    U, V, W = o.fields()
    # columns nx - 1 and nx carry the mass-flux correction (:722-726)
    ut = U[:-2, -1, 2]
    nx = ut.size
    lx = np.log(xg[:nx] + 1.0)
    slope = np.polyfit(lx, np.log(ut), 1)
    assert abs(slope[0] - m) < 1e-12
    # matched to 1 at x(1): the fit's value there
    assert abs(np.exp(np.polyval(slope, np.log(c.x[0] + 1.0))) - 1.0) < 1e-12
    assert np.all(V[:, -1, :] == 0.0)
    assert np.all(W[:, -1, :] == W[:, -2, :])
    q, scale = _net_flux(c, o)
    assert abs(q) < 1e-12 * scale
    o.close()


def test_random_inflow():
    from scipy import stats

    c, o = _case()
    amp = 0.05
    o.set_boundary(inflow_flag=4, amplitude=amp, seed=12345)
    o.compute_dt()
    o.step()
    U, V, W = o.fields()
    uin, vin, win = (c.tables[k] for k in ("U_inlet", "V_inlet", "W_inlet"))
    a_u = amp * (1.0 - uin)
    du = (U[0] - uin[:, None]) / np.where(a_u > 0, a_u, 1.0)[:, None]
    dv = (V[0] - vin[:, None]) / np.where(a_u[:vin.size] > 0, a_u[:vin.size], 1.0)[:, None]
    dw = (W[0] - win[:, None]) / np.where(a_u > 0, a_u, 1.0)[:, None]
    live = a_u > 1e-3  # rows inside the layer (the weight vanishes at the free stream)
    r = np.concatenate([du[live].ravel(), dv[live[:vin.size]].ravel(), dw[live].ravel()]) + 0.5
    assert r.min() >= -1e-9 and r.max() < 1.0 + 1e-9
    assert stats.kstest(r, "uniform").pvalue > 1e-3
    # once per step: a second boundary call of the same step keeps them
    first = U[0].copy()
    o.apply_bc(2, o.dt)
    assert np.array_equal(o.fields()[0][0], first)
    # the next step draws new values
    o.compute_dt()
    o.step()
    assert not np.allclose(o.fields()[0][0][live], first[live])
    o.close()


def test_unsupported_flags_rejected():
    c, o = _case()
    with pytest.raises(ValueError):
        o.set_boundary(inflow_flag=3)
    with pytest.raises(ValueError):
        o.set_boundary(top_flag=4)
    o.close()
