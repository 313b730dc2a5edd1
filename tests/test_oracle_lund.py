"""Pins of the oracle's Lund rescaling (inflow_flag 3 and 5,
compute_rescaled_inflow rescaled_inlet_bc.f90:60-300) by what the method
fixes:

- a field that does not vary in x is its own rescaling: the momentum
  thicknesses of the inlet and rescaling planes agree, so gamma = 1, and with
  the inlet delta set to the rescaling plane's delta99 (found here by the
  rule the reference states: the first row at or above 0.99 U_inf, linearly
  interpolated) the inner and outer coordinates coincide; the inflow plane
  must then equal the rescaling plane (flag 3), a wrong index, weight,
  interpolation or scaling shows;
- flag 5 puts only the rescaled fluctuations on the U/V/W_inlet means, with
  W's planes shifted by 16 in z (boundary_conditions.f90:531-545);
- the time-averaged means follow Umean_T <- dt/T Umean + (1 - dt/T) Umean_T
  (rescaled_inlet_bc.f90:111-115) from the field's means at set-up
  (initialization.f90:640-660).
"""
import numpy as np
import pytest

from . import cases, orc

IR = 15  # Lund_ix (1-based)


def _x_uniform_case():
    c = cases.laminar_small(nx=30, ny=14, nz=8, noise=0.03, seed=6, cfl=-1e-4)
    o = c.oracle()
    U, V, W = o.fields()
    # every x plane a copy of the rescaling plane
    U[:] = U[IR - 1][None]
    V[:] = V[IR - 1][None]
    W[:] = W[IR - 1][None]
    o.set_fields(U, V, W)
    o.set_old(U, V, W)
    return c, o, U, V, W


def _delta99(c, U):
    _, yg = orc.centers(c.y)
    m = U[IR - 1][:, 1:-1].mean(axis=1)
    level = 0.99 * m[-1]
    hit = np.nonzero(m >= level)[0]
    j = hit[0] if hit.size and hit[0] >= 1 else m.size // 2
    w1 = (level - m[j - 1]) / (m[j] - m[j - 1])
    return abs(w1 * yg[j] + (1 - w1) * yg[j - 1])


def test_x_uniform_field_is_its_own_rescaling():
    c, o, U, V, W = _x_uniform_case()
    o.set_boundary(inflow_flag=3, i_rescale=IR, delta_inlet=_delta99(c, U), t_resc=0.5)
    o.compute_dt()
    o.apply_bc(1, o.dt)
    U1, V1, W1 = o.fields()
    # rows 2 .. n-1 (the wall row and the top row are the wall and top
    # conditions' own, set after the inflow)
    for a, b in ((U1[0, 1:-1], U[IR - 1, 1:-1]), (V1[0, 1:-1], V[IR - 1, 1:-1]), (W1[0, 1:-1], W[IR - 1, 1:-1])):
        assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()
    o.close()


def test_flag5_fluctuations_on_inlet_means():
    c, o, U, V, W = _x_uniform_case()
    o.set_boundary(inflow_flag=5, i_rescale=IR, delta_inlet=_delta99(c, U), t_resc=0.5)
    o.compute_dt()
    o.apply_bc(1, o.dt)
    U1, V1, W1 = o.fields()
    uin, vin, win = (c.tables[k] for k in ("U_inlet", "V_inlet", "W_inlet"))
    nzg = U.shape[2]
    um = U[IR - 1][:, 1:nzg - 1].mean(axis=1)
    vm = V[IR - 1][:, 1:nzg - 1].mean(axis=1)
    fu = uin[:, None] + (U[IR - 1] - um[:, None])
    fv = vin[:, None] + (V[IR - 1] - vm[:, None])
    nz = W.shape[2]
    ks = (np.arange(nz) + 17) % nz  # 1-based ks = mod(k + 16, nz) + 1
    fw = win[:, None] + W[IR - 1][:, ks]
    # interior z planes: the z ghost planes are the periodic condition's copies
    for a, b in ((U1[0, 1:-1, 1:-1], fu[1:-1, 1:-1]), (V1[0, 1:-1, 1:-1], fv[1:-1, 1:-1]),
                 (W1[0, 1:-1, 1:-1], fw[1:-1, 1:-1])):
        assert np.abs(a - b).max() < 1e-13 * max(np.abs(b).max(), 1.0)
    o.close()


def test_time_averaged_means():
    c, o, U, V, W = _x_uniform_case()
    t_resc = 0.37
    o.set_boundary(inflow_flag=3, i_rescale=IR, delta_inlet=_delta99(c, U), t_resc=t_resc)
    m0 = o.lund_means()
    # the step-start field changes: its rescaling and inlet means are new
    U2, V2 = 1.1 * U, 0.9 * V
    o.set_old(U2, V2, W)
    o.compute_dt()
    dt = o.dt
    o.apply_bc(1, dt)
    m1 = o.lund_means()
    nzg = U.shape[2]
    new = (U2[IR - 1][:, 1:nzg - 1].mean(1), V2[IR - 1][:, 1:nzg - 1].mean(1), U2[0][:, 1:nzg - 1].mean(1),
           V2[0][:, 1:nzg - 1].mean(1))
    for a, b0, bn in zip(m1, m0, new):
        want = dt / t_resc * bn + (1 - dt / t_resc) * b0
        assert np.allclose(a, want, rtol=1e-13, atol=1e-16)
    o.close()


def test_lund_parameters_checked():
    c, o, U, V, W = _x_uniform_case()
    for bad in (dict(i_rescale=1, delta_inlet=0.01, t_resc=1.0), dict(i_rescale=IR, delta_inlet=0.0, t_resc=1.0),
                dict(i_rescale=IR, delta_inlet=0.01, t_resc=0.0)):
        with pytest.raises(ValueError):
            o.set_boundary(inflow_flag=3, **bad)
    o.close()
