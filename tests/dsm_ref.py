"""Vectorised numpy statement of the dynamic Smagorinsky models in index-map
form (test helper): z-averaged (LES = 2, subgrid.f90:261-414) and per cell
(LES = 3, sgs_dynamic_Smagorinsky_noaverage, :436-583, whose x fixes are
apply_periodic_bc_x, boundary_conditions.f90:295-314: centres 0 <- nxg-3,
nxg-2 <- 1, nxg-1 <- 2, faces 0 <- nx-2, nx-1 <- 1).

The reference fills the ghost entries of its work arrays in place
(apply_Neumann_bc_x, apply_periodic_bc_z) between steps.  This statement
instead materialises each work array once with every ghost entry read through
an index map: x clamped into 1..nxg-2 (0-based), and the z planes 0, nzg-2,
nzg-1 of a cell-centred array redirected to nzg-3, 1, 2 (faces: 0, nz-1 to
nz-2, 1).  It is a second derivation, independent of the loop-by-loop C
oracle; the two agreeing checks the maps the device kernels use.
"""
import numpy as np

from tests import orc

FIL = {(a, c): (4 / 9 if a == 0 and c == 0 else (1 / 9 if a == 0 or c == 0 else 1 / 36))
       for a in (-1, 0, 1) for c in (-1, 0, 1)}


def zmap_c(n3):
    k = np.arange(n3)
    m = k.copy()
    m[0], m[n3 - 2], m[n3 - 1] = n3 - 3, 1, 2
    return m


def zmap_f(n3):
    k = np.arange(n3)
    m = k.copy()
    m[0], m[n3 - 1] = n3 - 2, 1
    return m


def yfilter(G, jlo, jhi):
    """y part of filter_xzy on rows jlo..jhi (one-sided at both ends)."""
    out = np.zeros_like(G)
    out[:, jlo + 1:jhi] = G[:, jlo:jhi - 1] / 6 + 2 / 3 * G[:, jlo + 1:jhi] + G[:, jlo + 2:jhi + 1] / 6
    out[:, jlo] = 2 / 3 * G[:, jlo] + 1 / 3 * G[:, jlo + 1]
    out[:, jhi] = 1 / 3 * G[:, jhi - 1] + 2 / 3 * G[:, jhi]
    return out


def xzfilter(Y):
    """xz part of filter_xzy at every (i, k) with both neighbours present."""
    out = np.zeros_like(Y)
    n1, n3 = Y.shape[0], Y.shape[2]
    for (a, c), f in FIL.items():
        out[1:n1 - 1, :, 1:n3 - 1] += f * Y[1 + a:n1 - 1 + a, :, 1 + c:n3 - 1 + c]
    return out


def strain(x, y, z, xg, yg, zg, U, V, W):
    """compute_Sij on the interior cells of the P extent; returns (s[6], |S|)
    as full-extent arrays (zeros outside the interior)."""
    nxg, nyg, nzg = U.shape[0] + 1, U.shape[1], U.shape[2]
    shp = (nxg, nyg, nzg)
    I, J, K = slice(1, nxg - 1), slice(1, nyg - 1), slice(1, nzg - 1)
    s = [np.zeros(shp) for _ in range(6)]
    dx = np.diff(x)[:, None, None]
    dy = np.diff(y)[None, :, None]
    dz = np.diff(z)[None, None, :]
    s[0][I, J, K] = ((U[1:, :, :] - U[:-1, :, :]) / dx)[:, J, K]
    s[1][I, J, K] = ((V[:, 1:, :] - V[:, :-1, :]) / dy)[I, :, K]
    s[2][I, J, K] = ((W[:, :, 1:] - W[:, :, :-1]) / dz)[I, J, :]
    rxg = 1 / np.diff(xg)[:, None, None]
    ryg = 1 / np.diff(yg)[None, :, None]
    rzg = 1 / np.diff(zg)[None, None, :]
    # edges at 0-based (a, b, c) >= 1, stored at [a, b, c]
    E4, E5, E6 = np.zeros(shp), np.zeros(shp), np.zeros(shp)
    E4[1:, 1:, 1:] = 0.5 * ((U[:, 1:, 1:] - U[:, :-1, 1:]) * ryg[:, :, :] +
                            (V[1:, :, 1:] - V[:-1, :, 1:]) * rxg)
    E5[1:, 1:, 1:] = 0.5 * ((U[:, 1:, 1:] - U[:, 1:, :-1]) * rzg +
                            (W[1:, 1:, :] - W[:-1, 1:, :]) * rxg)
    E6[1:, 1:, 1:] = 0.5 * ((V[1:, :, 1:] - V[1:, :, :-1]) * rzg +
                            (W[1:, 1:, :] - W[1:, :-1, :]) * ryg)
    s[3][I, J, K] = 0.25 * (E4[1:-1, 1:-1, 1:-1] + E4[2:, 1:-1, 1:-1] + E4[1:-1, 2:, 1:-1] + E4[2:, 2:, 1:-1])
    s[4][I, J, K] = 0.25 * (E5[1:-1, 1:-1, 1:-1] + E5[2:, 1:-1, 1:-1] + E5[1:-1, 1:-1, 2:] + E5[2:, 1:-1, 2:])
    s[5][I, J, K] = 0.25 * (E6[1:-1, 1:-1, 1:-1] + E6[1:-1, 2:, 1:-1] + E6[1:-1, 1:-1, 2:] + E6[1:-1, 2:, 2:])
    mag = np.sqrt(sum((2.0 if c < 3 else 4.0) * s[c] ** 2 for c in range(6)))
    return s, mag


def xmap_periodic(n1, faces):
    m = np.arange(n1)
    if faces:
        m[0], m[n1 - 1] = n1 - 2, 1
    else:
        m[0], m[n1 - 2], m[n1 - 1] = n1 - 3, 1, 2
    return m


def dsm(x, y, z, U, V, W, nut_bc=0, noavg=False):
    nx, ny, nz = x.size, y.size, z.size
    nxg, nyg, nzg = nx + 1, ny + 1, nz + 1
    _, xg = orc.centers(x)
    _, yg = orc.centers(y)
    _, zg = orc.centers(z)
    ic = xmap_periodic(nxg, False) if noavg else np.clip(np.arange(nxg), 1, nxg - 2)
    iface = xmap_periodic(nx, True) if noavg else np.clip(np.arange(nx), 1, nx - 2)
    zc = zmap_c(nzg)
    Jall = np.arange(nyg)
    # cell-centred velocities (interpolation.f90 + the fixes of subgrid.f90:282-284)
    cu = 0.5 * (U[ic - 1] + U[ic])
    cv = np.zeros((nxg, nyg, nzg))
    cv[:, 1:nyg - 1] = 0.5 * (V[:, :-1] + V[:, 1:])
    cw = 0.5 * (W[:, :, zc - 1] + W[:, :, zc])
    remap = np.ix_(ic, Jall, zc)
    pu, pv, pw = cu[remap], cv[remap], cw[remap]
    prods = [pu * pu, pv * pv, pw * pw, pu * pv, pu * pw, pv * pw]
    jlo, jhi = 1, nyg - 2
    filt = lambda G: xzfilter(yfilter(G, jlo, jhi))
    hp = [filt(p) for p in prods]
    hu, hv, hw = filt(cu), filt(cv), filt(cw)
    L = [hp[0] - hu * hu, hp[1] - hv * hv, hp[2] - hw * hw, hp[3] - hu * hv, hp[4] - hu * hw, hp[5] - hv * hw]
    # unfiltered strain, |S| S_ij filtered
    s, S = strain(x, y, z, xg, yg, zg, U, V, W)
    hs = [filt((S * s[c])[remap]) for c in range(6)]
    # test-filtered velocities with their fixes
    Uf = xzfilter(yfilter(U, 0, nyg - 1))[np.ix_(iface, np.arange(nyg), zc)]
    Vf = xzfilter(yfilter(V, 0, ny - 1))[np.ix_(ic, np.arange(ny), zc)]
    Wf = xzfilter(yfilter(W, 0, nyg - 1))[np.ix_(ic, Jall, zmap_f(nz))]
    sh, Sh = strain(x, y, z, xg, yg, zg, Uf, Vf, Wf)
    M = [4.0 * Sh * sh[c] - hs[c] for c in range(6)]
    LM = M[0] * L[0] + M[1] * L[1] + M[2] * L[2] + 2 * (M[3] * L[3] + M[4] * L[4] + M[5] * L[5])
    MM = M[0] ** 2 + M[1] ** 2 + M[2] ** 2 + 2 * (M[3] ** 2 + M[4] ** 2 + M[5] ** 2)
    ks = slice(1, nzg - 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        nut = np.zeros((nxg, nyg, nzg))
        if noavg:
            core = -0.5 * LM / MM * S
            nut[:, 1:nyg - 1] = core[np.ix_(ic, np.arange(1, nyg - 1), zc)]
        else:
            avg = LM[:, :, ks].mean(axis=2) / MM[:, :, ks].mean(axis=2)
            core = -0.5 * avg[:, :, None] * S
            nut[1:nxg - 1, 1:nyg - 1, 1:nzg - 1] = core[1:nxg - 1, 1:nyg - 1, 1:nzg - 1]
    nut = np.fmax(nut, 0.0)
    sg = -1.0 if nut_bc == 1 else 1.0
    nut[:, 0] = sg * nut[:, 1]
    nut[:, -1] = sg * nut[:, -2]
    return nut
