"""numpy statement of the learned closure's data path (test helper): the box
filter of a DNS onto an LES grid refined from it by (fx, fy, fz) cells, the
exact SGS stress of each LES cell, the cell features (velocity-gradient
tensor, wall distance) and the least-squares eddy-viscosity target.

Arrays are Fortran-ordered with the reference's extents, indices 0-based:
U[nx, nyg, nzg], V[nxg, ny, nzg], W[nxg, nyg, nz].
"""
import numpy as np

from tests import orc


def zmap_c(nzg):
    m = np.arange(nzg)
    m[0], m[nzg - 2], m[nzg - 1] = nzg - 3, 1, 2
    return m


def zmap_f(nz):
    m = np.arange(nz)
    m[0], m[nz - 1] = nz - 2, 1
    return m


def coarsen(Uf, Vf, Wf, yf, fx, fy, fz, nc):
    """Face box averages of the DNS on the LES grid (nx, ny, nz) = nc."""
    nx, ny, nz = nc
    nxg, nyg, nzg = nx + 1, ny + 1, nz + 1
    dy = np.diff(yf)  # dy[j-1] = y(j) - y(j-1)
    Uc = np.zeros((nx, nyg, nzg), order="F")
    Vc = np.zeros((nxg, ny, nzg), order="F")
    Wc = np.zeros((nxg, nyg, nz), order="F")
    zc = zmap_c(nzg)

    def rows(J):  # fine rows of coarse cell row J (J >= 1) and their dy weights
        j = np.arange(fy * (J - 1) + 1, fy * J + 1)
        return j, dy[j - 1]

    def planes(K):
        return np.arange(fz * (K - 1) + 1, fz * K + 1)

    def xcells(I):
        if I == 0:
            return np.array([0])
        if I == nxg - 1:
            return np.array([Vf.shape[0] - 1])
        return np.arange(fx * (I - 1) + 1, fx * I + 1)

    for K in range(nzg):
        k = planes(zc[K])
        for I in range(nx):
            i = fx * I
            for J in range(nyg):
                if J == nyg - 1:
                    Uc[I, J, K] = Uf[i, -1, k].mean()
                else:
                    j, w = rows(max(J, 1))
                    v = (w[:, None] * Uf[i][np.ix_(j, k)]).sum() / (w.sum() * fz)
                    Uc[I, J, K] = -v if J == 0 else v
        for I in range(nxg):
            ii = xcells(I)
            for J in range(ny):
                Vc[I, J, K] = Vf[np.ix_(ii, [fy * J], k)].mean()
    zf = zmap_f(nz)
    for K in range(nz):
        k = fz * zf[K]
        for I in range(nxg):
            ii = xcells(I)
            for J in range(nyg):
                if J == nyg - 1:
                    Wc[I, J, K] = Wf[ii, -1, k].mean()
                else:
                    j, w = rows(max(J, 1))
                    v = (w[None, :] * Wf[np.ix_(ii, j, [k])][:, :, 0]).sum() / (w.sum() * len(ii))
                    Wc[I, J, K] = -v if J == 0 else v
    return Uc, Vc, Wc


def sgs_stress(Uf, Vf, Wf, yf, fx, fy, fz, nc):
    """tau_ab of each interior LES cell (I, J, K in [1, nxg-2] x [1, nyg-2] x
    [1, nzg-3]) from the DNS cell-centred velocities of its block; array
    [6, nxg-2, nyg-2, nzg-3] in the order 11 22 33 12 13 23."""
    nx, ny, nz = nc
    nxg, nyg, nzg = nx + 1, ny + 1, nz + 1
    dy = np.diff(yf)
    u = 0.5 * (Uf[:-1] + Uf[1:])          # cells 1..nxf-1 -> index i-1
    v = 0.5 * (Vf[:, :-1] + Vf[:, 1:])
    w = 0.5 * (Wf[:, :, :-1] + Wf[:, :, 1:])
    out = np.zeros((6, nxg - 2, nyg - 2, nzg - 3))
    for I in range(1, nxg - 1):
        ii = np.arange(fx * (I - 1) + 1, fx * I + 1)
        for J in range(1, nyg - 1):
            jj = np.arange(fy * (J - 1) + 1, fy * J + 1)
            wy = dy[jj - 1]
            for K in range(1, nzg - 2):
                kk = np.arange(fz * (K - 1) + 1, fz * K + 1)
                bu = u[np.ix_(ii - 1, jj, kk)]
                bv = v[np.ix_(ii, jj - 1, kk)]
                bw = w[np.ix_(ii, jj, kk - 1)]
                wt = np.broadcast_to(wy[None, :, None], bu.shape)
                avg = lambda a: (wt * a).sum() / wt.sum()
                mu, mv, mw = avg(bu), avg(bv), avg(bw)
                out[:, I - 1, J - 1, K - 1] = [avg(bu * bu) - mu * mu, avg(bv * bv) - mv * mv,
                                               avg(bw * bw) - mw * mw, avg(bu * bv) - mu * mv,
                                               avg(bu * bw) - mu * mw, avg(bv * bw) - mv * mw]
    return out


def grad_tensor(x, y, z, U, V, W):
    """g[a*3+b] = d u_a / d x_b at every interior cell (i, j, k in
    [1, nxg-2] x [1, nyg-2] x [1, nzg-2]); array [9, nxg-2, nyg-2, nzg-2]."""
    _, xg = orc.centers(x)
    _, yg = orc.centers(y)
    _, zg = orc.centers(z)
    nxg, nyg, nzg = x.size + 1, y.size + 1, z.size + 1
    I, J, K = np.arange(1, nxg - 1), np.arange(1, nyg - 1), np.arange(1, nzg - 1)
    ix = lambda *a: np.ix_(*a)
    g = np.zeros((9, I.size, J.size, K.size))
    g[0] = (U[ix(I, J, K)] - U[ix(I - 1, J, K)]) / (x[I] - x[I - 1])[:, None, None]
    g[4] = (V[ix(I, J, K)] - V[ix(I, J - 1, K)]) / (y[J] - y[J - 1])[None, :, None]
    g[8] = (W[ix(I, J, K)] - W[ix(I, J, K - 1)]) / (z[K] - z[K - 1])[None, None, :]
    for da in (0, 1):
        for db in (0, 1):
            a, b = I + da, J + db
            g[1] += (U[ix(a - 1, b, K)] - U[ix(a - 1, b - 1, K)]) / (yg[b] - yg[b - 1])[None, :, None]
            g[3] += (V[ix(a, b - 1, K)] - V[ix(a - 1, b - 1, K)]) / (xg[a] - xg[a - 1])[:, None, None]
            a, c = I + da, K + db
            g[2] += (U[ix(a - 1, J, c)] - U[ix(a - 1, J, c - 1)]) / (zg[c] - zg[c - 1])[None, None, :]
            g[6] += (W[ix(a, J, c - 1)] - W[ix(a - 1, J, c - 1)]) / (xg[a] - xg[a - 1])[:, None, None]
            b, c = J + da, K + db
            g[5] += (V[ix(I, b - 1, c)] - V[ix(I, b - 1, c - 1)]) / (zg[c] - zg[c - 1])[None, None, :]
            g[7] += (W[ix(I, b, c - 1)] - W[ix(I, b - 1, c - 1)]) / (yg[b] - yg[b - 1])[None, :, None]
    for q in (1, 3, 2, 6, 5, 7):
        g[q] *= 0.25
    return g


def features(x, y, z, U, V, W, delta99):
    """features [10, ...], S [6, ...], |S|, Delta^2 |S| on the interior cells."""
    g = grad_tensor(x, y, z, U, V, W)
    gn = np.sqrt((g ** 2).sum(0))
    f = np.zeros((10,) + gn.shape)
    with np.errstate(invalid="ignore", divide="ignore"):
        f[:9] = np.where(gn > 0, g / gn, 0.0)
    _, yg = orc.centers(y)
    f[9] = (yg[1:-1] / delta99)[None, :, None]
    S = np.stack([g[0], g[4], g[8], 0.5 * (g[1] + g[3]), 0.5 * (g[2] + g[6]), 0.5 * (g[5] + g[7])])
    smag = np.sqrt(2 * (S[0] ** 2 + S[1] ** 2 + S[2] ** 2) + 4 * (S[3] ** 2 + S[4] ** 2 + S[5] ** 2))
    nxg, nyg, nzg = x.size + 1, y.size + 1, z.size + 1
    vol = (np.diff(x)[:, None, None] * np.diff(y)[None, :, None] * np.diff(z)[None, None, :])
    vol = vol[: nxg - 2, : nyg - 2, : nzg - 2]
    scale = np.cbrt(vol) ** 2 * smag
    return f, S, smag, scale


def target(S, tau, scale):
    ss = S[0] ** 2 + S[1] ** 2 + S[2] ** 2 + 2 * (S[3] ** 2 + S[4] ** 2 + S[5] ** 2)
    ts = tau[0] * S[0] + tau[1] * S[1] + tau[2] * S[2] + 2 * (tau[3] * S[3] + tau[4] * S[4] + tau[5] * S[5])
    return -ts / (2 * ss) / scale
