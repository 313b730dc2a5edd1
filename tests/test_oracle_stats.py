"""Pins of the oracle's compute_statistics restatement (statistics.f90:44-80,
125-135) by closed forms.

Fields linear in y (cell centres yg) interpolate exactly to the y faces,
fields linear in xg average exactly to the x faces, so every mean is a known
function of the plane values.  The z dependence is chosen so that the plane
ranges show: U, V and P are summed over planes 2..nzg-1, which include the
periodic duplicate nzg-1 of plane 2 (a cosine of the plane index has zero
mean over the nzp distinct planes, so only the duplicate survives); W over
faces 1..nz-2 (a linear function of the face index tells that range from
2..nz-1); nu_t_mean keeps column nxg and row nyg at 0."""
import numpy as np

from . import cases, orc

NX, NY, NZ = 9, 7, 8


def _oracle():
    x, y, z = cases.tanh_grid(NX, NY, NZ, 0.3, 0.05, 0.02, alpha=2.0)
    o = orc.Oracle(x, y, z, 1e-3)
    _, xg = orc.centers(x)
    _, yg = orc.centers(y)
    return o, x, y, z, xg, yg


def test_statistics_closed_forms():
    o, x, y, z, xg, yg = _oracle()
    nxg, nyg, nzg = NX + 1, NY + 1, NZ + 1
    nzp = NZ - 2
    kk = np.arange(1, nzg + 1)  # 1-based plane index
    cz = np.cos(2 * np.pi * (kk - 2) / nzp)  # periodic in the plane index
    a, b, c = 0.7, 3.0, 0.05
    U = a + b * yg[None, :, None] + c * cz[None, None, :] + 0.0 * np.zeros((NX, nyg, nzg))
    d0, e0 = 0.01, 0.2
    V = d0 + e0 * xg[:, None, None] + np.zeros((nxg, NY, nzg))
    f0, g0, h0 = -0.02, 0.4, 0.003
    kw = np.arange(1, NZ + 1)
    W = f0 + g0 * yg[None, :, None] + h0 * kw[None, None, :] + np.zeros((nxg, nyg, NZ))
    p0, p1, p2 = 0.1, -0.01, 0.02
    ii = np.arange(1, nxg + 1)
    P = p0 + p1 * kk[None, None, :] + p2 * ii[:, None, None] + np.zeros((nxg, nyg, nzg))
    n0, n1 = 1e-4, 2e-5
    NT = n0 + n1 * ii[:, None, None] * kk[None, None, :] + np.zeros((nxg, nyg, nzg))
    o.set_fields(*[np.asfortranarray(v) for v in (U, V, W)])
    o.set(3, np.asfortranarray(P))
    o.set(4, np.asfortranarray(NT))
    st = o.statistics()

    ks = kk[1:nzg - 1]  # planes 2..nzg-1
    m1 = cz[ks - 1].mean()  # = 1 / (nzg - 2): only the duplicate plane
    m2 = (cz[ks - 1] ** 2).mean()
    assert abs(m1 - 1.0 / (nzg - 2)) < 1e-15
    Aj = a + b * y  # U interpolated to the y faces
    Bi = d0 + e0 * 0.5 * (xg[:-1] + xg[1:])  # V averaged to the x faces
    exp = {
        "Umean": np.broadcast_to(Aj + c * m1, (NX, NY)),
        "U2mean": np.broadcast_to(Aj ** 2 + 2 * Aj * c * m1 + c * c * m2, (NX, NY)),
        "Vmean": np.broadcast_to(Bi[:, None], (NX, NY)),
        "V2mean": np.broadcast_to(Bi[:, None] ** 2, (NX, NY)),
        "UVmean": Bi[:, None] * (Aj + c * m1)[None, :],
    }
    kf = kw[: NZ - 2]  # faces 1..nz-2
    Fj = f0 + g0 * y
    exp["Wmean"] = np.broadcast_to(Fj + h0 * kf.mean(), (NX, NY))
    exp["W2mean"] = np.broadcast_to(Fj ** 2 + 2 * Fj * h0 * kf.mean() + h0 * h0 * (kf ** 2).mean(), (NX, NY))
    Pi = p0 + p2 * ii[:NX]
    exp["Pmean"] = np.broadcast_to((Pi + p1 * ks.mean())[:, None], (NX, NY))
    exp["P2mean"] = np.broadcast_to((Pi ** 2 + 2 * Pi * p1 * ks.mean() + p1 * p1 * (ks ** 2).mean())[:, None],
                                    (NX, NY))
    for k, v in exp.items():
        assert np.allclose(st[k], v, rtol=1e-13, atol=1e-15), k
    ntm = st["nu_t_mean"]
    assert ntm.shape == (nxg, nyg)
    assert np.allclose(ntm[:NX, :NY], (n0 + n1 * ii[:NX] * ks.mean())[:, None], rtol=1e-13, atol=0)
    assert np.all(ntm[NX, :] == 0) and np.all(ntm[:, NY] == 0)


def test_statistics_cf_consistent_with_wall_stats():
    """UVmean(:, 1) and the wall means of U are what the Cf formula uses
    (statistics.f90:137-160): with nu_t = 0, Cf = 2 (-UVmean(:,1) + nu dU/dy)."""
    o, x, y, z, xg, yg = _oracle()
    rng = np.random.default_rng(4)
    shapes = o.shapes
    F = [np.asfortranarray(rng.standard_normal(shapes[c]) * 0.1 + (1.0 if c == 0 else 0.0)) for c in range(3)]
    o.set_fields(*F)
    o.set(4, np.zeros(shapes[4], order="F"))
    st = o.statistics()
    U = F[0]
    u1 = U[:, 0, 1:-1].mean(axis=1)
    u2 = U[:, 1, 1:-1].mean(axis=1)
    dudy = (u2 - u1) / (yg[1] - yg[0])
    cf = 2 * (-st["UVmean"][:, 0] + 1e-3 * dudy)
    assert np.allclose(o.wall_stats(), cf, rtol=1e-12, atol=1e-14)
