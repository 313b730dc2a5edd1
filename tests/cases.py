"""Workload builders shared by the tests (CPU side, numpy + oracle).

A case is the reference's laminar Blasius set-up on either the init_flow
sinh grid (input_output.f90:130-279) or a tanh-stretched y grid
y_j = Ly*(1 - tanh(alpha*(1-eta_j))/tanh(alpha)), with x starting at 1 and
uniform z, optionally with seeded uniform noise on the interior velocity.
"""
import os

import numpy as np

from . import orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_blasius(path=None):
    """Reference fixture laminar_test/blasius_solution.dat: n, eta[n], f[n], f'[n]."""
    path = path or os.path.join(GOLDEN, "blasius_solution.dat")
    with open(path) as fh:
        n = int(fh.readline())
        vals = np.array(fh.read().split(), dtype=np.float64)
    assert vals.size == 3 * n
    return vals[:n], vals[n:2 * n], vals[2 * n:]


def load_temporal_modes(path=None):
    """Reference fixture transition_test/temporal_modes_inlet.dat (layout read
    at boundary_conditions.f90:1669-1722, little-endian stream)."""
    path = path or os.path.join(GOLDEN, "temporal_modes_inlet.dat")
    b = np.fromfile(path, dtype=np.uint8)
    ny, n, m = np.frombuffer(b[:12].tobytes(), "<i4")
    beta, omega = np.frombuffer(b[12:28].tobytes(), "<f8")
    d = np.frombuffer(b[28:].tobytes(), "<f8")
    o = 0

    def take(cnt):
        nonlocal o
        v = d[o:o + cnt]
        o += cnt
        return v

    ymesh = take(ny)
    zmode = take(n) * beta
    tmode = take(m) * omega
    q = {}
    for c in ("u", "v", "w"):
        re = take(ny * n * m).reshape((ny, n, m), order="F")
        im = take(ny * n * m).reshape((ny, n, m), order="F")
        q[c] = re + 1j * im
    assert o == d.size
    return dict(ny=int(ny), n=int(n), m=int(m), beta=float(beta), omega=float(omega), ymesh=ymesh,
                zmode=zmode, tmode=tmode, q=q)


def tanh_grid(nx, ny, nz, Lx, Ly, Lz, alpha=2.0, x0=1.0):
    x = x0 + Lx * np.arange(nx) / (nx - 1)
    eta = np.arange(ny) / (ny - 1)
    y = Ly * (1.0 - np.tanh(alpha * (1.0 - eta)) / np.tanh(alpha))
    y[0] = 0.0
    z = Lz * np.arange(nz) / (nz - 2)
    return x, y, z


def blasius_fields(x, y, z, nu, eta, f, df):
    """init_flow's Blasius field (input_output.f90:206-259) on arbitrary grids."""
    nx, ny, nz = x.size, y.size, z.size
    _, xg = orc.centers(x)
    _, yg = orc.centers(y)

    def interp(e, tab):
        # first jj with eta(jj) > e, linear between jj-1 and jj; beyond the table -> None
        jj = np.searchsorted(eta, e, side="right")
        ok = jj < eta.size
        jj = np.clip(jj, 1, eta.size - 1)
        w1 = (e - eta[jj - 1]) / (eta[jj] - eta[jj - 1])
        w0 = 1.0 - w1
        return w0 * tab[jj - 1] + w1 * tab[jj], ok

    U = np.ones((nx, ny + 1, nz + 1), order="F")
    for ii in range(nx):
        e = yg * np.sqrt(1.0 / (2.0 * nu * x[ii]))
        val, ok = interp(e, df)
        U[ii, :, :] = np.where(ok, val, 1.0)[:, None]
    V = np.zeros((nx + 1, ny, nz + 1), order="F")
    g = eta * df - f
    for ii in range(nx + 1):
        rex = xg[ii] / nu
        e = y * np.sqrt(1.0 / (2.0 * nu * xg[ii]))
        val, ok = interp(e, g)
        V[ii, :, :] = (1.0 / np.sqrt(2.0 * rex) * np.where(ok, val, g[-1]))[:, None]
    W = np.zeros((nx + 1, ny + 1, nz), order="F")
    return U, V, W


def add_noise(U, V, W, amp, seed):
    rng = np.random.default_rng(seed)
    for a in (U, V, W):
        a[1:-1, 1:-1, 1:-1] += amp * (2.0 * rng.random(a[1:-1, 1:-1, 1:-1].shape) - 1.0)


class Case:
    def __init__(self, nx, ny, nz, Lx, Ly, Lz, nu, alpha=2.0, grid="tanh", noise=0.0, seed=0, scheme=3,
                 les=0, cfl=-1e-4, modes=False, dPdx=0.0, dPdz=0.0):
        self.eta, self.f, self.df = load_blasius()
        if grid == "sinh":
            x, y, z, U, V, W = orc.init_flow(nx, ny, nz, Lx, Ly, Lz, alpha, nu, self.eta, self.f, self.df)
        else:
            x, y, z = tanh_grid(nx, ny, nz, Lx, Ly, Lz, alpha)
            U, V, W = blasius_fields(x, y, z, nu, self.eta, self.f, self.df)
        if noise:
            add_noise(U, V, W, noise, seed)
        self.x, self.y, self.z = x, y, z
        self.U0, self.V0, self.W0 = U, V, W
        self.nu, self.scheme, self.les, self.cfl = nu, scheme, les, cfl
        self.dPdx, self.dPdz = dPdx, dPdz
        self.tables = orc.blasius_tables(x, y, nu, self.eta, self.f, self.df)
        self.modes = None
        if modes:
            tm = load_temporal_modes()
            _, yg = orc.centers(y)
            ntot = tm["n"] * tm["m"]
            qs = {}
            for c, grid_ in (("u", yg), ("v", y), ("w", yg)):
                q = tm["q"][c].reshape((tm["ny"], ntot), order="F")
                re, im = orc.interp_modes(tm["ymesh"], q.real, q.imag, grid_)
                qs[c] = (re + 1j * im).reshape((grid_.size, tm["n"], tm["m"]), order="F")
            self.modes = dict(zmode=tm["zmode"], tmode=tm["tmode"], qu=qs["u"], qv=qs["v"], qw=qs["w"])

    def oracle(self):
        o = orc.Oracle(self.x, self.y, self.z, self.nu, self.dPdx, self.dPdz, self.scheme, self.les, self.cfl)
        o.set_bc_tables(self.tables)
        if self.modes is not None:
            o.set_temporal_modes(**self.modes)
        o.set_fields(self.U0, self.V0, self.W0)
        o.periodic()
        return o


def laminar_small(**kw):
    """Scaled-down laminar Blasius case: x from 1, Re_x0 = 1/nu."""
    args = dict(nx=24, ny=16, nz=8, Lx=0.15, Ly=0.015, Lz=0.006, nu=4.898844e-6, alpha=2.0)
    args.update(kw)
    return Case(**args)


def oracle_step_p(make_oracle, fields, t):
    """The P that one oracle step records (projection.f90:75-85) when it
    starts from the given U, V, W at time t.  A HIP step from the same state
    is held to this: P after many steps also carries the P response to the
    round-off differences of the two states (the pressure of a nearly steady
    flow is a small residual of div U*, of which the previous projection's
    rounding is a visible part)."""
    o = make_oracle()
    o.set_fields(*fields)
    o.t = t
    o.compute_dt()
    o.step()
    P = o.get(3)
    o.close()
    return P
