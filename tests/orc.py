"""ctypes view of the CPU oracle (oracle/liborc.so) — test infrastructure.

The oracle is the checker: it restates the reference Fortran solver in FP64 C
(see oracle/orc.h for the per-function reference citations).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# ORC_LIB selects another build of the oracle (the sanitizer build of oracle/Makefile)
LIB_PATH = os.environ.get("ORC_LIB") or os.path.join(ROOT, "oracle", "liborc.so")

_lib = None

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "orc_create": (_P, [_I, _I, _I, _P, _P, _P, _D, _D, _D, _I, _I, _D]),
            "orc_destroy": (None, [_P]),
            "orc_set_bc_tables": (None, [_P] + [_P] * 6),
            "orc_set_temporal_modes": (None, [_P, _I, _I] + [_P] * 8),
            "orc_field_size": (_I, [_P, _I]),
            "orc_set_field": (None, [_P, _I, _P]),
            "orc_get_field": (None, [_P, _I, _P]),
            "orc_set_time": (None, [_P, _D]),
            "orc_get_time": (_D, [_P]),
            "orc_apply_periodic_z": (None, [_P]),
            "orc_compute_dt": (_D, [_P]),
            "orc_step": (None, [_P]),
            "orc_set_boundary": (_I, [_P, _I, _I, _D, ctypes.c_ulonglong, _D, _D, _D, _D, _I, _D, _D, _D]),
            "orc_lund_means": (None, [_P, _P, _P, _P, _P]),
            "orc_set_wall_model": (_I, [_P, _I, _I, _D, _D, _D, _D, _D, _I, _D]),
            "orc_wall_model": (None, [_P]),
            "orc_get_wall_model": (None, [_P, _P, _P, _P, _P]),
            "orc_wall_beta": (_D, [_P]),
            "orc_rhs": (None, [_P, _P, _P, _P]),
            "orc_apply_bc": (None, [_P, _I, _D]),
            "orc_set_old": (None, [_P, _P, _P, _P]),
            "orc_projection": (None, [_P, _I, _D]),
            "orc_poisson": (None, [_P, _P]),
            "orc_max_divergence": (_D, [_P]),
            "orc_wall_stats": (None, [_P, _P]),
            "orc_monitor": (None, [_P, _P]),
            "orc_eddy_viscosity": (None, [_P]),
            "orc_set_nut_bc": (None, [_P, _I]),
            "orc_get_kxx_kzz": (None, [_P, _P, _P]),
            "orc_get_dyy": (None, [_P, _P, _P, _P]),
            "orc_get_dt": (_D, [_P]),
            "orc_blasius_tables": (None, [_I, _I, _P, _P, _P, _P, _D, _I] + [_P] * 11),
            "orc_interp_modes": (None, [_I, _P, _I, _P, _P, _I, _P, _P, _P]),
            "orc_init_flow": (None, [_I, _I, _I, _D, _D, _D, _D, _D, _I] + [_P] * 9),
            "orc_centers": (None, [_I, _P, _P, _P]),
            "orc_fft_once": (None, [_I, _I, _P, _P]),
            "orc_statistics": (None, [_P, _P, _P]),
            "orc_set_threads": (None, [_I]),
            "orc_get_threads": (_I, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def centers(x):
    x = f64(x)
    n = x.size
    xm = np.zeros(n - 1)
    xg = np.zeros(n + 1)
    lib().orc_centers(n, _p(x), _p(xm), _p(xg))
    return xm, xg


def blasius_tables(x, y, nu, eta, f, df):
    x, y, eta, f, df = map(f64, (x, y, eta, f, df))
    nx, ny = x.size, y.size
    _, xg = centers(x)
    _, yg = centers(y)
    out = dict(
        U_inlet=np.zeros(ny + 1), V_inlet=np.zeros(ny), W_inlet=np.zeros(ny + 1),
        U_outlet=np.zeros(ny + 1), V_outlet=np.zeros(ny), U_top=np.zeros(nx),
        V_top=np.zeros(nx + 1), W_top=np.zeros(nx + 1))
    lib().orc_blasius_tables(nx, ny, _p(x), _p(y), _p(xg), _p(yg), nu, eta.size, _p(eta), _p(f), _p(df),
                             *[_p(out[k]) for k in ("U_inlet", "V_inlet", "W_inlet", "U_outlet",
                                                    "V_outlet", "U_top", "V_top", "W_top")])
    return out


def interp_modes(ymesh, q_re, q_im, ygrid):
    """q_re/q_im: (ny_inlet, ntot) Fortran-ordered (j fastest) -> (nout, ntot)."""
    ymesh, ygrid = f64(ymesh), f64(ygrid)
    q_re = np.asfortranarray(q_re, dtype=np.float64)
    q_im = np.asfortranarray(q_im, dtype=np.float64)
    ntot = q_re.shape[1]
    o_re = np.zeros((ygrid.size, ntot), order="F")
    o_im = np.zeros((ygrid.size, ntot), order="F")
    lib().orc_interp_modes(ymesh.size, _p(ymesh), ntot, q_re.ctypes.data, q_im.ctypes.data, ygrid.size,
                           _p(ygrid), o_re.ctypes.data, o_im.ctypes.data)
    return o_re, o_im


def init_flow(nx, ny, nz, Lx, Ly, Lz, alpha, nu, eta, f, df):
    eta, f, df = map(f64, (eta, f, df))
    x, y, z = np.zeros(nx), np.zeros(ny), np.zeros(nz)
    U = np.zeros((nx, ny + 1, nz + 1), order="F")
    V = np.zeros((nx + 1, ny, nz + 1), order="F")
    W = np.zeros((nx + 1, ny + 1, nz), order="F")
    lib().orc_init_flow(nx, ny, nz, Lx, Ly, Lz, alpha, nu, eta.size, _p(eta), _p(f), _p(df), _p(x), _p(y),
                        _p(z), U.ctypes.data, V.ctypes.data, W.ctypes.data)
    return x, y, z, U, V, W


class Oracle:
    """One FP64 oracle state; fields are Fortran-ordered numpy arrays."""

    def __init__(self, x, y, z, nu, dPdx=0.0, dPdz=0.0, scheme=3, les=0, cfl=-1e-4):
        self.x, self.y, self.z = f64(x), f64(y), f64(z)
        self.nx, self.ny, self.nz = self.x.size, self.y.size, self.z.size
        self.h = lib().orc_create(self.nx, self.ny, self.nz, _p(self.x), _p(self.y), _p(self.z), nu, dPdx,
                                  dPdz, scheme, les, cfl)
        if not self.h:
            raise ValueError(f"oracle: unsupported configuration (les={les})")
        self.shapes = [(self.nx, self.ny + 1, self.nz + 1), (self.nx + 1, self.ny, self.nz + 1),
                       (self.nx + 1, self.ny + 1, self.nz), (self.nx + 1, self.ny + 1, self.nz + 1),
                       (self.nx + 1, self.ny + 1, self.nz + 1)]

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bc_tables(self, t):
        a = [f64(t[k]) for k in ("U_inlet", "V_inlet", "W_inlet", "U_top", "V_top", "W_top")]
        lib().orc_set_bc_tables(self.h, *[_p(v) for v in a])

    def set_temporal_modes(self, zmode, tmode, qu, qv, qw):
        """q*: complex arrays (nyg|ny, n_modes, m_modes)."""
        zmode, tmode = f64(zmode), f64(tmode)
        arrs = []
        for q in (qu, qv, qw):
            arrs.append(np.asfortranarray(q.real, dtype=np.float64))
            arrs.append(np.asfortranarray(q.imag, dtype=np.float64))
        lib().orc_set_temporal_modes(self.h, zmode.size, tmode.size, _p(zmode), _p(tmode),
                                     *[a.ctypes.data for a in arrs])

    def set(self, field, a):
        a = np.asfortranarray(a, dtype=np.float64)
        assert a.shape == self.shapes[field], (a.shape, self.shapes[field])
        lib().orc_set_field(self.h, field, a.ctypes.data)

    def get(self, field):
        a = np.zeros(self.shapes[field], order="F")
        lib().orc_get_field(self.h, field, a.ctypes.data)
        return a

    def set_fields(self, U, V, W):
        self.set(0, U)
        self.set(1, V)
        self.set(2, W)

    def fields(self):
        return self.get(0), self.get(1), self.get(2)

    @property
    def t(self):
        return lib().orc_get_time(self.h)

    @t.setter
    def t(self, v):
        lib().orc_set_time(self.h, v)

    def periodic(self):
        lib().orc_apply_periodic_z(self.h)

    def compute_dt(self):
        return lib().orc_compute_dt(self.h)

    def step(self):
        lib().orc_step(self.h)

    def set_boundary(self, inflow_flag=1, top_flag=0, amplitude=0.0, seed=0, vbs_max=0.0, x_bs=0.0, sigma_bs=1.0,
                     phi_bs=0.0, i_rescale=0, beta_hartree=0.0, delta_inlet=0.0, t_resc=0.0):
        rc = lib().orc_set_boundary(self.h, inflow_flag, top_flag, amplitude, seed, vbs_max, x_bs, sigma_bs, phi_bs,
                                    i_rescale, beta_hartree, delta_inlet, t_resc)
        if rc != 0:
            raise ValueError("orc_set_boundary: unsupported flags")

    def lund_means(self):
        nyg, ny = self.shapes[0][1], self.shapes[1][1]
        a, b, c, d = np.zeros(nyg), np.zeros(ny), np.zeros(nyg), np.zeros(ny)
        lib().orc_lund_means(self.h, _p(a), _p(b), _p(c), _p(d))
        return a, b, c, d

    def set_wall_model(self, iwall_model=0, istress_model=0, alpha_mean_x=0.0, alpha_mean_y=0.0, alpha_mean_z=0.0,
                       alpha_std=0.0, freq_mult=0.0, iwall_model_nut=0, frac_vis=0.0):
        rc = lib().orc_set_wall_model(self.h, iwall_model, istress_model, alpha_mean_x, alpha_mean_y, alpha_mean_z,
                                      alpha_std, freq_mult, iwall_model_nut, frac_vis)
        if rc != 0:
            raise ValueError("orc_set_wall_model: unsupported model")

    def wall_model(self):
        lib().orc_wall_model(self.h)

    def wall_beta(self):
        return lib().orc_wall_beta(self.h)

    def get_wall_model(self):
        """alpha_x(nx), alpha_y(nxg), alpha_z(nxg), utau_model(nx)"""
        nx, nxg = self.shapes[0][0], self.shapes[1][0]
        a, b, c, d = np.zeros(nx), np.zeros(nxg), np.zeros(nxg), np.zeros(nx)
        lib().orc_get_wall_model(self.h, _p(a), _p(b), _p(c), _p(d))
        return a, b, c, d

    def rhs(self):
        ru = np.zeros(self.shapes[0], order="F")
        rv = np.zeros(self.shapes[1], order="F")
        rw = np.zeros(self.shapes[2], order="F")
        lib().orc_rhs(self.h, ru.ctypes.data, rv.ctypes.data, rw.ctypes.data)
        return ru, rv, rw

    def set_old(self, Uo, Vo, Wo):
        a = [np.asfortranarray(v, dtype=np.float64) for v in (Uo, Vo, Wo)]
        lib().orc_set_old(self.h, *[v.ctypes.data for v in a])

    def apply_bc(self, rk_step, dt):
        lib().orc_apply_bc(self.h, rk_step, dt)

    def projection(self, rk_step, dt):
        lib().orc_projection(self.h, rk_step, dt)

    def poisson(self, rhs):
        r = np.asfortranarray(rhs, dtype=np.float64).copy(order="F")
        assert r.shape == self.shapes[3]
        lib().orc_poisson(self.h, r.ctypes.data)
        return r

    def max_divergence(self):
        return lib().orc_max_divergence(self.h)

    def wall_stats(self):
        cf = np.zeros(self.nx)
        lib().orc_wall_stats(self.h, _p(cf))
        return cf

    def statistics(self):
        """compute_statistics' z means: dict of the nine (nx, ny) arrays and
        nu_t_mean (nxg, nyg)"""
        means = np.zeros(9 * self.nx * self.ny)
        ntm = np.zeros((self.nx + 1, self.ny + 1), order="F")
        lib().orc_statistics(self.h, _p(means), ntm.ctypes.data)
        names = ("Umean", "Vmean", "Wmean", "U2mean", "V2mean", "W2mean", "UVmean", "Pmean", "P2mean")
        m = means.reshape((9, self.ny, self.nx))
        out = {k: np.asfortranarray(m[q].T) for q, k in enumerate(names)}
        out["nu_t_mean"] = ntm
        return out

    def monitor(self):
        out = np.zeros(8)
        lib().orc_monitor(self.h, _p(out))
        return out

    def eddy_viscosity(self):
        """compute_eddy_viscosity on the current fields; returns nu_t."""
        lib().orc_eddy_viscosity(self.h)
        return self.get(4)

    def set_nut_bc(self, dirichlet):
        lib().orc_set_nut_bc(self.h, int(dirichlet))

    def kxx_kzz(self):
        kxx = np.zeros(self.nx - 2)
        kzz = np.zeros(self.nz - 2)
        lib().orc_get_kxx_kzz(self.h, _p(kxx), _p(kzz))
        return kxx, kzz

    def dyy(self):
        nm = self.ny - 1
        d, lo, up = np.zeros(nm), np.zeros(nm), np.zeros(nm)
        lib().orc_get_dyy(self.h, _p(d), _p(lo), _p(up))
        return d, lo[:nm - 1], up[:nm - 1]

    @property
    def dt(self):
        return lib().orc_get_dt(self.h)
