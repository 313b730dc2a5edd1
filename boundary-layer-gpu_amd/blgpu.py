"""Python binding of the blgpu C ABI (include/blgpu.h) via ctypes.

Fields are numpy arrays in the reference's Fortran layout (x fastest, full
allocated extent).  The library must be built in-tree
(`make -C boundary-layer-gpu_amd`); importing this module never falls back to
another implementation.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libblgpu.so")

FIELD_U, FIELD_V, FIELD_W, FIELD_P, FIELD_NUT = range(5)
PREC_FP64, PREC_MIXED = 0, 1

_D, _I, _P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)


class BLConfig(ctypes.Structure):
    _fields_ = [("nx", _I), ("ny", _I), ("nz", _I), ("x", _DP), ("y", _DP), ("z", _DP), ("nu", _D),
                ("dPdx", _D), ("dPdz", _D), ("cfl", _D), ("scheme", _I), ("les", _I), ("precision", _I),
                ("U_inlet", _DP), ("V_inlet", _DP), ("W_inlet", _DP), ("U_top", _DP), ("V_top", _DP),
                ("W_top", _DP), ("n_modes", _I), ("m_modes", _I), ("zmode", _DP), ("tmode", _DP),
                ("qu_re", _DP), ("qu_im", _DP), ("qv_re", _DP), ("qv_im", _DP), ("qw_re", _DP),
                ("qw_im", _DP), ("nut_bc", _I)]


class BLBoundary(ctypes.Structure):
    """bl_boundary: the general dispatcher's flags (boundary_conditions.f90:22-121)."""
    _fields_ = [("inflow_flag", _I), ("top_flag", _I), ("amplitude", _D), ("seed", ctypes.c_ulonglong),
                ("vbs_max", _D), ("x_bs", _D), ("sigma_bs", _D), ("phi_bs", _D), ("i_rescale", _I),
                ("beta_hartree", _D), ("delta_inlet", _D), ("t_resc", _D)]


class BLWallParams(ctypes.Structure):
    """bl_wall_params: the wall model (.turbb WM, TauwModel, alphas)."""
    _fields_ = [("iwall_model", _I), ("istress_model", _I), ("alpha_mean_x", _D), ("alpha_mean_y", _D),
                ("alpha_mean_z", _D), ("alpha_std", _D), ("freq_mult", _D), ("iwall_model_nut", _I),
                ("frac_vis", _D)]


# every entry point declared in include/blgpu.h with its ctypes signature
SIGNATURES = {
    "bl_create": (_I, [ctypes.POINTER(BLConfig), ctypes.POINTER(_P)]),
    "bl_destroy": (None, [_P]),
    "bl_last_error": (ctypes.c_char_p, []),
    "bl_upload": (_I, [_P, _I, _P]),
    "bl_download": (_I, [_P, _I, _P]),
    "bl_set_time": (_I, [_P, _D]),
    "bl_set_boundary": (_I, [_P, _P]),
    "bl_set_wall_model": (_I, [_P, _P]),
    "bl_compute_wall_model": (_I, [_P]),
    "bl_get_wall_model": (_I, [_P, _P, _P, _P, _P]),
    "bl_get_time": (_I, [_P, _DP]),
    "bl_compute_dt": (_I, [_P, _DP]),
    "bl_step": (_I, [_P]),
    "bl_run": (_I, [_P, _I]),
    "bl_synchronize": (_I, [_P]),
    "bl_eddy_viscosity": (_I, [_P]),
    "bl_max_divergence": (_I, [_P, _DP]),
    "bl_wall_stats": (_I, [_P, _P]),
    "bl_monitor": (_I, [_P, _P]),
    "bl_statistics": (_I, [_P, _P, _P, _P]),
    "bl_max_divergence_noalloc": (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P]),
    "bl_fftz_create": (_I, [_I, ctypes.POINTER(_P)]),
    "bl_fftz_destroy": (None, [_P]),
    "bl_fft_z": (_I, [_P, _I, _P, _P]),
    "bl_ifft_z": (_I, [_P, _I, _P, _P]),
    "bl_snapshot_begin": (_I, [_P]),
    "bl_snapshot_end": (_I, [_P] + [ctypes.POINTER(_DP)] * 5),
    "bl_rhs": (_I, [_P, _P, _P, _P]),
    "bl_set_old": (_I, [_P, _P, _P, _P]),
    "bl_apply_bc": (_I, [_P, _I, _D]),
    "bl_projection": (_I, [_P, _I, _D]),
    "bl_poisson": (_I, [_P, _P]),
    "bl_profile_enable": (_I, [_P, _I]),
    "bl_profile_select": (_I, [_P, ctypes.c_ulonglong]),
    "bl_profile_read": (_I, [_P, _P, _P, _I]),
    "bl_profile_count": (_I, []),
    "bl_profile_name": (ctypes.c_char_p, [_I]),
    "bl_blasius_profile": (_I, [_I, _D, _D, _I, _P, _P, _P]),
    "bl_blasius_tables": (_I, [_I, _I, _P, _P, _D, _I, _P, _P, _P] + [_P] * 6),
    "bl_blasius_field": (_I, [_I, _I, _I, _P, _P, _D, _I, _P, _P, _P, _P, _P, _P]),
    "bl_add_noise": (_I, [_I, _I, _I, _D, ctypes.c_ulonglong, _P, _P, _P]),
    "bl_grid_tanh": (_I, [_I, _I, _I, _D, _D, _D, _D, _D, _P, _P, _P]),
    "bl_mlp_create": (_I, [ctypes.c_longlong, _I, ctypes.POINTER(_P)]),
    "bl_mlp_destroy": (None, [_P]),
    "bl_mlp_set_params": (_I, [_P, _P]),
    "bl_mlp_get_params": (_I, [_P, _P]),
    "bl_mlp_forward": (_I, [_P, _P, _I, _P]),
    "bl_mlp_gradient": (_I, [_P, _P, _P, _I, _P, _P]),
    "bl_mlp_add_rows": (_I, [_P, _P, ctypes.c_longlong]),
    "bl_mlp_get_rows": (_I, [_P, ctypes.c_longlong, ctypes.c_longlong, _P]),
    "bl_mlp_count": (_I, [_P, ctypes.POINTER(ctypes.c_longlong)]),
    "bl_mlp_clear": (_I, [_P]),
    "bl_mlp_train": (_I, [_P, _I, _I, ctypes.c_float, ctypes.c_ulonglong, _P]),
    "bl_closure_samples": (_I, [_P, _P, _P, _I, _I, _I, _D, _D]),
    "bl_set_closure": (_I, [_P, _P, _D, _D]),
    "bl_closure_fallback": (_I, [_P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    "bl_closure_infer": (_I, [_P, _P, ctypes.c_longlong, _P]),
}

MLP_IN, MLP_HIDDEN, MLP_PARAMS, MLP_ROW = 10, 64, 9089, 28

_lib = None


class BLError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BLError(f"blgpu library not built: {LIB_PATH} (run make -C boundary-layer-gpu_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = lib().bl_last_error().decode(errors="replace")
        raise BLError(f"{what} failed ({rc}): {msg}")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(_DP)


# ----------------------------------------------------------------- set-up ----
def blasius_profile(n=1000, eta_max=20.0, fpp0=0.46960, sub=64):
    eta, f, df = np.zeros(n), np.zeros(n), np.zeros(n)
    _check(lib().bl_blasius_profile(n, eta_max, fpp0, sub, eta.ctypes.data, f.ctypes.data, df.ctypes.data),
           "bl_blasius_profile")
    return eta, f, df


def blasius_tables(x, y, nu, eta, f, df):
    x, y, eta, f, df = map(_f64, (x, y, eta, f, df))
    nx, ny = x.size, y.size
    t = dict(U_inlet=np.zeros(ny + 1), V_inlet=np.zeros(ny), W_inlet=np.zeros(ny + 1), U_top=np.zeros(nx),
             V_top=np.zeros(nx + 1), W_top=np.zeros(nx + 1))
    _check(lib().bl_blasius_tables(nx, ny, x.ctypes.data, y.ctypes.data, nu, eta.size, eta.ctypes.data,
                                   f.ctypes.data, df.ctypes.data,
                                   *[t[k].ctypes.data for k in ("U_inlet", "V_inlet", "W_inlet", "U_top",
                                                                "V_top", "W_top")]), "bl_blasius_tables")
    return t


def blasius_field(x, y, z, nu, eta, f, df):
    x, y, eta, f, df = map(_f64, (x, y, eta, f, df))
    nx, ny, nz = x.size, y.size, len(z)
    U = np.zeros((nx, ny + 1, nz + 1), order="F")
    V = np.zeros((nx + 1, ny, nz + 1), order="F")
    W = np.zeros((nx + 1, ny + 1, nz), order="F")
    _check(lib().bl_blasius_field(nx, ny, nz, x.ctypes.data, y.ctypes.data, nu, eta.size, eta.ctypes.data,
                                  f.ctypes.data, df.ctypes.data, U.ctypes.data, V.ctypes.data, W.ctypes.data),
           "bl_blasius_field")
    return U, V, W


def add_noise(U, V, W, amp, seed):
    for a in (U, V, W):
        assert a.flags.f_contiguous and a.dtype == np.float64
    nx, nyg, nzg = U.shape
    _check(lib().bl_add_noise(nx, nyg - 1, nzg - 1, amp, seed, U.ctypes.data, V.ctypes.data, W.ctypes.data),
           "bl_add_noise")


def grid_tanh(nx, ny, nz, Lx, Ly, Lz, alpha=2.0, x0=1.0):
    x, y, z = np.zeros(nx), np.zeros(ny), np.zeros(nz)
    _check(lib().bl_grid_tanh(nx, ny, nz, Lx, Ly, Lz, alpha, x0, x.ctypes.data, y.ctypes.data, z.ctypes.data),
           "bl_grid_tanh")
    return x, y, z


# ----------------------------------------------------------------- solver ----
class Solver:
    """One device-resident solver (bl_create ... bl_destroy)."""

    def __init__(self, x, y, z, nu, tables, dPdx=0.0, dPdz=0.0, scheme=3, les=0, cfl=-1e-4, modes=None,
                 precision=PREC_FP64, nut_bc=0):
        self._keep = []
        keep = lambda a: (self._keep.append(_f64(a)), self._keep[-1])[1]
        x, y, z = keep(x), keep(y), keep(z)
        self.nx, self.ny, self.nz = x.size, y.size, z.size
        c = BLConfig()
        c.nx, c.ny, c.nz = self.nx, self.ny, self.nz
        c.x, c.y, c.z = _ptr(x), _ptr(y), _ptr(z)
        c.nu, c.dPdx, c.dPdz, c.cfl = nu, dPdx, dPdz, cfl
        c.scheme, c.les, c.precision = scheme, les, precision
        c.nut_bc = nut_bc
        for k in ("U_inlet", "V_inlet", "W_inlet", "U_top", "V_top", "W_top"):
            setattr(c, k, _ptr(keep(tables[k])))
        if modes is not None:
            c.n_modes = len(modes["zmode"])
            c.m_modes = len(modes["tmode"])
            c.zmode = _ptr(keep(modes["zmode"]))
            c.tmode = _ptr(keep(modes["tmode"]))
            for name, key in (("qu", "qu"), ("qv", "qv"), ("qw", "qw")):
                q = np.asfortranarray(modes[key])
                re = self._keep.append(np.asfortranarray(q.real, dtype=np.float64)) or self._keep[-1]
                im = self._keep.append(np.asfortranarray(q.imag, dtype=np.float64)) or self._keep[-1]
                setattr(c, name + "_re", _ptr(re))
                setattr(c, name + "_im", _ptr(im))
        self.h = _P()
        _check(lib().bl_create(ctypes.byref(c), ctypes.byref(self.h)), "bl_create")
        self.shapes = [(self.nx, self.ny + 1, self.nz + 1), (self.nx + 1, self.ny, self.nz + 1),
                       (self.nx + 1, self.ny + 1, self.nz), (self.nx + 1, self.ny + 1, self.nz + 1),
                       (self.nx + 1, self.ny + 1, self.nz + 1)]

    def close(self):
        if getattr(self, "h", None):
            lib().bl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, field, a):
        a = np.asfortranarray(a, dtype=np.float64)
        assert a.shape == self.shapes[field], (a.shape, self.shapes[field])
        _check(lib().bl_upload(self.h, field, a.ctypes.data), "bl_upload")

    def download(self, field):
        a = np.zeros(self.shapes[field], order="F")
        _check(lib().bl_download(self.h, field, a.ctypes.data), "bl_download")
        return a

    def set_fields(self, U, V, W):
        self.upload(FIELD_U, U)
        self.upload(FIELD_V, V)
        self.upload(FIELD_W, W)

    def fields(self):
        return self.download(FIELD_U), self.download(FIELD_V), self.download(FIELD_W)

    @property
    def t(self):
        v = _D()
        _check(lib().bl_get_time(self.h, ctypes.byref(v)), "bl_get_time")
        return v.value

    @t.setter
    def t(self, v):
        _check(lib().bl_set_time(self.h, v), "bl_set_time")

    def set_boundary(self, inflow_flag=1, top_flag=0, amplitude=0.0, seed=0, vbs_max=0.0, x_bs=0.0, sigma_bs=1.0,
                     phi_bs=0.0, i_rescale=0, beta_hartree=0.0, delta_inlet=0.0, t_resc=0.0):
        """bl_set_boundary: inflow_flag 1-5, top_flag 0-3 and their parameters
        (.turbb Amplitude; Vmax, x0, sigma, phi, Lund_ix, Lund_deltai, Lund_T)."""
        b = BLBoundary(inflow_flag, top_flag, amplitude, seed, vbs_max, x_bs, sigma_bs, phi_bs, i_rescale,
                       beta_hartree, delta_inlet, t_resc)
        _check(lib().bl_set_boundary(self.h, ctypes.byref(b)), "bl_set_boundary")

    def set_wall_model(self, iwall_model=0, istress_model=0, alpha_mean_x=0.0, alpha_mean_y=0.0, alpha_mean_z=0.0,
                       alpha_std=0.0, freq_mult=0.0, iwall_model_nut=0, frac_vis=0.0):
        """bl_set_wall_model: WM 0, 4, 5, 6, 8, 9, 11, 12, 15, TauwModel 0/1,
        WMnutflag 0/2 and WMnut (compute_wall_model wallmodel.f90:25-144)."""
        p = BLWallParams(iwall_model, istress_model, alpha_mean_x, alpha_mean_y, alpha_mean_z, alpha_std, freq_mult,
                         iwall_model_nut, frac_vis)
        _check(lib().bl_set_wall_model(self.h, ctypes.byref(p)), "bl_set_wall_model")

    def compute_wall_model(self):
        _check(lib().bl_compute_wall_model(self.h), "bl_compute_wall_model")

    def get_wall_model(self):
        """alpha_x(nx), alpha_y(nxg), alpha_z(nxg), utau_model(nx)"""
        nx = self.nx
        a, b, c, d = np.zeros(nx), np.zeros(nx + 1), np.zeros(nx + 1), np.zeros(nx)
        _check(lib().bl_get_wall_model(self.h, a.ctypes.data, b.ctypes.data, c.ctypes.data, d.ctypes.data),
               "bl_get_wall_model")
        return a, b, c, d

    def compute_dt(self):
        v = _D()
        _check(lib().bl_compute_dt(self.h, ctypes.byref(v)), "bl_compute_dt")
        return v.value

    def step(self):
        _check(lib().bl_step(self.h), "bl_step")

    def run(self, n):
        _check(lib().bl_run(self.h, n), "bl_run")

    def synchronize(self):
        _check(lib().bl_synchronize(self.h), "bl_synchronize")

    def eddy_viscosity(self):
        _check(lib().bl_eddy_viscosity(self.h), "bl_eddy_viscosity")

    def max_divergence(self):
        v = _D()
        _check(lib().bl_max_divergence(self.h, ctypes.byref(v)), "bl_max_divergence")
        return v.value

    def wall_stats(self):
        cf = np.zeros(self.nx)
        _check(lib().bl_wall_stats(self.h, cf.ctypes.data), "bl_wall_stats")
        return cf

    def monitor(self):
        out = np.zeros(8)
        _check(lib().bl_monitor(self.h, out.ctypes.data), "bl_monitor")
        return out

    STAT_NAMES = ("Umean", "Vmean", "Wmean", "U2mean", "V2mean", "W2mean", "UVmean", "Pmean", "P2mean")

    def statistics(self):
        """compute_statistics' z means (bl_statistics): a dict of the nine
        (nx, ny) arrays, nu_t_mean (nxg, nyg) and Cf (nx)."""
        nx, ny = self.nx, self.ny
        means = np.zeros(9 * nx * ny)
        ntm = np.zeros((nx + 1, ny + 1), order="F")
        cf = np.zeros(nx)
        _check(lib().bl_statistics(self.h, means.ctypes.data, ntm.ctypes.data, cf.ctypes.data), "bl_statistics")
        m = means.reshape((9, ny, nx))
        out = {k: np.asfortranarray(m[q].T) for q, k in enumerate(self.STAT_NAMES)}
        out["nu_t_mean"] = ntm
        out["Cf"] = cf
        return out

    def snapshot_begin(self):
        _check(lib().bl_snapshot_begin(self.h), "bl_snapshot_begin")

    def snapshot_end(self):
        """U, V, W, P, nu_t of the last snapshot_begin (copies of the
        library's pinned buffer)."""
        ptrs = [_DP() for _ in range(5)]
        _check(lib().bl_snapshot_end(self.h, *[ctypes.byref(q) for q in ptrs]), "bl_snapshot_end")
        out = []
        for q, c in zip(ptrs, (0, 1, 2, 3, 4)):
            shape = self.shapes[c]
            a = np.ctypeslib.as_array(q, shape=(int(np.prod(shape)),))
            out.append(np.array(a.reshape(shape, order="F"), order="F", copy=True))
        return tuple(out)

    def rhs(self):
        r = [np.zeros(self.shapes[c], order="F") for c in range(3)]
        _check(lib().bl_rhs(self.h, *[a.ctypes.data for a in r]), "bl_rhs")
        return r

    def set_old(self, Uo, Vo, Wo):
        a = [np.asfortranarray(v, dtype=np.float64) for v in (Uo, Vo, Wo)]
        _check(lib().bl_set_old(self.h, *[v.ctypes.data for v in a]), "bl_set_old")

    def apply_bc(self, rk_step, dt):
        _check(lib().bl_apply_bc(self.h, rk_step, dt), "bl_apply_bc")

    def projection(self, rk_step, dt):
        _check(lib().bl_projection(self.h, rk_step, dt), "bl_projection")

    def closure_fallback(self):
        """(cells evaluated, cells that kept the DSM value) in the last evaluation."""
        c, f = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().bl_closure_fallback(self.h, ctypes.byref(c), ctypes.byref(f)), "bl_closure_fallback")
        return c.value, f.value

    def profile(self, on=True, only=None):
        """Bracket kernel launches with HIP events; `only`: the class names to
        bracket (default all)."""
        mask = (1 << 64) - 1
        if only is not None:
            names = [lib().bl_profile_name(i).decode() for i in range(lib().bl_profile_count())]
            mask = 0
            for n in only:
                mask |= 1 << names.index(n)
        _check(lib().bl_profile_select(self.h, mask), "bl_profile_select")
        _check(lib().bl_profile_enable(self.h, 1 if on else 0), "bl_profile_enable")

    def profile_read(self):
        n = lib().bl_profile_count()
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        _check(lib().bl_profile_read(self.h, ms.ctypes.data, cnt.ctypes.data, n), "bl_profile_read")
        return {lib().bl_profile_name(i).decode(): (float(ms[i]), int(cnt[i])) for i in range(n)}

    def poisson(self, rhs):
        r = np.asfortranarray(rhs, dtype=np.float64).copy(order="F")
        assert r.shape == self.shapes[FIELD_P]
        _check(lib().bl_poisson(self.h, r.ctypes.data), "bl_poisson")
        return r


class FFTZ:
    """The fftz module's plan (bl_fftz_*): real DFT along the second axis of
    (n1, N) arrays, r2c unnormalised and c2r divided by N."""

    def __init__(self, N):
        self.N = int(N)
        p = _P()
        _check(lib().bl_fftz_create(self.N, ctypes.byref(p)), "bl_fftz_create")
        self.p = p

    def close(self):
        if getattr(self, "p", None):
            lib().bl_fftz_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, U):
        U = np.asfortranarray(U, dtype=np.float64)
        n1, N = U.shape
        assert N == self.N
        out = np.zeros((2, n1, N // 2 + 1), order="F")
        _check(lib().bl_fft_z(self.p, n1, U.ctypes.data, out.ctypes.data), "bl_fft_z")
        return out[0] + 1j * out[1]

    def inverse(self, Uh):
        Uh = np.asarray(Uh, dtype=np.complex128)
        n1, K = Uh.shape
        assert K == self.N // 2 + 1
        buf = np.zeros((2, n1, K), order="F")
        buf[0] = Uh.real
        buf[1] = Uh.imag
        out = np.zeros((n1, self.N), order="F")
        _check(lib().bl_ifft_z(self.p, n1, buf.ctypes.data, out.ctypes.data), "bl_ifft_z")
        return out


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Closure:
    """Learned subgrid closure: device dataset, FP32 master parameters, Adam
    (bl_mlp_*)."""

    def __init__(self, capacity, batch_max=65536):
        self.h = _P()
        _check(lib().bl_mlp_create(int(capacity), int(batch_max), ctypes.byref(self.h)), "bl_mlp_create")
        self.batch_max = batch_max

    def close(self):
        if getattr(self, "h", None):
            lib().bl_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        t = np.zeros(MLP_PARAMS, dtype=np.float32)
        _check(lib().bl_mlp_get_params(self.h, t.ctypes.data), "bl_mlp_get_params")
        return t

    @params.setter
    def params(self, theta):
        t = _f32(theta)
        assert t.size == MLP_PARAMS
        _check(lib().bl_mlp_set_params(self.h, t.ctypes.data), "bl_mlp_set_params")

    def forward(self, X):
        X = _f32(X)
        y = np.zeros(X.shape[0], dtype=np.float32)
        _check(lib().bl_mlp_forward(self.h, X.ctypes.data, X.shape[0], y.ctypes.data), "bl_mlp_forward")
        return y

    def gradient(self, X, t):
        X, t = _f32(X), _f32(t)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        _check(lib().bl_mlp_gradient(self.h, X.ctypes.data, t.ctypes.data, X.shape[0], g.ctypes.data,
                                     ctypes.byref(loss)), "bl_mlp_gradient")
        return g, loss.value

    def add_rows(self, rows):
        r = _f32(rows)
        assert r.ndim == 2 and r.shape[1] == MLP_ROW
        _check(lib().bl_mlp_add_rows(self.h, r.ctypes.data, r.shape[0]), "bl_mlp_add_rows")

    def rows(self, first=0, n=None):
        n = self.count - first if n is None else n
        r = np.zeros((n, MLP_ROW), dtype=np.float32)
        _check(lib().bl_mlp_get_rows(self.h, int(first), int(n), r.ctypes.data), "bl_mlp_get_rows")
        return r

    @property
    def count(self):
        n = ctypes.c_longlong()
        _check(lib().bl_mlp_count(self.h, ctypes.byref(n)), "bl_mlp_count")
        return n.value

    def clear(self):
        _check(lib().bl_mlp_clear(self.h), "bl_mlp_clear")

    def infer_bf16(self, X):
        """The BF16-MFMA inference kernel on feature rows."""
        X = _f32(X)
        y = np.zeros(X.shape[0], dtype=np.float32)
        _check(lib().bl_closure_infer(self.h, X.ctypes.data, X.shape[0], y.ctypes.data), "bl_closure_infer")
        return y

    def samples_from(self, fine, coarse, fx, fy, fz, delta99, smin=0.0):
        """Rows from the DNS solver `fine` filtered onto the LES solver `coarse`."""
        _check(lib().bl_closure_samples(fine.h, coarse.h, self.h, fx, fy, fz, delta99, smin), "bl_closure_samples")

    def install(self, solver, delta99, nsigma=3.0):
        _check(lib().bl_set_closure(solver.h, self.h, delta99, nsigma), "bl_set_closure")

    def train(self, steps, batch, lr=1e-3, seed=0, history=True):
        hist = np.zeros(steps, dtype=np.float32) if history else None
        _check(lib().bl_mlp_train(self.h, int(steps), int(batch), float(lr), int(seed),
                                  hist.ctypes.data if history else None), "bl_mlp_train")
        return hist
