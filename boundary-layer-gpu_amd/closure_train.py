"""Training pipeline and a-priori judge of the learned subgrid closure (host
side, over the blgpu C ABI).

A DNS solver on a grid that refines the LES grid by (fx, fy, fz) cells is
stepped; every `every` steps its field is box-filtered onto the LES solver
(bl_closure_samples), which appends one training row per LES cell: the
features the LES evaluates at run time and the least-squares eddy viscosity
of the exact SGS stress.  The network is trained on the device
(bl_mlp_train) and installed in an LES = 4 solver (bl_set_closure).

The a-priori judge compares, on the weighted rows of a held-out filtered
snapshot, the deviatoric SGS stress of each model, tau_model = -2 nu_t S,
with the exact one: relative Frobenius error and the correlation of the
off-diagonal components.  Models: the trained network, the dynamic
Smagorinsky model on the same filtered field (subgrid.f90:261-414), the
constant-coefficient Smagorinsky model (Cs = 0.1), no model, and the
per-cell least-squares optimum (the floor of any eddy-viscosity model).
"""
import numpy as np

ROW_T, ROW_W, ROW_SCALE, ROW_S, ROW_TAU = 10, 11, 12, 13, 19
NIN, NH = 10, 64


def init_params(seed=0):
    """FP32 initial parameters in the bl_mlp layout: uniform weights with
    half the He bound sqrt(6 / fan_in), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for shape in ((NH, NIN), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH,), (1,)):
        n = int(np.prod(shape))
        if len(shape) == 2 or shape == (NH,) and len(parts) == 6:
            fan_in = shape[1] if len(shape) == 2 else NH
            lim = 0.5 * np.sqrt(6.0 / fan_in)
            parts.append(rng.uniform(-lim, lim, size=n))
        else:
            parts.append(np.zeros(n))
    return np.concatenate(parts).astype(np.float32)


def lesgrid_pair(blgpu, nc, factors, box, alpha=2.0, x0=1.0):
    """Coinciding LES and DNS grids from the same tanh formula."""
    (nx, ny, nz), (fx, fy, fz) = nc, factors
    Lx, Ly, Lz = box
    fine = blgpu.grid_tanh(fx * (nx - 1) + 1, fy * (ny - 1) + 1, fz * (nz - 2) + 2, Lx, Ly, Lz, alpha, x0)
    coarse = blgpu.grid_tanh(nx, ny, nz, Lx, Ly, Lz, alpha, x0)
    return fine, coarse


def collect(fine, coarse, closure, factors, delta99, snapshots, every, smin=0.0, log=None):
    """Step the DNS and append the filtered rows of `snapshots` snapshots."""
    fx, fy, fz = factors
    for s in range(snapshots):
        for _ in range(every):
            fine.compute_dt()
            fine.step()
        closure.samples_from(fine, coarse, fx, fy, fz, delta99, smin)
        if log:
            log(f"snapshot {s + 1}/{snapshots}: rows {closure.count}")


def _dev(t):
    """deviatoric part of symmetric tensors stored as (11, 22, 33, 12, 13, 23) rows"""
    tr = (t[:, 0] + t[:, 1] + t[:, 2]) / 3.0
    d = t.copy()
    d[:, :3] -= tr[:, None]
    return d


def _fro(t):
    return np.sqrt((t[:, :3] ** 2).sum(1) + 2.0 * (t[:, 3:] ** 2).sum(1))


def stress_error(nu, rows):
    """relative Frobenius error of tau_dev = -2 nu S against the exact deviatoric
    stress, and the correlation of the off-diagonal components"""
    S = rows[:, ROW_S:ROW_S + 6].astype(np.float64)
    tau = _dev(rows[:, ROW_TAU:ROW_TAU + 6].astype(np.float64))
    model = -2.0 * nu[:, None] * S
    err = np.sqrt((_fro(model - tau) ** 2).sum() / (_fro(tau) ** 2).sum())
    a, b = model[:, 3:].ravel(), tau[:, 3:].ravel()
    corr = float(np.corrcoef(a, b)[0, 1]) if a.std() > 0 and b.std() > 0 else 0.0
    return float(err), corr


def apriori(closure, rows, nu_dsm=None, cs=0.1):
    """Stress errors of each model on weighted rows (see the module docstring)."""
    w = rows[:, ROW_W] > 0
    r = rows[w]
    scale = r[:, ROW_SCALE].astype(np.float64)
    # Delta^2 |S| = scale, and constant Smagorinsky nu = (Cs Delta)^2 |S|
    out = {}
    c = closure.forward(r[:, :10]) if r.shape[0] <= closure.batch_max else np.concatenate(
        [closure.forward(r[q:q + closure.batch_max, :10]) for q in range(0, r.shape[0], closure.batch_max)])
    out["learned"] = stress_error(np.maximum(c.astype(np.float64), 0.0) * scale, r)
    out["smagorinsky_cs%.2f" % cs] = stress_error(cs * cs * scale, r)
    out["no_model"] = stress_error(np.zeros_like(scale), r)
    out["least_squares_floor"] = stress_error(r[:, ROW_T].astype(np.float64) * scale, r)
    if nu_dsm is not None:
        out["dynamic_smagorinsky"] = stress_error(nu_dsm[w].astype(np.float64), r)
    return {k: dict(rel_error=v[0], corr_offdiag=v[1]) for k, v in out.items()}


def dsm_at_rows(les_solver):
    """nu_t of the dynamic Smagorinsky model on the solver's current field at
    the LES cells in row order (I fastest over [1, nxg-2], then J over
    [1, nyg-2], then K over [1, nzg-3]); the solver must have les = 2."""
    les_solver.eddy_viscosity()
    nut = les_solver.download(4)
    return nut[1:-1, 1:-1, 1:-2].ravel(order="F")
