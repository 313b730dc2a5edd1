"""Benchmark: ns/cell/step of the MI355X per-time-step path at 814x257x193.

Workload (the reference's throughput configuration): tanh-stretched grid
(alpha = 2), Lx = 400 theta, Ly = Lz = 40 theta, x from 1, nu for
Re_theta = 300 at the inlet, Blasius initial field plus uniform noise of
amplitude 0.01 (seed 0), RK3, FP64, dt recomputed every step exactly as
compute_dt does (time_integration.f90:528-591).  compute_dt takes the max of
dx/|u| (time_integration.f90:553-555), so a positive CFL always selects the
viscous bound dt = CFL*dy_min^2/nu; at CFL = 0.8 that puts nu*dt/dy_min^2 = 0.8
past RK3's diffusive stability limit (about 0.63) and the reference's own
algorithm diverges within ~15 steps.  The bench therefore runs CFL = 0.5; the
work per step does not depend on dt.  Inputs are resident
in HBM before the timed region.  One "step" = compute_dt + one RK3 step.

Prints one JSON line.  `roofline` describes the kernel with the largest
share of GPU time, measured with HIP events on the solver's stream during the
timed steps; `cpu_baseline` is the FP64 C oracle timed on this host for one
RK3 step of a z-slab of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

THETA = 1.469653e-3  # theta_in for Re_theta = 300 with x0 = 1
NU = 4.898844e-6  # 1 / Re_x0, Re_x0 = (300/0.664)^2
CFL = 0.5  # see the module docstring: 0.8 is unstable for the reference algorithm
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def workload(nx, ny, nz, seed=0, noise=0.01):
    from tests.blgpu_import import blgpu

    x, y, z = blgpu.grid_tanh(nx, ny, nz, 400 * THETA, 40 * THETA, 40 * THETA, 2.0, 1.0)
    eta, f, df = blgpu.blasius_profile()
    tables = blgpu.blasius_tables(x, y, NU, eta, f, df)
    U, V, W = blgpu.blasius_field(x, y, z, NU, eta, f, df)
    blgpu.add_noise(U, V, W, noise, seed)
    # force periodicity of the initial field (initialization.f90:244-246)
    for a in (U, V):
        a[:, :, 0] = a[:, :, -3]
        a[:, :, -2] = a[:, :, 1]
        a[:, :, -1] = a[:, :, 2]
    W[:, :, 0] = W[:, :, -2]
    W[:, :, -1] = W[:, :, 1]
    return x, y, z, tables, (U, V, W)


def kernel_units(name, nx, ny, nz):
    """Algorithmic work of one launch: ('flop', F) or ('byte', B)."""
    N, M, NM = nx - 2, nz - 2, ny - 1
    cells = nx * ny * nz
    if name in ("xform_x_fwd", "xform_x_inv"):
        return "flop", 2.0 * N * N * NM * M
    if name in ("xform_z_fwd", "xform_z_inv"):
        return "flop", 2.0 * M * M * N * NM
    per_cell_passes = {
        # one pass = one read or write of an N_cell FP64 array
        "rk_stage": 11.0,  # mean over RK3 stages of (read U,V,W[,Uo,F1]; write U*[,F1])
        "divergence": 4.0,
        "thomas": 3.0,
        "project": 7.0,
        "pressure": 2.0,
    }
    if name in per_cell_passes:
        return "byte", per_cell_passes[name] * 8.0 * cells
    return "byte", None


def cpu_baseline(nz_sample):
    from tests import orc

    nx, ny = 814, 257
    x, y, z, tables, (U, V, W) = workload(nx, ny, nz_sample)
    o = orc.Oracle(x, y, z, NU, 0.0, 0.0, 3, 0, CFL)
    o.set_bc_tables(tables)
    o.set_fields(U, V, W)
    o.compute_dt()
    o.step()  # warm-up (page-in, thread pool)
    t0 = time.perf_counter()
    o.compute_dt()
    o.step()
    el = time.perf_counter() - t0
    cells = nx * ny * nz_sample
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=el / cells * 1e9, unit="ns/cell/step", cores=cores, kind="port",
                sample=f"FP64 C oracle, 1 timed RK3 step (after 1 warm-up) of an {nx}x{ny}x{nz_sample} "
                       f"z-slab of the same workload, OpenMP on {cores} threads")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=814)
    ap.add_argument("--ny", type=int, default=257)
    ap.add_argument("--nz", type=int, default=193)
    ap.add_argument("--cpu-nz", type=int, default=17, help="z extent of the CPU-baseline sample (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    from tests.blgpu_import import blgpu

    nx, ny, nz = args.nx, args.ny, args.nz
    print(f"bench: building workload {nx}x{ny}x{nz}", file=sys.stderr, flush=True)
    x, y, z, tables, (U, V, W) = workload(nx, ny, nz)
    s = blgpu.Solver(x, y, z, NU, tables, scheme=3, les=0, cfl=CFL)
    s.set_fields(U, V, W)
    del U, V, W
    dt = s.compute_dt()
    print(f"bench: dt = {dt:.6e}; warm-up {args.warmup} steps", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        s.compute_dt()
        s.step()
    s.synchronize()
    s.profile(True)
    print(f"bench: timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.compute_dt()
        s.step()
    s.synchronize()
    el = time.perf_counter() - t0
    prof = s.profile_read()
    s.profile(False)
    div = s.max_divergence()
    cf = s.wall_stats()
    if not np.all(np.isfinite(cf)):
        raise RuntimeError("non-finite Cf after the timed steps")
    cells = nx * ny * nz
    ms_per_step = el / args.steps * 1e3
    value = el / args.steps / cells * 1e9

    tot = {k: v[0] for k, v in prof.items() if v[1] > 0}
    gpu_ms = sum(tot.values())
    dom = max(tot, key=tot.get)
    dom_ms, dom_n = prof[dom]
    avg_s = dom_ms / dom_n * 1e-3
    kind, units = kernel_units(dom, nx, ny, nz)
    if kind == "flop":
        achieved = units / avg_s / 1e12
        roof = dict(bound="mfma", achieved=achieved, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=achieved / FP64_PEAK_TFLOPS, traffic=None)
    else:
        achieved = (units / avg_s / 1e9) if units else None
        roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=(achieved / HBM_PEAK_GBS) if achieved else None, traffic=None)
    roof.update(kernel=dom, avg_launch_ms=dom_ms / dom_n, share_of_gpu_time=dom_ms / gpu_ms)
    kernels = {k: dict(ms_per_step=v[0] / args.steps, launches_per_step=v[1] / args.steps)
               for k, v in prof.items() if v[1] > 0}

    cpu = None
    if args.cpu_nz > 0:
        print(f"bench: CPU baseline sample nz={args.cpu_nz}", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_nz)

    out = {
        "metric": "ns/cell/step, 814x257x193 RK3 DNS (FP64 vs mixed); LES Cf error at equal cost",
        "value": value,
        "unit": "ns/cell/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": value / 2.09,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{nx}x{ny}x{nz} tanh(alpha=2) Blasius + 1% noise seed 0, RK3, CFL={CFL}, FP64",
                   "nx": nx, "ny": ny, "nz": nz, "dt": dt, "precision": "fp64",
                   "max_divergence_after": div},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
