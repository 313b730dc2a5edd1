"""The learned-closure study at the sizes of the training and LES
configurations: a noisy-Blasius FP64 DNS on a grid that refines the LES grid
by (4, 2, 4) cells provides filtered snapshots; the network is trained on
the device; the held-out snapshot judges it a priori against the dynamic
Smagorinsky model (subgrid.f90:261-414), the constant-coefficient model, no
model and the per-cell least-squares floor of any eddy-viscosity model; then
LES = 4 (learned, dynamic Smagorinsky fallback) and LES = 2 run side by side
from the held-out filtered field in the mixed-precision form, and their Cf
and mean U(y) are compared with the DNS continued to the same time.

    python closure_study.py [--les 257 97 64] [--snapshots 40] [--every 50]
                            [--les-steps 5000] [--out result.json]

The default LES grid is 257x97x64 and the DNS 1025x193x250: the DNS x
transform has 1023 = 31 x 33 points, where 256 LES points would give the
prime 1019, which the x transform does not take.

Prints one JSON object (and writes it to --out).
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name):
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


THETA = 1.469653e-3  # theta_in for Re_theta = 300 with x0 = 1
NU = 4.898844e-6


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--les", type=int, nargs=3, default=[257, 97, 64], metavar=("NX", "NY", "NZ"))
    ap.add_argument("--box", type=float, nargs=3, default=[400.0, 40.0, 50.0], metavar=("LX", "LY", "LZ"),
                    help="box in units of theta_in")
    ap.add_argument("--noise", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--snapshots", type=int, default=40, help="training snapshots (one more is held out)")
    ap.add_argument("--every", type=int, default=50, help="DNS steps between snapshots")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--les-steps", type=int, default=5000)
    ap.add_argument("--nsigma", type=float, default=3.0)
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    blgpu = _load("blgpu")
    T = _load("closure_train")

    def log(*m):
        print(*m, file=sys.stderr, flush=True)

    nc, fac = tuple(a.les), (4, 2, 4)
    box = tuple(v * THETA for v in a.box)
    (xf, yf, zf), (xc, yc, zc) = T.lesgrid_pair(blgpu, nc, fac, box)
    eta, f, df = blgpu.blasius_profile()
    Uf, Vf, Wf = blgpu.blasius_field(xf, yf, zf, NU, eta, f, df)
    blgpu.add_noise(Uf, Vf, Wf, a.noise, a.seed)
    fine = blgpu.Solver(xf, yf, zf, NU, blgpu.blasius_tables(xf, yf, NU, eta, f, df), cfl=a.cfl)
    fine.set_fields(Uf, Vf, Wf)
    del Uf, Vf, Wf
    tab_c = blgpu.blasius_tables(xc, yc, NU, eta, f, df)
    judge = blgpu.Solver(xc, yc, zc, NU, tab_c, les=2, cfl=a.cfl)
    d99 = 7.5 * THETA
    ncell = (nc[0] - 1) * (nc[1] - 1) * (nc[2] - 2)
    cl = blgpu.Closure(capacity=(a.snapshots + 1) * ncell, batch_max=a.batch)
    log(f"study: DNS {xf.size}x{yf.size}x{zf.size} ({xf.size * yf.size * zf.size / 1e6:.1f} M cells), "
        f"LES {nc}, {a.snapshots} snapshots every {a.every} steps")
    t0 = time.perf_counter()
    T.collect(fine, judge, cl, fac, d99, snapshots=a.snapshots + 1, every=a.every, smin=1e-3 / THETA,
              log=lambda m: log("study:", m))
    t_dns = time.perf_counter() - t0
    rows = cl.rows()
    train, held = rows[: a.snapshots * ncell], rows[a.snapshots * ncell:]
    nu_dsm = T.dsm_at_rows(judge)
    cl.clear()
    cl.add_rows(train)
    cl.params = T.init_params(0)
    t0 = time.perf_counter()
    hist = cl.train(a.train_steps, a.batch, lr=1e-3, seed=0)
    t_train = time.perf_counter() - t0
    ap_after = T.apriori(cl, held, nu_dsm)
    log("study: a priori", json.dumps(ap_after))
    # a posteriori: both LES from the held-out filtered field
    Uc, Vc, Wc = judge.fields()
    t_start = judge.t
    runs = {}
    for name, les in (("dynamic_smagorinsky", 2), ("learned", 4)):
        s = blgpu.Solver(xc, yc, zc, NU, tab_c, les=les, cfl=a.cfl, precision=blgpu.PREC_MIXED)
        s.set_fields(Uc, Vc, Wc)
        s.t = t_start
        if les == 4:
            cl.install(s, d99, a.nsigma)
        s.compute_dt()
        s.step()
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.les_steps - 1):
            s.compute_dt()
            s.step()
        s.synchronize()
        el = time.perf_counter() - t0
        # per-kernel-class times over 20 more steps (HIP events, outside the timed steps)
        s.profile(True)
        for _ in range(20):
            s.compute_dt()
            s.step()
        s.synchronize()
        kern = {k: v[0] / 20 for k, v in s.profile_read().items() if v[1] > 0}
        s.profile(False)
        runs[name] = dict(solver=s, ms=el / (a.les_steps - 1) * 1e3, kernels=kern)
        log(f"study: LES {name}: {runs[name]['ms']:.3f} ms/step, t = {s.t:.5f}")
    t_end = runs["learned"]["solver"].t
    n_dns = 0
    while fine.t < t_end - 1e-12:
        fine.compute_dt()
        fine.step()
        n_dns += 1
        if n_dns % 1000 == 0:
            log(f"study: DNS continued {n_dns} steps, t = {fine.t:.5f} of {t_end:.5f}")
    cf_dns = fine.wall_stats()[:: fac[0]]
    T.collect(fine, judge, cl, fac, d99, snapshots=1, every=0, smin=1e-3 / THETA)
    prof = lambda U: U[1:-1, 1:-1, 1:-2].mean(axis=(0, 2))
    u_dns = prof(judge.fields()[0])
    cells = nc[0] * nc[1] * nc[2]
    out = dict(grid_les=list(nc), grid_dns=[xf.size, yf.size, zf.size], noise=a.noise, seed=a.seed,
               snapshots=a.snapshots, every=a.every, dns_seconds=t_dns,
               training=dict(rows=int(train.shape[0]), weighted=int((train[:, T.ROW_W] > 0).sum()),
                             steps=a.train_steps, batch=a.batch, seconds=t_train,
                             loss_first=float(hist[:20].mean()), loss_last=float(hist[-20:].mean())),
               apriori_heldout=ap_after, les_steps=a.les_steps, les_time_span=t_end - t_start,
               dns_steps_to_match=n_dns)
    inner = slice(2, -2)
    for name, r in runs.items():
        s = r["solver"]
        cf = s.wall_stats()
        d = dict(ms_per_step=r["ms"], ns_per_cell_step=r["ms"] * 1e6 / cells, kernels_ms_per_step=r["kernels"],
                 cf_rel_l2_vs_dns=float(np.linalg.norm(cf[inner] - cf_dns[inner]) / np.linalg.norm(cf_dns[inner])),
                 mean_u_rel_l2_vs_filtered_dns=float(np.linalg.norm(prof(s.fields()[0]) - u_dns) /
                                                     np.linalg.norm(u_dns)))
        if name == "learned":
            d["cells"], d["fallback_cells"] = s.closure_fallback()
        r["cf"] = cf
        out[name] = d
    cl_, cd_ = runs["learned"]["cf"][inner], runs["dynamic_smagorinsky"]["cf"][inner]
    out["cf_rel_l2_learned_vs_dsm"] = float(np.linalg.norm(cl_ - cd_) / np.linalg.norm(cd_))
    out["cost_ratio_learned_over_dsm"] = out["learned"]["ms_per_step"] / out["dynamic_smagorinsky"]["ms_per_step"]
    ap_ = ap_after
    gain = (ap_["no_model"]["rel_error"] - ap_["learned"]["rel_error"])
    room = (ap_["no_model"]["rel_error"] - ap_["least_squares_floor"]["rel_error"])
    out["share_of_eddy_viscosity_floor_captured"] = gain / room if room > 0 else None
    for r in runs.values():
        r["solver"].close()
    cl.close()
    judge.close()
    fine.close()
    txt = json.dumps(out)
    print(txt, flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(txt + "\n")


if __name__ == "__main__":
    main()
