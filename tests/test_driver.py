"""Fortran driver (bin/bl_driver): the reference's main program over the C ABI.

The driver reads a .turbb parameter file (params.f90), builds or reads the
initial field (input_output.f90:130-487), runs the time loop of main.f90:84-113
through bl_compute_dt/bl_step, and writes the statistics text files and the
binary snapshots in the reference's formats (input_output.f90:496-746).  The
GPU tests compare its snapshots and statistics with the FP64 oracle run on the
same case; the CPU tests cover parameter checking and the clean failure when
no device is present.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests import orc
from tests.cases import GOLDEN, Case, load_temporal_modes
from tests.snapshot import read_snapshot, read_stats, write_initial

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "boundary-layer-gpu_amd", "bin", "bl_driver")

needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="bl_driver not built (run build())")

BASE = dict(nxyz="40 20 8", CFL="-2e-4", nu="1e-5", RKscheme="3", LES="0", WM="0", dPdx="0.0", dPdz="0.0",
            top_flag="0", nsteps="20", nsave="20", nstats="10", nmonitor="10", init_rand="1",
            boxsize="0.2 0.02 0.01 3.0", init_step="0", fileout="./data/BL_lam", inflow_flag="1",
            inflow_file="blasius_solution.dat", timeinflow_file="**TO DO**")


def write_case(d, **over):
    p = dict(BASE)
    p.update(over)
    shutil.copy(os.path.join(GOLDEN, "blasius_solution.dat"), os.path.join(d, "blasius_solution.dat"))
    with open(os.path.join(d, "case.turbb"), "w") as fh:
        fh.write("! driver test case\n")
        for k, v in p.items():
            if v is not None:
                fh.write(f"{k} = {v}\n")
    return p


def run_driver(d, timeout=120):
    print(f"bl_driver in {d}", flush=True)
    return subprocess.run([DRIVER, "-i", "case.turbb"], cwd=d, capture_output=True, text=True, timeout=timeout)


def stats_from_fields(x, y, nu, U, V, W, P):
    """Means of statistics.f90:44-176 over the interior z planes."""
    nx, ny = x.size, y.size
    _, yg = orc.centers(y)
    wy0 = (yg[1:] - y) / (yg[1:] - yg[:-1])
    wy1 = 1.0 - wy0
    nzg = U.shape[2]
    nz = W.shape[2]
    ks = slice(1, nzg - 1)
    tu = wy0[None, :, None] * U[:, :ny, ks] + wy1[None, :, None] * U[:, 1:ny + 1, ks]
    tv = 0.5 * (V[:nx, :, ks] + V[1:nx + 1, :, ks])
    wx = 0.5 * (W[:nx, :, :nz - 2] + W[1:nx + 1, :, :nz - 2])
    wf = wy0[None, :, None] * wx[:, :ny] + wy1[None, :, None] * wx[:, 1:ny + 1]
    return dict(Umean=tu.mean(2), Vmean=tv.mean(2), Wmean=wf.mean(2), U2mean=(tu ** 2).mean(2),
                V2mean=(tv ** 2).mean(2), W2mean=(wf ** 2).mean(2), Pmean=P[:nx, :ny, ks].mean(2),
                P2mean=(P[:nx, :ny, ks] ** 2).mean(2))


def check_snapshot(snap, o, x, y, z, t, step, rtol=1e-10):
    assert snap["step"] == step
    assert snap["t"] == pytest.approx(t, rel=1e-13)
    for name, a in (("x", x), ("y", y), ("z", z)):
        np.testing.assert_array_equal(snap[name], a)
    np.testing.assert_allclose(snap["xm"], 0.5 * (x[1:] + x[:-1]), rtol=0, atol=1e-15)
    U, V, W = o.fields()
    scale = np.abs(U).max()  # velocity scale (W is exactly zero in a laminar case)
    for name, ref in (("U", U), ("V", V), ("W", W)):
        got = snap[name]
        assert got.shape == ref.shape[:2] + (ref.shape[2] - 1,)
        err = np.abs(got - ref[:, :, :-1]).max() / scale
        assert err < rtol, (name, err)
    P = o.get(3)
    assert snap["P"].shape == P.shape
    assert np.abs(snap["P"] - P).max() <= rtol * np.abs(P).max() + 1e-14
    assert not np.any(snap["nu_t"])


@pytest.mark.gpu
@needs_driver
def test_driver_fresh_start_matches_oracle(tmp_path):
    d = str(tmp_path)
    write_case(d)
    r = run_driver(d)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Maximum divergence" in r.stdout

    case = Case(40, 20, 8, 0.2, 0.02, 0.01, 1e-5, alpha=3.0, grid="sinh", scheme=3, cfl=-2e-4)
    o = case.oracle()
    stats_steps = {1, 10, 20}
    for istep in range(1, 21):
        o.compute_dt()
        o.step()
        if istep in stats_steps:
            s = read_stats(os.path.join(d, "data", f"BL_lam.{istep:08d}.stats.txt"), 40, 20)
            assert s["istep"] == istep
            assert s["t"] == pytest.approx(o.t, abs=1e-8)
            # values are written with F15.8: compare to that resolution
            np.testing.assert_allclose(s["Cf"], o.wall_stats(), rtol=0, atol=1e-8)
            np.testing.assert_allclose(s["x"], case.x, rtol=0, atol=1e-8)
            ref = stats_from_fields(case.x, case.y, case.nu, *o.fields(), o.get(3))
            for k, v in ref.items():
                np.testing.assert_allclose(s[k], v, rtol=0, atol=2e-8, err_msg=k)
    snap = read_snapshot(os.path.join(d, "data", "BL_lam.00000020"))
    check_snapshot(snap, o, case.x, case.y, case.z, o.t, 20)
    assert os.path.islink(os.path.join(d, "data", "BL_lam.restart"))


@pytest.mark.gpu
@needs_driver
def test_driver_dispatcher_flags(tmp_path):
    """The .turbb keys of the general dispatcher (params.f90:122-162):
    top_flag = 1 with Vmax, x0, sigma, phi and Lund_ix, and inflow_flag = 4
    with Amplitude, reach bl_set_boundary; the snapshot after 20 steps
    matches the oracle with the same flags (seed 0)."""
    d = str(tmp_path)
    write_case(d, top_flag="1", Vmax="0.05", x0="1.1", sigma="0.02", phi="0.0", Lund_ix="3", inflow_flag="4",
               Amplitude="0.02")
    r = run_driver(d)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    case = Case(40, 20, 8, 0.2, 0.02, 0.01, 1e-5, alpha=3.0, grid="sinh", scheme=3, cfl=-2e-4)
    o = case.oracle()
    o.set_boundary(inflow_flag=4, amplitude=0.02, seed=0, top_flag=1, vbs_max=0.05, x_bs=1.1, sigma_bs=0.02,
                   phi_bs=0.0, i_rescale=3)
    for _ in range(20):
        o.compute_dt()
        o.step()
    snap = read_snapshot(os.path.join(d, "data", "BL_lam.00000020"))
    U, V, W = o.fields()
    scale = np.abs(U).max()
    for name, ref in (("U", U), ("V", V), ("W", W)):
        err = np.abs(snap[name] - ref[:, :, :-1]).max() / scale
        print("driver dispatcher flags", name, err, flush=True)
        assert err < 1e-10, name
    o.close()


@pytest.mark.gpu
@needs_driver
def test_driver_wall_model(tmp_path):
    """The wall-model keys (params.f90:154-155, 198-206): WM = 5 with
    alphas (alpha_mean_x, _y, _z, alpha_std, freq_mult) and dPdx reach
    bl_set_wall_model; the Robin / pseudo-Robin wall runs every stage and the
    snapshot after 20 steps matches the oracle with the same model."""
    d = str(tmp_path)
    write_case(d, WM="5", TauwModel="1", alphas="3e-4 2e-4 4e-4 0.3 0.2", dPdx="4e-4")
    r = run_driver(d)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    case = Case(40, 20, 8, 0.2, 0.02, 0.01, 1e-5, alpha=3.0, grid="sinh", scheme=3, cfl=-2e-4, dPdx=4e-4)
    o = case.oracle()
    o.set_wall_model(5, istress_model=1, alpha_mean_x=3e-4, alpha_mean_y=2e-4, alpha_mean_z=4e-4, alpha_std=0.3,
                     freq_mult=0.2)
    for _ in range(20):
        o.compute_dt()
        o.step()
    snap = read_snapshot(os.path.join(d, "data", "BL_lam.00000020"))
    U, V, W = o.fields()
    assert np.abs(V[:, 0, :]).max() > 0
    scale = np.abs(U).max()
    for name, ref in (("U", U), ("V", V), ("W", W)):
        err = np.abs(snap[name] - ref[:, :, :-1]).max() / scale
        print("driver wall model", name, err, flush=True)
        assert err < 1e-10, name
    o.close()


@pytest.mark.gpu
@needs_driver
def test_driver_lund_fluctuation_inflow(tmp_path):
    """inflow_flag = 5 from a restart file: the turbulent tables of
    compute_turbulent_solution_for_bc (boundary_conditions.f90:1371-1425: the
    inlet profile read from the inflow_file stream, U_top = 1, the 1/7-law
    V_top) and the Lund keys (Lund_ix, Lund_deltai, Lund_T) reach the
    library; the snapshot after 6 steps matches the oracle with the same
    tables and flags."""
    d = str(tmp_path)
    case = Case(40, 20, 8, 0.2, 0.02, 0.01, 1e-5, alpha=3.0, grid="sinh", noise=0.03, seed=8, scheme=3,
                cfl=-2e-4)
    t0 = 0.05
    write_initial(os.path.join(d, "start.bin"), t0, case.nu, 10, case.x, case.y, case.z, case.U0, case.V0, case.W0)
    nyg, ny = case.y.size + 1, case.y.size
    uin = case.tables["U_inlet"] * 1.01
    vin = case.tables["V_inlet"] * 0.9
    with open(os.path.join(d, "profile.bin"), "wb") as fh:
        fh.write(np.array([ny], "<i4").tobytes())
        fh.write(np.asarray(uin, "<f8").tobytes())
        fh.write(np.asarray(vin, "<f8").tobytes())
    write_case(d, init_rand="0", filein="start.bin", init_step="-45", nsteps="6", nsave="6", nstats="6",
               nmonitor="6", inflow_flag="5", inflow_file="profile.bin", Lund_ix="20", Lund_deltai="0.004",
               Lund_T="0.4")
    r = run_driver(d)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    o = orc.Oracle(case.x, case.y, case.z, case.nu, 0.0, 0.0, 3, 0, -2e-4)
    _, xg = orc.centers(case.x)
    tab = dict(U_inlet=uin, V_inlet=vin, W_inlet=np.zeros(nyg), U_top=np.ones(case.x.size),
               V_top=0.0160 * (xg / case.nu) ** (-1.0 / 7.0), W_top=np.zeros(xg.size))
    o.set_bc_tables(tab)
    o.set_fields(case.U0, case.V0, case.W0)
    o.periodic()
    o.t = t0
    o.set_boundary(inflow_flag=5, i_rescale=20, delta_inlet=0.004, t_resc=0.4)
    for _ in range(6):
        o.compute_dt()
        o.step()
    snap = read_snapshot(os.path.join(d, "data", "BL_lam.00000016"))
    U, V, W = o.fields()
    scale = np.abs(U).max()
    for name, ref in (("U", U), ("V", V), ("W", W)):
        err = np.abs(snap[name] - ref[:, :, :-1]).max() / scale
        print("driver lund", name, err, flush=True)
        assert err < 1e-10, name
    o.close()


@pytest.mark.gpu
@needs_driver
def test_driver_mixed_precision(tmp_path):
    """The optional `precision = 1` key runs the mixed-precision path: the
    snapshot after 20 steps agrees with the FP64 oracle to FP32 accuracy."""
    d = str(tmp_path)
    write_case(d, precision="1")
    r = run_driver(d)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    case = Case(40, 20, 8, 0.2, 0.02, 0.01, 1e-5, alpha=3.0, grid="sinh", scheme=3, cfl=-2e-4)
    o = case.oracle()
    for istep in range(1, 21):
        o.compute_dt()
        o.step()
    snap = read_snapshot(os.path.join(d, "data", "BL_lam.00000020"))
    U = o.fields()[0]
    for name, ref in zip("UVW", o.fields()):
        e = np.linalg.norm(snap[name] - ref[:, :, :-1]) / np.linalg.norm(U[:, :, :-1])
        print("driver mixed", name, e, flush=True)
        assert e < 1e-4, name


@pytest.mark.gpu
@needs_driver
def test_driver_restart_rk2_temporal_modes(tmp_path):
    """Restart from a full-plane input file (init_rand = 0), step taken from
    the file (init_step = -45), RK2, temporal inflow modes read from the
    reference's transition-test mode file."""
    d = str(tmp_path)
    case = Case(40, 20, 8, 0.2, 0.02, 0.01, 1e-5, alpha=3.0, grid="sinh", noise=0.02, seed=5, scheme=2,
                cfl=-1e-4, modes=True)
    t0 = 0.125
    write_initial(os.path.join(d, "start.bin"), t0, case.nu, 100, case.x, case.y, case.z, case.U0, case.V0,
                  case.W0)
    shutil.copy(os.path.join(GOLDEN, "temporal_modes_inlet.dat"), os.path.join(d, "modes.dat"))
    write_case(d, init_rand="0", filein="start.bin", init_step="-45", RKscheme="2", CFL="-1e-4", nsteps="6",
               nsave="3", nstats="3", nmonitor="3", timeinflow_file="modes.dat")
    r = run_driver(d)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Reading init step from file" in r.stdout

    o = case.oracle()
    o.t = t0
    for istep in range(1, 7):
        o.compute_dt()
        o.step()
        if istep in (3, 6):
            snap = read_snapshot(os.path.join(d, "data", f"BL_lam.{100 + istep:08d}"))
            check_snapshot(snap, o, case.x, case.y, case.z, o.t, 100 + istep)


@needs_driver
def test_driver_requires_input_file(tmp_path):
    r = subprocess.run([DRIVER], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert "No input file provided" in r.stdout + r.stderr


@needs_driver
@pytest.mark.parametrize("key,val,msg", [
    ("RKscheme", "1", "only RKscheme = 2 or 3"),
    ("WM", "1", "WM must be 0, 4, 5, 6, 8, 9, 11, 12 or 15"),
    ("LES", "5", "LES must be 0, 1, 2 or 3"),
    ("inflow_flag", "3", "Lund inflow (3, 5) starts from a flow field file"),
    ("top_flag", "5", "top_flag"),
    ("precision", "2", "precision must be 0"),
])
def test_driver_rejects_unbuilt_options(tmp_path, key, val, msg):
    d = str(tmp_path)
    write_case(d, **{key: val})
    r = run_driver(d, timeout=60)
    assert msg in r.stdout + r.stderr


@needs_driver
def test_driver_rejects_filein_with_random_init(tmp_path):
    d = str(tmp_path)
    write_case(d, filein="x.bin")
    r = run_driver(d, timeout=60)
    assert "init_rand = 1, but filein exists" in r.stdout + r.stderr


@needs_driver
@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_driver_fails_cleanly_without_device(tmp_path):
    d = str(tmp_path)
    write_case(d)
    r = run_driver(d, timeout=60)
    assert r.returncode != 0
    assert "blgpu error in bl_create" in r.stdout + r.stderr


def test_snapshot_roundtrip(tmp_path):
    """write_initial and read_snapshot agree on the header layout."""
    case = Case(12, 8, 6, 0.2, 0.02, 0.01, 1e-5, alpha=3.0, grid="sinh")
    p = os.path.join(str(tmp_path), "s.bin")
    write_initial(p, 0.5, 1e-5, 7, case.x, case.y, case.z, case.U0, case.V0, case.W0)
    b = open(p, "rb").read()
    hdr = 16 + 8 + 6 * 4 + 8 * (2 * (12 + 8 + 6) - 3)
    assert len(b) == hdr + 3 * 12 + 8 * (case.U0.size + case.V0.size + case.W0.size)
    # the first field block starts with the U extents
    assert tuple(np.frombuffer(b, "<i4", 3, hdr)) == case.U0.shape
    tm = load_temporal_modes()
    assert tm["n"] > 0
