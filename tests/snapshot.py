"""Reader and writer of the reference's binary snapshot format (test helper).

Writer side (input_output.f90:496-666, single rank): t, nu (f64), -73, step
(i32), then six (n:i32, coords:f64[n]) blocks for x, y, z, xm, ym, zm, then
per field (n1,n2,n3:i32) and the data, with U, V, W and nu_t one z plane short
of their declared n3 and P complete.
Reader side (input_output.f90:288-487, single rank): the same header, then
full planes for U, V and W located by offsets from the end of the header.
"""
import numpy as np


def read_snapshot(path):
    b = open(path, "rb").read()
    o = 0

    def take(dt, n):
        nonlocal o
        a = np.frombuffer(b, dtype=dt, count=n, offset=o)
        o += a.nbytes
        return a

    t, nu = take("<f8", 2)
    first = take("<i4", 1)[0]
    step = None
    if first == -73:
        step = int(take("<i4", 1)[0])
        first = take("<i4", 1)[0]
    grids = {}
    n = first
    for name in ("x", "y", "z", "xm", "ym", "zm"):
        if name != "x":
            n = take("<i4", 1)[0]
        grids[name] = take("<f8", n).copy()
    fields = {}
    for name, short in (("U", 1), ("V", 1), ("W", 1), ("nu_t", 1), ("P", 0)):
        n1, n2, n3 = take("<i4", 3)
        n3w = n3 - short
        fields[name] = take("<f8", n1 * n2 * n3w).reshape((n1, n2, n3w), order="F").copy()
    assert o == len(b), (o, len(b))
    return dict(t=float(t), nu=float(nu), step=step, **grids, **fields)


def write_initial(path, t, nu, step, x, y, z, U, V, W):
    """File in the layout read_input_data reads: full planes."""
    xm = 0.5 * (x[1:] + x[:-1])
    ym = 0.5 * (y[1:] + y[:-1])
    zm = 0.5 * (z[1:] + z[:-1])
    with open(path, "wb") as fh:
        fh.write(np.array([t, nu], "<f8").tobytes())
        fh.write(np.array([-73, step], "<i4").tobytes())
        for a in (x, y, z, xm, ym, zm):
            fh.write(np.array([a.size], "<i4").tobytes())
            fh.write(np.asarray(a, "<f8").tobytes())
        for a in (U, V, W):
            fh.write(np.array(a.shape, "<i4").tobytes())
            fh.write(np.asfortranarray(a, dtype="<f8").tobytes(order="F"))


def read_stats(path, nx, ny):
    """Text statistics file (input_output.f90:703-740)."""
    lines = open(path).read().split("\n")
    head = lines[0].split()
    vals = np.array(" ".join(lines[1:]).split(), dtype=np.float64)
    o = 0
    out = {}
    for name, cnt in (("Cf", nx), ("x", nx), ("y", ny)):
        out[name] = vals[o:o + cnt]
        o += cnt
    for name in ("Umean", "Vmean", "Wmean", "U2mean", "V2mean", "W2mean", "Pmean", "P2mean"):
        out[name] = vals[o:o + nx * ny].reshape(nx, ny)
        o += nx * ny
    assert o == vals.size
    out["t"], out["nu"] = float(head[1]), float(head[2])
    out["istep"] = int(head[-1])
    return out
