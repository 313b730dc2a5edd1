"""bl_max_divergence_noalloc (no device allocation; its working memory is a
file-scope device array) against bl_max_divergence on the same fields:
check_divergence (projection.f90:391-413) of caller-owned device arrays in
the reference's layout.  Two calls in a row, the second on a smaller grid,
and a grid past the scratch limit, which the host rejects before any launch.
The device arrays are the caller's: allocated here through the HIP runtime
that the library itself uses (the one already mapped into the process)."""
import ctypes
import sys

import numpy as np
import pytest

from . import cases
from .blgpu_import import blgpu, solver_for

pytestmark = pytest.mark.gpu

_D = ctypes.c_double
_P = ctypes.c_void_p
_I = ctypes.c_int


def noalloc_fn():
    f = blgpu.lib().bl_max_divergence_noalloc
    f.restype = _I
    f.argtypes = [_P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P]
    return f


def hip_runtime():
    """the HIP runtime library the solver library loaded"""
    blgpu.lib()
    paths = [ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln and "/" in ln]
    assert paths, "libamdhip64 is not mapped"
    h = ctypes.CDLL(paths[0])
    h.hipMalloc.argtypes = [ctypes.POINTER(_P), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [_P, _P, ctypes.c_size_t, _I]
    h.hipFree.argtypes = [_P]
    h.hipDeviceSynchronize.argtypes = []
    for fn in (h.hipMalloc, h.hipMemcpy, h.hipFree, h.hipDeviceSynchronize):
        fn.restype = _I
    return h


class DevArray:
    """a device copy of a host float64 array (Fortran order), or n zeros"""

    def __init__(self, hip, a=None, n=None):
        self.hip = hip
        self.a = np.asfortranarray(a, dtype=np.float64).ravel(order="F") if a is not None else np.zeros(n)
        self.p = _P()
        assert hip.hipMalloc(ctypes.byref(self.p), self.a.nbytes) == 0
        assert hip.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.a)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = _P()


def noalloc_div(hip, U, V, W, x, y, z):
    t = [DevArray(hip, a) for a in (U, V, W, x, y, z)]
    out = DevArray(hip, np.array([-1.0]))
    rc = noalloc_fn()(*[a.p for a in t[:3]], x.size, y.size, z.size, *[a.p for a in t[3:]], out.p, None)
    assert rc == 0, blgpu.lib().bl_last_error().decode()
    assert hip.hipDeviceSynchronize() == 0
    v = float(out.get()[0])
    for a in t + [out]:
        a.free()
    return v


def test_noalloc_matches_max_divergence():
    hip = hip_runtime()
    for shape, seed in (((70, 41, 23), 1), ((24, 16, 9), 2)):
        nx, ny, nz = shape
        c = cases.laminar_small(nx=nx, ny=ny, nz=nz, noise=0.01, seed=seed)
        s, _ = solver_for(c)
        ref = s.max_divergence()
        got = noalloc_div(hip, *s.fields(), c.x, c.y, c.z)
        print("noalloc", shape, "max div", got, "bl_max_divergence", ref, flush=True)
        sys.stdout.flush()
        assert got == pytest.approx(ref, rel=1e-15, abs=0.0)
        s.close()


def test_noalloc_rejects_oversize_grid():
    hip = hip_runtime()
    small = DevArray(hip, n=8)
    out = DevArray(hip, n=1)
    p = small.p
    # a 4000 x 1000 x 2000 grid needs more reduction partials than the
    # scratch holds: rejected on the host, nothing is launched
    rc = noalloc_fn()(p, p, p, 4000, 1000, 2000, p, p, p, out.p, None)
    msg = blgpu.lib().bl_last_error().decode()
    print("oversize:", rc, msg, flush=True)
    small.free()
    out.free()
    assert rc == 1 and "scratch" in msg
