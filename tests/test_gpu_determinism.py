"""Run-to-run reproducibility: two solvers built from the same state give
bit-identical fields after RK3 steps, in both precisions and with the DSM
eddy viscosity.  Every kernel writes each output element from one lane and
reduces in a fixed order (no floating-point atomics), so nothing depends on
scheduling.  Regression: the x transforms once wrote the outputs of the
self-mirrored frequencies (k1 = 0 and k1 = N1/2) from two lanes with values
equal only up to rounding, and the result depended on which write landed
last."""
import sys

import numpy as np
import pytest

from . import cases
from .blgpu_import import blgpu, solver_for

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision,les", [(blgpu.PREC_FP64, 0), (blgpu.PREC_MIXED, 0), (blgpu.PREC_FP64, 2)])
def test_bitwise_reproducible(precision, les):
    c = cases.laminar_small(nx=26, ny=14, nz=10, noise=0.02, seed=3, scheme=3, les=les)
    say("determinism case", precision, les, c.x.size, c.y.size, c.z.size)
    o = c.oracle()
    runs = []
    for _ in range(2):
        s, _ = solver_for(c, o, precision=precision)
        for _ in range(3):
            s.compute_dt()
            s.step()
        runs.append(s.fields() + (s.download(blgpu.FIELD_P),))
    for name, a, b in zip("UVWP", *runs):
        say(name, "max |diff|", float(np.abs(a - b).max()))
        assert np.array_equal(a, b), name
