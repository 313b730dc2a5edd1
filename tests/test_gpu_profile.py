"""Per-kernel HIP-event profile (bl_profile_enable / bl_profile_select / bl_profile_read).

An RK3 step launches 3 RK stages and 3 projections (each with one x
transform, one Thomas solve and one velocity correction, which also records
the pressure in the first); the profile counts those launches per class,
and a selected profile counts only the selected classes while the step's
results stay the same."""
import sys

import numpy as np
import pytest

from . import cases
from .blgpu_import import solver_for

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_profile_counts_and_selection():
    c = cases.laminar_small(nx=24, ny=16, nz=8, noise=0.01, seed=1)
    say("profile case", c.x.size, c.y.size, c.z.size)
    s_all, o = solver_for(c)
    s_sel, _ = solver_for(c, o)
    s_all.profile(True)
    s_sel.profile(True, only=["rk_stage"])
    for s in (s_all, s_sel):
        s.compute_dt()
        s.step()
    p_all, p_sel = s_all.profile_read(), s_sel.profile_read()
    say("all", {k: v[1] for k, v in p_all.items() if v[1]})
    say("selected", {k: v[1] for k, v in p_sel.items() if v[1]})
    assert p_all["rk_stage"][1] == 3
    assert p_all["xform_x_fwd"][1] == 3 and p_all["thomas"][1] == 3 and p_all["project"][1] == 3
    assert all(v[0] >= 0.0 for v in p_all.values())
    assert p_sel["rk_stage"][1] == 3 and p_sel["rk_stage"][0] > 0.0
    assert all(v[1] == 0 for k, v in p_sel.items() if k != "rk_stage")
    for a, b in zip(s_all.fields(), s_sel.fields()):
        assert np.array_equal(a, b)
    s_all.profile(False)
    s_sel.profile(False)
