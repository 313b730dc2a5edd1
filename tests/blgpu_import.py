"""Import helper for the product's Python binding (boundary-layer-gpu_amd/blgpu.py)."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "boundary-layer-gpu_amd")


def load():
    spec = importlib.util.spec_from_file_location("blgpu", os.path.join(PKG, "blgpu.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


blgpu = load()


def solver_for(case, oracle=None, **kw):
    """Build a device solver with the same configuration and initial state as
    case.oracle()."""
    o = oracle if oracle is not None else case.oracle()
    args = dict(dPdx=case.dPdx, dPdz=case.dPdz, scheme=case.scheme, les=case.les, cfl=case.cfl,
                modes=case.modes)
    args.update(kw)
    s = blgpu.Solver(case.x, case.y, case.z, case.nu, case.tables, **args)
    s.set_fields(*o.fields())
    s.t = o.t
    return s, o
