"""Write golden input/output vectors from the reference's own numpy plugins.

Run in the build container only (the reference tree is not available where
the tests run):

    python tests/golden/make_golden.py /root/reference

It loads pysweep/equations/heat.py and pysweep/equations/euler.py by file
path.  Those modules import mpi4py, h5py and pysweep.equations.sodShock at
module level; none of them is used by ``set_globals``/``step``, and mpi4py
and h5py are not installed here, so empty placeholder modules are entered in
``sys.modules`` before the import.  Only ``set_globals`` and ``step`` are
called.  Outputs (small .npz fixtures, inputs and expected outputs only):

* heat_step.npz  -- forward-Euler and RK2 stages on a seeded random field;
* euler_step.npz -- intermediate and final RK2 stages on (a) a seeded random
  perturbation of a uniform state and (b) a coarse isentropic vortex.
"""
import importlib.util
import os
import sys
import types

import numpy as np


def _placeholders():
    for name in ("mpi4py", "mpi4py.MPI", "h5py", "pysweep", "pysweep.equations",
                 "pysweep.equations.sodShock"):
        if name not in sys.modules:
            sys.modules[name] = types.ModuleType(name)
    sys.modules["mpi4py"].MPI = sys.modules["mpi4py.MPI"]
    sys.modules["mpi4py"].rc = types.SimpleNamespace(recv_mprobe=False)
    sys.modules["pysweep"].equations = sys.modules["pysweep.equations"]
    sys.modules["pysweep.equations"].sodShock = sys.modules["pysweep.equations.sodShock"]


def _load(path, name):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main(ref_root):
    _placeholders()
    eq = os.path.join(ref_root, "pysweep", "equations")
    out = os.path.dirname(os.path.abspath(__file__))
    rng = np.random.default_rng(20240611)

    # ---------------- heat ----------------
    heat = _load(os.path.join(eq, "heat.py"), "ref_heat")
    n = 12
    alpha, dx = 0.7, 1.0 / n
    dt = 0.1 * dx * dx / alpha
    iidx = [(i, j) for i in range(n) for j in range(n)]
    base = rng.uniform(-1.0, 1.0, size=(3, 1, n, n))
    rec = {"alpha": alpha, "dx": dx, "dt": dt}
    for scheme, tag in ((True, "fe"), (False, "rk2")):
        heat.set_globals(0.0, 1.0, dt, dx, dx, alpha, scheme)
        for gts in ((1,) if scheme else (1, 2)):
            st = base.copy()
            heat.step(st, iidx, 1, gts)
            rec[f"{tag}_g{gts}_in"] = base
            rec[f"{tag}_g{gts}_out"] = st[2]
    np.savez_compressed(os.path.join(out, "heat_step.npz"), **rec)

    # ---------------- euler ----------------
    euler = _load(os.path.join(eq, "euler.py"), "ref_euler")
    gamma = 1.4
    rec = {"gamma": gamma}
    cases = {}
    n = 12
    dx = 10.0 / n
    dt = 0.1 * dx
    q = np.empty((3, 4, n, n))
    for lvl in range(3):
        q[lvl, 0] = 1.0 + 0.2 * rng.uniform(-1, 1, (n, n))
        q[lvl, 1] = 0.3 + 0.2 * rng.uniform(-1, 1, (n, n))
        q[lvl, 2] = -0.2 + 0.2 * rng.uniform(-1, 1, (n, n))
        q[lvl, 3] = 2.5 + 0.3 * rng.uniform(-1, 1, (n, n))
    cases["random"] = (q, dx, dt)
    n = 16
    dx = 10.0 / n
    dt = 0.1 * dx
    xs = np.linspace(-5, 5, n, endpoint=False)
    v = np.empty((4, n, n))
    for i, x in enumerate(xs):
        for j, y in enumerate(xs):
            v[:, i, j] = euler.analytical(x, y, 0.0, gamma=gamma)
    q = np.stack([v, v, v])
    cases["vortex"] = (q, dx, dt)
    for tag, (q, dx, dt) in cases.items():
        n = q.shape[-1]
        iidx = [(i, j) for i in range(n) for j in range(n)]
        euler.set_globals(0.0, 1.0, dt, dx, dx, gamma)
        rec[f"{tag}_dx"] = dx
        rec[f"{tag}_dt"] = dt
        rec[f"{tag}_in"] = q
        for gts in (1, 2):
            st = q.copy()
            euler.step(st, iidx, 1, gts)
            rec[f"{tag}_g{gts}_out"] = st[2]
    np.savez_compressed(os.path.join(out, "euler_step.npz"), **rec)
    print("wrote", os.path.join(out, "heat_step.npz"), os.path.join(out, "euler_step.npz"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
