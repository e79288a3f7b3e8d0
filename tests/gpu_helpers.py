"""Shared helpers of the GPU tests: run the product Solver and the oracle on
the same seeded inputs."""
import numpy as np

from oracle import euler as o_euler
from oracle import heat as o_heat
from oracle import plugins as o_plugins
from oracle import swept as o_swept
from pysweep_amd import Solver
from pysweep_amd.equations import path as eq_path


def gpu_solve(name, ic, globals_, k, s, b, swept=True, exact=True, history=False, tuning=None):
    solver = Solver(ic, sendWarning=False)
    solver.simulation = swept
    solver.operating = k
    solver.intermediate = s
    solver.blocksize = (b, b, 1)
    solver.share = 1
    solver.globals = list(globals_)
    solver.setCPU(eq_path(name + ".py"))
    solver.setGPU(eq_path(name + ".cu"))
    solver.exact = exact
    solver.history = history
    solver.tuning = dict(tuning or {})
    solver.loadCPUModule()
    solver()
    return solver


def swept_kernel(phases):
    """Solver tuning that runs a swept solve on the five phase kernels
    (phases=True) or on the band kernel, whatever the grid size."""
    return {"phases": True} if phases else {"band": True}


def heat_case(n, fe, steps, seed=0):
    dx = 1.0 / n
    dt = 0.1 * dx * dx
    ic = o_heat.initial_condition(n, seed=seed)
    g = [0.0, steps * dt, dt, dx, dx, 1.0, fe]
    c = o_heat.courant_py(1.0, dt, dx)
    step = (lambda q, qb, gg: o_heat.step_py(q[0], None if qb is None else qb[0], gg, c, fe)[None])
    return ic, g, step


def euler_case(n, steps, eps=5.0, xc=0.0, yc=0.0, sq=o_euler.sq_mul):
    dx = 10.0 / n
    dt = 0.1 * dx
    ic = o_euler.initial_condition(n, eps=eps, xc=xc, yc=yc)
    g = [0.0, steps * dt, dt, dx, dx, 1.4]
    step = lambda q, qb, gg: o_euler.step_py(q, qb, gg, 1.4, dt, dx, dx, sq=sq)
    return ic, g, step


def oracle_stages(step, ic, s, stages):
    ring, fin = o_swept.standard_solve(step, ic, s, stages)
    return ring.get(fin)


def normwise(a, b):
    """max |a-b| / max |b| per variable (the parity measure, tolerance 1e-12)."""
    return max(float(np.abs(a[v] - b[v]).max() / max(np.abs(b[v]).max(), 1e-300)) for v in range(a.shape[0]))


__all__ = ["gpu_solve", "heat_case", "euler_case", "oracle_stages", "normwise", "o_plugins", "o_swept",
           "o_euler", "o_heat", "swept_kernel"]
