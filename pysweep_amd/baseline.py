"""One-at-a-time plain solves: the comparison point of the batched and
served engines.  Each problem, in arrival order, gets its own ring and runs
every stage with the standard kernel (one launch per stage, no batching)."""
import time

import numpy

from pysweep_amd import device, hip, plugins, schedule
from pysweep_amd.batch import PLUGINS
from pysweep_amd.equations import path as eq_path

_cpu = {}


def solve_plain(p):
    py, cu64, cu32, nv, ops, _ = PLUGINS[p.plugin]
    cu = eq_path(cu64 if p.precision == "fp64" else cu32)
    real = "double" if p.precision == "fp64" else "float"
    nvv, nx, ny = p.ic.shape
    cpu = _cpu.get(p.plugin)
    if cpu is None:
        cpu = _cpu[p.plugin] = plugins.load_cpu_module(eq_path(py), name=f"pysweep_plain_{p.plugin}")
    run = device.StandardRun(cu, nv, ops, p.its, nx, ny, real=real)
    run.set_globals(cpu, p.globals)
    run.upload_ic(p.ic)
    N = schedule.requested_steps(p.globals[0], p.globals[1], p.globals[2])
    stages = p.its * N
    run.run(1, stages + 1)
    out = run.level(p.its - 1 + stages)
    run.close()
    return out


def serve_one_at_a_time(stream, clock=time.perf_counter):
    """stream: list of (arrival seconds, Problem).  Returns list of
    (arrival, finish, result) with the requests served in arrival order."""
    hip.ensure_init(0)
    t0 = clock()
    out = []
    for arrival, p in sorted(stream, key=lambda x: x[0]):
        now = clock() - t0
        if now < arrival:
            time.sleep(arrival - now)
        res = solve_plain(p)
        out.append((arrival, clock() - t0, res))
    return out
