"""Verbose solve summary and log.yaml records (core/io.py:10-33)."""
import os
from datetime import datetime

import numpy
import yaml


def solver_print(solver):
    """The verbose summary of a solve (the reference prints its own banner,
    core/io.py:102-132; this one is laid out for one GPU)."""
    method = "swept" if solver.simulation else "standard"
    run = getattr(solver, "_run", None)
    spec = getattr(run, "spec", None)
    if solver.simulation and getattr(spec, "kind", None) == "band":
        method += f" (band kernel, {spec.levels} levels per launch, {run.rows} rows per wavefront)"
    elif solver.simulation and spec is not None:
        method += " (phase kernels)"
    rows = [("method", f"{method}, blocksize {solver.blocksize[0]}, {solver.operating} operating points "
                       f"({2 * solver.operating + 1}-point cross), {solver.intermediate} intermediate steps"),
            ("placement", f"share {solver.share} of the block rows on the GPU, the rest on the host"
                          if solver.share < 1 else "whole grid on one GPU"),
            ("state", f"{getattr(solver, 'arrayShape', None)} {_dtype_name(solver.dtype)}"),
            ("plugins", f"{solver.cpuStr} | {solver.gpuStr}")]
    if len(solver.globals) >= 3:
        t0, tf, dt = solver.globals[:3]
        rows.append(("time", f"t0 {t0}, tf {tf}, dt {dt}"))
    if solver.globals[3:]:
        rows.append(("plugin globals", ", ".join(str(g) for g in solver.globals[3:])))
    rows.append(("output", solver.output or "none"))
    if solver.exid:
        rows.append(("excluded devices", ", ".join(str(e) for e in solver.exid)))
    width = max(len(k) for k, _ in rows)
    return "\npysweep_amd solve\n" + "".join(f"  {k:<{width}}  {v}\n" for k, v in rows)


def _dtype_name(dtype):
    try:
        return numpy.dtype(dtype).name
    except TypeError:
        return str(dtype)


def short_entry(solver, clocktime):
    """getShortEntry (core/io.py:29-33) plus GCUPS and the device solve time."""
    shape = list(getattr(solver, "arrayShape", ()))
    data = {"runtime": float(clocktime), "swept": bool(solver.simulation), "blocksize": int(solver.blocksize[0]),
            "share": float(solver.share), "cpu": str(solver.cpuStr), "array_shape": shape,
            "time_per_step": float(clocktime / max(shape[0] - 1, 1)) if shape else None}
    if getattr(solver, "solve_seconds", None):
        data["solve_seconds"] = float(solver.solve_seconds)
        data["gcups"] = float(solver.gcups)
    return {datetime.now().strftime("%m/%d/%Y-%H:%M:%S.%f"): data}


def update_log_file(solver, clocktime, path="log.yaml"):
    """Append a record to log.yaml (core/io.py:10-22)."""
    prev = {}
    if os.path.isfile(path):
        with open(path) as f:
            prev = yaml.safe_load(f) or {}
    prev.update(short_entry(solver, clocktime))
    with open(path, "w") as f:
        yaml.safe_dump(prev, f)
