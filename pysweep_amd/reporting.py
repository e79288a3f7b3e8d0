"""Verbose banner and log.yaml records (core/io.py:10-33, 102-132)."""
import os
from datetime import datetime

import yaml


def solver_print(solver):
    s = "\nPysweep simulation properties:\n"
    s += "\tsimulation type: {}\n".format("swept" if solver.simulation else "standard")
    s += "\tnumber of processes: 1\n"
    s += "\tgpu source: {}\n".format(solver.gpuStr)
    s += "\tcpu source: {}\n".format(solver.cpuStr)
    s += "\toutput file: {}\n".format(solver.output)
    s += "\tverbose: {}\n".format(solver.verbose)
    s += "\tArray shape: {}\n".format(getattr(solver, "arrayShape", None))
    s += "\tdtype: {}\n".format(solver.dtype)
    s += "\tshare: {}\n".format(solver.share)
    s += "\tblocksize: {}\n".format(solver.blocksize[0])
    s += "\tstencil size: {}\n".format(int(solver.operating * 2 + 1))
    s += "\tintermediate time steps: {}\n".format(solver.intermediate)
    if len(solver.globals) >= 3:
        s += "\ttime data (t0,tf,dt): ({},{},{})\n".format(*solver.globals[:3])
    if solver.globals[5:]:
        s += "\tglobals: [" + ",".join(str(g) for g in solver.globals) + "]\n"
    if solver.exid:
        s += "\texcluded gpus: [" + ",".join(str(e) for e in solver.exid) + "]\n"
    return s


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
