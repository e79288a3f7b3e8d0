"""YAML input (the reference's io.yamlManager, core/io.py:35-75).

Keys and defaults: swept True, verbose False, intermediate_steps 1,
operating_points 1, blocksize 32 (must be a multiple of 2*operating_points;
stored as (b, b, 1)), share 0.9, exid [], globals [], cpu "", gpu ""
(made absolute), filename "output.hdf5", dtype "float64".  Extra keys of
this build: exact (False), history (False).
"""
import os

import numpy
import yaml


def yamlManager(solver):
    with open(solver.yamlFileName) as f:
        doc = yaml.safe_load(f) or {}
    solver.yamlFile = doc

    def get(key, default):
        return doc[key] if key in doc else default

    solver.simulation = get("swept", True)
    solver.verbose = get("verbose", False)
    solver.intermediate = get("intermediate_steps", 1)
    solver.operating = get("operating_points", 1)
    b = get("blocksize", 32)
    assert b % (2 * solver.operating) == 0, ("Invalid blocksize, blocksize must satisfy BS = 2*ops*k and the "
                                             "architectural limit where k is any integer factor.")
    solver.blocksize = (b, b, 1)
    solver.share = get("share", 0.9)
    solver.exid = get("exid", [])
    solver.globals = list(get("globals", []))
    solver.cpuStr = solver.cpu = get("cpu", "")
    solver.gpuStr = solver.gpu = get("gpu", "")
    if solver.cpu:
        solver.cpu = os.path.abspath(solver.cpu)
    if solver.gpu:
        solver.gpu = os.path.abspath(solver.gpu)
    solver.output = get("filename", "output.hdf5")
    solver.dtypeStr = get("dtype", "float64")
    solver.dtype = numpy.dtype(solver.dtypeStr)
    solver.exact = get("exact", False)
    solver.history = get("history", False)
    if solver.cpu:
        solver.loadCPUModule()
