"""HDF5 output in the reference's layout (core/io.py:78-90) and initial
conditions read from dataset "data" (core/solver.py:89-101).  Files are read
back with h5py where an interpreter has it (the reference's own reader)."""
import os
import subprocess
import types

import numpy as np
import pytest

from pysweep_amd import hdf5io

CONDA_PY = "/opt/conda/bin/python3.9"
HAVE_H5PY_PY = os.path.exists(CONDA_PY) and subprocess.run(
    [CONDA_PY, "-c", "import h5py"], capture_output=True).returncode == 0

pytestmark = pytest.mark.skipif(hdf5io.native() is None, reason="HDF5 C library not installed")


def _solver(tmp_path):
    rng = np.random.default_rng(5)
    return types.SimpleNamespace(blocksize=(16, 16, 1), share=1.0, exid=[], globals=[0.0, 0.5, 0.01, 0.1, 0.1, 1.4],
                                 runtime=1.25, data=rng.standard_normal((3, 4, 16, 8)), dtype=np.dtype("float64"))


def test_output_layout_round_trip(tmp_path):
    s = _solver(tmp_path)
    path = hdf5io.write_output(s, str(tmp_path / "out.hdf5"))
    assert path.endswith(".hdf5")
    assert np.array_equal(hdf5io.read_dataset(path, "data"), s.data)
    assert np.array_equal(hdf5io.read_dataset(path, "globals"), np.array(s.globals))
    assert hdf5io.read_dataset(path, "blocksize").tolist() == [16.0]


@pytest.mark.skipif(not HAVE_H5PY_PY, reason="no interpreter with h5py")
def test_output_read_by_h5py(tmp_path):
    s = _solver(tmp_path)
    path = hdf5io.write_output(s, str(tmp_path / "out.hdf5"))
    code = ("import h5py, numpy, sys\n"
            "f = h5py.File(sys.argv[1], 'r')\n"
            "print(sorted(f.keys()))\n"
            "for k in sorted(f.keys()): print(k, f[k].dtype, f[k].shape)\n"
            "numpy.save(sys.argv[2], f['data'][...])\n")
    npy = str(tmp_path / "data.npy")
    out = subprocess.run([CONDA_PY, "-c", code, path, npy], capture_output=True, text=True, check=True).stdout
    lines = out.strip().splitlines()
    assert lines[0] == "['blocksize', 'clocktime', 'data', 'exid', 'globals', 'share']"
    assert "blocksize int64 (1,)" in lines and "share float64 (1,)" in lines and "exid float64 (0,)" in lines
    assert "globals float64 (6,)" in lines and "clocktime float64 (1,)" in lines
    assert "data float64 (3, 4, 16, 8)" in lines
    assert np.array_equal(np.load(npy), s.data)


@pytest.mark.skipif(not HAVE_H5PY_PY, reason="no interpreter with h5py")
def test_initial_conditions_from_reference_style_file(tmp_path):
    """An IC file as the reference writes it (h5py's default float32 dataset
    "data", equations/euler.py:33-56) reads back as float64."""
    path = str(tmp_path / "ic.hdf5")
    code = ("import h5py, numpy, sys\n"
            "x = numpy.linspace(-5, 5, 12).reshape(1, 3, 4).repeat(4, 0)\n"
            "f = h5py.File(sys.argv[1], 'w'); d = f.create_dataset('data', (4, 3, 4)); d[...] = x; f.close()\n")
    subprocess.run([CONDA_PY, "-c", code, path], check=True)
    ic = hdf5io.read_initial_conditions(path)
    want = np.linspace(-5, 5, 12).reshape(1, 3, 4).repeat(4, 0).astype(np.float32).astype(np.float64)
    assert ic.dtype == np.float64 and np.array_equal(ic, want)


@pytest.mark.gpu
def test_solver_writes_every_step_in_reference_layout(device, tmp_path):
    """Solver with write_output: dataset data holds the full steps, the first
    being the initial condition and each equal to the oracle's level."""
    from gpu_helpers import heat_case, oracle_stages
    from pysweep_amd import Solver
    from pysweep_amd.equations import path as eq_path
    ic, g, step = heat_case(32, True, 12, seed=4)
    s = Solver(ic, sendWarning=False)
    s.simulation, s.operating, s.intermediate, s.blocksize, s.share = True, 1, 1, (8, 8, 1), 1
    s.globals = list(g)
    s.setCPU(eq_path("heat.py"))
    s.setGPU(eq_path("heat.cu"))
    s.exact = True
    s.write_output = True
    s.output = str(tmp_path / "heatOutput.hdf5")
    s.loadCPUModule()
    s()
    data = hdf5io.read_dataset(s.output, "data")
    assert data.shape == (13, 1, 32, 32)
    assert np.array_equal(data[0], ic)
    for i in (1, 6, 12):
        assert np.array_equal(data[i], oracle_stages(step, ic, 1, i)), i
    assert hdf5io.read_dataset(s.output, "clocktime")[0] > 0
