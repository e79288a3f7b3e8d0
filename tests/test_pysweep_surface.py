"""The reference's import surface (pysweep/__init__.py:1-5, setup.py:17-19):
``import pysweep`` gives Solver, the plugin pairs with their
initial-condition helpers, and the command line; a caller written against
the reference runs unchanged."""
import os

import numpy as np
import pytest
import yaml


def test_package_exports():
    import pysweep
    from pysweep_amd.solver import Solver
    assert pysweep.Solver is Solver
    from pysweep.core.solver import Solver as S2
    assert S2 is Solver
    for name in ("example", "euler", "heat", "checker", "half"):
        mod = getattr(pysweep.equations, name)
        assert callable(mod.set_globals) and callable(mod.step) and callable(mod.createInitialConditions)
        for ext in (".py", ".cu"):
            assert os.path.isfile(pysweep.utils.commandline.getEqnPath(name + ext))
    assert callable(pysweep.utils.commandline.commandLine)


def test_plugin_files_load_by_path_and_step_on_cpu():
    """Solver.loadCPUModule imports the plugin file by path (core/solver.py:83-87);
    the loaded heat step is the forward-Euler 5-point update."""
    from pysweep.utils.commandline import getEqnPath
    from pysweep_amd.plugins import load_cpu_module
    heat = load_cpu_module(getEqnPath("heat.py"))
    heat.set_globals(0.0, 1.0, 0.1, 1.0, 1.0, 1.0, True)
    rng = np.random.default_rng(3)
    st = np.zeros((2, 1, 6, 6))
    st[0, 0] = rng.random((6, 6))
    heat.step(st, [(i, j) for i in range(6) for j in range(6)], 0, 1)
    u = st[0, 0]
    lap = np.roll(u, -1, 0) + np.roll(u, 1, 0) + np.roll(u, -1, 1) + np.roll(u, 1, 1) - 4 * u
    assert np.allclose(st[1, 0], u + 0.1 * lap, rtol=0, atol=1e-15)


def test_heat_helpers_match_the_decaying_mode(tmp_path):
    import pysweep
    from pysweep_amd import hdf5io
    heat = pysweep.equations.heat
    n, alpha, t = 16, 0.5, 0.01
    u, X, Y = heat.analytical(n, n, t, alpha=alpha)
    assert u.shape == (1, n, n) and np.allclose(X, np.arange(n) / n)
    want = np.outer(np.sin(2 * np.pi * X), np.sin(2 * np.pi * Y)) * np.exp(-8 * np.pi ** 2 * alpha * t)
    assert np.allclose(u[0], want, rtol=0, atol=1e-15)
    # the mode is an eigenfunction of the Laplacian: u_t = alpha * lap(u)
    h = 1e-4
    x, y = 0.3, 0.7
    dudt = (heat.analyticalEquation(x, y, t + h, alpha) - heat.analyticalEquation(x, y, t - h, alpha)) / (2 * h)
    lap = -8 * np.pi ** 2 * heat.analyticalEquation(x, y, t, alpha)
    assert abs(dudt - alpha * lap) < 1e-5 * abs(lap)  # central difference, O(h^2)
    f = heat.createInitialConditions(n, n, alpha=alpha, filename=str(tmp_path / "heatConditions.hdf5"))
    data = hdf5io.read_initial_conditions(f)
    x = np.linspace(0, 1, n)
    ref = np.outer(np.sin(2 * np.pi * x), np.sin(2 * np.pi * x)).astype(np.float32)
    assert data.shape == (1, n, n) and np.array_equal(data[0], ref.astype(np.float64))
    T, dt = heat.getFinalTime(100, 10, alpha=123.0)  # alpha is not used (as in the reference)
    assert dt == pytest.approx(0.1 * 0.01 ** 2 / 0.01) and T == pytest.approx(10 * dt)


def test_euler_helpers_give_the_advected_vortex(tmp_path):
    import pysweep
    from pysweep_amd import hdf5io
    eu = pysweep.equations.euler
    g = 1.4
    # far field: rho = 1, u = v = 1, p = 1/gamma
    q = eu.analytical(1e3, -1e3, 0.0)
    assert np.allclose(q, (1.0, 1.0, 1.0, 1 / g / (g - 1) + 1.0), rtol=0, atol=1e-12)
    # the vortex moves with the free stream (1, 1)
    a = np.array(eu.analytical(0.3, -0.2, 0.0))
    b = np.array(eu.analytical(0.3 + 0.5, -0.2 + 0.5, 0.5))
    assert np.allclose(a, b, rtol=0, atol=1e-13)
    # the centre density (1 + dT)^(1/(g-1)), dT = -(g-1) beta^2/2
    beta = np.sqrt(2 / g) * 5 * np.sqrt(2) / 4 / np.pi * np.exp(0.5)
    assert abs(eu.analytical(0.0, 0.0, 0.0)[0] - (1 - (g - 1) * beta ** 2 / 2) ** (1 / (g - 1))) < 1e-14
    n = 32
    arr = eu.getAnalyticalArray(n, n, 0.0)
    assert arr.shape == (4, n, n)
    xs = np.linspace(-5, 5, n, endpoint=False)
    assert np.allclose(arr[:, 5, 9], eu.analytical(xs[5], xs[9], 0.0), rtol=0, atol=1e-15)
    f = eu.createInitialConditions(n, n, filename=str(tmp_path / "eulerConditions.hdf5"))
    data = hdf5io.read_initial_conditions(f)
    xe = np.linspace(-5, 5, n)
    assert data.shape == (4, n, n)
    assert np.array_equal(data[:, 3, 7], np.float32(eu.analytical(xe[3], xe[7], 0.0)).astype(np.float64))
    T, dt = eu.getFinalTime(100, 7)
    assert dt == pytest.approx(0.01) and T == pytest.approx(0.07)


def test_known_answer_plugin_conditions(tmp_path):
    import pysweep
    from pysweep_amd import hdf5io
    d = hdf5io.read_initial_conditions(pysweep.equations.checker.createInitialConditions(
        2, 4, 6, filename=str(tmp_path / "c.hdf5")))
    assert d.shape == (2, 4, 6) and d[0, 0, 0] == 1 and d[1, 1, 1] == 1 and d[0, 0, 1] == 0 and d[1, 1, 0] == 0
    d = hdf5io.read_initial_conditions(pysweep.equations.half.createInitialConditions(
        3, 5, filename=str(tmp_path / "h.hdf5")))
    assert d.shape == (1, 3, 5) and d[0, 2, 4] == 14 / 100 and d[0, 1, 0] == 5 / 100
    d = hdf5io.read_initial_conditions(pysweep.equations.example.createInitialConditions(
        2, 8, 8, filename=str(tmp_path / "e.hdf5")))
    assert d.shape == (2, 8, 8) and not d.any()


def _write_yaml(path, swept, b, k, s, globals_, name):
    from pysweep.utils.commandline import getEqnPath
    doc = {"swept": swept, "filename": "out.hdf5", "verbose": False, "blocksize": b, "share": 1,
           "dtype": "float64", "globals": list(globals_), "intermediate_steps": s, "operating_points": k,
           "cpu": getEqnPath(name + ".py"), "gpu": getEqnPath(name + ".cu")}
    with open(path, "w") as f:
        yaml.safe_dump(doc, f)
    return str(path)


@pytest.mark.gpu
@pytest.mark.parametrize("swept", [True, False], ids=["swept", "standard"])
def test_reference_calling_style_euler(device, tmp_path, swept):
    """The reference's way in: an IC file from the plugin, a YAML input,
    ``pysweep.Solver(file, yaml)()``; the final state equals the oracle's
    bitwise (the reference's step quantization, not exact steps)."""
    import pysweep
    from gpu_helpers import oracle_stages
    from oracle import euler as o_euler
    n, steps, b = 64, 10, 16
    ic_file = pysweep.equations.euler.createInitialConditions(n, n, filename=str(tmp_path / "ic.hdf5"))
    dx = 10 / (n - 1)
    dt = 0.1 * dx
    yml = _write_yaml(tmp_path / "euler.yaml", swept, b, 2, 2, [0, dt * steps, dt, dx, dx, 1.4], "euler")
    solver = pysweep.Solver(ic_file, yml)
    solver()
    from pysweep_amd import hdf5io
    ic = hdf5io.read_initial_conditions(ic_file)
    want = oracle_stages(lambda q, qb, g: o_euler.step_py(q, qb, g, 1.4, dt, dx, dx), ic, 2, solver.stages)
    assert np.array_equal(solver.final_state, want)


@pytest.mark.gpu
def test_reference_calling_style_heat(device, tmp_path):
    import pysweep
    from gpu_helpers import oracle_stages
    from oracle import heat as o_heat
    n, steps, b = 64, 20, 8
    ic_file = pysweep.equations.heat.createInitialConditions(n, n, alpha=1, filename=str(tmp_path / "ic.hdf5"))
    dx = 1 / (n - 1)
    dt = float(0.1 * dx ** 2)
    yml = _write_yaml(tmp_path / "heat.yaml", True, b, 1, 1, [0, dt * steps, dt, dx, dx, 1, True], "heat")
    solver = pysweep.Solver(ic_file, yml)
    solver()
    from pysweep_amd import hdf5io
    ic = hdf5io.read_initial_conditions(ic_file)
    c = o_heat.courant_py(1, dt, dx)
    want = oracle_stages(lambda q, qb, g: o_heat.step_py(q[0], None, g, c, True)[None], ic, 1, solver.stages)
    assert np.array_equal(solver.final_state, want)
